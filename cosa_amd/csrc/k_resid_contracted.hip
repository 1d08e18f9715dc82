// k_resid_contracted.hip -- the one-pass Euler residual kernels of
// k_resid.hip built with multiply-add contraction (the build compiles this
// file with -ffp-contract=fast): the fast FP64 mode of
// cosahip_set_precision (mode 3).  Every product and sum is still binary64;
// a contracted a*b+c rounds once instead of twice, so the residual differs
// from the reference's in the last bits (tests/test_gpu_contracted.py holds
// it to the FP64 parity bar).
#define COSAHIP_CONTRACTED 1
#include "k_resid.hip"
