// k_resid_contracted.hip -- the one-pass Euler residual kernels of
// k_resid.hip built with multiply-add contraction and reciprocal
// multiplication (the build compiles this file with -ffp-contract=fast
// -freciprocal-math): the fast FP64 mode of cosahip_set_precision (mode 3).
// Every operation is still binary64; a contracted a*b+c rounds once instead
// of twice and quotients sharing a divisor multiply by its one reciprocal,
// so the residual differs from the reference's in the last bits
// (tests/test_gpu_contracted.py holds it to the FP64 parity bar).
#define COSAHIP_CONTRACTED 1
#include "k_resid.hip"
