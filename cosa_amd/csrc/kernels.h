// kernels.h -- launchers of the libcosahip device kernels (one per reference
// routine family).  All launch on the solve's stream and return hipError_t.
#pragma once
#include "cosahip_internal.h"

namespace cosahip {

// ---- state conversions and array utilities (k_core.hip) ----
hipError_t launch_cons2prim(Solve& s, Level& L, double* q);
hipError_t launch_prim2cons(Solve& s, Level& L, double* q);
hipError_t launch_copy_interior(Solve& s, Level& L, double* dst, const double* src);
hipError_t launch_add_interior(Solve& s, Level& L, double* dst, const double* src, double fact);
hipError_t launch_zero_res(Solve& s, Level& L, double* res);        // 0..imax box
hipError_t launch_zero_res_shell(Solve& s, Level& L, double* res);  // first halo layer only
hipError_t launch_update(Solve& s, Level& L, double ar);
hipError_t launch_src_hb(Solve& s, Level& L);
hipError_t launch_rms(Solve& s, Level& L, const double* res, const double* rhs, int use_rhs);
hipError_t launch_prtest(Solve& s, Level& L, const double* q, int stage_seq);
hipError_t launch_bc(Solve& s, Level& L, double* q);
hipError_t launch_cut(Solve& s, Level& L, double* q);
hipError_t launch_edges_corners(Solve& s, Level& L, double* q);
hipError_t launch_rest1(Solve& s, Level& L1, Level& L2, const double* q1, double* q2);
hipError_t launch_rest2(Solve& s, Level& L1, Level& L2, double* rhs2, const double* res1);
hipError_t launch_sub_box(Solve& s, Level& L, double* dst, const double* src);  // dst -= src on 0..imax
hipError_t launch_prol(Solve& s, Level& L1, Level& L2, double* q1, const double* q2);
hipError_t launch_edges(Solve& s, Level& L, double* q);
hipError_t launch_cut_ncomp(Solve& s, Level& L, double* a, int ncomp);
hipError_t launch_rest_scalar(Solve& s, Level& L1, Level& L2, const double* a1, double* a2);

// ---- spectral radii, residual, smoothing (k_flux.hip) ----
hipError_t launch_lambda_deltat(Solve& s, Level& L);
hipError_t launch_residual(Solve& s, Level& L, const double* q, double* res);
hipError_t launch_cirs(Solve& s, Level& L);
hipError_t launch_flux_dir(Solve& s, Level& L, int idir, const double* q, double* res);
hipError_t launch_h0_last(Solve& s, Level& L, const double* q);

// ---- viscous fluxes and Wilcox k-omega (k_turb.hip) ----
hipError_t launch_residual_viscous(Solve& s, Level& L, const double* q, double* res);
hipError_t launch_tur_vis(Solve& s, Level& L, const double* q);
hipError_t launch_fixq(Solve& s, Level& L);
hipError_t launch_turb_limits(Solve& s, Level& L, double ar);

}  // namespace cosahip
