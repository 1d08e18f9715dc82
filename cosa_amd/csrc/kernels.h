// kernels.h -- launchers of the libcosahip device kernels (one per reference
// routine family).  All launch on the solve's stream and return hipError_t.
#pragma once
#include "cosahip_internal.h"

namespace cosahip {

// ---- state conversions and array utilities (k_core.hip) ----
// shadow: also write the FP32 copy of the interior when the FP32-residual
// mode is in effect and q is the level's Q
// qold: also copy the conservative interior there first (copy_array)
hipError_t launch_cons2prim(Solve& s, Level& L, double* q, bool shadow = false, double* qold = nullptr);
hipError_t launch_halo_f32(Solve& s, Level& L);
hipError_t launch_to_f32(Solve& s, const double* src, float* dst, ll n);
// halo_only: the face-halo strips only (the interior is about to be rewritten)
hipError_t launch_prim2cons(Solve& s, Level& L, double* q, bool halo_only = false);
hipError_t launch_copy_interior(Solve& s, Level& L, double* dst, const double* src);
hipError_t launch_add_interior(Solve& s, Level& L, double* dst, const double* src, double fact);
hipError_t launch_zero_res(Solve& s, Level& L, double* res);        // 0..imax box
hipError_t launch_zero_res_shell(Solve& s, Level& L, double* res);  // first halo layer only
hipError_t launch_update(Solve& s, Level& L, double ar);
int rkim_group(int m);
hipError_t launch_lim_res_ome_rkim(Solve& s, Level& L);
hipError_t launch_rms_hb_term(Solve& s, Level& L, double* work1);
// dual time stepping: res += src_uns(q) (then, sub_rhs, res -= rhs);
// q1 = q and rhs = rhs_uns(q1, q2) on level 1; work1 = the rkimp norm term
hipError_t launch_src_uns(Solve& s, Level& L, int sub_rhs);
hipError_t launch_rhs_uns(Solve& s, Level& L);
hipError_t launch_rms_dual_term(Solve& s, Level& L, double* work1);
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
hipError_t launch_edges(Solve& s, Level& L, double* q, float* q32 = nullptr);
hipError_t launch_cut_ncomp(Solve& s, Level& L, double* a, int ncomp);
hipError_t launch_rest_scalar(Solve& s, Level& L1, Level& L2, const double* a1, double* a2);

// inputs of the residual kernels in precision T: the level's FP64 fields,
// or their FP32 copies in the FP32-residual mode
template <typename T>
struct ResIn {
  const T* q;       // primitive state with halos
  const T* s[3];    // si, sj, sk
  const T* mut;     // eddy viscosity (k-omega)
  const T* xd[9];   // metric derivatives xideri .. zetaderk
  const T* const* dxd;   // device table of xd
};

// ---- spectral radii, residual, smoothing (k_flux.hip) ----
hipError_t launch_lambda_deltat(Solve& s, Level& L);
hipError_t launch_residual(Solve& s, Level& L, const double* q, double* res);
// the specialised inviscid residual (k_resid.hip), variant = ResidVariant
// other than RESID_GENERIC; hipErrorInvalidValue for an unknown variant
hipError_t launch_residual_tile(Solve& s, Level& L, int variant, const double* q, const double* si, const double* sj,
                                const double* sk, double* res);
hipError_t launch_residual_tile(Solve& s, Level& L, int variant, const float* q, const float* si, const float* sj,
                                const float* sk, double* res);
// the same kernels built with multiply-add contraction (fast FP64 mode,
// k_resid_contracted.hip); RESID_GENERIC runs as RESID_TILE7 there
hipError_t launch_residual_tile_contracted(Solve& s, Level& L, int variant, const double* q, const double* si,
                                           const double* sj, const double* sk, double* res);
// work items of one block in the tile table of a TI x TJ residual tile
// (k segments of RT_KSEG planes); 31-cell rows put one face per lane of a
// 32-lane row, 32-cell rows (TK_T32) spread the 33 faces of a row over the
// work-group
constexpr int RT_TI = 31;    // cells per tile along i of TK_FUSED / TK_FUSED15
constexpr int RT_KSEG = 32;  // k-planes per work-group
ll residual_tile_items(int TI, int TJ, int imax, int jmax, int kmax, int nsnap);
// the RK update fused into the last CIRS sweep (mean flow): q = qold - ar
// res dtvol per cell, then (prim) the next stage's cons2prim; q32: the
// FP32 shadow of the primitive state, or null
struct CirsUpdate {
  double* q;
  const double* qold;
  float* q32;
  double ar;
  int prim;
};
hipError_t launch_cirs(Solve& s, Level& L, const CirsUpdate* up = nullptr);
// vpre: viscous face fluxes of direction idir - 1 whose difference is added
// to res first (nullptr: none)
template <typename T>
hipError_t launch_flux_dir(Solve& s, Level& L, int idir, const ResIn<T>& in, double* res, const T* vpre = nullptr);
template <typename T>
hipError_t launch_h0_last(Solve& s, Level& L, const ResIn<T>& in);
// the inviscid residual of all three directions of a viscous / k-omega
// level in one pass (k_resid_visc.hip), with the viscous face fluxes of
// direction d read from vf + (d - 1) vslot and differenced after the
// inviscid ones of that direction; vf null: inviscid only
template <typename T>
hipError_t launch_residual_tile_visc(Solve& s, Level& L, const ResIn<T>& in, double* res, const T* vf, ll vslot);

// ---- viscous fluxes and Wilcox k-omega (k_turb.hip) ----
hipError_t launch_residual_viscous(Solve& s, Level& L, const double* q, double* res);
hipError_t launch_sst_first_mut(Solve& s, Level& L);
hipError_t launch_tur_vis(Solve& s, Level& L, const double* q);
hipError_t launch_fixq(Solve& s, Level& L);
hipError_t launch_turb_limits(Solve& s, Level& L, double ar);

// ---- forces (k_force.hip) ----
// ---- grid motion and metric set-up (k_grid.hip) ----
hipError_t launch_grid_metrics(Solve& s, Level& L, const GridTiles& gt, int* dbad, int do_volex, int do_bedges,
                               int do_vmetrix);
hipError_t launch_grid_coarsen(Solve& s, Level& F, Level& C, const GridTiles& gtc);
hipError_t launch_grid_pitch(Solve& s, Level& L, const GridTiles& gt, const double* pose_dev, double dtheta0,
                             double omegas, double xh, double yh, const double* base);

// device buffers of one force evaluation: the wall-face runs, the per-face
// terms (12, faces, 0:2N), per-block sums (12, nsurface, 0:2N, blocks), the
// coefficients (12, 0:2N, nsurface) and the geometry (13 values, then the
// moment centres (3, 0:2N))
struct ForceBufs {
  const WallRun* runs;
  int nruns, nfaces;
  double *terms, *sums, *coef;
  const double *geom, *momr;
};
hipError_t launch_force(Solve& s, Level& L, int nsurface, const ForceBufs& fb);

}  // namespace cosahip
