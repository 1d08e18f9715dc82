// cosahip_internal.h -- device layout helpers and solve state of libcosahip.
//
// Device arrays keep the reference's flat layout (src/input.f90:4878-5131):
// whole-level allocations with blocks concatenated at the setoffs offsets,
// i fastest, so one block's P3 array is
//   a(-1:imax+1,-1:jmax+1,-1:kmax+1, ncomp, 0:2N)
// and uploads/downloads are single copies.  Kernels address blocks through
// a small per-level table of Blk records and iterate over per-level "tile"
// tables (block, first work item) so one launch covers every block of a
// level whatever their sizes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "cosahip.h"

namespace cosahip {

typedef long long ll;

// one block of one level, as seen by kernels
struct Blk {
  int imax, jmax, kmax, nbc;   // node counts; number of BC patches
  ll o3, o2, o1, o0;           // off_p3, off_p2, off_p1, off_0 (points)
};

// run-time constants every kernel needs (passed by value)
struct Par {
  int npde, nsnap, nsnap_met, lmet, hbmove, moving;
  int ilimf, irsop, nl_crs;
  double gamma, rkap, epslimf;
  double gam_gm1;   // gamma / (gamma - 1), the same correctly rounded quotient the flux evaluates
  double cfl, rat, psirs;
  int efx_setup_f, efx_all_f, eig_limit_f, eig_lim_all_f;
  int entfx_f_w[3];
  double cntrpy_f, entfxctf_f, eig_cutoff_f;
  double machfs, alpha, beta, rhoinf, pinf, ainf;
  double qfst[5];
  double omega;
  double qmin1, qmin5;
  // viscous flow and Wilcox k-omega
  int kom, viscous, foturb, ilimt, turcmp, lim_prodtke, lim_prodomega, wallwilc, prd_all_lvls, nlevel;
  int efx_setup_t, efx_all_t, eig_limit_t, eig_lim_all_t;
  int entfx_t_w[3];
  double epslimt, cflt, ratt, cntrpy_t, entfxctf_t, eig_cutoff_t;
  double tkefar, omefar, reyno, pranl, prant, stemp, cdff, epst, cmu, sigk, sigw, bkw, bstrkw, gkw, prdlim, roughk;
  double qmin6, qmin7;
  // k-omega model: kom above is set for every model; sst = 0 Wilcox, 1 BSL,
  // 2 SST, with the BSL/SST constants of input (src/input.f90:2523-2541)
  int sst;
  double sst_a1, sigk1, sigk2, sigw1, sigw2, b1, b2, g1, g2;
  // dual time stepping: on, the rkimp update, rkimp without IRS on one
  // level (the physical-time term only in the update); dual_fact 1 on the
  // first step of a backward-Euler start, else 1.5; dual_dt the physical step
  int dualt, dual_rkimp, dual_plain;
  double dual_fact, dual_dt;
};

__host__ __device__ inline ll n3of(const Blk& b) { return (ll)(b.imax + 3) * (b.jmax + 3) * (b.kmax + 3); }
__host__ __device__ inline ll n2of(const Blk& b) { return (ll)(b.imax + 2) * (b.jmax + 2) * (b.kmax + 2); }
__host__ __device__ inline ll n1of(const Blk& b) { return (ll)(b.imax + 1) * (b.jmax + 1) * (b.kmax + 1); }
// P3 point index of (i,j,k), i,j,k from -1
__host__ __device__ inline ll i3(const Blk& b, int i, int j, int k) {
  return (ll)(i + 1) + (ll)(b.imax + 3) * ((ll)(j + 1) + (ll)(b.jmax + 3) * (ll)(k + 1));
}
// P2 point index (0-based from 0)
__host__ __device__ inline ll i2(const Blk& b, int i, int j, int k) {
  return (ll)i + (ll)(b.imax + 2) * ((ll)j + (ll)(b.jmax + 2) * (ll)k);
}
// P1 point index
__host__ __device__ inline ll i1(const Blk& b, int i, int j, int k) {
  return (ll)i + (ll)(b.imax + 1) * ((ll)j + (ll)(b.jmax + 1) * (ll)k);
}

// work item tiles: (block, first item) -- one GPU workgroup per tile
struct Tile {
  int block, base;
};

enum TileKind {
  TK_CELLS = 0,   // interior cells (imax-1)(jmax-1)(kmax-1), i fastest
  TK_ROWS_I,      // i-direction flux rows: (jmax-1)(kmax-1) rows x segments
  TK_LINES_J,     // (imax-1)(kmax-1) lines along j
  TK_LINES_K,     // (imax-1)(jmax-1) lines along k
  TK_LINES_I,     // (jmax-1)(kmax-1) lines along i, 64 per tile
  TK_CELLS_P1,    // 0..imax x 0..jmax x 0..kmax
  TK_COARSE,      // cells of the next coarser level, per block (restriction)
  TK_PROL,        // 2x2x2 prolongation groups 0..imax-1 step 2, per block
  TK_HALO,        // prtest region: interior + face halo strips
  TK_FUSED,       // fused residual: (31 x 7)-cell (i,j) tiles x 32-plane k segments
  TK_LINES_I_T,   // (jmax-1)(kmax-1) lines along i, TS per tile (thread per line)
  TK_LINES_IL,    // (jmax-1)(kmax-1) lines along i, il_lpw per tile (whole lines resident in LDS)
  TK_LINES_IV,    // (jmax-1)(kmax-1) lines along i, 256/npde per tile (thread per line and variable)
  TK_SEGS_J,      // MSEG-cell segments of the (imax-1)(kmax-1) lines along j, i fastest
  TK_SEGS_K,      // MSEG-cell segments of the (imax-1)(jmax-1) lines along k, i fastest
  TK_FACES_I,     // i faces imax x (jmax-1) x (kmax-1)
  TK_FACES_J,     // j faces (imax-1) x jmax x (kmax-1)
  TK_FACES_K,     // k faces (imax-1) x (jmax-1) x kmax
  TK_FUSED15,     // residual tiles of 31 x 15 cells x 32-plane k segments (k_residual_tile, TJ = 15)
  TK_SHELL,       // the halo shell of 0..imax x 0..jmax x 0..kmax (the box without the interior cells)
  TK_T32,         // residual tiles of 32 x 7 cells x 32-plane k segments (RESID_TILE32_W4, RESID_SVA32)
  TK_NKIND
};

// inviscid residual kernels (the same bits; chosen per level by
// cosahip_autotune or COSAHIP_RESIDUAL at create time)
enum ResidVariant {
  RESID_GENERIC = 0,   // k_residual_fused: run-time scheme flags, register k-window, 31 x 7 tiles
  RESID_TILE7,         // k_residual_tile: specialised, k-window from LDS / L2, 31 x 7 tiles
  RESID_TILE7_REG,     // k_residual_tile: specialised, register k-window, 31 x 7 tiles
  RESID_TILE15,        // k_residual_tile: specialised, k-window from LDS / L2, 31 x 15 tiles
  RESID_TILE7_W4,      // RESID_TILE7 held to 128 VGPRs (4 waves per SIMD)
  RESID_SVA7,          // k_residual_sva: van Albada slope once per cell, 31 x 7 tiles (else as RESID_TILE7)
  RESID_SVA15,         // k_residual_sva, 31 x 15 tiles (else as RESID_TILE15)
  RESID_SVA7_W4,       // RESID_SVA7 held to 128 VGPRs (4 waves per SIMD)
  RESID_TILE32_W4,     // RESID_TILE7_W4 on 32 x 7 tiles (the 33 i-faces of a row spread over the work-group)
  RESID_SVA32,         // RESID_SVA7 on 32 x 7 tiles (else as RESID_TILE32_W4)
  RESID_NVARIANT
};

// a BC patch or one directional cut record, flattened for kernels
struct PatchRec {
  int block;       // 0-based
  int type;        // BC type (cuts: -1)
  int t[10];       // bctopo(1:10)
  int count;       // face cells of the patch (transverse range product)
  int n2;          // length of the first transverse range
};
struct CutRec {
  int b1, b2;      // 0-based
  int t[21];       // cutopo(1:21)
  int count, n2;   // (len(ic2)+1)*(len(ic3)+1), len(ic2)+1
};

// one wall BC patch of one block for the force integration: faces in the
// reference's loop order (i3 outer, i2 inner)
struct WallRun {
  int block;      // 0-based
  int surface;    // 1-based
  int visc;       // 1: the viscous part is integrated too
  int idir, inrout;
  int lo2, n2, lo3;
  int first, count;   // face range in the term array
};

// grid set-up work lists of one level (k_grid.hip): tile tables of the
// face families, cells, halo cells, edge cells and nodes, and the cut
// records of cut_v / cut_x in a flat form
enum GridTileKind {
  GT_CELLS = 0,   // interior cells, snapshot 0
  GT_FI,          // i faces x snapshots of the metrics
  GT_FJ,          // j faces x snapshots
  GT_FK,          // k faces x snapshots
  GT_VOLEX,       // face-halo cells of vol
  GT_EDGES,       // edge-halo cells of vol (12 edges)
  GT_NODEHALO,    // face-halo nodes x snapshots
  GT_NODES,       // nodes 1..imax x 1..jmax x 1..kmax x snapshots
  GT_NODES0,      // nodes 1..imax x 1..jmax x 1..kmax, snapshot 0
  GT_NKIND
};
struct GeomCut {
  int b1, b2;                  // 0-based blocks
  int ibcpt, inr;              // halo index of block 1, interior index of block 2
  int ic1, ic2, ic3, jc1, jc2, jc3;   // 0-based directions
  int str1[3], sg1[3], str2[3], sg2[3];
  int count, n2;
};
struct GridPtrs {
  const double *x, *y, *z, *xd, *yd, *zd;
};
struct GridTiles {
  bool built = false;
  std::vector<Tile> h[GT_NKIND];
  Tile* t[GT_NKIND] = {};
  int n[GT_NKIND] = {};
  GeomCut* cut_v = nullptr;
  GeomCut* cut_x = nullptr;
  int ncut = 0, maxcut_v = 1, maxcut_x = 1;
};

struct Level {
  int nl = 0, nblocks = 0, ncuts = 0;
  std::vector<Blk> hblk;
  std::vector<int32_t> imax, jmax, kmax, nbcs, bctopo, cutopo;
  Blk* dblk = nullptr;
  ll n3tot = 0, n2tot = 0, n1tot = 0, n0tot = 0, ncell = 0;
  double* f[COSAHIP_NFIELDS] = {};
  double* dq = nullptr;       // HB term (qhb) and scratch
  double* work1 = nullptr;    // copy of res for the HB norm
  double* dqc = nullptr;      // SST: (P3, 9, nsnap) cell-centre velocity differences of the stage
  int cirs_mode = 0;          // CIRS i-sweep picked by cosahip_autotune: 0 none, 1 lds, 2 iv, 3 line
  int resid_mode = RESID_GENERIC;   // inviscid residual kernel (ResidVariant)
  int il_lpw = 0;             // k_cirs_il: lines per work-group (0: lines too long for its LDS budget)
  int il_stride = 0;          // k_cirs_il: odd LDS row stride >= the longest i-line of the level
  double* il_rtab = nullptr;  // k_cirs_il: multipliers and reciprocal divisors, 2 il_stride per block
  double* jk_rtab[2] = {};    // k_cirs_jk (irsop 1), j and k lines: r(a) and 1/d(a), a = 1..L, 2 (JK_RT) per block
  double* cirs_r = nullptr;   // Thomas multipliers r(j), one per cell
  double* cirs_rt = nullptr;  // the same for k, omega (k-omega)
  double* rms_part = nullptr; // per-tile partial sums of squares
  unsigned char* sbc = nullptr;  // strong-BC face flags (6 faces per block)
  std::vector<ll> sbc_off;    // per block: offset of its 6 face flag planes
  ll* sbc_doff = nullptr;
  std::vector<Tile> htile[TK_NKIND];
  Tile* tile[TK_NKIND] = {};
  int ntile[TK_NKIND] = {};
  std::vector<PatchRec> hpatch;
  PatchRec* dpatch = nullptr;
  int maxpatch = 0;
  // BC work tiles (patch, first face item), patches in order, and the first
  // tile of each patch (ptile_first[np] = number of tiles)
  std::vector<Tile> hptile;
  std::vector<int> ptile_first;
  Tile* ptile = nullptr;
  std::vector<CutRec> hcut;
  CutRec* dcut = nullptr;
  int maxcut = 0;
  bool thin = false;          // a block has < 2 cells in some direction
  // viscous / k-omega scratch
  double* vflux = nullptr;    // face fluxes (imax, jmax, kmax, npde, 0:2N), off_0 offsets; 3 slots
  double* tbeta[3] = {};      // CIRS coefficients of k, omega (P1, 0:2N)
  double* h0tab = nullptr;    // per (block, snapshot, direction): h0l, h0r of the block's last face
  double** dxd = nullptr;     // device table of the nine metric-derivative fields
  // FP32-residual mode: single-precision shadows read by the residual kernel
  float* q32 = nullptr;       // primitive q, same layout as Q
  float* m32[3] = {};         // si, sj, sk
  bool m32_ok = false;        // m32 holds the current metrics
  float* xd32[9] = {};        // metric derivatives (viscous)
  float** dxd32 = nullptr;    // device table of xd32
  bool xd32_ok = false;
  // k-omega sources: cell-centre averages of the nine metric derivatives
  // (xi_x .. zeta_z), (P3, 9, nsnap_met), computed once per metric upload
  double* cmet = nullptr;
  float* cmet32 = nullptr;
  bool cmet_ok = false, cmet32_ok = false;
  float* mut32 = nullptr;     // eddy viscosity (k-omega)
  bool mut32_ok = false;
  // force integration: wall runs of the last (nsurface, include_inviscid)
  // request, per-face terms, per-block sums, coefficients
  int force_key = -1;
  std::vector<struct WallRun> hwall;
  struct WallRun* dwall = nullptr;
  int nwall_faces = 0;
  double* fterms = nullptr;   // (12, faces, 0:2N)
  double* fsums = nullptr;    // (12, nsurface, 0:2N, blocks)
  double* fcoef = nullptr;    // (12, 0:2N, nsurface)
  double* fgeom = nullptr;    // direction vectors, lengths, Mach, moment centres
  // device grid set-up (cosahip_grid_*)
  GridTiles gt;
  double* xbase = nullptr;    // base pose of a pitching grid: x, y, z of snapshot 0 (3 x n2tot)
  double* gpose = nullptr;    // per-snapshot dcos(dtheta), dsin(dtheta), dcos(omegas zeit)
  int* gbad = nullptr;        // non-positive face area flag
};

struct TimedLaunch {
  hipEvent_t a, b;
  int fam;
  ll items;
};

struct Solve {
  cosahip_config cfg;
  Par par;
  int device = 0;
  double* scratch = nullptr;   // this GPU's fixed force scratch array (cosahip_force_noalloc)
  bool fuse_update = true;     // RK update and next cons2prim in the last CIRS sweep (COSAHIP_FUSED_UPDATE=0: off)
  hipStream_t stream = nullptr;
  std::vector<Level> lev;     // index nl-1
  double dhb[131 * 131];      // dhb(i,j) column-major
  double* ddhb = nullptr;
  unsigned long long* dflag = nullptr;   // prtest first offender key
  double* hpinned = nullptr;             // small pinned buffer for results
  // per-block freestream (independent solves batched as blocks of one
  // solve): qfst(1:5) of bc_far0 per block on the device, and the Mach
  // number and angles per block for the force coefficients
  double* dqfst = nullptr;
  std::vector<double> bqfst, bflow;      // 5 and 3 (machfs, alpha, beta) per block
  bool block_rms = false;                // copy every tile's norm partials with the level-1 norms
  double* hpart = nullptr;               // pinned: the partials (ntile x 8)
  ll hpart_n = 0;
  double resrms[9] = {};                 // resrms(1:8), index 0 unused
  std::string err;
  int32_t errinfo[8] = {};
  // residual kernel choice: fused single pass (default) or one pass per
  // direction (COSAHIP_RESIDUAL=split in the environment at create time)
  bool split_residual = false;
  int resid_force = -1;   // COSAHIP_RESIDUAL=fused|tile7|tile7reg|tile15|tile7w4|sva7|sva15|sva7w4: that kernel on every level
  // viscous face fluxes after the inviscid pass of each direction: face
  // fluxes stored then differenced (default), flux and difference in one
  // marching kernel (COSAHIP_VISCOUS=march), or fused into the inviscid line
  // march (COSAHIP_VISCOUS=fused); or the face fluxes of all three
  // directions stored, then one tile pass for the inviscid fluxes and every
  // difference (COSAHIP_VISCOUS=tile), all at create time
  bool fuse_viscous = false;
  bool viscous_faces = true;
  bool viscous_tile = false;
  // CIRS i-sweep: one thread per (line, variable) on levels with fewer than
  // CIRS_IV_MAX_LINES lines along i, one thread per line with LDS-staged
  // chunks on larger ones; COSAHIP_CIRS_I=lds|iv at create time forces one
  int cirs_i_mode = 0;   // 0 auto, 1 lds, 2 iv, 3 line
  int il_lds = 24576;    // LDS bytes per work-group of the line-resident CIRS i-sweep
  // CIRS j/k sweeps of irsop 1: the forward pass kept in chunks of 4 cells
  // (k_cirs_jkc, the default; measured fastest) or 8 (COSAHIP_CIRS_JK=chunk8
  // at create time), or stored whole (0, COSAHIP_CIRS_JK=plain)
  int cirs_jk_chunk = 4;
  // residual precision: 0 FP64, 1 FP32 residual / FP64 state, 2 FP32 until
  // the level-1 density residual rms falls below prec_tail_rms, FP64 after,
  // 3 FP64 with multiply-add contraction (Euler residual)
  int prec_mode = 0;
  double prec_tail_rms = 0;
  bool prec_tail = false;
  bool fp32() const { return prec_mode == 1 || (prec_mode == 2 && !prec_tail); }
  // asynchronous multigrid cycles (cosahip_cycle_enqueue / _wait): a cycle
  // is in flight; the cycle's launch sequence captured as a HIP graph
  // (replayed while its key -- the residual precision -- is unchanged;
  // COSAHIP_GRAPH=0 at create time runs every cycle eagerly)
  bool cycle_pending = false;
  // asynchronous checkpoint: a device snapshot of q (and mut) taken in
  // stream order, copied to the host on a second stream
  hipStream_t ck_stream = nullptr;
  hipEvent_t ck_snap = nullptr, ck_done = nullptr;
  double* ck_buf = nullptr;
  ll ck_cap = 0;
  bool ck_pending = false;
  int mgit = 1;                 // cycle counter of the last smooth (CFL ramping)
  int dual_prop = 0;            // dual time stepping: two-state start (dual_prop_start)
  bool use_graph = true;
  hipGraphExec_t gexec = nullptr;
  int gkey = -1;
  // instrumentation
  bool timing = false;
  std::vector<TimedLaunch> pending;
  std::vector<hipEvent_t> evpool;
  double fam_ms[COSAHIP_NFAM] = {};
  ll fam_launch[COSAHIP_NFAM] = {}, fam_items[COSAHIP_NFAM] = {};
};

// bracket a launch sequence with timing events when enabled
struct TimeScope {
  Solve& s;
  int fam;
  ll items;
  hipEvent_t a = nullptr;
  TimeScope(Solve& s_, int fam_, ll items_);
  ~TimeScope();
};

// tile sizes
constexpr int TS = 256;        // items per tile, 1-D kernels
constexpr int JK_RT = 1025;    // per-block stride of the j/k CIRS tables (lines up to 1024 cells)
constexpr int ROWSEG = 63;     // cells per i-row segment (64 faces)
constexpr int MSEG = 16;       // cells per segment of a j / k line march
constexpr ll CIRS_IV_MAX_LINES = 262144;   // i-lines (x snapshots) of a level below which k_cirs_iv is used

}  // namespace cosahip
