// api.hip -- the C ABI of libcosahip (include/cosahip.h): solve contexts,
// level definition, host<->device transfers and the drivers smooth /
// restrict / prolong, which issue the kernels in the order of COSA's own
// smooth() (src/cosa.f90:1594-2032), restrict() (2672-2754) and prolong()
// (3167-3249).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>

#include "kernels.h"

using namespace cosahip;

#define CK(call)                                                                  \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess) return fail(s, COSAHIP_ERR_DEVICE, #call, hipGetErrorString(e_)); \
  } while (0)

static hipError_t scratch_address(int device, double** out);

static int fail(Solve* s, int code, const char* what, const char* detail = "") {
  if (s) {
    s->err = std::string(what) + (detail && *detail ? std::string(": ") + detail : std::string());
  }
  return code;
}

// ---------------------------------------------------------------------------
// small device helpers

__global__ void k_fill_u64(unsigned long long* p, unsigned long long v, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = v;
}

// ---------------------------------------------------------------------------
// instrumentation: HIP events recorded on the solve's stream around launches

static hipEvent_t pool_event(Solve& s) {
  if (!s.evpool.empty()) {
    hipEvent_t e = s.evpool.back();
    s.evpool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

namespace cosahip {
TimeScope::TimeScope(Solve& s_, int fam_, ll items_) : s(s_), fam(fam_), items(items_) {
  if (!s.timing) return;
  a = pool_event(s);
  if (a) (void)hipEventRecord(a, s.stream);
}
TimeScope::~TimeScope() {
  if (!s.timing || !a) return;
  hipEvent_t b = pool_event(s);
  if (!b) return;
  (void)hipEventRecord(b, s.stream);
  s.pending.push_back(TimedLaunch{a, b, fam, items});
}
}  // namespace cosahip

static void drain_timing(Solve& s) {
  if (s.pending.empty()) return;
  (void)hipStreamSynchronize(s.stream);
  for (auto& t : s.pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
      s.fam_ms[t.fam] += ms;
      s.fam_launch[t.fam] += 1;
      s.fam_items[t.fam] += t.items;
    }
    s.evpool.push_back(t.a);
    s.evpool.push_back(t.b);
  }
  s.pending.clear();
}

static void free_level(Level& L) {
  for (int f = 0; f < COSAHIP_NFIELDS; f++)
    if (L.f[f]) (void)hipFree(L.f[f]);
  void* ptrs[] = {L.dblk, L.dq, L.work1, L.cirs_r, L.cirs_rt, L.rms_part, L.sbc, L.sbc_doff, L.dpatch, L.dcut,
                  L.vflux, L.tbeta[0], L.tbeta[1], L.tbeta[2], L.h0tab, L.dxd, L.q32, L.m32[0], L.m32[1], L.m32[2],
                  L.dxd32, L.mut32, L.dwall, L.fterms, L.fsums, L.fcoef, L.fgeom, L.dqc, L.il_rtab, L.cmet,
                  L.cmet32, L.ptile, L.jk_rtab[0], L.jk_rtab[1]};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (float* p : L.xd32)
    if (p) (void)hipFree(p);
  for (int k = 0; k < TK_NKIND; k++)
    if (L.tile[k]) (void)hipFree(L.tile[k]);
  for (int k = 0; k < GT_NKIND; k++)
    if (L.gt.t[k]) (void)hipFree(L.gt.t[k]);
  void* gptrs[] = {L.gt.cut_v, L.gt.cut_x, L.xbase, L.gpose, L.gbad};
  for (void* p : gptrs)
    if (p) (void)hipFree(p);
  L = Level();
}

static Par make_par(const cosahip_config& c) {
  Par p;
  memset(&p, 0, sizeof(p));
  p.npde = c.npde;
  p.nsnap = 2 * c.nharms + 1;
  p.nsnap_met = 2 * c.nharms * c.hbmove + 1;
  p.lmet = c.lmet;
  p.hbmove = c.hbmove;
  p.moving = c.moving;
  p.ilimf = c.ilimf;
  p.irsop = c.irsop;
  p.nl_crs = c.nl_crs;
  p.gamma = c.gamma;
  p.gam_gm1 = c.gamma / (c.gamma - 1);
  p.rkap = c.rkap;
  p.epslimf = c.epslimf;
  p.cfl = c.cfl;
  p.rat = c.rat;
  p.psirs = c.psirs;
  p.efx_setup_f = c.efx_setup_f;
  p.efx_all_f = c.efx_all_f;
  p.eig_limit_f = c.eig_limit_f;
  p.eig_lim_all_f = c.eig_lim_all_f;
  for (int i = 0; i < 3; i++) p.entfx_f_w[i] = c.entfx_f_w[i];
  p.cntrpy_f = c.cntrpy_f;
  p.entfxctf_f = c.entfxctf_f;
  p.eig_cutoff_f = c.eig_cutoff_f;
  p.machfs = c.machfs;
  p.alpha = c.alpha;
  p.beta = c.beta;
  p.rhoinf = c.rhoinf;
  p.pinf = c.pinf;
  p.ainf = c.ainf;
  // freestream of bc_far0, aircraft/vawt branch (src/bc.f90:145-152)
  p.qfst[0] = c.rhoinf;
  p.qfst[1] = c.machfs * cos(c.alpha) * cos(c.beta);
  p.qfst[2] = c.machfs * sin(c.alpha) * cos(c.beta);
  p.qfst[3] = c.machfs * sin(c.beta);
  p.qfst[4] = c.pinf;
  p.omega = c.omega;
  p.qmin1 = c.qmin[0];
  p.qmin5 = c.qmin[4];
  p.kom = c.kom || c.kom_bsl || c.kom_sst;
  p.sst = c.kom_sst ? 2 : (c.kom_bsl ? 1 : 0);
  {
    const double kappa = 0.41, bstr = 0.09;
    p.sst_a1 = 0.31;
    p.sigk1 = c.kom_sst ? 0.85 : 0.5;
    p.sigk2 = 1.0;
    p.sigw1 = 0.5;
    p.sigw2 = 0.856;
    p.b1 = 3.0 / 40.0;
    p.b2 = 0.0828;
    p.g1 = p.b1 / bstr - p.sigw1 * (kappa * kappa) / sqrt(bstr);
    p.g2 = p.b2 / bstr - p.sigw2 * (kappa * kappa) / sqrt(bstr);
  }
  p.viscous = c.viscous;
  p.foturb = c.foturb;
  p.ilimt = c.ilimt;
  p.turcmp = c.turcmp;
  p.lim_prodtke = c.lim_prodtke;
  p.lim_prodomega = c.lim_prodomega;
  p.wallwilc = c.wallwilc;
  p.prd_all_lvls = c.prd_all_lvls;
  p.nlevel = c.nlevel;
  p.efx_setup_t = c.efx_setup_t;
  p.efx_all_t = c.efx_all_t;
  p.eig_limit_t = c.eig_limit_t;
  p.eig_lim_all_t = c.eig_lim_all_t;
  for (int i = 0; i < 3; i++) p.entfx_t_w[i] = c.entfx_t_w[i];
  p.epslimt = c.epslimt;
  p.cflt = c.cflt;
  p.ratt = c.ratt;
  p.cntrpy_t = c.cntrpy_t;
  p.entfxctf_t = c.entfxctf_t;
  p.eig_cutoff_t = c.eig_cutoff_t;
  p.tkefar = c.tkefar;
  p.omefar = c.omefar;
  p.reyno = c.reyno;
  p.pranl = c.pranl;
  p.prant = c.prant;
  p.stemp = c.stemp;
  p.cdff = c.cdff;
  p.epst = c.epst;
  p.cmu = c.cmu;
  p.sigk = c.sigk;
  p.sigw = c.sigw;
  p.bkw = c.bkw;
  p.bstrkw = c.bstrkw;
  p.gkw = c.gkw;
  p.prdlim = c.prdlim;
  p.roughk = c.roughk;
  p.qmin6 = c.qmin[5];
  p.qmin7 = c.qmin[6];
  return p;
}

// any of the k-omega models (Wilcox, BSL, SST)
static bool komega(const cosahip_config& c) { return c.kom || c.kom_bsl || c.kom_sst; }

static int check_config(Solve* s, const cosahip_config& c) {
  if ((c.kom != 0) + (c.kom_bsl != 0) + (c.kom_sst != 0) > 1)
    return fail(s, COSAHIP_ERR_ARG, "at most one of kom, kom_bsl, kom_sst");
  if (komega(c) && !c.viscous) return fail(s, COSAHIP_ERR_ARG, "k-omega requires viscous");
  if (c.npde != (komega(c) ? 7 : 5)) return fail(s, COSAHIP_ERR_ARG, "npde must be 5 (Euler, laminar) or 7 (k-omega)");
  if (c.rkimp && rkim_group(2 * c.nharms + 1) < 1)
    return fail(s, COSAHIP_ERR_UNSUPPORTED, "rkimp: 2N+1 snapshots exceed the per-cell LDS system (N <= 40)");
  if (komega(c) && (c.prd_all_lvls || c.lim_corr[0] || c.lim_corr[1] || c.lim_corr[2]))
    return fail(s, COSAHIP_ERR_UNSUPPORTED, "k-omega prd_all_lvls / lim_corr options are not built");
  if (komega(c) && c.ho_rest) return fail(s, COSAHIP_ERR_UNSUPPORTED, "k-omega with ho_rest (bc_res) is not built");
  if (c.ilimt < 0 || c.ilimt > 4) return fail(s, COSAHIP_ERR_ARG, "ilimt out of range 0..4");
  if (c.ramping < 0 || c.ramping > 2) return fail(s, COSAHIP_ERR_ARG, "ramping must be 0, 1 or 2");
  if (c.ramping && c.irsop == 0)
    return fail(s, COSAHIP_ERR_ARG, "CFL ramping cannot be used without IRS (src/input.f90:4217-4220)");
  if (c.viscous && (c.reyno <= 0 || c.pranl <= 0 || c.prant <= 0))
    return fail(s, COSAHIP_ERR_ARG, "viscous runs need reyno, pranl, prant > 0");
  if (c.nharms < 0 || c.nharms > 65) return fail(s, COSAHIP_ERR_ARG, "nharms out of range 0..65");
  if (c.nharms > 0 && !c.harbal) return fail(s, COSAHIP_ERR_ARG, "nharms > 0 requires harbal");
  if (c.lmet != 4 && c.lmet != 7) return fail(s, COSAHIP_ERR_ARG, "lmet must be 4 or 7");
  if (c.moving && c.lmet != 7) return fail(s, COSAHIP_ERR_ARG, "moving grids need lmet = 7");
  if (c.nlevel < 1 || c.nlevel > 4) return fail(s, COSAHIP_ERR_ARG, "nlevel out of range 1..4");
  if (c.irsop < 0 || c.irsop == 4 || c.irsop > 6) return fail(s, COSAHIP_ERR_ARG, "irsop must be 0, 1, 2, 3, 5 or 6");
  if (c.cutcirs < 0 || c.cutcirs > 1) return fail(s, COSAHIP_ERR_ARG, "cutcirs must be 0 or 1");
  if (c.ilimf < 0 || c.ilimf > 4) return fail(s, COSAHIP_ERR_ARG, "ilimf out of range 0..4");
  if (c.nstage < 1 || c.nstage > 5) return fail(s, COSAHIP_ERR_ARG, "nstage out of range 1..5");
  if (c.iupdt < 1 || c.iprint < 1) return fail(s, COSAHIP_ERR_ARG, "iupdt and iprint must be >= 1");
  return COSAHIP_OK;
}

// ---------------------------------------------------------------------------
extern "C" {

const char* cosahip_build_info(void) {
  return "libcosahip " "gfx950 FP64, ABI " "1";
}

int cosahip_create(const cosahip_config* cfg, int device, cosahip_solve** out) {
  if (!cfg || !out) return COSAHIP_ERR_ARG;
  *out = nullptr;
  Solve* s = new (std::nothrow) Solve();
  if (!s) return COSAHIP_ERR_NOMEM;
  s->cfg = *cfg;
  int rc = check_config(s, *cfg);
  if (rc) {
    fprintf(stderr, "cosahip_create: %s\n", s->err.c_str());
    delete s;
    return rc;
  }
  s->par = make_par(*cfg);
  const char* rk = getenv("COSAHIP_RESIDUAL");
  s->split_residual = rk && strcmp(rk, "split") == 0;
  if (rk && !s->split_residual)
    s->resid_force = strcmp(rk, "fused") == 0 ? RESID_GENERIC : strcmp(rk, "tile7") == 0 ? RESID_TILE7
                     : strcmp(rk, "tile7reg") == 0 ? RESID_TILE7_REG : strcmp(rk, "tile15") == 0 ? RESID_TILE15
                     : strcmp(rk, "tile7w4") == 0 ? RESID_TILE7_W4 : strcmp(rk, "sva7") == 0 ? RESID_SVA7
                     : strcmp(rk, "sva15") == 0 ? RESID_SVA15 : strcmp(rk, "sva7w4") == 0 ? RESID_SVA7_W4
                     : strcmp(rk, "tile32w4") == 0 ? RESID_TILE32_W4 : strcmp(rk, "sva32") == 0 ? RESID_SVA32 : -1;
  const char* fk = getenv("COSAHIP_FUSED_UPDATE");
  s->fuse_update = !(fk && strcmp(fk, "0") == 0);
  const char* gk = getenv("COSAHIP_GRAPH");
  s->use_graph = !(gk && strcmp(gk, "0") == 0);
  const char* ck = getenv("COSAHIP_CIRS_I");
  s->cirs_i_mode = !ck ? 0 : strcmp(ck, "lds") == 0 ? 1 : strcmp(ck, "iv") == 0 ? 2 : strcmp(ck, "line") == 0 ? 3 : 0;
  const char* lk = getenv("COSAHIP_CIRS_IL_LDS");
  if (lk) s->il_lds = std::min(65536, std::max(4096, atoi(lk)));
  const char* vk = getenv("COSAHIP_VISCOUS");
  s->fuse_viscous = vk && strcmp(vk, "fused") == 0;
  s->viscous_faces = !(vk && (strcmp(vk, "march") == 0 || strcmp(vk, "fused") == 0));
  s->viscous_tile = vk && strcmp(vk, "tile") == 0;
  const char* jk = getenv("COSAHIP_CIRS_JK");
  s->cirs_jk_chunk = !jk ? 4 : strcmp(jk, "plain") == 0 ? 0 : strcmp(jk, "chunk8") == 0 ? 8 : 4;
  if (device >= 0) {
    if (hipSetDevice(device) != hipSuccess) {
      delete s;
      return COSAHIP_ERR_DEVICE;
    }
  }
  if (hipGetDevice(&s->device) != hipSuccess || hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&s->dflag, 64) != hipSuccess || hipHostMalloc((void**)&s->hpinned, 64 * sizeof(double)) != hipSuccess ||
      hipMalloc(&s->ddhb, sizeof(double) * 131 * 131) != hipSuccess) {
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return COSAHIP_ERR_DEVICE;
  }
  // the fixed scratch array of the no-allocation force path on this device
  if (scratch_address(s->device, &s->scratch) != hipSuccess) {
    fprintf(stderr, "cosahip_create: hipGetSymbolAddress of the force scratch array failed\n");
    s->scratch = nullptr;
  }
  s->lev.resize(cfg->nlevel);
  *out = reinterpret_cast<cosahip_solve*>(s);
  return COSAHIP_OK;
}

int cosahip_destroy(cosahip_solve* h) {
  Solve* s = reinterpret_cast<Solve*>(h);
  if (!s) return COSAHIP_OK;
  (void)hipSetDevice(s->device);
  (void)hipStreamSynchronize(s->stream);
  drain_timing(*s);
  if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
  for (hipEvent_t e : s->evpool) (void)hipEventDestroy(e);
  for (auto& L : s->lev) free_level(L);
  if (s->dflag) (void)hipFree(s->dflag);
  if (s->ddhb) (void)hipFree(s->ddhb);
  if (s->hpinned) (void)hipHostFree(s->hpinned);
  if (s->hpart) (void)hipHostFree(s->hpart);
  if (s->dqfst) (void)hipFree(s->dqfst);
  if (s->ck_stream) {
    (void)hipStreamSynchronize(s->ck_stream);
    (void)hipStreamDestroy(s->ck_stream);
  }
  if (s->ck_snap) (void)hipEventDestroy(s->ck_snap);
  if (s->ck_done) (void)hipEventDestroy(s->ck_done);
  if (s->ck_buf) (void)hipFree(s->ck_buf);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return COSAHIP_OK;
}

const char* cosahip_last_error(const cosahip_solve* h) {
  const Solve* s = reinterpret_cast<const Solve*>(h);
  return s ? s->err.c_str() : "null solve";
}

int cosahip_error_info(const cosahip_solve* h, int32_t info[8]) {
  const Solve* s = reinterpret_cast<const Solve*>(h);
  if (!s || !info) return COSAHIP_ERR_ARG;
  memcpy(info, s->errinfo, sizeof(s->errinfo));
  return COSAHIP_OK;
}

int cosahip_set_hb_operator(cosahip_solve* h, int nsnap, const double* dhb) {
  Solve* s = reinterpret_cast<Solve*>(h);
  if (!s || !dhb) return COSAHIP_ERR_ARG;
  if (nsnap != s->par.nsnap) return fail(s, COSAHIP_ERR_ARG, "dhb size must be 2*nharms+1");
  memcpy(s->dhb, dhb, sizeof(double) * nsnap * nsnap);
  CK(hipSetDevice(s->device));
  CK(hipMemcpy(s->ddhb, s->dhb, sizeof(double) * nsnap * nsnap, hipMemcpyHostToDevice));
  return COSAHIP_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// level definition

template <class T>
static hipError_t upload_vec(T** d, const std::vector<T>& h) {
  size_t n = std::max<size_t>(h.size(), 1);
  hipError_t e = hipMalloc(d, n * sizeof(T));
  if (e != hipSuccess) return e;
  if (!h.empty()) e = hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
  return e;
}

static void add_tiles(std::vector<Tile>& v, int block, ll items, int per) {
  for (ll b0 = 0; b0 < items; b0 += per) v.push_back(Tile{block, (int)b0});
}

// k_cirs_il (CIRS i-sweep, constant coefficient of irsop 1): how many whole
// lines of all npde variables one work-group keeps in LDS, within
// s->il_lds bytes (two tables, lines, per-row offsets; default 24 KB, so that
// six work-groups share a CU), raised to what one line needs for long lines
// (at most 64 KB).

static void plan_cirs_il(Solve* s, Level& L) {
  L.il_lpw = 0;
  L.il_stride = 0;
  if (s->par.irsop != 1 || !L.nblocks) return;
  int lmax = 1;
  for (const Blk& B : L.hblk) lmax = std::max(lmax, B.imax - 1);
  const int S = lmax | 1;
  const int nv = s->par.npde;
  const int budget = std::max(s->il_lds, std::min(65536, 16 * S + nv * (8 * S + 16)));
  const int lpw = std::min((budget - 16 * S) / (nv * (8 * S + 16)), 256 / nv);
  if (lpw < 1) return;
  L.il_lpw = lpw;
  L.il_stride = S;
}

// Per block, the Thomas multipliers r(a) = c(a) / (b(a) - r(a-1) a(a)) of
// the constant coefficient epsc along an i-line (tridi, src/cosa.f90:
// 2418-2443), the recurrence of fill_rtab in k_flux.hip evaluated on the
// host with the same operations (correctly rounded divisions, no
// contraction), so the values are the same; and the reciprocals 1 / d(a)
// of the forward-pass divisors d(1) = b, d(a) = b - r(a-1) a(a), rounded
// to nearest.  Layout: block b at 2 il_stride b, r then 1/d.
static int upload_cirs_il_table(Solve* s, Level& L) {
  if (L.il_rtab) (void)hipFree(L.il_rtab);
  L.il_rtab = nullptr;
  if (!L.il_lpw) return COSAHIP_OK;
  const double epsc = 0.25 * (s->par.rat * s->par.rat - 1.0);
  const size_t S = L.il_stride;
  std::vector<double> tab((size_t)L.nblocks * 2 * S, 0.0);
  for (int b = 0; b < L.nblocks; b++) {
    const int n = L.hblk[b].imax - 1;
    double rp = 0.0;
    for (int a = 1; a <= n; a++) {
      const double ca = a == 1 ? 0.0 : -epsc, cb = 1 + 2 * epsc, cc = a == n ? 0.0 : -epsc;
      const double d = a == 1 ? cb : cb - rp * ca;
      const double r = a == 1 ? cc / cb : cc / (cb - rp * ca);
      tab[2 * S * b + a - 1] = r;
      tab[2 * S * b + S + a - 1] = 1.0 / d;
      rp = r;
    }
  }
  CK(upload_vec(&L.il_rtab, tab));
  return COSAHIP_OK;
}

// The same tables for the j and k line sweeps of k_cirs_jk (irsop 1, lines
// up to JK_RT - 1 cells): block b of direction d at 2 JK_RT b, r(a) at a,
// 1/d(a) at JK_RT + a, a = 1..L.
static int upload_cirs_jk_tables(Solve* s, Level& L) {
  for (int d = 0; d < 2; d++) {
    if (L.jk_rtab[d]) (void)hipFree(L.jk_rtab[d]);
    L.jk_rtab[d] = nullptr;
  }
  if (s->par.irsop != 1) return COSAHIP_OK;
  const double epsc = 0.25 * (s->par.rat * s->par.rat - 1.0);
  for (int d = 0; d < 2; d++) {
    std::vector<double> tab((size_t)L.nblocks * 2 * JK_RT, 0.0);
    for (int b = 0; b < L.nblocks; b++) {
      const int n = (d == 0 ? L.hblk[b].jmax : L.hblk[b].kmax) - 1;
      if (n >= JK_RT) continue;
      double rp = 0.0;
      for (int a = 1; a <= n; a++) {
        const double ca = a == 1 ? 0.0 : -epsc, cb = 1 + 2 * epsc, cc = a == n ? 0.0 : -epsc;
        const double dd = a == 1 ? cb : cb - rp * ca;
        const double r = a == 1 ? cc / cb : cc / (cb - rp * ca);
        tab[2 * JK_RT * b + a] = r;
        tab[2 * JK_RT * b + JK_RT + a] = 1.0 / dd;
        rp = r;
      }
    }
    CK(upload_vec(&L.jk_rtab[d], tab));
  }
  return COSAHIP_OK;
}

static int build_tables(Solve* s, Level& L, const Level* Lc) {
  const int ns = s->par.nsnap;
  for (int k = 0; k < TK_NKIND; k++) L.htile[k].clear();
  plan_cirs_il(s, L);
  for (int b = 0; b < L.nblocks; b++) {
    const Blk& B = L.hblk[b];
    const ll im = B.imax - 1, jm = B.jmax - 1, km = B.kmax - 1;
    add_tiles(L.htile[TK_CELLS], b, ns * im * jm * km, TS);
    const ll nseg = (im + ROWSEG - 1) / ROWSEG;
    add_tiles(L.htile[TK_ROWS_I], b, ns * jm * km * nseg, 4);
    add_tiles(L.htile[TK_LINES_J], b, ns * im * km, TS);
    add_tiles(L.htile[TK_LINES_K], b, ns * im * jm, TS);
    add_tiles(L.htile[TK_SEGS_J], b, ns * im * km * ((jm + MSEG - 1) / MSEG), TS);
    add_tiles(L.htile[TK_SEGS_K], b, ns * im * jm * ((km + MSEG - 1) / MSEG), TS);
    add_tiles(L.htile[TK_LINES_I], b, ns * jm * km, 64);
    add_tiles(L.htile[TK_LINES_I_T], b, ns * jm * km, TS);
    add_tiles(L.htile[TK_LINES_IV], b, ns * jm * km, 256 / s->par.npde);
    if (L.il_lpw) add_tiles(L.htile[TK_LINES_IL], b, ns * jm * km, L.il_lpw);
    if (s->par.viscous) {
      add_tiles(L.htile[TK_FACES_I], b, ns * (im + 1) * jm * km, TS);
      add_tiles(L.htile[TK_FACES_J], b, ns * im * (jm + 1) * km, TS);
      add_tiles(L.htile[TK_FACES_K], b, ns * im * jm * (km + 1), TS);
    }
    add_tiles(L.htile[TK_CELLS_P1], b, ns * (ll)(B.imax + 1) * (B.jmax + 1) * (B.kmax + 1), TS);
    add_tiles(L.htile[TK_SHELL], b, ns * (2 * (ll)(B.jmax + 1) * (B.kmax + 1) + im * (2 * (ll)(B.kmax + 1) + 2 * jm)),
              TS);
    add_tiles(L.htile[TK_HALO], b, ns * ((ll)(B.imax + 3) * jm * km + 4 * im * km + 4 * im * jm), TS);
    add_tiles(L.htile[TK_FUSED], b, residual_tile_items(RT_TI, 7, B.imax, B.jmax, B.kmax, ns), 1);
    add_tiles(L.htile[TK_FUSED15], b, residual_tile_items(RT_TI, 15, B.imax, B.jmax, B.kmax, ns), 1);
    add_tiles(L.htile[TK_T32], b, residual_tile_items(32, 7, B.imax, B.jmax, B.kmax, ns), 1);
    if (Lc) {
      const Blk& C = Lc->hblk[b];
      add_tiles(L.htile[TK_COARSE], b, ns * (ll)(C.imax - 1) * (C.jmax - 1) * (C.kmax - 1), TS);
      add_tiles(L.htile[TK_PROL], b, ns * (ll)((B.imax + 1) / 2) * ((B.jmax + 1) / 2) * ((B.kmax + 1) / 2), TS);
    }
  }
  for (int k = 0; k < TK_NKIND; k++) {
    if (L.tile[k]) (void)hipFree(L.tile[k]);
    L.tile[k] = nullptr;
    L.ntile[k] = (int)L.htile[k].size();
    CK(upload_vec(&L.tile[k], L.htile[k]));
  }
  int rc = upload_cirs_jk_tables(s, L);
  if (rc) return rc;
  return upload_cirs_il_table(s, L);
}

// Host replica of k_cut's index arithmetic: every halo target and interior
// source of the cut must lie inside its block's P3 extents.
static bool cut_in_bounds(const Blk& b1, const Blk& b2, const int32_t* t, int n2, int count) {
  const int idir1 = t[1], inout1 = t[2], idir2 = t[10], inout2 = t[11];
  int istr1[4], iend1[4], istr2[4], iend2[4], iord[4];
  istr1[1] = t[3]; iend1[1] = t[4]; istr1[2] = t[5]; iend1[2] = t[6]; istr1[3] = t[7]; iend1[3] = t[8];
  istr2[1] = t[12]; iend2[1] = t[13]; istr2[2] = t[14]; iend2[2] = t[15]; istr2[3] = t[16]; iend2[3] = t[17];
  iord[1] = t[18]; iord[2] = t[19]; iord[3] = t[20];
  for (int m = 1; m <= 3; m++)
    if (iord[m] < 1 || iord[m] > 3) return false;
  const int ijk1[4] = {0, b1.imax, b1.jmax, b1.kmax}, ijk2[4] = {0, b2.imax, b2.jmax, b2.kmax};
  const int ibcpt = inout1 == 1 ? 0 : ijk1[idir1], ibcinc = inout1 == 1 ? -1 : 1;
  const int inr = inout2 == 1 ? 1 : ijk2[idir2] - 1, inrinc = inout2 == 1 ? 1 : -1;
  auto cyc = [](int d, int m) { return (d - 1 + m - 1) % 3 + 1; };
  const int ic1 = cyc(idir1, 1), ic2 = cyc(idir1, 2), ic3 = cyc(idir1, 3);
  const int jc1 = iord[ic1], jc2 = iord[ic2], jc3 = iord[ic3];
  const int sg1c2 = iend1[ic2] > istr1[ic2] ? 1 : -1, sg1c3 = iend1[ic3] > istr1[ic3] ? 1 : -1;
  const int sg2c2 = iend2[jc2] > istr2[jc2] ? 1 : -1, sg2c3 = iend2[jc3] > istr2[jc3] ? 1 : -1;
  auto inside = [](const Blk& b, int i, int j, int k) {
    return i >= -1 && i <= b.imax + 1 && j >= -1 && j <= b.jmax + 1 && k >= -1 && k <= b.kmax + 1;
  };
  for (int item = 0; item < count; item++) {
    const int i2_ = item % n2, i3_ = item / n2;
    const int a2 = istr1[ic2] + sg1c2 * i2_, a3 = istr1[ic3] + sg1c3 * i3_;
    int x1[4] = {0, 0, 0, 0};
    x1[ic1] = ibcpt;
    x1[ic2] = a2;
    x1[ic3] = a3;
    int x2[4] = {0, x1[1], x1[2], x1[3]};
    x2[ic1] += ibcinc;
    const int c2 = istr2[jc2] + sg2c2 * i2_, c3 = istr2[jc3] + sg2c3 * i3_;
    int y1[4] = {0, 0, 0, 0};
    y1[jc1] = inr;
    y1[jc2] = c2;
    y1[jc3] = c3;
    int y2[4] = {0, y1[1], y1[2], y1[3]};
    y2[jc1] += inrinc;
    if (!inside(b1, x1[1], x1[2], x1[3]) || !inside(b1, x2[1], x2[2], x2[3]) || !inside(b2, y1[1], y1[2], y1[3]) ||
        !inside(b2, y2[1], y2[2], y2[3]))
      return false;
  }
  return true;
}

static ll field_size(const Solve* s, const Level& L, int field);

// FP32 shadows of the level's primitive q and face metrics (FP32-residual mode)
static int alloc_fp32(Solve* s, Level& L) {
  if (L.q32) return COSAHIP_OK;
  const ll nq = field_size(s, L, COSAHIP_Q), nm = field_size(s, L, COSAHIP_SI);
  if (hipMalloc(&L.q32, nq * sizeof(float)) != hipSuccess) return fail(s, COSAHIP_ERR_NOMEM, "device allocation failed");
  for (int d = 0; d < 3; d++)
    if (hipMalloc(&L.m32[d], nm * sizeof(float)) != hipSuccess)
      return fail(s, COSAHIP_ERR_NOMEM, "device allocation failed");
  L.m32_ok = false;
  if (s->par.viscous) {
    const ll nx = field_size(s, L, COSAHIP_XIDERI);
    for (int m = 0; m < 9; m++)
      if (hipMalloc(&L.xd32[m], nx * sizeof(float)) != hipSuccess)
        return fail(s, COSAHIP_ERR_NOMEM, "device allocation failed");
    std::vector<float*> t(L.xd32, L.xd32 + 9);
    if (hipMalloc(&L.dxd32, 9 * sizeof(float*)) != hipSuccess)
      return fail(s, COSAHIP_ERR_NOMEM, "device allocation failed");
    CK(hipMemcpy(L.dxd32, t.data(), 9 * sizeof(float*), hipMemcpyHostToDevice));
    L.xd32_ok = false;
  }
  if (s->par.kom) {
    if (hipMalloc(&L.mut32, field_size(s, L, COSAHIP_MUT) * sizeof(float)) != hipSuccess)
      return fail(s, COSAHIP_ERR_NOMEM, "device allocation failed");
    L.mut32_ok = false;
  }
  return COSAHIP_OK;
}
static bool strong_bc(int t) {
  return t == 13 || t == 14 || t / 100 == 15 || t == 22 || t == 23 || t == 92 || t == 93 || t == 5;
}

extern "C" int cosahip_define_level(cosahip_solve* h, int nl, int nblocks, const int32_t* imax, const int32_t* jmax,
                                    const int32_t* kmax, const int32_t* nbcs, const int32_t* bctopo, int ncuts,
                                    const int32_t* cutopo) {
  Solve* s = reinterpret_cast<Solve*>(h);
  if (!s) return COSAHIP_ERR_ARG;
  if (s->cycle_pending) return fail(s, COSAHIP_ERR_STATE, "define_level: an enqueued cycle has not been waited for");
  if (nl < 1 || nl > (int)s->lev.size()) return fail(s, COSAHIP_ERR_ARG, "level out of range");
  if (nblocks < 1 || !imax || !jmax || !kmax || !nbcs) return fail(s, COSAHIP_ERR_ARG, "bad block arrays");
  CK(hipSetDevice(s->device));
  if (nl > 1 && s->lev[0].nblocks && s->lev[0].nblocks != nblocks)
    return fail(s, COSAHIP_ERR_ARG, "every level has the blocks of level 1");
  if (nl == 1) {
    // per-block freestream, initially the configuration's for every block
    s->bqfst.assign(5 * (size_t)nblocks, 0.0);
    s->bflow.assign(3 * (size_t)nblocks, 0.0);
    for (int b = 0; b < nblocks; b++) {
      for (int v = 0; v < 5; v++) s->bqfst[5 * b + v] = s->par.qfst[v];
      s->bflow[3 * b] = s->cfg.machfs;
      s->bflow[3 * b + 1] = s->cfg.alpha;
      s->bflow[3 * b + 2] = s->cfg.beta;
    }
    if (s->dqfst) (void)hipFree(s->dqfst);
    s->dqfst = nullptr;
    if (hipMalloc(&s->dqfst, 5 * (size_t)nblocks * sizeof(double)) != hipSuccess)
      return fail(s, COSAHIP_ERR_NOMEM, "device allocation failed");
    CK(hipMemcpy(s->dqfst, s->bqfst.data(), 5 * (size_t)nblocks * sizeof(double), hipMemcpyHostToDevice));
  }
  Level& L = s->lev[nl - 1];
  // a captured cycle graph points at the buffers about to be freed
  if (s->gexec) {
    CK(hipStreamSynchronize(s->stream));
    CK(hipGraphExecDestroy(s->gexec));
    s->gexec = nullptr;
  }
  free_level(L);
  L.nl = nl;
  L.nblocks = nblocks;
  L.ncuts = ncuts;
  L.imax.assign(imax, imax + nblocks);
  L.jmax.assign(jmax, jmax + nblocks);
  L.kmax.assign(kmax, kmax + nblocks);
  L.nbcs.assign(nbcs, nbcs + nblocks);
  ll npatch = 0;
  for (int b = 0; b < nblocks; b++) {
    if (imax[b] < 2 || jmax[b] < 2 || kmax[b] < 2) return fail(s, COSAHIP_ERR_ARG, "block with fewer than 2 nodes");
    npatch += nbcs[b];
  }
  if (npatch && !bctopo) return fail(s, COSAHIP_ERR_ARG, "bctopo missing");
  L.bctopo.assign(bctopo, bctopo + 10 * npatch);
  if (ncuts && !cutopo) return fail(s, COSAHIP_ERR_ARG, "cutopo missing");
  L.cutopo.assign(cutopo, cutopo + 21 * (ll)ncuts);
  // blocks and offsets (setoffs)
  ll o3 = 0, o2 = 0, o1 = 0, o0 = 0;
  for (int b = 0; b < nblocks; b++) {
    Blk B;
    B.imax = imax[b];
    B.jmax = jmax[b];
    B.kmax = kmax[b];
    B.nbc = nbcs[b];
    B.o3 = o3;
    B.o2 = o2;
    B.o1 = o1;
    B.o0 = o0;
    o3 += n3of(B);
    o2 += n2of(B);
    o1 += n1of(B);
    o0 += (ll)B.imax * B.jmax * B.kmax;
    L.ncell += (ll)(B.imax - 1) * (B.jmax - 1) * (B.kmax - 1);
    if (B.imax - 1 < 2 || B.jmax - 1 < 2 || B.kmax - 1 < 2) L.thin = true;
    L.hblk.push_back(B);
  }
  L.n3tot = o3;
  L.n2tot = o2;
  L.n1tot = o1;
  L.n0tot = o0;
  // BC patches and strong-BC face flags
  ll pi = 0;
  L.sbc_off.resize(nblocks);
  ll soff = 0;
  for (int b = 0; b < nblocks; b++) {
    const Blk& B = L.hblk[b];
    L.sbc_off[b] = soff;
    soff += 2 * (ll)(B.jmax + 1) * (B.kmax + 1) + 2 * (ll)(B.imax + 1) * (B.kmax + 1) + 2 * (ll)(B.imax + 1) * (B.jmax + 1);
  }
  std::vector<unsigned char> sbc(std::max<ll>(soff, 1), 0);
  for (int b = 0; b < nblocks; b++) {
    const Blk& B = L.hblk[b];
    for (int ib = 0; ib < nbcs[b]; ib++, pi++) {
      const int32_t* t = bctopo + 10 * pi;
      PatchRec r;
      memset(&r, 0, sizeof(r));
      r.block = b;
      r.type = t[0];
      for (int m = 0; m < 10; m++) r.t[m] = t[m];
      const int idir = t[1];
      if (idir < 1 || idir > 3) return fail(s, COSAHIP_ERR_ARG, "bctopo: idir out of range");
      const int ic2 = (idir - 1 + 1) % 3 + 1, ic3 = (idir - 1 + 2) % 3 + 1;
      const int n2 = t[2 * ic2 + 2] - t[2 * ic2 + 1] + 1, n3c = t[2 * ic3 + 2] - t[2 * ic3 + 1] + 1;
      if (n2 < 0 || n3c < 0) return fail(s, COSAHIP_ERR_ARG, "bctopo: empty or reversed range");
      const int ext[4] = {0, B.imax - 1, B.jmax - 1, B.kmax - 1};
      if (t[2 * ic2 + 1] < 1 || t[2 * ic2 + 2] > ext[ic2] || t[2 * ic3 + 1] < 1 || t[2 * ic3 + 2] > ext[ic3])
        return fail(s, COSAHIP_ERR_ARG, "bctopo: range outside the block");
      r.count = n2 * n3c;
      r.n2 = n2 > 0 ? n2 : 1;
      const int ty = t[0];
      if (ty != 30 && ty != 14 && ty != 80 && ty != 4 && !(ty / 100 == 15 && s->par.viscous)) {
        char msg[112];
        snprintf(msg, sizeof msg, "BC type %d is not built in this version (30, 14, 80, 4 and, viscous, 15xx are)",
                 ty);
        return fail(s, COSAHIP_ERR_UNSUPPORTED, msg);
      }
      L.hpatch.push_back(r);
      L.maxpatch = std::max(L.maxpatch, r.count);
      if (strong_bc(ty)) {
        const int face = 2 * (idir - 1) + (t[2] == 1 ? 0 : 1);
        ll off = L.sbc_off[b];
        ll w;
        // flag byte: bit 0 strong BC; bits 1/2 first / last cell of the patch
        // along the lower-numbered transverse direction, bits 3/4 along the
        // other one (one-sided differences of bq_der at patch ends)
        int r1, r2;   // bctopo slots of the two transverse ranges
        if (idir == 1) {
          off += (face - 0) * (ll)(B.jmax + 1) * (B.kmax + 1);
          w = B.jmax + 1;
          r1 = 5; r2 = 7;
        } else if (idir == 2) {
          off += 2 * (ll)(B.jmax + 1) * (B.kmax + 1) + (face - 2) * (ll)(B.imax + 1) * (B.kmax + 1);
          w = B.imax + 1;
          r1 = 3; r2 = 7;
        } else {
          off += 2 * (ll)(B.jmax + 1) * (B.kmax + 1) + 2 * (ll)(B.imax + 1) * (B.kmax + 1) +
                 (face - 4) * (ll)(B.imax + 1) * (B.jmax + 1);
          w = B.imax + 1;
          r1 = 3; r2 = 5;
        }
        for (int c2 = t[r2]; c2 <= t[r2 + 1]; c2++)
          for (int c1 = t[r1]; c1 <= t[r1 + 1]; c1++)
            sbc[off + c1 + w * c2] = (unsigned char)(1 | (c1 == t[r1]) << 1 | (c1 == t[r1 + 1]) << 2 |
                                                     (c2 == t[r2]) << 3 | (c2 == t[r2 + 1]) << 4);
      }
    }
  }
  // cut records
  for (int ic = 0; ic < ncuts; ic++) {
    const int32_t* t = cutopo + 21 * (ll)ic;
    CutRec r;
    memset(&r, 0, sizeof(r));
    r.b1 = t[0] - 1;
    r.b2 = t[9] - 1;
    if (r.b1 < 0 || r.b1 >= nblocks || r.b2 < 0 || r.b2 >= nblocks) return fail(s, COSAHIP_ERR_ARG, "cutopo: block");
    for (int m = 0; m < 21; m++) r.t[m] = t[m];
    const int idir1 = t[1];
    if (idir1 < 1 || idir1 > 3 || t[10] < 1 || t[10] > 3) return fail(s, COSAHIP_ERR_ARG, "cutopo: idir");
    const int ic2 = (idir1 - 1 + 1) % 3 + 1, ic3 = (idir1 - 1 + 2) % 3 + 1;
    const int len2 = abs(t[2 * ic2 + 2] - t[2 * ic2 + 1]), len3 = abs(t[2 * ic3 + 2] - t[2 * ic3 + 1]);
    r.n2 = len2 + 1;
    r.count = (len2 + 1) * (len3 + 1);
    if (!cut_in_bounds(L.hblk[r.b1], L.hblk[r.b2], t, r.n2, r.count))
      return fail(s, COSAHIP_ERR_ARG, "cutopo: cut maps outside a block");
    L.hcut.push_back(r);
    L.maxcut = std::max(L.maxcut, r.count);
  }
  CK(upload_vec(&L.dblk, L.hblk));
  CK(upload_vec(&L.dpatch, L.hpatch));
  // one work tile per TS face items of each patch (all snapshots), so a BC
  // launch has exactly the work-groups its patches need
  L.hptile.clear();
  L.ptile_first.assign(1, 0);
  for (size_t ip = 0; ip < L.hpatch.size(); ip++) {
    add_tiles(L.hptile, (int)ip, (ll)L.hpatch[ip].count * s->par.nsnap, TS);
    L.ptile_first.push_back((int)L.hptile.size());
  }
  CK(upload_vec(&L.ptile, L.hptile));
  CK(upload_vec(&L.dcut, L.hcut));
  CK(upload_vec(&L.sbc, sbc));
  CK(upload_vec(&L.sbc_doff, L.sbc_off));
  // fields
  const Par& p = s->par;
  for (int f = 0; f < COSAHIP_NFIELDS; f++) {
    const ll sz = std::max<ll>(field_size(s, L, f), 1);
    if (hipMalloc(&L.f[f], sz * sizeof(double)) != hipSuccess)
      return fail(s, COSAHIP_ERR_NOMEM, "device allocation failed");
    CK(hipMemset(L.f[f], 0, sz * sizeof(double)));
  }
  if (p.viscous) {
    // one slot of face fluxes per direction (the tile arrangement keeps all
    // three; the others use the first)
    const ll nv = 3 * L.n0tot * p.npde * p.nsnap;
    if (hipMalloc(&L.vflux, nv * sizeof(double)) != hipSuccess)
      return fail(s, COSAHIP_ERR_NOMEM, "device allocation failed");
    CK(hipMemset(L.vflux, 0, nv * sizeof(double)));
  }
  if (p.kom) {
    for (int d = 0; d < 3; d++) {
      if (hipMalloc(&L.tbeta[d], L.n1tot * p.nsnap * sizeof(double)) != hipSuccess)
        return fail(s, COSAHIP_ERR_NOMEM, "device allocation failed");
      CK(hipMemset(L.tbeta[d], 0, L.n1tot * p.nsnap * sizeof(double)));
    }
    if (hipMalloc(&L.h0tab, (ll)nblocks * p.nsnap * 6 * sizeof(double)) != hipSuccess)
      return fail(s, COSAHIP_ERR_NOMEM, "device allocation failed");
  }
  if (p.kom && p.viscous) {
    const ll nc = L.n3tot * 9 * p.nsnap_met;
    if (hipMalloc(&L.cmet, nc * sizeof(double)) != hipSuccess || hipMalloc(&L.cmet32, nc * sizeof(float)) != hipSuccess)
      return fail(s, COSAHIP_ERR_NOMEM, "device allocation failed");
  }
  if (p.sst == 2) {
    const ll nd = L.n3tot * 9 * p.nsnap;
    if (hipMalloc(&L.dqc, nd * sizeof(double)) != hipSuccess)
      return fail(s, COSAHIP_ERR_NOMEM, "device allocation failed");
    CK(hipMemset(L.dqc, 0, nd * sizeof(double)));
  }
  if (p.viscous) {
    std::vector<double*> xdp(L.f + COSAHIP_XIDERI, L.f + COSAHIP_XIDERI + 9);
    CK(upload_vec(&L.dxd, xdp));
  }
  const ll nq = L.n3tot * p.npde * p.nsnap;
  if (hipMalloc(&L.dq, nq * sizeof(double)) != hipSuccess || hipMalloc(&L.work1, nq * sizeof(double)) != hipSuccess ||
      hipMalloc(&L.cirs_r, L.n3tot * p.nsnap * sizeof(double)) != hipSuccess ||
      (p.kom && hipMalloc(&L.cirs_rt, L.n3tot * p.nsnap * sizeof(double)) != hipSuccess))
    return fail(s, COSAHIP_ERR_NOMEM, "device allocation failed");
  CK(hipMemset(L.dq, 0, nq * sizeof(double)));
  CK(hipMemset(L.work1, 0, nq * sizeof(double)));
  int rc = build_tables(s, L, nl < (int)s->lev.size() && s->lev[nl].nblocks ? &s->lev[nl] : nullptr);
  if (rc) return rc;
  if (nl >= 2 && s->lev[nl - 2].nblocks) {
    rc = build_tables(s, s->lev[nl - 2], &L);
    if (rc) return rc;
  }
  CK(hipMalloc(&L.rms_part, ((ll)L.ntile[TK_CELLS] + 1) * 8 * sizeof(double)));
  if (s->prec_mode == 1 || s->prec_mode == 2) return alloc_fp32(s, L);
  return COSAHIP_OK;
}

// ---------------------------------------------------------------------------
// transfers

static ll field_size(const Solve* s, const Level& L, int field) {
  const Par& p = s->par;
  switch (field) {
    case COSAHIP_Q: case COSAHIP_QOLD: case COSAHIP_RES: case COSAHIP_RHS: return L.n3tot * p.npde * p.nsnap;
    case COSAHIP_SI: case COSAHIP_SJ: case COSAHIP_SK: return L.n2tot * p.lmet * p.nsnap_met;
    case COSAHIP_VOL: return L.n1tot;
    case COSAHIP_DTVOL: case COSAHIP_BETAXI: case COSAHIP_BETAETA: case COSAHIP_BETAZETA: return L.n1tot * p.nsnap;
    case COSAHIP_XDOT: case COSAHIP_YDOT: case COSAHIP_ZDOT: return p.moving ? L.n2tot * p.nsnap_met : 0;
    case COSAHIP_MUT: case COSAHIP_PRD: case COSAHIP_DIVVEL: case COSAHIP_DELPLS:
      return p.kom ? L.n3tot * p.nsnap : 0;
    case COSAHIP_DIST: return p.viscous ? L.n1tot : 0;
    case COSAHIP_DTVOLT: return p.kom ? L.n1tot * p.nsnap : 0;
    case COSAHIP_X: case COSAHIP_Y: case COSAHIP_Z: return L.n2tot * p.nsnap_met;
    case COSAHIP_TTRM: return p.sst ? L.n3tot * 3 * p.nsnap : 0;
    case COSAHIP_Q1: case COSAHIP_Q2: return p.dualt && L.nl == 1 ? L.n3tot * p.npde : 0;
  }
  if (field >= COSAHIP_XIDERI && field <= COSAHIP_ZETADERK) return p.viscous ? L.n0tot * 3 * p.nsnap_met : 0;
  return -1;
}

static Level* get_level(Solve* s, int nl) {
  if (!s || nl < 1 || nl > (int)s->lev.size() || !s->lev[nl - 1].nblocks) return nullptr;
  return &s->lev[nl - 1];
}

extern "C" int64_t cosahip_field_size(const cosahip_solve* h, int nl, int field) {
  const Solve* s = reinterpret_cast<const Solve*>(h);
  if (!s || nl < 1 || nl > (int)s->lev.size() || !s->lev[nl - 1].nblocks) return -1;
  if (field < 0 || field >= COSAHIP_NFIELDS) return -1;
  return field_size(s, s->lev[nl - 1], field);
}

extern "C" int cosahip_upload(cosahip_solve* h, int nl, int field, const double* host, int64_t count) {
  Solve* s = reinterpret_cast<Solve*>(h);
  Level* L = get_level(s, nl);
  if (!L) return fail(s, COSAHIP_ERR_STATE, "upload: level not defined");
  if (field < 0 || field >= COSAHIP_NFIELDS || !host) return fail(s, COSAHIP_ERR_ARG, "upload: bad field");
  if (count != field_size(s, *L, field)) return fail(s, COSAHIP_ERR_ARG, "upload: count does not match the field size");
  if (!count) return COSAHIP_OK;
  CK(hipSetDevice(s->device));
  CK(hipStreamSynchronize(s->stream));
  CK(hipMemcpy(L->f[field], host, count * sizeof(double), hipMemcpyHostToDevice));
  if (field == COSAHIP_SI || field == COSAHIP_SJ || field == COSAHIP_SK) L->m32_ok = false;
  if (field >= COSAHIP_XIDERI && field <= COSAHIP_ZETADERK) L->xd32_ok = L->cmet_ok = L->cmet32_ok = false;
  if (field == COSAHIP_MUT) L->mut32_ok = false;
  // a captured cycle replays the conversions of the state it was captured
  // in: new metrics or mut need a fresh capture
  if (s->gexec && ((field >= COSAHIP_SI && field <= COSAHIP_SK) ||
                   (field >= COSAHIP_XIDERI && field <= COSAHIP_ZETADERK) || field == COSAHIP_MUT)) {
    CK(hipGraphExecDestroy(s->gexec));
    s->gexec = nullptr;
  }
  return COSAHIP_OK;
}

// ---------------------------------------------------------------------------
// grid motion and metric set-up on the device (src/cosa.f90:297-322 per
// level: movegrid, vols, imetrix, volex + cutman_v, vol_edges, coordex +
// cutman_x, vmetrix; coarsen, src/cosa.f90:160-166)

// cut_v / cut_x records in the flat form k_cut_layer reads (src/grid.f90:
// 612-838, 242-524): ranges, running directions and the halo / interior
// indices of either kind
static GeomCut geom_cut(const CutRec& r, const Level& L, int kind) {
  static const int CYC[3][3] = {{1, 2, 3}, {2, 3, 1}, {3, 1, 2}};
  const int32_t* t = r.t;
  const Blk& B1 = L.hblk[r.b1];
  const Blk& B2 = L.hblk[r.b2];
  const int ijk1[3] = {B1.imax, B1.jmax, B1.kmax}, ijk2[3] = {B2.imax, B2.jmax, B2.kmax};
  const int idir1 = t[1], inout1 = t[2], idir2 = t[10], inout2 = t[11];
  int str1[3] = {t[3], t[5], t[7]}, end1[3] = {t[4], t[6], t[8]};
  int str2[3] = {t[12], t[14], t[16]}, end2[3] = {t[13], t[15], t[17]};
  GeomCut g;
  g.b1 = r.b1;
  g.b2 = r.b2;
  if (kind == 0) {
    g.ibcpt = inout1 == 1 ? 0 : ijk1[idir1 - 1];
    g.inr = inout2 == 1 ? 1 : ijk2[idir2 - 1] - 1;
  } else {
    g.ibcpt = inout1 == 1 ? 0 : ijk1[idir1 - 1] + 1;
    g.inr = inout2 == 1 ? 2 : ijk2[idir2 - 1] - 1;
    for (int l = 1; l <= 3; l++) {
      if (l != idir1) (end1[l - 1] > str1[l - 1] ? end1[l - 1] : str1[l - 1]) += 1;
      if (l != idir2) (end2[l - 1] > str2[l - 1] ? end2[l - 1] : str2[l - 1]) += 1;
    }
  }
  int len[3];
  for (int l = 0; l < 3; l++) {
    len[l] = std::abs(end1[l] - str1[l]);
    g.str1[l] = str1[l];
    g.str2[l] = str2[l];
    g.sg1[l] = end1[l] > str1[l] ? 1 : -1;
    g.sg2[l] = end2[l] > str2[l] ? 1 : -1;
  }
  g.ic1 = CYC[idir1 - 1][0] - 1;
  g.ic2 = CYC[idir1 - 1][1] - 1;
  g.ic3 = CYC[idir1 - 1][2] - 1;
  g.jc1 = t[18 + g.ic1] - 1;
  g.jc2 = t[18 + g.ic2] - 1;
  g.jc3 = t[18 + g.ic3] - 1;
  g.n2 = len[g.ic2] + 1;
  g.count = g.n2 * (len[g.ic3] + 1);
  return g;
}

// every index k_cut_layer forms must lie inside its block's extents (P1
// cells 0..imax for kind 0, P2 nodes 0..imax+1 for kind 1)
static bool geom_cut_in_bounds(const GeomCut& g, const Level& L, int kind) {
  const Blk& B1 = L.hblk[g.b1];
  const Blk& B2 = L.hblk[g.b2];
  const int hi1[3] = {B1.imax + kind, B1.jmax + kind, B1.kmax + kind};
  const int hi2[3] = {B2.imax + kind, B2.jmax + kind, B2.kmax + kind};
  const int n3 = g.count / g.n2;
  for (int corner = 0; corner < 4; corner++) {
    const int i2 = (corner & 1) ? g.n2 - 1 : 0, i3 = (corner & 2) ? n3 - 1 : 0;
    int p1[3], p2[3];
    p1[g.ic1] = g.ibcpt;
    p1[g.ic2] = g.str1[g.ic2] + g.sg1[g.ic2] * i2;
    p1[g.ic3] = g.str1[g.ic3] + g.sg1[g.ic3] * i3;
    p2[g.jc1] = g.inr;
    p2[g.jc2] = g.str2[g.jc2] + g.sg2[g.jc2] * i2;
    p2[g.jc3] = g.str2[g.jc3] + g.sg2[g.jc3] * i3;
    for (int d = 0; d < 3; d++)
      if (p1[d] < 0 || p1[d] > hi1[d] || p2[d] < 0 || p2[d] > hi2[d]) return false;
  }
  return true;
}

static int build_grid_tiles(Solve* s, Level& L) {
  GridTiles& gt = L.gt;
  if (gt.built) return COSAHIP_OK;
  const int nsm = s->par.nsnap_met;
  for (int b = 0; b < L.nblocks; b++) {
    const Blk& B = L.hblk[b];
    const ll I = B.imax, J = B.jmax, K = B.kmax;
    add_tiles(gt.h[GT_CELLS], b, (I - 1) * (J - 1) * (K - 1), TS);
    add_tiles(gt.h[GT_FI], b, nsm * I * (J - 1) * (K - 1), TS);
    add_tiles(gt.h[GT_FJ], b, nsm * (I - 1) * J * (K - 1), TS);
    add_tiles(gt.h[GT_FK], b, nsm * (I - 1) * (J - 1) * K, TS);
    add_tiles(gt.h[GT_VOLEX], b, 2 * ((I - 1) * (K - 1) + (J - 1) * (K - 1) + (I - 1) * (J - 1)), TS);
    add_tiles(gt.h[GT_EDGES], b, 4 * ((K - 1) + (J - 1) + (I - 1)), TS);
    add_tiles(gt.h[GT_NODEHALO], b, nsm * 2 * (I * K + J * K + I * J), TS);
    add_tiles(gt.h[GT_NODES], b, nsm * I * J * K, TS);
    add_tiles(gt.h[GT_NODES0], b, I * J * K, TS);
  }
  for (int k = 0; k < GT_NKIND; k++) {
    gt.n[k] = (int)gt.h[k].size();
    CK(upload_vec(&gt.t[k], gt.h[k]));
  }
  std::vector<GeomCut> cv, cx;
  for (const CutRec& r : L.hcut) {
    cv.push_back(geom_cut(r, L, 0));
    cx.push_back(geom_cut(r, L, 1));
    if (!geom_cut_in_bounds(cv.back(), L, 0) || !geom_cut_in_bounds(cx.back(), L, 1))
      return fail(s, COSAHIP_ERR_ARG, "grid: a cut record maps outside a block");
    gt.maxcut_v = std::max(gt.maxcut_v, cv.back().count);
    gt.maxcut_x = std::max(gt.maxcut_x, cx.back().count);
  }
  gt.ncut = (int)cv.size();
  CK(upload_vec(&gt.cut_v, cv));
  CK(upload_vec(&gt.cut_x, cx));
  if (!L.gbad) CK(hipMalloc(&L.gbad, sizeof(int)));
  gt.built = true;
  return COSAHIP_OK;
}

// the metrics of a level changed: derived device data and a captured cycle
// graph are stale
static int metrics_changed(Solve* s, Level& L) {
  L.m32_ok = false;
  L.xd32_ok = false;
  L.cmet_ok = false;
  L.cmet32_ok = false;
  if (s->gexec) {
    CK(hipGraphExecDestroy(s->gexec));
    s->gexec = nullptr;
  }
  return COSAHIP_OK;
}

extern "C" int cosahip_grid_metrics(cosahip_solve* h, int nl) {
  Solve* s = reinterpret_cast<Solve*>(h);
  Level* Lp = get_level(s, nl);
  if (!Lp) return fail(s, COSAHIP_ERR_STATE, "grid_metrics: level not defined");
  if (s->cycle_pending) return fail(s, COSAHIP_ERR_STATE, "grid_metrics: a cycle is in flight");
  CK(hipSetDevice(s->device));
  Level& L = *Lp;
  int rc = build_grid_tiles(s, L);
  if (rc) return rc;
  const cosahip_config& c = s->cfg;
  const int do_volex = c.viscous || c.nlevel > 1;
  const int do_bedges = nl < c.nlevel && !c.bln_prol;
  const int do_vmetrix = c.viscous;
  CK(hipMemsetAsync(L.gbad, 0, sizeof(int), s->stream));
  CK(launch_grid_metrics(*s, L, L.gt, L.gbad, do_volex, do_bedges, do_vmetrix));
  int bad = 0;
  CK(hipMemcpyAsync(&bad, L.gbad, sizeof(int), hipMemcpyDeviceToHost, s->stream));
  CK(hipStreamSynchronize(s->stream));
  rc = metrics_changed(s, L);
  if (rc) return rc;
  if (bad) return fail(s, COSAHIP_ERR_ARG, "grid_metrics: a face area is not positive (imetrix_b stops here)");
  return COSAHIP_OK;
}

extern "C" int cosahip_grid_coarsen(cosahip_solve* h, int nl) {
  Solve* s = reinterpret_cast<Solve*>(h);
  Level* F = get_level(s, nl);
  Level* C = get_level(s, nl + 1);
  if (!F || !C) return fail(s, COSAHIP_ERR_STATE, "grid_coarsen: levels nl and nl + 1 must be defined");
  if (s->cycle_pending) return fail(s, COSAHIP_ERR_STATE, "grid_coarsen: a cycle is in flight");
  if (F->nblocks != C->nblocks) return fail(s, COSAHIP_ERR_ARG, "grid_coarsen: block counts differ");
  for (int b = 0; b < F->nblocks; b++) {
    const Blk& fb = F->hblk[b];
    const Blk& cb = C->hblk[b];
    if (2 * cb.imax - 1 > fb.imax || 2 * cb.jmax - 1 > fb.jmax || 2 * cb.kmax - 1 > fb.kmax)
      return fail(s, COSAHIP_ERR_ARG, "grid_coarsen: a coarse block is not the fine block's every second node");
  }
  CK(hipSetDevice(s->device));
  int rc = build_grid_tiles(s, *C);
  if (rc) return rc;
  CK(launch_grid_coarsen(*s, *F, *C, C->gt));
  CK(hipStreamSynchronize(s->stream));
  return COSAHIP_OK;
}

extern "C" int cosahip_grid_pitch(cosahip_solve* h, int nl, double dtheta0, double omegas, double xh, double yh,
                                  const double* pose) {
  Solve* s = reinterpret_cast<Solve*>(h);
  Level* Lp = get_level(s, nl);
  if (!Lp) return fail(s, COSAHIP_ERR_STATE, "grid_pitch: level not defined");
  if (!s->par.moving) return fail(s, COSAHIP_ERR_ARG, "grid_pitch: the configuration is not moving");
  if (!pose) return fail(s, COSAHIP_ERR_ARG, "grid_pitch: pose is null");
  if (s->cycle_pending) return fail(s, COSAHIP_ERR_STATE, "grid_pitch: a cycle is in flight");
  CK(hipSetDevice(s->device));
  Level& L = *Lp;
  int rc = build_grid_tiles(s, L);
  if (rc) return rc;
  const int nsm = s->par.nsnap_met;
  if (!L.xbase) {
    // the base pose: snapshot 0 of X, Y, Z as first uploaded (x0 of movegrid)
    CK(hipMalloc(&L.xbase, 3 * L.n2tot * sizeof(double)));
    const int F[3] = {COSAHIP_X, COSAHIP_Y, COSAHIP_Z};
    for (int m = 0; m < 3; m++)
      for (int b = 0; b < L.nblocks; b++) {
        const Blk& B = L.hblk[b];
        CK(hipMemcpyAsync(L.xbase + m * L.n2tot + B.o2, L.f[F[m]] + B.o2 * nsm, n2of(B) * sizeof(double),
                          hipMemcpyDeviceToDevice, s->stream));
      }
  }
  if (!L.gpose) CK(hipMalloc(&L.gpose, 3 * (size_t)nsm * sizeof(double)));
  CK(hipMemcpyAsync(L.gpose, pose, 3 * (size_t)nsm * sizeof(double), hipMemcpyHostToDevice, s->stream));
  CK(launch_grid_pitch(*s, L, L.gt, L.gpose, dtheta0, omegas, xh, yh, L.xbase));
  CK(hipStreamSynchronize(s->stream));
  return COSAHIP_OK;
}

// Wall runs of surfaces 1..nsurface (types 1500 + s - 1, and type 14 as
// surface 1 when include_inviscid) in the order force_b visits them: blocks,
// then patches, then (i3, i2).  emit(run) is called once per run, in order;
// returns an error code, *nfaces the number of wall faces.
template <typename Emit>
static int enum_wall_runs(Solve* s, const Level& L, int nsurface, int include_inviscid, int* nfaces, Emit emit) {
  const Par& p = s->par;
  int nf = 0, off = 0;
  for (int b = 0; b < L.nblocks; b++) {
    for (int ib = 0; ib < L.nbcs[b]; ib++) {
      const int32_t* t = L.bctopo.data() + 10 * (ll)(off + ib);
      int surf = 0, visc = 0;
      if (t[0] >= 1500 && t[0] < 1500 + nsurface) {
        surf = t[0] - 1499;
        visc = p.viscous ? 1 : 0;
      } else if (include_inviscid && t[0] == 14) {
        surf = 1;
      }
      if (!surf) continue;
      const int idir = t[1];
      if (idir < 1 || idir > 3) return fail(s, COSAHIP_ERR_ARG, "force: bad patch direction");
      const int ic2 = idir % 3 + 1, ic3 = (idir + 1) % 3 + 1;
      const int lo[3] = {t[3], t[5], t[7]}, hi[3] = {t[4], t[6], t[8]};
      const int n2 = hi[ic2 - 1] - lo[ic2 - 1] + 1, n3 = hi[ic3 - 1] - lo[ic3 - 1] + 1;
      if (n2 <= 0 || n3 <= 0) continue;
      WallRun w;
      w.block = b;
      w.surface = surf;
      w.visc = visc;
      w.idir = idir;
      w.inrout = t[2];
      w.lo2 = lo[ic2 - 1];
      w.n2 = n2;
      w.lo3 = lo[ic3 - 1];
      w.first = nf;
      w.count = n2 * n3;
      int rc = emit(w);
      if (rc) return rc;
      nf += n2 * n3;
    }
    off += L.nbcs[b];
  }
  *nfaces = nf;
  return COSAHIP_OK;
}

static int build_wall_runs(Solve* s, Level& L, int nsurface, int include_inviscid) {
  const int key = nsurface * 2 + (include_inviscid ? 1 : 0);
  if (L.force_key == key) return COSAHIP_OK;
  std::vector<WallRun> runs;
  int nf = 0;
  int rc = enum_wall_runs(s, L, nsurface, include_inviscid, &nf, [&](const WallRun& w) {
    runs.push_back(w);
    return COSAHIP_OK;
  });
  if (rc) return rc;
  void* old[] = {L.dwall, L.fterms, L.fsums, L.fcoef, L.fgeom};
  for (void* q : old)
    if (q) CK(hipFree(q));
  L.dwall = nullptr;
  L.fterms = L.fsums = L.fcoef = L.fgeom = nullptr;
  L.force_key = -1;
  const int nsnap = s->par.nsnap;
  if (hipMalloc(&L.dwall, std::max<size_t>(runs.size(), 1) * sizeof(WallRun)) != hipSuccess ||
      hipMalloc(&L.fterms, std::max<ll>((ll)nf * nsnap * 12, 1) * sizeof(double)) != hipSuccess ||
      hipMalloc(&L.fsums, (ll)L.nblocks * nsnap * nsurface * 12 * sizeof(double)) != hipSuccess ||
      hipMalloc(&L.fcoef, (ll)nsnap * nsurface * 12 * sizeof(double)) != hipSuccess ||
      hipMalloc(&L.fgeom, (13 + 3 * (ll)nsnap) * sizeof(double)) != hipSuccess)
    return fail(s, COSAHIP_ERR_NOMEM, "device allocation failed");
  if (!runs.empty()) CK(hipMemcpy(L.dwall, runs.data(), runs.size() * sizeof(WallRun), hipMemcpyHostToDevice));
  L.hwall = runs;
  L.nwall_faces = nf;
  L.force_key = key;
  return COSAHIP_OK;
}

// direction vectors of lift, drag and side force (src/output.f90:4451-4462),
// reference lengths, Mach number and moment centres: g[0 .. 13 + 3 nsnap)
static void force_geometry(const cosahip_config& c, int nsnap, const double* lngth, const double* momr, double* g) {
  const double alpha = c.alpha, beta = c.beta;
  const double den = std::sqrt(std::sin(beta) * std::sin(beta) + std::cos(alpha) * std::cos(alpha) *
                                                                     (std::cos(beta) * std::cos(beta)));
  g[0] = -std::sin(alpha) * std::cos(alpha) * (std::cos(beta) * std::cos(beta)) / den;
  g[1] = den;
  g[2] = std::sin(alpha) * std::sin(beta) * std::cos(beta) / den;
  g[3] = std::cos(alpha) * std::cos(beta);
  g[4] = std::sin(alpha) * std::cos(beta);
  g[5] = std::sin(beta);
  g[6] = -std::sin(beta) / den;
  g[7] = 0.0;
  g[8] = std::cos(alpha) * std::cos(beta) / den;
  g[9] = lngth[0];
  g[10] = lngth[1];
  g[11] = lngth[2];
  g[12] = c.machfs > 1e-14 ? c.machfs : 1.0;
  for (int m = 0; m < 3 * nsnap; m++) g[13 + m] = momr[m];
}

extern "C" int cosahip_force(cosahip_solve* h, int nl, int nsurface, int include_inviscid, const double* lngth,
                             const double* momr, double* coef) {
  Solve* s = reinterpret_cast<Solve*>(h);
  Level* Lp = get_level(s, nl);
  if (!Lp) return fail(s, COSAHIP_ERR_STATE, "force: level not defined");
  CK(hipSetDevice(s->device));
  Level& L = *Lp;
  if (nsurface < 1 || nsurface > 100 || !lngth || !momr || !coef)
    return fail(s, COSAHIP_ERR_ARG, "force: nsurface must be 1..100 and lngth, momr, coef non-null");
  int rc = build_wall_runs(s, L, nsurface, include_inviscid);
  if (rc) return rc;
  const int nsnap = s->par.nsnap;
  std::vector<double> g(13 + 3 * (size_t)nsnap);
  force_geometry(s->cfg, nsnap, lngth, momr, g.data());
  CK(hipMemcpyAsync(L.fgeom, g.data(), g.size() * sizeof(double), hipMemcpyHostToDevice, s->stream));
  const ForceBufs fb{L.dwall, (int)L.hwall.size(), L.nwall_faces, L.fterms, L.fsums, L.fcoef, L.fgeom, L.fgeom + 13};
  CK(launch_force(*s, L, nsurface, fb));
  CK(hipMemcpyAsync(coef, L.fcoef, (size_t)12 * nsnap * nsurface * sizeof(double), hipMemcpyDeviceToHost,
                    s->stream));
  CK(hipStreamSynchronize(s->stream));
  return COSAHIP_OK;
}

// ---------------------------------------------------------------------------
// force coefficients without device allocation: the buffers of one force
// evaluation live in a fixed device array of the library.  One array per
// GPU, shared by every solve on that GPU and kept until the process ends;
// calls that overlap on one GPU race on it.

static constexpr long long FORCE_SCRATCH_DOUBLES = 1ll << 20;   // 8 MiB
static __device__ __align__(256) double scratch[FORCE_SCRATCH_DOUBLES];
static constexpr int MAX_DEVICES = 64;
static double* g_scratch[MAX_DEVICES];   // device address of `scratch` on each GPU
static std::mutex g_scratch_mu;

// the address of this GPU's copy of `scratch` (the current device must be
// `device`); looked up once per device at solve creation
static hipError_t scratch_address(int device, double** out) {
  if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[device]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[device] = static_cast<double*>(ptr);
  }
  *out = g_scratch[device];
  return hipSuccess;
}

extern "C" int cosahip_force_noalloc(cosahip_solve* h, int nl, int nsurface, int include_inviscid,
                                     const double* lngth, const double* momr, double* coef) {
  Solve* s = reinterpret_cast<Solve*>(h);
  Level* Lp = get_level(s, nl);
  if (!Lp) return fail(s, COSAHIP_ERR_STATE, "force_noalloc: level not defined");
  if (nsurface < 1 || nsurface > 100 || !lngth || !momr || !coef)
    return fail(s, COSAHIP_ERR_ARG, "force_noalloc: nsurface must be 1..100 and lngth, momr, coef non-null");
  Level& L = *Lp;
  if (!s->scratch) return fail(s, COSAHIP_ERR_STATE, "force_noalloc: no scratch address for this device");
  CK(hipSetDevice(s->device));
  const int nsnap = s->par.nsnap;
  // count the runs and faces first: the layout must fit before anything is written
  int nruns = 0, nf = 0;
  int rc = enum_wall_runs(s, L, nsurface, include_inviscid, &nf, [&](const WallRun&) {
    nruns++;
    return COSAHIP_OK;
  });
  if (rc) return rc;
  // layout in doubles, each part 256-byte aligned: runs, geometry, coefficients, sums, terms
  auto up = [](ll n) { return (n + 31) / 32 * 32; };
  const ll n_runs = up(((ll)nruns * (ll)sizeof(WallRun) + 7) / 8);
  const ll n_geom = up(13 + 3 * (ll)nsnap);
  const ll n_coef = up((ll)nsnap * nsurface * 12);
  const ll n_sums = up((ll)L.nblocks * nsnap * nsurface * 12);
  const ll n_terms = up((ll)nf * nsnap * 12);
  const ll need = n_runs + n_geom + n_coef + n_sums + n_terms;
  if (need > FORCE_SCRATCH_DOUBLES) {
    char msg[160];
    snprintf(msg, sizeof msg, "force_noalloc: needs %lld doubles of scratch, the maximum is %lld", need,
             FORCE_SCRATCH_DOUBLES);
    return fail(s, COSAHIP_ERR_ARG, msg);
  }
  double* base = s->scratch;
  WallRun* d_runs = reinterpret_cast<WallRun*>(base);
  double* d_geom = base + n_runs;
  double* d_coef = d_geom + n_geom;
  double* d_sums = d_coef + n_coef;
  double* d_terms = d_sums + n_sums;
  // runs through a small host buffer, a chunk at a time (no host or device allocation)
  WallRun chunk[64];
  int nc = 0, done = 0;
  int nf2 = 0;
  hipError_t err = hipSuccess;
  auto flush = [&]() {
    if (nc && err == hipSuccess) {
      err = hipMemcpyAsync(d_runs + done, chunk, nc * sizeof(WallRun), hipMemcpyHostToDevice, s->stream);
      if (err == hipSuccess) err = hipStreamSynchronize(s->stream);
    }
    done += nc;
    nc = 0;
  };
  rc = enum_wall_runs(s, L, nsurface, include_inviscid, &nf2, [&](const WallRun& w) {
    chunk[nc++] = w;
    if (nc == 64) flush();
    return COSAHIP_OK;
  });
  if (rc) return rc;
  flush();
  CK(err);
  double g[13 + 3 * 131];
  if (nsnap > 131) return fail(s, COSAHIP_ERR_ARG, "force_noalloc: at most 131 snapshots");
  force_geometry(s->cfg, nsnap, lngth, momr, g);
  CK(hipMemcpyAsync(d_geom, g, (13 + 3 * (size_t)nsnap) * sizeof(double), hipMemcpyHostToDevice, s->stream));
  const ForceBufs fb{d_runs, nruns, nf, d_terms, d_sums, d_coef, d_geom, d_geom + 13};
  CK(launch_force(*s, L, nsurface, fb));
  CK(hipMemcpyAsync(coef, d_coef, (size_t)12 * nsnap * nsurface * sizeof(double), hipMemcpyDeviceToHost,
                    s->stream));
  CK(hipStreamSynchronize(s->stream));
  return COSAHIP_OK;
}

// ---------------------------------------------------------------------------
// blocks as independent solves (serving: requests of one grid batched as the
// blocks of one solve; no counterpart in the reference)

extern "C" int cosahip_set_block_freestream(cosahip_solve* h, int block, double machfs, double alpha,
                                            double beta) {
  Solve* s = reinterpret_cast<Solve*>(h);
  if (!s) return COSAHIP_ERR_ARG;
  const int nb = (int)s->bqfst.size() / 5;
  if (block < 1 || block > nb) return fail(s, COSAHIP_ERR_ARG, "set_block_freestream: block out of range");
  if (s->cycle_pending) return fail(s, COSAHIP_ERR_STATE, "set_block_freestream: a cycle is in flight");
  const cosahip_config& c = s->cfg;
  // viscous terms, the wall omega, mut and the viscous forces scale with the
  // configuration's Mach/Reynolds ratio (one value for the whole solve)
  if (s->par.viscous && machfs != c.machfs)
    return fail(s, COSAHIP_ERR_UNSUPPORTED,
                "set_block_freestream: a viscous solve keeps one Mach number (viscous scaling M/Re is per solve)");
  CK(hipSetDevice(s->device));
  double* f = s->bqfst.data() + 5 * (block - 1);
  // as the configuration's freestream (bc_far0, src/bc.f90:145-152)
  f[0] = c.rhoinf;
  f[1] = machfs * cos(alpha) * cos(beta);
  f[2] = machfs * sin(alpha) * cos(beta);
  f[3] = machfs * sin(beta);
  f[4] = c.pinf;
  s->bflow[3 * (block - 1)] = machfs;
  s->bflow[3 * (block - 1) + 1] = alpha;
  s->bflow[3 * (block - 1) + 2] = beta;
  CK(hipStreamSynchronize(s->stream));
  CK(hipMemcpy(s->dqfst + 5 * (block - 1), f, 5 * sizeof(double), hipMemcpyHostToDevice));
  return COSAHIP_OK;
}

extern "C" int cosahip_set_block_rms(cosahip_solve* h, int on) {
  Solve* s = reinterpret_cast<Solve*>(h);
  if (!s) return COSAHIP_ERR_ARG;
  if (!s->lev[0].nblocks) return fail(s, COSAHIP_ERR_STATE, "set_block_rms: level 1 not defined");
  CK(hipSetDevice(s->device));
  const ll n = (ll)s->lev[0].ntile[TK_CELLS] * 8;
  if (on && !s->hpart) {
    if (hipHostMalloc((void**)&s->hpart, std::max<ll>(n, 8) * sizeof(double)) != hipSuccess)
      return fail(s, COSAHIP_ERR_NOMEM, "pinned allocation failed");
    s->hpart_n = n;
    memset(s->hpart, 0, std::max<ll>(n, 8) * sizeof(double));
  }
  s->block_rms = on != 0;
  return COSAHIP_OK;
}

// resrms(1:npde+1) of every block from the level-1 partial sums of the last
// norm evaluation (tiles never span blocks; each block's tiles summed in
// order), as resg_rms / resg_rms_hb with the block's own cell count
extern "C" int cosahip_block_rms(cosahip_solve* h, double* out) {
  Solve* s = reinterpret_cast<Solve*>(h);
  if (!s || !out) return COSAHIP_ERR_ARG;
  if (!s->block_rms || !s->hpart) return fail(s, COSAHIP_ERR_STATE, "block_rms: not enabled");
  if (s->cycle_pending) return fail(s, COSAHIP_ERR_STATE, "block_rms: a cycle is in flight");
  CK(hipStreamSynchronize(s->stream));
  const Level& L = s->lev[0];
  const int np = s->par.npde;
  std::vector<double> sum((size_t)L.nblocks * 8, 0.0);
  for (int t = 0; t < L.ntile[TK_CELLS]; t++) {
    const int b = L.htile[TK_CELLS][t].block;
    for (int v = 0; v < np; v++) sum[8 * (size_t)b + v] = sum[8 * (size_t)b + v] + s->hpart[(ll)t * 8 + v];
  }
  for (int b = 0; b < L.nblocks; b++) {
    const Blk& B = L.hblk[b];
    const double nc = (double)((ll)(B.imax - 1) * (B.jmax - 1) * (B.kmax - 1));
    double tot = 0;
    for (int v = 0; v < np; v++) tot = tot + sum[8 * (size_t)b + v];
    double* o = out + 8 * (size_t)b;
    if (s->cfg.harbal) {
      const double d5 = s->par.nsnap;
      for (int v = 0; v < np; v++) o[v] = sqrt(sum[8 * (size_t)b + v] / nc / d5);
      o[np] = sqrt(tot / nc / d5 / np);
    } else {
      for (int v = 0; v < np; v++) o[v] = sqrt(sum[8 * (size_t)b + v] / nc);
      o[np] = sqrt(tot / nc / np);
    }
    for (int v = np + 1; v < 8; v++) o[v] = 0.0;
  }
  return COSAHIP_OK;
}

// force coefficients of every block with the block's own Mach number and
// angles: force_b's sums per block (k_force_terms, k_force_sums), then the
// projection of k_force_coef without the block sum (one block's share of
// cosahip_force, the same arithmetic)
extern "C" int cosahip_force_blocks(cosahip_solve* h, int nl, int nsurface, int include_inviscid,
                                    const double* lngth, const double* momr, double* coef) {
  Solve* s = reinterpret_cast<Solve*>(h);
  Level* Lp = get_level(s, nl);
  if (!Lp) return fail(s, COSAHIP_ERR_STATE, "force_blocks: level not defined");
  if (nsurface < 1 || nsurface > 100 || !lngth || !momr || !coef)
    return fail(s, COSAHIP_ERR_ARG, "force_blocks: nsurface must be 1..100 and lngth, momr, coef non-null");
  Level& L = *Lp;
  const int nsnap = s->par.nsnap;
  std::vector<double> tmp((size_t)12 * nsnap * nsurface);
  int rc = cosahip_force(h, nl, nsurface, include_inviscid, lngth, momr, tmp.data());   // fills L.fsums
  if (rc) return rc;
  std::vector<double> S((size_t)L.nblocks * nsnap * nsurface * 12);
  CK(hipMemcpy(S.data(), L.fsums, S.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int b = 0; b < L.nblocks; b++) {
    const double mach0 = s->bflow[3 * b], alpha = s->bflow[3 * b + 1], beta = s->bflow[3 * b + 2];
    const double den = std::sqrt(std::sin(beta) * std::sin(beta) + std::cos(alpha) * std::cos(alpha) *
                                                                       (std::cos(beta) * std::cos(beta)));
    const double aax = -std::sin(alpha) * std::cos(alpha) * (std::cos(beta) * std::cos(beta)) / den, aay = den,
                 aaz = std::sin(alpha) * std::sin(beta) * std::cos(beta) / den;
    const double ssx = std::cos(alpha) * std::cos(beta), ssy = std::sin(alpha) * std::cos(beta),
                 ssz = std::sin(beta);
    const double bbx = -std::sin(beta) / den, bby = 0.0, bbz = std::cos(alpha) * std::cos(beta) / den;
    const double l1 = lngth[0], l2 = lngth[1], l3 = lngth[2];
    const double mach = mach0 > 1e-14 ? mach0 : 1.0;
    for (int sf = 0; sf < nsurface; sf++)
      for (int n = 0; n < nsnap; n++) {
        const double* fp = S.data() + (((size_t)b * nsnap + n) * nsurface + sf) * 12;
        const double* fv = fp + 6;
        double* o = coef + (((size_t)b * nsurface + sf) * nsnap + n) * 12;
        const double tot0 = 0.0;
        o[0] = tot0 + (fp[0] * aax + fp[1] * aay + fp[2] * aaz) * 2 / (l1 * l2 * (mach * mach));
        o[1] = tot0 + (fp[0] * ssx + fp[1] * ssy + fp[2] * ssz) * 2 / (l1 * l2 * (mach * mach));
        o[2] = tot0 + (fp[0] * bbx + fp[1] * bby + fp[2] * bbz) * 2 / (l1 * l2 * (mach * mach));
        o[3] = tot0 + fp[3] * 2 / (l1 * l2 * l3 * (mach * mach));
        o[4] = tot0 + fp[4] * 2 / (l1 * l2 * l3 * (mach * mach));
        o[5] = tot0 + fp[5] * 2 / (l1 * l2 * l3 * (mach * mach));
        o[6] = tot0 + (fv[0] * aax + fv[1] * aay + fv[2] * aaz) * 2 / (l1 * l2 * (mach * mach));
        o[7] = tot0 + (fv[0] * ssx + fv[1] * ssy + fv[2] * ssz) * 2 / (l1 * l2 * (mach * mach));
        o[8] = tot0 + (fv[0] * bbx + fv[1] * bby + fv[2] * bbz) * 2 / (l1 * l2 * (mach * mach));
        o[9] = tot0 + fv[3] * 2 / (l1 * l2 * l3 * (mach * mach));
        o[10] = tot0 + fv[4] * 2 / (l1 * l2 * l3 * (mach * mach));
        o[11] = tot0 + fv[5] * 2 / (l1 * l2 * l3 * (mach * mach));
      }
  }
  return COSAHIP_OK;
}

extern "C" int cosahip_set_precision(cosahip_solve* h, int mode, double tail_rms) {
  Solve* s = reinterpret_cast<Solve*>(h);
  if (!s) return COSAHIP_ERR_ARG;
  if (mode < 0 || mode > 3) return fail(s, COSAHIP_ERR_ARG, "set_precision: mode must be 0, 1, 2 or 3");
  if (mode == 2 && !(tail_rms > 0)) return fail(s, COSAHIP_ERR_ARG, "set_precision: mode 2 needs tail_rms > 0");
  if (mode && s->split_residual && !s->par.viscous && s->par.npde == 5)
    return fail(s, COSAHIP_ERR_ARG, "set_precision: the FP32 residual needs the fused residual kernel");
  if (mode == 3 && (s->par.viscous || s->par.npde != 5))
    return fail(s, COSAHIP_ERR_ARG, "set_precision: the contracted FP64 residual (mode 3) covers the Euler residual");
  CK(hipSetDevice(s->device));
  s->prec_mode = mode;
  s->prec_tail_rms = mode == 2 ? tail_rms : 0;
  s->prec_tail = false;
  if (mode == 1 || mode == 2)
    for (Level& L : s->lev)
      if (L.nblocks) {
        int rc = alloc_fp32(s, L);
        if (rc) return rc;
      }
  return COSAHIP_OK;
}

extern "C" int cosahip_precision_in_use(const cosahip_solve* h) {
  const Solve* s = reinterpret_cast<const Solve*>(h);
  if (!s) return COSAHIP_ERR_ARG;
  return s->fp32() ? 32 : 64;
}

extern "C" int cosahip_download(cosahip_solve* h, int nl, int field, double* host, int64_t count) {
  Solve* s = reinterpret_cast<Solve*>(h);
  Level* L = get_level(s, nl);
  if (!L) return fail(s, COSAHIP_ERR_STATE, "download: level not defined");
  if (field < 0 || field >= COSAHIP_NFIELDS || !host) return fail(s, COSAHIP_ERR_ARG, "download: bad field");
  if (count != field_size(s, *L, field)) return fail(s, COSAHIP_ERR_ARG, "download: count does not match the field size");
  if (!count) return COSAHIP_OK;
  CK(hipSetDevice(s->device));
  CK(hipStreamSynchronize(s->stream));
  CK(hipMemcpy(host, L->f[field], count * sizeof(double), hipMemcpyDeviceToHost));
  return COSAHIP_OK;
}

// Checkpoint without stopping the solve: snapshot in stream order on the
// device, then the device-to-host copy on the checkpoint stream, which runs
// while the solve's stream goes on with the next cycles.
extern "C" int cosahip_checkpoint_begin(cosahip_solve* h, int nl, double* host_q, int64_t nq, double* host_mut,
                                        int64_t nmut) {
  Solve* s = reinterpret_cast<Solve*>(h);
  Level* L = get_level(s, nl);
  if (!L) return fail(s, COSAHIP_ERR_STATE, "checkpoint_begin: level not defined");
  if (s->ck_pending) return fail(s, COSAHIP_ERR_STATE, "checkpoint_begin: the previous checkpoint was not waited for");
  if (!host_q || nq != field_size(s, *L, COSAHIP_Q))
    return fail(s, COSAHIP_ERR_ARG, "checkpoint_begin: q count does not match the field size");
  const bool want_mut = s->par.kom != 0;
  if (want_mut && (!host_mut || nmut != field_size(s, *L, COSAHIP_MUT)))
    return fail(s, COSAHIP_ERR_ARG, "checkpoint_begin: k-omega needs mut with the field's count");
  if (!want_mut && (host_mut || nmut)) return fail(s, COSAHIP_ERR_ARG, "checkpoint_begin: no mut without k-omega");
  CK(hipSetDevice(s->device));
  if (!s->ck_stream) {
    CK(hipStreamCreateWithFlags(&s->ck_stream, hipStreamNonBlocking));
    CK(hipEventCreateWithFlags(&s->ck_snap, hipEventDisableTiming));
    CK(hipEventCreateWithFlags(&s->ck_done, hipEventDisableTiming));
  }
  const ll need = nq + (want_mut ? nmut : 0);
  if (need > s->ck_cap) {
    if (s->ck_buf) CK(hipFree(s->ck_buf));
    s->ck_buf = nullptr;
    s->ck_cap = 0;
    if (hipMalloc(&s->ck_buf, need * sizeof(double)) != hipSuccess)
      return fail(s, COSAHIP_ERR_NOMEM, "checkpoint_begin: device allocation failed");
    s->ck_cap = need;
  }
  CK(hipMemcpyAsync(s->ck_buf, L->f[COSAHIP_Q], nq * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
  if (want_mut)
    CK(hipMemcpyAsync(s->ck_buf + nq, L->f[COSAHIP_MUT], nmut * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
  CK(hipEventRecord(s->ck_snap, s->stream));
  CK(hipStreamWaitEvent(s->ck_stream, s->ck_snap, 0));
  CK(hipMemcpyAsync(host_q, s->ck_buf, nq * sizeof(double), hipMemcpyDeviceToHost, s->ck_stream));
  if (want_mut)
    CK(hipMemcpyAsync(host_mut, s->ck_buf + nq, nmut * sizeof(double), hipMemcpyDeviceToHost, s->ck_stream));
  CK(hipEventRecord(s->ck_done, s->ck_stream));
  s->ck_pending = true;
  return COSAHIP_OK;
}

extern "C" int cosahip_checkpoint_wait(cosahip_solve* h) {
  Solve* s = reinterpret_cast<Solve*>(h);
  if (!s) return COSAHIP_ERR_ARG;
  if (!s->ck_pending) return fail(s, COSAHIP_ERR_STATE, "checkpoint_wait: no checkpoint in flight");
  CK(hipSetDevice(s->device));
  s->ck_pending = false;
  CK(hipEventSynchronize(s->ck_done));
  return COSAHIP_OK;
}

extern "C" int cosahip_host_alloc(int64_t bytes, void** out) {
  if (!out || bytes <= 0) return COSAHIP_ERR_ARG;
  *out = nullptr;
  if (hipHostMalloc(out, (size_t)bytes, hipHostMallocDefault) != hipSuccess) return COSAHIP_ERR_NOMEM;
  return COSAHIP_OK;
}

extern "C" int cosahip_host_free(void* p) {
  if (p && hipHostFree(p) != hipSuccess) return COSAHIP_ERR_DEVICE;
  return COSAHIP_OK;
}

extern "C" int cosahip_synchronize(cosahip_solve* h) {
  Solve* s = reinterpret_cast<Solve*>(h);
  if (!s) return COSAHIP_ERR_ARG;
  CK(hipStreamSynchronize(s->stream));
  return COSAHIP_OK;
}

// ---------------------------------------------------------------------------
// whole-level reference routines

static double* qfield(Level& L, int field) {
  if (field == COSAHIP_Q || field == COSAHIP_QOLD) return L.f[field];
  return nullptr;
}

#define LEVEL_OR_FAIL(s, nl, Lp)                                                   \
  Solve* s = reinterpret_cast<Solve*>(h);                                          \
  Level* Lp = get_level(s, nl);                                                    \
  if (!Lp) return fail(s, COSAHIP_ERR_STATE, "level not defined");                 \
  CK(hipSetDevice(s->device));

extern "C" int cosahip_cons2prim(cosahip_solve* h, int nl, int field) {
  LEVEL_OR_FAIL(s, nl, L);
  double* q = qfield(*L, field);
  if (!q) return fail(s, COSAHIP_ERR_ARG, "cons2prim: field must be Q or QOLD");
  CK(launch_cons2prim(*s, *L, q));
  return COSAHIP_OK;
}
extern "C" int cosahip_prim2cons(cosahip_solve* h, int nl, int field) {
  LEVEL_OR_FAIL(s, nl, L);
  double* q = qfield(*L, field);
  if (!q) return fail(s, COSAHIP_ERR_ARG, "prim2cons: field must be Q or QOLD");
  CK(launch_prim2cons(*s, *L, q));
  return COSAHIP_OK;
}
extern "C" int cosahip_bc(cosahip_solve* h, int nl, int field) {
  LEVEL_OR_FAIL(s, nl, L);
  double* q = qfield(*L, field);
  if (!q) return fail(s, COSAHIP_ERR_ARG, "bc: field must be Q or QOLD");
  CK(launch_bc(*s, *L, q));
  return COSAHIP_OK;
}
extern "C" int cosahip_cutman_q(cosahip_solve* h, int nl, int field) {
  LEVEL_OR_FAIL(s, nl, L);
  double* q = qfield(*L, field);
  if (!q && field == COSAHIP_RES) q = L->f[COSAHIP_RES];
  if (!q) return fail(s, COSAHIP_ERR_ARG, "cutman_q: field must be Q, QOLD or RES");
  CK(launch_cut(*s, *L, q));
  return COSAHIP_OK;
}
extern "C" int cosahip_edges_corners(cosahip_solve* h, int nl, int field) {
  LEVEL_OR_FAIL(s, nl, L);
  double* q = qfield(*L, field);
  if (!q) return fail(s, COSAHIP_ERR_ARG, "edges_corners: field must be Q or QOLD");
  CK(launch_edges_corners(*s, *L, q));
  return COSAHIP_OK;
}

static int decode_prtest(Solve* s, int nl, unsigned long long key) {
  const int kind = (int)(key & 3);
  const int loc = (int)((key >> 2) & 0x7FFFFFF);
  const unsigned long long pos = (key >> 29) & ((1ull << 26) - 1);
  const int reg = (int)(pos % 5);
  const int bn = (int)(pos / 5);
  const int nsnap = s->par.nsnap;
  const int b = bn / nsnap, n = bn % nsnap;
  Level& L = s->lev[nl - 1];
  const Blk& B = L.hblk[b];
  const int im = B.imax - 1, jm = B.jmax - 1;
  int i = 0, j = 0, k = 0, x = loc;
  if (reg == 0) {
    i = x % (B.imax + 3) - 1;
    x /= (B.imax + 3);
    j = x % jm + 1;
    k = x / jm + 1;
  } else if (reg <= 2) {
    i = x % im + 1;
    x /= im;
    int jj = x % 2;
    k = x / 2 + 1;
    j = reg == 1 ? jj - 1 : B.jmax + jj;
  } else {
    i = x % im + 1;
    x /= im;
    j = x % jm + 1;
    int kk = x / jm;
    k = reg == 3 ? kk - 1 : B.kmax + kk;
  }
  int32_t info[8] = {nl, b + 1, i, j, k, (int)(key >> 55), kind, n};
  memcpy(s->errinfo, info, sizeof(info));
  char msg[160];
  snprintf(msg, sizeof msg, "%s <= qmin at level %d block %d (i,j,k) = (%d,%d,%d) snapshot %d",
           kind == 1 ? "pressure" : "density", nl, b + 1, i, j, k, n);
  return fail(s, COSAHIP_ERR_POSITIVITY, msg);
}

static const unsigned long long NOFLAG = ~0ull;
static const int RMS_ASYNC_SLOT = 16;   // pinned slots of the async cycle's norms
static const int FLAG_ASYNC_SLOT = 32;  // and of its per-level positivity flags

// prtest first-offender key of level nl (one flag slot per level)
static int read_flag(Solve* s, int nl, unsigned long long* key) {
  CK(hipMemcpyAsync(s->hpinned, s->dflag + (nl - 1), sizeof(unsigned long long), hipMemcpyDeviceToHost, s->stream));
  CK(hipStreamSynchronize(s->stream));
  memcpy(key, s->hpinned, sizeof(*key));
  return COSAHIP_OK;
}

extern "C" int cosahip_prtest(cosahip_solve* h, int nl, int iter) {
  LEVEL_OR_FAIL(s, nl, L);
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s->stream, s->dflag + (nl - 1), NOFLAG, 1);
  CK(launch_prtest(*s, *L, L->f[COSAHIP_Q], iter));
  unsigned long long key;
  int rc = read_flag(s, nl, &key);
  if (rc) return rc;
  if (key != NOFLAG) return decode_prtest(s, nl, key);
  return COSAHIP_OK;
}
// CFL ramping of k-omega runs (bdeltat, src/cosa.f90:4760-4820): cfl and
// cflt of the solve from the cycle counter, kept between calls as the
// reference keeps its globals; ramping1 refreshes at cycle 1 and every
// 100th, ramping2 at every call.  Same arithmetic as the reference.
static void cfl_ramp(Solve* s, int mgit) {
  const cosahip_config& c = s->cfg;
  Par& p = s->par;
  if (!komega(c)) return;
  const int* nc = c.ncycr;
  if (c.ramping == 1 && (mgit == 1 || mgit / 100 * 100 == mgit)) {
    const double sqrt3 = sqrt(3.0);
    const double s33 = sqrt3 * sqrt3 * sqrt3;
    const double dcfli1 = (c.vcfli[1] - c.vcfli[0]) * s33, dcfli2 = c.vcfli[2] - c.vcfli[1];
    const double dcflit1 = (c.vcflit[1] - c.vcflit[0]) * s33, dcflit2 = c.vcflit[2] - c.vcflit[1];
    if (mgit <= nc[1]) {
      p.cfl = c.vcfli[0];
      p.cflt = c.vcflit[0];
    } else if (mgit <= nc[2]) {
      const double x = (double)(mgit - nc[1]) / (nc[2] - nc[1]) / sqrt3;
      p.cfl = c.vcfli[0] + dcfli1 * (x * x * x);
      p.cflt = c.vcflit[0] + dcflit1 * (x * x * x);
    } else if (mgit <= nc[3]) {
      p.cfl = c.vcfli[1] + dcfli2 * ((double)(mgit - nc[2]) / (nc[3] - nc[2]));
      p.cflt = c.vcflit[1] + dcflit2 * ((double)(mgit - nc[2]) / (nc[3] - nc[2]));
    } else {
      p.cfl = c.vcfli[2];
      p.cflt = c.vcflit[2];
    }
  } else if (c.ramping == 2) {
    const double dcfli1 = c.vcfli[1] - c.vcfli[0], dcfli2 = c.vcfli[2] - c.vcfli[1];
    const double dcflit1 = c.vcflit[1] - c.vcflit[0], dcflit2 = c.vcflit[2] - c.vcflit[1];
    if (mgit <= nc[0]) {
      p.cfl = c.vcfli[0];
      p.cflt = c.vcflit[0];
    } else if (mgit <= nc[1]) {
      p.cfl = c.vcfli[0] + dcfli1 * ((double)(mgit - nc[0]) / (nc[1] - nc[0]));
      p.cflt = c.vcflit[0] + dcflit1 * ((double)(mgit - nc[0]) / (nc[1] - nc[0]));
    } else if (mgit <= nc[2]) {
      p.cfl = c.vcfli[1];
      p.cflt = c.vcflit[1];
    } else if (mgit <= nc[3]) {
      p.cfl = c.vcfli[1] + dcfli2 * ((double)(mgit - nc[2]) / (nc[3] - nc[2]));
      p.cflt = c.vcflit[1] + dcflit2 * ((double)(mgit - nc[2]) / (nc[3] - nc[2]));
    } else {
      p.cfl = c.vcfli[2];
      p.cflt = c.vcflit[2];
    }
  }
}

extern "C" int cosahip_lambda_deltat(cosahip_solve* h, int nl) {
  LEVEL_OR_FAIL(s, nl, L);
  cfl_ramp(s, s->mgit);
  CK(launch_lambda_deltat(*s, *L));
  return COSAHIP_OK;
}
extern "C" int cosahip_residual(cosahip_solve* h, int nl) {
  LEVEL_OR_FAIL(s, nl, L);
  if (s->fp32() && L->q32) CK(launch_to_f32(*s, L->f[COSAHIP_Q], L->q32, field_size(s, *L, COSAHIP_Q)));
  CK(launch_zero_res(*s, *L, L->f[COSAHIP_RES]));
  CK(launch_residual(*s, *L, L->f[COSAHIP_Q], L->f[COSAHIP_RES]));
  return COSAHIP_OK;
}
extern "C" int cosahip_src_uns(cosahip_solve* h, int nl) {
  LEVEL_OR_FAIL(s, nl, L);
  if (!s->par.dualt) return fail(s, COSAHIP_ERR_STATE, "src_uns: dual time stepping is off (cosahip_set_dual_time)");
  CK(launch_src_uns(*s, *L, 0));
  return COSAHIP_OK;
}

// ---------------------------------------------------------------------------
// dual time stepping (src/cosa.f90:437-672)

static void drop_graph(Solve* s) {
  if (s->gexec) {
    (void)hipGraphExecDestroy(s->gexec);
    s->gexec = nullptr;
  }
}

extern "C" int cosahip_set_dual_time(cosahip_solve* h, double dt, int dual_prop_start) {
  LEVEL_OR_FAIL(s, 1, L);
  if (s->cycle_pending) return fail(s, COSAHIP_ERR_STATE, "set_dual_time: an enqueued cycle has not been waited for");
  if (s->cfg.harbal || s->par.nsnap != 1)
    return fail(s, COSAHIP_ERR_ARG, "set_dual_time: dual time stepping excludes harmonic balance");
  if (!(dt > 0)) return fail(s, COSAHIP_ERR_ARG, "set_dual_time: dt must be > 0");
  CK(hipSetDevice(s->device));
  CK(hipStreamSynchronize(s->stream));
  const bool was = s->par.dualt != 0;
  s->par.dualt = 1;
  s->par.dual_dt = dt;
  s->par.dual_rkimp = s->cfg.rkimp ? 1 : 0;
  s->par.dual_plain = s->cfg.rkimp && s->cfg.irsop == 0 && s->cfg.nlevel == 1;
  s->par.dual_fact = dual_prop_start ? 1.5 : 1.0;
  s->dual_prop = dual_prop_start ? 1 : 0;
  if (!was) {
    const ll n = field_size(s, *L, COSAHIP_Q1);
    for (int f : {COSAHIP_Q1, COSAHIP_Q2}) {
      if (L->f[f]) CK(hipFree(L->f[f]));
      L->f[f] = nullptr;
      if (hipMalloc(&L->f[f], n * sizeof(double)) != hipSuccess) {
        s->par.dualt = 0;
        return fail(s, COSAHIP_ERR_NOMEM, "device allocation failed");
      }
      CK(hipMemset(L->f[f], 0, n * sizeof(double)));
    }
  }
  drop_graph(s);
  return COSAHIP_OK;
}

extern "C" int cosahip_dual_begin(cosahip_solve* h, int itime) {
  LEVEL_OR_FAIL(s, 1, L);
  if (!s->par.dualt) return fail(s, COSAHIP_ERR_STATE, "dual_begin: dual time stepping is off (cosahip_set_dual_time)");
  if (s->cycle_pending) return fail(s, COSAHIP_ERR_STATE, "dual_begin: an enqueued cycle has not been waited for");
  if (itime < 1) return fail(s, COSAHIP_ERR_ARG, "dual_begin: itime >= 1");
  CK(hipSetDevice(s->device));
  const double fact = (itime == 1 && !s->dual_prop) ? 1.0 : 1.5;
  if (fact != s->par.dual_fact) drop_graph(s);   // the kernels take the factor by value
  s->par.dual_fact = fact;
  CK(launch_rhs_uns(*s, *L));
  return COSAHIP_OK;
}

extern "C" int cosahip_dual_end(cosahip_solve* h) {
  LEVEL_OR_FAIL(s, 1, L);
  if (!s->par.dualt) return fail(s, COSAHIP_ERR_STATE, "dual_end: dual time stepping is off (cosahip_set_dual_time)");
  if (s->cycle_pending) return fail(s, COSAHIP_ERR_STATE, "dual_end: an enqueued cycle has not been waited for");
  CK(hipSetDevice(s->device));
  CK(launch_copy_interior(*s, *L, L->f[COSAHIP_Q2], L->f[COSAHIP_Q1]));
  return COSAHIP_OK;
}

extern "C" int cosahip_src_hb(cosahip_solve* h, int nl) {
  LEVEL_OR_FAIL(s, nl, L);
  if (!s->cfg.harbal) return COSAHIP_OK;
  CK(launch_src_hb(*s, *L));
  return COSAHIP_OK;
}

// rkim without IRS on one grid level: src_hb is not added to the residual
// (src/cosa.f90:1916-1921); the norm adds the HB term itself
static bool rkim_plain(const cosahip_config& c) { return c.harbal && c.rkimp && c.irsop == 0 && c.nlevel == 1; }

// resg_rms partial sums of level L, copied into pinned slot `slot` (no
// wait).  res_minus_rhs: res already holds res - rhs (the dual-time finest
// level inside smooth), so the norm is of res itself -- the same bits.
static int rms_enqueue(Solve* s, Level& L, int slot, bool res_minus_rhs = false) {
  const int np = s->par.npde;
  if (s->cfg.harbal) {
    CK(launch_copy_interior(*s, L, L.work1, L.f[COSAHIP_RES]));
    if (rkim_plain(s->cfg)) CK(launch_rms_hb_term(*s, L, L.work1));
    CK(launch_rms(*s, L, L.work1, nullptr, 0));
  } else if (s->par.dualt && s->par.dual_plain) {
    CK(launch_rms_dual_term(*s, L, L.work1));
    CK(launch_rms(*s, L, L.work1, nullptr, 0));
  } else if (res_minus_rhs) {
    CK(launch_rms(*s, L, L.f[COSAHIP_RES], nullptr, 0));
  } else {
    CK(launch_rms(*s, L, L.f[COSAHIP_RES], L.f[COSAHIP_RHS], 1));
  }
  CK(hipMemcpyAsync(s->hpinned + slot, L.rms_part + (ll)L.ntile[TK_CELLS] * 8, np * sizeof(double),
                    hipMemcpyDeviceToHost, s->stream));
  if (s->block_rms && L.nl == 1 && s->hpart && s->hpart_n >= (ll)L.ntile[TK_CELLS] * 8)
    CK(hipMemcpyAsync(s->hpart, L.rms_part, (ll)L.ntile[TK_CELLS] * 8 * sizeof(double), hipMemcpyDeviceToHost,
                      s->stream));
  return COSAHIP_OK;
}

// resrms(1:npde+1) from the sums in pinned slot `slot` (after the stream
// has reached the copy)
static void rms_finish(Solve* s, Level& L, int slot, double* out) {
  const int np = s->par.npde;
  double sum[8] = {0};
  for (int v = 0; v < np; v++) sum[v] = s->hpinned[slot + v];
  double tot = 0;
  for (int v = 0; v < np; v++) tot = tot + sum[v];
  const double nc = (double)L.ncell;
  if (s->cfg.harbal) {
    const double d5 = s->par.nsnap;
    for (int v = 0; v < np; v++) out[v] = sqrt(sum[v] / nc / d5);
    out[np] = sqrt(tot / nc / d5 / np);
  } else {
    for (int v = 0; v < np; v++) out[v] = sqrt(sum[v] / nc);
    out[np] = sqrt(tot / nc / np);
  }
}

static int rms(Solve* s, Level& L, double* out, bool res_minus_rhs = false) {
  int rc = rms_enqueue(s, L, 0, res_minus_rhs);
  if (rc) return rc;
  CK(hipStreamSynchronize(s->stream));
  rms_finish(s, L, 0, out);
  return COSAHIP_OK;
}

extern "C" int cosahip_resg_rms(cosahip_solve* h, int nl, double* resrms) {
  LEVEL_OR_FAIL(s, nl, L);
  if (!resrms) return fail(s, COSAHIP_ERR_ARG, "resrms null");
  return rms(s, *L, resrms);
}
extern "C" int cosahip_add_array(cosahip_solve* h, int nl, int dst, int src, double fact) {
  LEVEL_OR_FAIL(s, nl, L);
  if (dst < 0 || dst > COSAHIP_RHS || src < 0 || src > COSAHIP_RHS) return fail(s, COSAHIP_ERR_ARG, "add_array: field");
  CK(launch_add_interior(*s, *L, L->f[dst], L->f[src], fact));
  return COSAHIP_OK;
}
extern "C" int cosahip_copy_array(cosahip_solve* h, int nl, int dst, int src) {
  LEVEL_OR_FAIL(s, nl, L);
  if (dst < 0 || dst > COSAHIP_RHS || src < 0 || src > COSAHIP_RHS) return fail(s, COSAHIP_ERR_ARG, "copy_array: field");
  CK(launch_copy_interior(*s, *L, L->f[dst], L->f[src]));
  return COSAHIP_OK;
}
extern "C" int cosahip_cirs(cosahip_solve* h, int nl) {
  LEVEL_OR_FAIL(s, nl, L);
  CK(launch_cirs(*s, *L));
  return COSAHIP_OK;
}
extern "C" int cosahip_tur_vis(cosahip_solve* h, int nl, int field) {
  LEVEL_OR_FAIL(s, nl, L);
  if (!s->par.kom) return fail(s, COSAHIP_ERR_STATE, "tur_vis: not a k-omega configuration");
  double* q = qfield(*L, field);
  if (!q) return fail(s, COSAHIP_ERR_ARG, "tur_vis: field must be Q or QOLD");
  CK(launch_tur_vis(*s, *L, q));
  return COSAHIP_OK;
}
extern "C" int cosahip_fixq(cosahip_solve* h, int nl) {
  LEVEL_OR_FAIL(s, nl, L);
  if (!s->par.kom) return fail(s, COSAHIP_ERR_STATE, "fixq: not a k-omega configuration");
  CK(launch_fixq(*s, *L));
  return COSAHIP_OK;
}
extern "C" int cosahip_turb_limits(cosahip_solve* h, int nl, double ar) {
  LEVEL_OR_FAIL(s, nl, L);
  if (!s->par.kom) return fail(s, COSAHIP_ERR_STATE, "turb_limits: not a k-omega configuration");
  CK(launch_turb_limits(*s, *L, ar));
  return COSAHIP_OK;
}
extern "C" int cosahip_update(cosahip_solve* h, int nl, double ar) {
  LEVEL_OR_FAIL(s, nl, L);
  CK(launch_update(*s, *L, ar));
  return COSAHIP_OK;
}

// ---------------------------------------------------------------------------
// smooth (src/cosa.f90:1594-2032)

// The stage loop of smooth.  async: the launch sequence of one V-cycle
// enqueued without host synchronisation (cosahip_cycle_enqueue): the
// level-1 norms go to a pinned slot and are read, with the convergence test
// and the positivity flags, by cosahip_cycle_wait.
static int smooth_core(Solve* s, int nl, int from_fine, int to_coarse, int niter, int mgit, bool async,
                       int32_t* done) {
  Level& L = s->lev[nl - 1];
  const cosahip_config& c = s->cfg;
  if (done) *done = 0;
  if (c.rkimp && !c.harbal && !s->par.dualt)
    return fail(s, COSAHIP_ERR_STATE, "rkimp without harmonic balance needs dual time stepping (cosahip_set_dual_time)");
  double* q = L.f[COSAHIP_Q];
  const int niterp = to_coarse ? niter + 1 : niter;
  s->mgit = mgit;
  if (!async) hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s->stream, s->dflag + (nl - 1), NOFLAG, 1);
  int seq = 0;
  bool returned_early = false;
  // the RK update folded into the last CIRS sweep, and with it the next
  // stage's cons2prim (mean flow with cirs: the state is the same bits)
  const bool fuse_update = s->fuse_update && !komega(c) && s->par.irsop >= 1 && s->par.irsop != 4 &&
                           s->par.irsop <= 6 && c.cutcirs != 1 && !c.rkimp;
  bool q_prim = false;   // the previous stage left q in primitive variables
  for (int iter = 1; iter <= niterp && !returned_early; iter++) {
    const ll cs = L.ncell * s->par.nsnap;
    // qold = q (copy_array) is folded into the first stage's cons2prim,
    // which reads the same conservative interior just after it
    const bool copy_qold = iter <= niter;
    if (copy_qold && q_prim) {
      TimeScope t(*s, COSAHIP_FAM_OTHER, cs);
      CK(launch_copy_interior(*s, L, L.f[COSAHIP_QOLD], q));
    }
    for (int irk = 1; irk <= c.nstage; irk++, seq++) {
      if (!q_prim) {
        TimeScope t(*s, COSAHIP_FAM_CONVERT, cs);
        CK(launch_cons2prim(*s, L, q, true, irk == 1 && copy_qold ? L.f[COSAHIP_QOLD] : nullptr));
      }
      q_prim = false;
      {
        TimeScope t(*s, COSAHIP_FAM_HALO, cs);
        CK(launch_bc(*s, L, q));
        CK(launch_cut(*s, L, q));
        CK(launch_halo_f32(*s, L));
      }
      if (komega(c) && mgit == 1 && iter == 1 && irk == 1 && from_fine && !c.prd_all_lvls) {
        // first visit of a coarse level: mut from the restricted state
        TimeScope t(*s, COSAHIP_FAM_OTHER, cs);
        if (c.kom_sst) CK(launch_sst_first_mut(*s, L));
        else CK(launch_tur_vis(*s, L, q));
      }
      {
        TimeScope t(*s, COSAHIP_FAM_OTHER, cs);
        CK(launch_prtest(*s, L, q, seq));
      }
      if (iter % c.iupdt == 0 && irk == 1) {
        TimeScope t(*s, COSAHIP_FAM_LAMBDA, cs);
        cfl_ramp(s, mgit);
        CK(launch_lambda_deltat(*s, L));
      }
      {
        TimeScope t(*s, COSAHIP_FAM_RESIDUAL, cs);
        CK(launch_residual(*s, L, q, L.f[COSAHIP_RES]));
      }
      // prim2cons_b; with the update folded into the CIRS sweep of a stage
      // that leaves q primitive, nothing reads the conservative state before
      // the sweep rewrites every interior cell from qold (and the next
      // stage's BC fill rewrites the face halos), unless the HB or dual-time
      // source takes it; on the last such stage the sweep writes the
      // conservative interior, so only the face halos are converted
      const bool fused_stage = fuse_update && iter <= niter && !c.harbal && !s->par.dualt;
      if (!(fused_stage && irk < c.nstage)) {
        TimeScope t(*s, COSAHIP_FAM_CONVERT, cs);
        CK(launch_prim2cons(*s, L, q, fused_stage));
      }
      if (c.harbal && !rkim_plain(c)) {
        TimeScope t(*s, COSAHIP_FAM_OTHER, cs);
        CK(launch_src_hb(*s, L));
      }
      // dual time (src/cosa.f90:1910-1915, 1938-1941): src_uns, and on the
      // finest level the overall residual res - rhs, in one pass; without
      // IRS on one level, rkimp leaves src_uns out and subtracts after the norm
      const bool dual = s->par.dualt != 0;
      const bool dual_src = dual && !s->par.dual_plain;
      if (dual_src) {
        TimeScope t(*s, COSAHIP_FAM_OTHER, cs);
        CK(launch_src_uns(*s, L, nl == 1));
      }
      const bool sub_done = dual_src && nl == 1;
      if (async && nl == 1 && irk == 1 && iter != niter + 1) {
        int rc = rms_enqueue(s, L, RMS_ASYNC_SLOT, sub_done);
        if (rc) return rc;
      } else if (nl == 1 && irk == 1 && iter != niter + 1 &&
                 (mgit == 1 || mgit % c.iprint == 0 || mgit == c.ncycle)) {
        double rr[8];
        int rc = rms(s, L, rr, sub_done);
        if (rc) return rc;
        for (int v = 0; v <= s->par.npde; v++) s->resrms[v + 1] = rr[v];
        if (s->prec_mode == 2 && !s->prec_tail && s->resrms[1] < s->prec_tail_rms) s->prec_tail = true;
      }
      if (dual && nl == 1 && !sub_done) {
        TimeScope t(*s, COSAHIP_FAM_OTHER, cs);
        CK(launch_add_interior(*s, L, L.f[COSAHIP_RES], L.f[COSAHIP_RHS], -1.0));
      }
      if (from_fine && iter == 1 && irk == 1) {
        TimeScope t(*s, COSAHIP_FAM_OTHER, cs);
        CK(launch_add_interior(*s, L, L.f[COSAHIP_RHS], L.f[COSAHIP_RES], 1.0));
      }
      if (nl > 1) {
        TimeScope t(*s, COSAHIP_FAM_OTHER, cs);
        CK(launch_add_interior(*s, L, L.f[COSAHIP_RES], L.f[COSAHIP_RHS], -1.0));
      }
      if (iter == niter + 1) {
        if (komega(c)) {
          TimeScope t(*s, COSAHIP_FAM_OTHER, cs);
          CK(launch_turb_limits(*s, L, c.alphas[irk - 1]));
        }
        returned_early = true;
        break;
      }
      if (fuse_update) {
        // the last stage leaves q conservative: the next iteration's qold
        // copy and the transfers between levels read it
        const int prim = irk < c.nstage;
        const CirsUpdate up{q, L.f[COSAHIP_QOLD], prim && s->fp32() ? L.q32 : nullptr, c.alphas[irk - 1], prim};
        TimeScope t(*s, COSAHIP_FAM_CIRS, cs);
        CK(launch_cirs(*s, L, &up));
        q_prim = prim != 0;
      } else {
        {
          TimeScope t(*s, COSAHIP_FAM_CIRS, cs);
          CK(launch_cirs(*s, L));
        }
        {
          TimeScope t(*s, COSAHIP_FAM_UPDATE, cs);
          CK(launch_update(*s, L, c.alphas[irk - 1]));
        }
      }
      if (komega(c)) {
        TimeScope t(*s, COSAHIP_FAM_OTHER, cs);
        if (c.ifixq == 1) CK(launch_fixq(*s, L));
        if (irk == c.nstage && (nl == 1 || !c.prd_all_lvls)) CK(launch_tur_vis(*s, L, q));
      }
    }
    if (returned_early) break;
    if (async) continue;   // convergence is tested by cosahip_cycle_wait
    double errmax = s->resrms[1];
    for (int v = 2; v <= 5; v++) errmax = std::max(errmax, s->resrms[v]);
    if (errmax < c.toler) {
      if (done) *done = 1;
      break;
    }
  }
  return COSAHIP_OK;
}

extern "C" int cosahip_smooth(cosahip_solve* h, int nl, int from_fine, int to_coarse, int niter, int mgit,
                              int32_t* done, double* resrms) {
  LEVEL_OR_FAIL(s, nl, Lp);
  (void)Lp;
  if (s->cycle_pending) return fail(s, COSAHIP_ERR_STATE, "smooth: an enqueued cycle has not been waited for");
  int rc = smooth_core(s, nl, from_fine, to_coarse, niter, mgit, false, done);
  if (rc) return rc;
  unsigned long long key;
  rc = read_flag(s, nl, &key);
  if (rc) return rc;
  if (resrms)
    for (int v = 0; v < 8; v++) resrms[v] = s->resrms[v + 1];
  if (key != NOFLAG) return decode_prtest(s, nl, key);
  return COSAHIP_OK;
}

// ---------------------------------------------------------------------------
// restrict / prolong

static int transfer_ready(Solve* s, int nl) {
  if (nl < 1 || nl + 1 > (int)s->lev.size() || !s->lev[nl - 1].nblocks || !s->lev[nl].nblocks)
    return fail(s, COSAHIP_ERR_STATE, "restrict/prolong: both levels must be defined");
  Level& L1 = s->lev[nl - 1];
  Level& L2 = s->lev[nl];
  if (L1.nblocks != L2.nblocks) return fail(s, COSAHIP_ERR_ARG, "levels have different block counts");
  for (int b = 0; b < L1.nblocks; b++)
    if ((L1.hblk[b].imax - 1) != 2 * (L2.hblk[b].imax - 1) || (L1.hblk[b].jmax - 1) != 2 * (L2.hblk[b].jmax - 1) ||
        (L1.hblk[b].kmax - 1) != 2 * (L2.hblk[b].kmax - 1))
      return fail(s, COSAHIP_ERR_ARG, "coarse level is not the 2:1 coarsening of the fine level");
  return COSAHIP_OK;
}

static int restrict_core(Solve* s, int nl) {
  Level& L1 = s->lev[nl - 1];
  Level& L2 = s->lev[nl];
  TimeScope t(*s, COSAHIP_FAM_OTHER, L2.ncell * s->par.nsnap);
  CK(launch_rest1(*s, L1, L2, L1.f[COSAHIP_Q], L2.f[COSAHIP_Q]));
  if (komega(s->cfg) && !s->cfg.prd_all_lvls) {
    // rest4 of the production, delpls and divergence (src/cosa.f90:2734-2739)
    CK(launch_rest_scalar(*s, L1, L2, L1.f[COSAHIP_PRD], L2.f[COSAHIP_PRD]));
    CK(launch_rest_scalar(*s, L1, L2, L1.f[COSAHIP_DELPLS], L2.f[COSAHIP_DELPLS]));
    CK(launch_rest_scalar(*s, L1, L2, L1.f[COSAHIP_DIVVEL], L2.f[COSAHIP_DIVVEL]));
  }
  CK(launch_rest2(*s, L1, L2, L2.f[COSAHIP_RHS], L1.f[COSAHIP_RES]));
  return COSAHIP_OK;
}

static int prolong_core(Solve* s, int nl) {
  Level& L1 = s->lev[nl - 1];
  Level& L2 = s->lev[nl];
  TimeScope t(*s, COSAHIP_FAM_OTHER, L1.ncell * s->par.nsnap);
  CK(launch_rest1(*s, L1, L2, L1.f[COSAHIP_Q], L2.f[COSAHIP_QOLD]));
  double* arrs[2] = {L2.f[COSAHIP_QOLD], L2.f[COSAHIP_Q]};
  for (double* a : arrs) {
    CK(launch_cons2prim(*s, L2, a));
    CK(launch_bc(*s, L2, a));
    CK(launch_cut(*s, L2, a));
    CK(launch_prim2cons(*s, L2, a));
    CK(launch_edges_corners(*s, L2, a));
  }
  CK(launch_sub_box(*s, L2, L2.f[COSAHIP_Q], L2.f[COSAHIP_QOLD]));
  CK(launch_prol(*s, L1, L2, L1.f[COSAHIP_Q], L2.f[COSAHIP_Q]));
  return COSAHIP_OK;
}

extern "C" int cosahip_restrict(cosahip_solve* h, int nl) {
  Solve* s = reinterpret_cast<Solve*>(h);
  if (!s) return COSAHIP_ERR_ARG;
  int rc = transfer_ready(s, nl);
  if (rc) return rc;
  if (s->cycle_pending) return fail(s, COSAHIP_ERR_STATE, "restrict: an enqueued cycle has not been waited for");
  CK(hipSetDevice(s->device));
  return restrict_core(s, nl);
}

extern "C" int cosahip_prolong(cosahip_solve* h, int nl) {
  Solve* s = reinterpret_cast<Solve*>(h);
  if (!s) return COSAHIP_ERR_ARG;
  int rc = transfer_ready(s, nl);
  if (rc) return rc;
  if (s->cycle_pending) return fail(s, COSAHIP_ERR_STATE, "prolong: an enqueued cycle has not been waited for");
  CK(hipSetDevice(s->device));
  return prolong_core(s, nl);
}

// ---------------------------------------------------------------------------
// asynchronous multigrid cycles

// One pass of mg's cycle loop (src/cosa.f90:822-912) with nstart = npre =
// npost = ncrs = 1, as launches only: fine smooth, the coarse levels of the
// V and the prolongations, then the per-level positivity flags are copied
// to pinned memory.
static int cycle_body(Solve* s, int mgit) {
  const int nlevel = s->cfg.nlevel;
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s->stream, s->dflag, NOFLAG, nlevel);
  int rc = smooth_core(s, 1, 0, nlevel > 1, 1, mgit, true, nullptr);
  if (!rc && nlevel > 1) {
    rc = restrict_core(s, 1);
    if (!rc) rc = smooth_core(s, 2, 1, nlevel > 2, 1, mgit, true, nullptr);
    if (!rc && nlevel > 2) {
      rc = restrict_core(s, 2);
      if (!rc) rc = smooth_core(s, 3, 1, 0, 1, mgit, true, nullptr);
      if (!rc) rc = prolong_core(s, 2);
      if (!rc) rc = smooth_core(s, 2, 0, 0, 1, mgit, true, nullptr);
    }
    if (!rc) rc = prolong_core(s, 1);
  }
  if (rc) return rc;
  CK(hipMemcpyAsync(s->hpinned + FLAG_ASYNC_SLOT, s->dflag, nlevel * sizeof(unsigned long long),
                    hipMemcpyDeviceToHost, s->stream));
  return COSAHIP_OK;
}

extern "C" int cosahip_cycle_enqueue(cosahip_solve* h, int mgit) {
  Solve* s = reinterpret_cast<Solve*>(h);
  if (!s) return COSAHIP_ERR_ARG;
  if (s->cycle_pending) return fail(s, COSAHIP_ERR_STATE, "cycle_enqueue: the previous cycle has not been waited for");
  const int nlevel = s->cfg.nlevel;
  if (nlevel < 1 || nlevel > 3 || mgit < 1) return fail(s, COSAHIP_ERR_ARG, "cycle_enqueue: 1..3 levels, mgit >= 1");
  for (int nl = 1; nl <= nlevel; nl++)
    if (!s->lev[nl - 1].nblocks) return fail(s, COSAHIP_ERR_STATE, "cycle_enqueue: every level must be defined");
  for (int nl = 1; nl < nlevel; nl++) {
    int rc = transfer_ready(s, nl);
    if (rc) return rc;
  }
  CK(hipSetDevice(s->device));
  // the first cycle differs (k-omega: mut of a coarse level on its first
  // visit) and runs eagerly; later cycles replay one captured graph
  // (a CFL ramp changes the kernels' arguments from cycle to cycle: no graph)
  const bool graph = s->use_graph && mgit >= 2 && !s->timing && !s->cfg.ramping;
  // the key names what the captured kernels take by value and launch
  const int key = (s->fp32() ? 32 : 64) + (s->par.dualt ? (s->par.dual_fact == 1.0 ? 1000 : 2000) : 0);
  if (graph && s->gexec && s->gkey == key) {
    CK(hipGraphLaunch(s->gexec, s->stream));
  } else if (graph) {
    if (s->gexec) {
      CK(hipGraphExecDestroy(s->gexec));
      s->gexec = nullptr;
    }
    // the lazy conversions (FP32 shadows, cell-centre metric averages) mark
    // themselves done when their launch is recorded; if the capture fails
    // they never ran, so their marks are restored
    std::vector<std::array<bool, 5>> marks;
    for (const Level& L : s->lev) marks.push_back({L.m32_ok, L.xd32_ok, L.cmet_ok, L.cmet32_ok, L.mut32_ok});
    auto unmark = [&]() {
      for (size_t m = 0; m < s->lev.size(); m++) {
        Level& L = s->lev[m];
        L.m32_ok = marks[m][0];
        L.xd32_ok = marks[m][1];
        L.cmet_ok = marks[m][2];
        L.cmet32_ok = marks[m][3];
        L.mut32_ok = marks[m][4];
      }
    };
    hipGraph_t g = nullptr;
    CK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeRelaxed));
    const int rc = cycle_body(s, mgit);
    const hipError_t e = hipStreamEndCapture(s->stream, &g);
    if (rc || e != hipSuccess) {
      unmark();
      if (e == hipSuccess && g) (void)hipGraphDestroy(g);
      if (rc) return rc;
      CK(e);
    }
    const hipError_t ei = hipGraphInstantiate(&s->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
      s->gexec = nullptr;
      unmark();
      CK(ei);
    }
    s->gkey = key;
    CK(hipGraphLaunch(s->gexec, s->stream));
  } else {
    int rc = cycle_body(s, mgit);
    if (rc) return rc;
  }
  s->cycle_pending = true;
  return COSAHIP_OK;
}

extern "C" int cosahip_autotune(cosahip_solve* h, int reps, int32_t* choice) {
  Solve* s = reinterpret_cast<Solve*>(h);
  if (!s) return COSAHIP_ERR_ARG;
  if (s->cycle_pending) return fail(s, COSAHIP_ERR_STATE, "autotune: an enqueued cycle has not been waited for");
  if (reps < 1 || reps > 100) return fail(s, COSAHIP_ERR_ARG, "autotune: reps out of range 1..100");
  CK(hipSetDevice(s->device));
  for (int m = 0; m < 8; m++)
    if (choice) choice[m] = 0;
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  auto timed = [&](auto&& fn, float& ms) -> int {
    if (fn() != hipSuccess) return fail(s, COSAHIP_ERR_DEVICE, "autotune: launch failed");   // warm-up
    CK(hipEventRecord(e0, s->stream));
    for (int r = 0; r < reps; r++)
      if (fn() != hipSuccess) return fail(s, COSAHIP_ERR_DEVICE, "autotune: launch failed");
    CK(hipEventRecord(e1, s->stream));
    CK(hipEventSynchronize(e1));
    CK(hipEventElapsedTime(&ms, e0, e1));
    return COSAHIP_OK;
  };
  int rc = COSAHIP_OK;
  for (int nl = 1; nl <= (int)s->lev.size() && nl <= 3 && rc == COSAHIP_OK; nl++) {
    Level& L = s->lev[nl - 1];
    if (!L.nblocks || s->par.irsop < 1 || s->par.irsop == 4 || s->par.irsop > 6 || s->cfg.cutcirs) continue;
    // the sweeps run on a copy of the residual (work1), the state is untouched
    const ll nq = L.n3tot * s->par.npde * s->par.nsnap;
    CK(hipMemcpyAsync(L.work1, L.f[COSAHIP_RES], nq * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
    std::swap(L.f[COSAHIP_RES], L.work1);
    float t[4] = {0, 0, 0, 0};
    const int last = L.ntile[TK_LINES_IL] ? 3 : 2;
    for (int mode = 1; mode <= last && rc == COSAHIP_OK; mode++) {
      L.cirs_mode = mode;
      rc = timed([&]() { return launch_cirs(*s, L); }, t[mode]);
    }
    std::swap(L.f[COSAHIP_RES], L.work1);
    L.cirs_mode = 1;
    for (int mode = 2; mode <= last; mode++)
      if (t[mode] < t[L.cirs_mode]) L.cirs_mode = mode;
    if (choice) choice[nl - 1] = L.cirs_mode;
  }
  // inviscid residual kernel per level: timed into work1 (q is only read)
  for (int nl = 1; nl <= (int)s->lev.size() && nl <= 3 && rc == COSAHIP_OK; nl++) {
    Level& L = s->lev[nl - 1];
    if (!L.nblocks || s->par.viscous || s->par.npde != 5 || s->split_residual || s->resid_force >= 0) {
      if (choice) choice[4 + nl - 1] = L.nblocks ? (s->resid_force >= 0 ? s->resid_force : L.resid_mode) : -1;
      continue;
    }
    const double* q = L.f[COSAHIP_Q];
    float t[RESID_NVARIANT] = {};
    const int keep = L.resid_mode;
    for (int v = 0; v < RESID_NVARIANT && rc == COSAHIP_OK; v++) {
      L.resid_mode = v;
      rc = timed([&]() { return launch_residual(*s, L, q, L.work1); }, t[v]);
    }
    L.resid_mode = keep;
    if (rc == COSAHIP_OK) {
      for (int v = 0; v < RESID_NVARIANT; v++)
        if (t[v] < t[L.resid_mode]) L.resid_mode = v;
    }
    if (choice) choice[4 + nl - 1] = L.resid_mode;
  }
  if (rc == COSAHIP_OK && s->par.viscous && !s->fuse_viscous && getenv("COSAHIP_VISCOUS") == nullptr &&
      !s->lev.empty() && s->lev[0].nblocks) {
    Level& L = s->lev[0];
    float t[3] = {0, 0, 0};
    const bool keep_faces = s->viscous_faces, keep_tile = s->viscous_tile;
    // on copies: the residual also fills the edge halos of the state it reads
    const ll nq = L.n3tot * s->par.npde * s->par.nsnap;
    CK(hipMemcpyAsync(L.dq, L.f[COSAHIP_Q], nq * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
    for (int f = 0; f <= 2 && rc == COSAHIP_OK; f++) {   // march, stored faces, tile pass
      s->viscous_faces = f >= 1;
      s->viscous_tile = f == 2;
      rc = timed([&]() { return launch_residual(*s, L, L.dq, L.work1); }, t[f]);
    }
    if (rc == COSAHIP_OK) {
      int best = 1;
      for (int f = 0; f <= 2; f++)
        if (t[f] < t[best]) best = f;
      s->viscous_faces = best >= 1;
      s->viscous_tile = best == 2;
    } else {
      s->viscous_faces = keep_faces;
      s->viscous_tile = keep_tile;
    }
    if (choice) choice[3] = s->viscous_tile ? 2 : (s->viscous_faces ? 1 : 0);
  } else if (choice) {
    choice[3] = s->viscous_tile ? 2 : (s->viscous_faces ? 1 : 0);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  // the launch sequence changed: a captured cycle graph is rebuilt
  if (s->gexec) {
    (void)hipGraphExecDestroy(s->gexec);
    s->gexec = nullptr;
  }
  return rc;
}

extern "C" int cosahip_cycle_ready(cosahip_solve* h) {
  Solve* s = reinterpret_cast<Solve*>(h);
  if (!s) return COSAHIP_ERR_ARG;
  if (!s->cycle_pending) return fail(s, COSAHIP_ERR_STATE, "cycle_ready: no cycle enqueued");
  CK(hipSetDevice(s->device));
  const hipError_t e = hipStreamQuery(s->stream);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  CK(e);
  return 0;
}

extern "C" int cosahip_cycle_wait(cosahip_solve* h, int32_t* done, double* resrms) {
  Solve* s = reinterpret_cast<Solve*>(h);
  if (!s) return COSAHIP_ERR_ARG;
  if (!s->cycle_pending) return fail(s, COSAHIP_ERR_STATE, "cycle_wait: no cycle enqueued");
  CK(hipSetDevice(s->device));
  CK(hipStreamSynchronize(s->stream));
  s->cycle_pending = false;
  double rr[8];
  rms_finish(s, s->lev[0], RMS_ASYNC_SLOT, rr);
  for (int v = 0; v <= s->par.npde; v++) s->resrms[v + 1] = rr[v];
  if (s->prec_mode == 2 && !s->prec_tail && s->resrms[1] < s->prec_tail_rms) s->prec_tail = true;
  double errmax = s->resrms[1];
  for (int v = 2; v <= 5; v++) errmax = std::max(errmax, s->resrms[v]);
  if (done) *done = errmax < s->cfg.toler ? 1 : 0;
  if (resrms)
    for (int v = 0; v < 8; v++) resrms[v] = s->resrms[v + 1];
  for (int nl = 1; nl <= s->cfg.nlevel; nl++) {
    unsigned long long key;
    memcpy(&key, s->hpinned + FLAG_ASYNC_SLOT + (nl - 1), sizeof(key));
    if (key != NOFLAG) return decode_prtest(s, nl, key);
  }
  return COSAHIP_OK;
}

// ---------------------------------------------------------------------------
// instrumentation API

extern "C" int cosahip_timing_enable(cosahip_solve* h, int on) {
  Solve* s = reinterpret_cast<Solve*>(h);
  if (!s) return COSAHIP_ERR_ARG;
  if (!on) drain_timing(*s);
  s->timing = on != 0;
  return COSAHIP_OK;
}

extern "C" int cosahip_timing_reset(cosahip_solve* h) {
  Solve* s = reinterpret_cast<Solve*>(h);
  if (!s) return COSAHIP_ERR_ARG;
  drain_timing(*s);
  for (int f = 0; f < COSAHIP_NFAM; f++) {
    s->fam_ms[f] = 0;
    s->fam_launch[f] = 0;
    s->fam_items[f] = 0;
  }
  return COSAHIP_OK;
}

extern "C" int cosahip_timing_read(cosahip_solve* h, int family, double* total_ms, int64_t* launches,
                                   int64_t* items) {
  Solve* s = reinterpret_cast<Solve*>(h);
  if (!s || family < 0 || family >= COSAHIP_NFAM) return COSAHIP_ERR_ARG;
  drain_timing(*s);
  if (total_ms) *total_ms = s->fam_ms[family];
  if (launches) *launches = s->fam_launch[family];
  if (items) *items = s->fam_items[family];
  return COSAHIP_OK;
}
