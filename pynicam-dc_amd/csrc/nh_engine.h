// Context, workspace and launch sequence of the dynamical-core step.
//
// Generic over an executor X that provides
//   X(int device, void* stream, Context* ctx);
//   void* alloc(size_t bytes);  void release(void* p);
//   void h2d(void* dst, const void* src, size_t bytes);   (synchronous)
//   void d2d(void* dst, const void* src, size_t bytes);   (stream ordered)
//   template <class F> void run(const F& f, long n);      (f(t) for t < n)
//   template <class F> void run_pts(const F& f, const Dims& d);
//                                                         (f(plane, x) on the point grid)
//   template <class F> void run_tiles(const F& f, int nblocks, int lds_doubles);
//                                                         ([f.phase0, barrier,] f.phase1, barrier, f.phase2)
//   int check(const char* what);                          (0 or error)
// so that the same sequence drives the GPU build and the host test build.
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "nh_step.h"
#include "nh_tracer.h"
#include "nh_point.h"
#include "nh_tile.h"
#include "nh_colstencil.h"
#include "nh_budget.h"
#include "nh_trcadv.h"
#include "../../include/nicamhip.h"

NH_BEGIN

// fields of the step workspace, each NF values of the field precision
enum WsField {
  W_PROG0 = 0,            // 6  state at the start of the large step
  W_DIAG = 6,             // 6  pre, tem, vx, vy, vz, w
  W_ETH = 12,             // 1
  W_GADV = 13,            // 4  momentum advection tendency
  W_HDA = 17,             // 6  hyperdiffusion pass 1
  W_HDB = 23,             // 6  hyperdiffusion pass 2
  W_DDT = 29,             // 3  divergence damping, first operator
  W_DD = 32,              // 3  divergence damping of the large-step state
  W_DDV = 35,             // 3  divergence damping of the split state
  W_GT = 38,              // 6  large-step tendencies of the small steps
  W_GETOT = 44,           // 1
  W_FRHOG = 45,           // 1
  W_ETHTOT = 46,          // 1
  W_PM = 47,              // 5  time-mean mass fluxes
  W_SPLIT = 52,           // 12 split state (6) + path1 momentum (3) + vi_main out (3)
  // tracer transport
  W_FLXV = 64,            // 1  vertical mass flux of the half steps
  W_RHOG1 = 65,           // 1  density after the first vertical half step
  W_RHOG2 = 66,           // 1  density after the horizontal step
  W_FA = 67,              // 3  arc mass fluxes (AI, AIJ, AJ)
  W_XC = 70,              // 9  arc mass centroids
  W_GQ = 79,              // 3  gradient of the mixing ratio
  W_QA = 82,              // 3  upwind face values (faces 0..2)
  W_QOUT = 85,            // 2  limiter outflow bounds
  W_RQ = 87,              // 1  tracer after the horizontal step
  W_HV = 88,              // 3  horizontal divergences of the small-step momentum
  W_NFIELDS = 91
};

// FP64 fields (every form): the tridiagonal matrix, the forward-sweep scratch
// of the solve; the budgets' block partials reuse them between steps
enum DwsField { D_MC = 0, D_MU = 1, D_ML = 2, D_U = 3, D_GM = 4, D_NFIELDS = 5 };

struct Context {
  Dims d;
  Consts c;
  Geo G;
  // owned device buffers
  real* geo_buf = nullptr;      // coefficients and profiles in the field precision
  double* geo64_buf = nullptr;  // FP64 geometry (profiles, areas, latitude/longitude)
  int* int_buf = nullptr;
  int* halo_dst = nullptr;      // level-0 field addresses of the halo map
  int* halo_src = nullptr;
  long halo_n = 0;
  real* ws = nullptr;
  double* dws = nullptr;
  // adjoint workspace (allocated by the first adjoint call)
  real* aws = nullptr;
  double* adws = nullptr;
  int adj_ck = 0;              // small-step checkpoints it holds
  bool halo_adj_ok = false;    // halo destinations unique and never sources
  bool geo_ready = false;
  int device = 0;
  void* prof = nullptr;   // executor-specific kernel timer state
  real* field(int f) const { return ws + (long)f * d.NF; }
  double* dfield(int f) const { return dws + (long)f * d.NF; }
};

inline void set_dims(Dims& d, const nh_dims* in) {
  d.g = in->gall_1d;
  d.kall = in->kall;
  d.kmin = in->kmin;
  d.kmax = in->kmax;
  d.lall = in->lall;
  d.nreg = (long)d.lall * d.g * d.g;
  d.R = d.nreg * d.kall;
  d.NF = d.R + 2L * d.kall * 6;
}

template <class T>
inline void set_consts_t(ConstsT<T>& c, const nh_consts* in) {
  c.RADIUS = (T)in->radius;
  c.OHM = (T)in->ohm;
  c.GRAV = (T)in->grav;
  c.Rdry = (T)in->rdry;
  c.CVdry = (T)in->cvdry;
  c.Rvap = (T)in->rvap;
  c.CVW_qv = (T)in->cvw_qv;
  c.ALPHA = (T)in->alpha;
  c.GAMMA_H = (T)in->gamma_h;
  c.HDIFF_FACT_RHO = (T)in->hdiff_fact_rho;
  c.ALPHA_D = (T)in->alpha_d;
}

inline void set_consts(Consts& c, const nh_consts* in) {
  set_consts_t<real>(c, in);
  set_consts_t<double>(c.d, in);
  c.THUBURN_LIM = in->thuburn_lim;
}

// geometry arrays in upload order: regional [comp][nreg] and pole [v][lp][...]
// arrays in the field precision, vertical profiles [kall] in both precisions,
// then the FP64-only arrays
constexpr int kNReg = 10, kNPl = 10, kNVert = 14, kNReg64 = 3, kNPl64 = 2;
constexpr long kRegComp[kNReg] = {21, 21, 7, 18, 18, 3, 6, 9, 1, 9};
constexpr long kPlSize[kNPl] = {36, 36, 12, 108, 36, 36, 36, 36, 2, 36};

inline long geo_total(const Dims& d) {
  long t = (long)kNVert * d.kall;
  for (int i = 0; i < kNReg; ++i) t += kRegComp[i] * d.nreg;
  for (int i = 0; i < kNPl; ++i) t += kPlSize[i];
  return t;
}

inline long geo64_total(const Dims& d) { return (long)kNVert * d.kall + kNReg64 * d.nreg + kNPl64 * 12L + 3L * d.kall; }

// copy n doubles to the device in the field precision
template <class X>
void h2d_real(X& x, real* dst, const double* src, long n) {
  if (sizeof(real) == sizeof(double)) {
    x.h2d(dst, src, n * sizeof(double));
    return;
  }
  std::vector<real> tmp((size_t)n);
  for (long i = 0; i < n; ++i) tmp[(size_t)i] = (real)src[i];
  x.h2d(dst, tmp.data(), n * sizeof(real));
}

template <class X>
int upload_geometry(X& x, Context& ctx, const nh_geometry* g) {
  const Dims& d = ctx.d;
  if (!ctx.geo_buf) {
    ctx.geo_buf = (real*)x.alloc(geo_total(d) * sizeof(real));
    ctx.geo64_buf = (double*)x.alloc(geo64_total(d) * sizeof(double));
    ctx.int_buf = (int*)x.alloc(d.lall * sizeof(int));
    if (!ctx.geo_buf || !ctx.geo64_buf || !ctx.int_buf) return NH_ERR_ALLOC;
  }
  const double* regsrc[kNReg] = {g->coef_div, g->coef_grad, g->coef_lap, g->coef_intp, g->coef_diff,
                                 g->grd_x,    g->tri_w,     g->arc_hn,   g->rarea,      g->arc_xr};
  const double* plsrc[kNPl] = {g->coef_div_pl, g->coef_grad_pl, g->coef_lap_pl, g->coef_intp_pl,
                               g->coef_diff_pl, g->grd_x_pl,    g->tri_w_pl,    g->arc_hn_pl,
                               g->rarea_pl,     g->arc_xr_pl};
  const double* vsrc[kNVert] = {g->afact, g->bfact, g->cfact, g->dfact, g->rdgz, g->rdgzh, g->c2w0,
                                g->c2w1,  g->w2c0,  g->w2c1,  g->phi,   g->dgz,  g->gz,   g->gzh};
  const double* reg64[kNReg64] = {g->area, g->lat, g->lon};
  const double* pl64[kNPl64] = {g->lat_pl, g->lon_pl};
  for (int i = 0; i < kNReg; ++i)
    if (!regsrc[i]) return NH_ERR_ARG;
  for (int i = 0; i < kNPl; ++i)
    if (!plsrc[i]) return NH_ERR_ARG;
  for (int i = 0; i < kNVert; ++i)
    if (!vsrc[i]) return NH_ERR_ARG;
  for (int i = 0; i < kNReg64; ++i)
    if (!reg64[i]) return NH_ERR_ARG;
  for (int i = 0; i < kNPl64; ++i)
    if (!pl64[i]) return NH_ERR_ARG;
  if (!g->sgp) return NH_ERR_ARG;
  Geo& G = ctx.G;
  const real** regdst[kNReg] = {&G.cdiv, &G.cgrad, &G.clap, &G.cintp, &G.cdiff,
                                &G.gx,   &G.twt,   &G.ahn,  &G.rarea, &G.axr};
  const real** pldst[kNPl] = {&G.cdiv_pl, &G.cgrad_pl, &G.clap_pl, &G.cintp_pl, &G.cdiff_pl,
                              &G.gx_pl,   &G.twt_pl,   &G.ahn_pl,  &G.rarea_pl, &G.axr_pl};
  VProf<real>& VR = G;
  const real** vdst[kNVert] = {&VR.afact, &VR.bfact, &VR.cfact, &VR.dfact, &VR.rdgz, &VR.rdgzh, &VR.c2w0,
                               &VR.c2w1,  &VR.w2c0,  &VR.w2c1,  &VR.phi,   &VR.dgz,  &VR.gz,    &VR.gzh};
  VProf<double>& VD = G.v64;
  const double** v64dst[kNVert] = {&VD.afact, &VD.bfact, &VD.cfact, &VD.dfact, &VD.rdgz, &VD.rdgzh, &VD.c2w0,
                                   &VD.c2w1,  &VD.w2c0,  &VD.w2c1,  &VD.phi,   &VD.dgz,  &VD.gz,    &VD.gzh};
  const double** reg64dst[kNReg64] = {&G.area, &G.lat, &G.lon};
  const double** pl64dst[kNPl64] = {&G.lat_pl, &G.lon_pl};
  real* p = ctx.geo_buf;
  for (int i = 0; i < kNReg; ++i) {
    h2d_real(x, p, regsrc[i], kRegComp[i] * d.nreg);
    *regdst[i] = p;
    p += kRegComp[i] * d.nreg;
  }
  for (int i = 0; i < kNPl; ++i) {
    h2d_real(x, p, plsrc[i], kPlSize[i]);
    *pldst[i] = p;
    p += kPlSize[i];
  }
  for (int i = 0; i < kNVert; ++i) {
    h2d_real(x, p, vsrc[i], d.kall);
    *vdst[i] = p;
    p += d.kall;
  }
  double* q = ctx.geo64_buf;
  for (int i = 0; i < kNVert; ++i) {
    x.h2d(q, vsrc[i], d.kall * sizeof(double));
    *v64dst[i] = q;
    q += d.kall;
  }
  for (int i = 0; i < kNReg64; ++i) {
    x.h2d(q, reg64[i], d.nreg * sizeof(double));
    *reg64dst[i] = q;
    q += d.nreg;
  }
  for (int i = 0; i < kNPl64; ++i) {
    x.h2d(q, pl64[i], 12 * sizeof(double));
    *pl64dst[i] = q;
    q += 12;
  }
  // basic state (zeros when absent)
  {
    const double* bsrc[3] = {g->pre_bs, g->tem_bs, g->rho_bs};
    const double** bdst[3] = {&G.pre_bs, &G.tem_bs, &G.rho_bs};
    std::vector<double> zero((size_t)d.kall, 0.0);
    G.bs_on = false;
    for (int i = 0; i < 3; ++i) {
      const double* src = bsrc[i] ? bsrc[i] : zero.data();
      for (int k = 0; k < d.kall; ++k) G.bs_on = G.bs_on || src[k] != 0.0;
      x.h2d(q, src, d.kall * sizeof(double));
      *bdst[i] = q;
      q += d.kall;
    }
  }
  x.h2d(ctx.int_buf, g->sgp, d.lall * sizeof(int));
  G.sgp = ctx.int_buf;
  ctx.geo_ready = true;
  return x.check("upload_geometry");
}

template <class X>
int upload_halo(X& x, Context& ctx, const int32_t* dst, const int32_t* src, long n) {
  const long ncol = ctx.d.nreg + 12;
  for (long h = 0; h < n; ++h) {
    if (dst[h] < 0 || dst[h] >= ncol || src[h] < 0 || src[h] >= ncol) return NH_ERR_ARG;
  }
  if (ctx.halo_dst) { x.release(ctx.halo_dst); ctx.halo_dst = nullptr; }
  if (ctx.halo_src) { x.release(ctx.halo_src); ctx.halo_src = nullptr; }
  ctx.halo_n = n;
  // the adjoint of the gather scatters destination bars onto sources in one
  // pass: that needs every destination once and no destination as a source
  {
    std::vector<char> role((size_t)ncol, 0);
    bool ok = true;
    for (long h = 0; h < n && ok; ++h) {
      if (role[dst[h]] & 1) ok = false;
      role[dst[h]] |= 1;
    }
    for (long h = 0; h < n && ok; ++h) ok = !(role[src[h]] & 1);
    ctx.halo_adj_ok = ok;
  }
  if (n == 0) return 0;
  // the kernels index halo slot and level in 32 bits and take level-0 field
  // addresses in 32 bits
  if (n * ctx.d.kall >= (1L << 31) || ctx.d.NF >= (1L << 31)) return NH_ERR_ARG;
  std::vector<int> db((size_t)n), sb((size_t)n);
  for (long h = 0; h < n; ++h) {
    long ks;
    db[(size_t)h] = (int)halo_colbase(ctx.d, dst[h], &ks);
    sb[(size_t)h] = (int)halo_colbase(ctx.d, src[h], &ks);
  }
  ctx.halo_dst = (int*)x.alloc(n * sizeof(int));
  ctx.halo_src = (int*)x.alloc(n * sizeof(int));
  if (!ctx.halo_dst || !ctx.halo_src) return NH_ERR_ALLOC;
  x.h2d(ctx.halo_dst, db.data(), n * sizeof(int));
  x.h2d(ctx.halo_src, sb.data(), n * sizeof(int));
  return x.check("upload_halo");
}

// in-place halo fill of up to 16 fields
template <class X>
void halo(X& x, const Context& ctx, real* const* f, int nf) {
  if (ctx.halo_n == 0) return;
  for (int q0 = 0; q0 < nf; q0 += 16) {
    HaloGather h;
    h.d = ctx.d;
    h.dst = ctx.halo_dst;
    h.src = ctx.halo_src;
    h.n = ctx.halo_n;
    h.nf = (nf - q0 < 16) ? nf - q0 : 16;
    for (int q = 0; q < h.nf; ++q) h.f[q] = f[q0 + q];
    x.run(h, ctx.halo_n * ctx.d.kall);
  }
}

template <class X>
void halo_block(X& x, const Context& ctx, real* base, int nf) {
  real* f[16];
  for (int q0 = 0; q0 < nf; q0 += 16) {
    int m = (nf - q0 < 16) ? nf - q0 : 16;
    for (int q = 0; q < m; ++q) f[q] = base + (long)(q0 + q) * ctx.d.NF;
    halo(x, ctx, f, m);
  }
}

inline long ncols(const Dims& d) { return d.nreg + 12; }

// level chunks of the column tiles (divergence damping, stencil operators)
constexpr int kDivDampChunks = 4;
constexpr int kColChunks = 6;

// ----------------------------------------------------------------------------
// 3-D divergence damping of (vx, vy, vz, w) into out (3 fields)
// ----------------------------------------------------------------------------
template <class X>
void divdamp(X& x, const Context& ctx, const real* vx, const real* vy, const real* vz, const real* w,
             real* out0, real* out1, real* out2) {
  const ColGeom cg = col_geom(ctx.d, kDivDampChunks);
  DivDampK<1> a;
  a.d = ctx.d; a.c = ctx.c; a.G = ctx.G; a.cg = cg;
  a.v[0] = vx; a.v[1] = vy; a.v[2] = vz; a.rw = w;
  for (int q = 0; q < 3; ++q) a.out[q] = ctx.field(W_DDT + q);
  x.run_cols(a, cg.nblocks, DivDampK<1>::lds_elems(cg));
  real* tmp[3] = {ctx.field(W_DDT), ctx.field(W_DDT + 1), ctx.field(W_DDT + 2)};
  halo(x, ctx, tmp, 3);
  DivDampK<2> b;
  b.d = ctx.d; b.c = ctx.c; b.G = ctx.G; b.cg = cg;
  for (int q = 0; q < 3; ++q) b.v[q] = tmp[q];
  b.rw = nullptr;
  b.out[0] = out0; b.out[1] = out1; b.out[2] = out2;
  x.run_cols(b, cg.nblocks, DivDampK<2>::lds_elems(cg));
}

// ----------------------------------------------------------------------------
// tracer transport with the time-mean mass fluxes of the last stage
// (mod_dynamics.py:2346-2461 -> mod_src_tracer.py:83-1228)
// ----------------------------------------------------------------------------
// pm: mass fluxes (rhog, rhogvx, rhogvy, rhogvz, rhogw) with field stride NF;
// rhog_in: density at the start of the step; frhog: density tendency
template <class X>
void tracer_transport(X& x, const Context& ctx, real* progq, int nq, double dt, const real* pm,
                      const real* rhog_in, const real* frhog) {
  const Dims& d = ctx.d;
  const long NF = d.NF, nc = ncols(d);
  real* rhog1 = ctx.field(W_RHOG1);
  real* rhog2 = ctx.field(W_RHOG2);
  real* flxv = ctx.field(W_FLXV);
  {
    TrcVert1 f;
    f.d = d; f.c = ctx.c; f.G = ctx.G;
    f.dt = (real)dt;
    f.rw_mean = pm + 4 * NF;
    f.rhog_in = rhog_in;
    f.frhog = frhog;
    f.flxv = flxv; f.rhog1 = rhog1;
    f.rhogq = progq; f.nq = nq;
    f.limiter = ctx.c.THUBURN_LIM;
    x.run(f, nc);
  }
  real* fa[3] = {ctx.field(W_FA), ctx.field(W_FA + 1), ctx.field(W_FA + 2)};
  real* xc[9];
  for (int q = 0; q < 9; ++q) xc[q] = ctx.field(W_XC + q);
  {
    TrcHFlux f;
    f.d = d; f.c = ctx.c; f.G = ctx.G;
    f.dt = (real)dt; f.pm = pm;
    for (int q = 0; q < 3; ++q) f.fa[q] = fa[q];
    for (int q = 0; q < 9; ++q) f.xc[q] = xc[q];
    x.run_pts(f, d);
  }
  {
    TrcRhog2 f;
    f.d = d; f.c = ctx.c; f.G = ctx.G;
    f.dt = (real)dt;
    for (int q = 0; q < 3; ++q) f.fa[q] = fa[q];
    f.rhog1 = rhog1; f.frhog = frhog; f.rhog2 = rhog2;
    x.run_pts(f, d);
  }
  real* gq[3] = {ctx.field(W_GQ), ctx.field(W_GQ + 1), ctx.field(W_GQ + 2)};
  real* qa[3] = {ctx.field(W_QA), ctx.field(W_QA + 1), ctx.field(W_QA + 2)};
  real* qout[2] = {ctx.field(W_QOUT), ctx.field(W_QOUT + 1)};
  real* rqh = ctx.field(W_RQ);
  for (int iq = 0; iq < nq; ++iq) {
    real* rq = progq + (long)iq * NF;
    {
      TrcGradQ f;
      f.d = d; f.G = ctx.G;
      f.rq = rq; f.rhog1 = rhog1;
      for (int q = 0; q < 3; ++q) f.gq[q] = gq[q];
      x.run_pts(f, d);
    }
    halo(x, ctx, gq, 3);
    {
      TrcRemap f;
      f.d = d; f.c = ctx.c; f.G = ctx.G;
      f.dt = (real)dt; f.rq = rq; f.rhog1 = rhog1;
      for (int q = 0; q < 3; ++q) { f.gq[q] = gq[q]; f.fa[q] = fa[q]; f.qa[q] = qa[q]; }
      for (int q = 0; q < 9; ++q) f.xc[q] = xc[q];
      x.run_pts(f, d);
    }
    if (ctx.c.THUBURN_LIM) {
      TrcQout f;
      f.d = d; f.c = ctx.c; f.G = ctx.G;
      f.dt = (real)dt; f.rq = rq; f.rhog1 = rhog1; f.frhog = frhog;
      for (int q = 0; q < 3; ++q) f.fa[q] = fa[q];
      f.qout[0] = qout[0]; f.qout[1] = qout[1];
      x.run_pts(f, d);
      halo(x, ctx, qout, 2);
    }
    {
      TrcApply f;
      f.d = d; f.c = ctx.c; f.G = ctx.G;
      f.dt = (real)dt; f.rq = rq; f.rhog1 = rhog1;
      for (int q = 0; q < 3; ++q) { f.fa[q] = fa[q]; f.qa[q] = qa[q]; }
      f.qout[0] = qout[0]; f.qout[1] = qout[1];
      f.out = rqh; f.limiter = ctx.c.THUBURN_LIM;
      for (int q = 0; q < 3; ++q) f.gq[q] = gq[q];
      for (int q = 0; q < 9; ++q) f.xc[q] = xc[q];
      x.run_pts(f, d);
    }
    {
      TrcVert2 f;
      f.d = d; f.c = ctx.c; f.G = ctx.G;
      f.dt = (real)dt; f.flxv = flxv; f.rhog2 = rhog2; f.frhog = frhog;
      f.in = rqh; f.out = rq;
      f.limiter = ctx.c.THUBURN_LIM;
      x.run(f, nc);
    }
  }
}

// offline tracer step (TIME_integ_type 'TRCADV', mod_dynamics.py:2030-2076):
// the fluxes are prog's own momentum, rhog is the initial and the mean density,
// the density tendency is zero; the step ends with the halo fill of progq
template <class X>
int tracer_step(X& x, const Context& ctx, const real* prog, real* progq, int nq, double dtl) {
  const Dims& d = ctx.d;
  real* frhog = ctx.field(W_FRHOG);
  x.run(FillZero{frhog}, d.NF);
  tracer_transport(x, ctx, progq, nq, dtl, prog, prog, frhog);
  halo_block(x, ctx, progq, nq);
  return x.check("tracer_step");
}

template <class X>
int trcadv_winds(X& x, const Context& ctx, real* prog, double time, int kind, const double* par) {
  TrcadvWindsP f;
  f.d = ctx.d; f.G = ctx.G; f.prog = prog; f.time = time; f.kind = kind;
  for (int i = 0; i < 4; ++i) f.par[i] = par ? par[i] : 0.0;
  x.run_pts(f, ctx.d);
  return x.check("trcadv_winds");
}

// global budgets (Embudget._store_var, mod_embudget.py:82-139): FP64 block
// partials on the device into the FP64 scratch, compensated host sum of them
template <class X>
int budget(X& x, const Context& ctx, const real* prog, const real* progq, double lhv, double out[6]) {
  const Dims& d = ctx.d;
  const int W = d.g - 2;
  const long npts = (long)d.lall * (d.kmax - d.kmin + 1) * W * W;
  const int nblocks = (int)((npts + 255) / 256);
  if ((long)nblocks * B_NTERMS * 2 > D_NFIELDS * d.NF) return NH_ERR_ARG;
  double* part = ctx.dfield(0);
  BudgetT f{d, ctx.c, ctx.G, prog, progq, lhv, npts, part};
  x.run_tiles(f, nblocks, B_NTERMS * 256);
  std::vector<double> h((size_t)nblocks * B_NTERMS * 2);
  x.d2h(h.data(), part, h.size() * sizeof(double));
  if (x.check("budget")) return NH_ERR_DEVICE;
  double sum[B_NTERMS];
  for (int m = 0; m < B_NTERMS; ++m) {
    double s = 0.0, c = 0.0;
    for (int b = 0; b < nblocks; ++b) {
      neu_add(s, c, h[((size_t)b * B_NTERMS + m) * 2]);
      neu_add(s, c, h[((size_t)b * B_NTERMS + m) * 2 + 1]);
    }
    sum[m] = s + c;
  }
  out[0] = sum[B_POT];
  out[1] = sum[B_INT];
  out[2] = sum[B_KIN];
  out[3] = sum[B_POT] + sum[B_INT] + sum[B_KIN];   // etot = pot + eint + ekin
  out[4] = sum[B_DRY];
  out[5] = sum[B_VAP];
  return 0;
}

// the same sums with both reduction stages on the device and a caller-provided
// FP64 scratch of ws_n values (block partials, then the six final pairs); no
// allocation.  NH_ERR_ARG when the grid needs more partials than ws_n holds.
template <class X>
int budget_ws(X& x, const Context& ctx, const real* prog, const real* progq, double lhv, double out[6],
              double* ws, long ws_n) {
  const Dims& d = ctx.d;
  const int W = d.g - 2;
  const long npts = (long)d.lall * (d.kmax - d.kmin + 1) * W * W;
  const int nblocks = (int)((npts + 255) / 256);
  const long npart = (long)nblocks * B_NTERMS * 2;
  if (npart + 2 * B_NTERMS > ws_n) return NH_ERR_ARG;
  double* part = ws;
  double* fin = ws + npart;
  BudgetT f{d, ctx.c, ctx.G, prog, progq, lhv, npts, part};
  x.run_tiles(f, nblocks, B_NTERMS * 256);
  BudgetSumT r{nblocks, part, fin};
  x.run_tiles(r, 1, 2 * B_NTERMS * 256);
  double h[2 * B_NTERMS];
  x.d2h(h, fin, sizeof(h));
  if (x.check("budget")) return NH_ERR_DEVICE;
  double sum[B_NTERMS];
  for (int m = 0; m < B_NTERMS; ++m) sum[m] = h[2 * m] + h[2 * m + 1];
  out[0] = sum[B_POT];
  out[1] = sum[B_INT];
  out[2] = sum[B_KIN];
  out[3] = sum[B_POT] + sum[B_INT] + sum[B_KIN];
  out[4] = sum[B_DRY];
  out[5] = sum[B_VAP];
  return 0;
}

// ----------------------------------------------------------------------------
// one large step (mod_dynamics.py:1725-2741 with the RK3/HEVI schedule)
// ----------------------------------------------------------------------------
// the small-step buffers: split state ps (input), path1 momentum dv and
// vi_main output we; the new split state is (we0, dv0, dv1, dv2, we1, we2)
struct SmallBufs {
  real* ps[6];
  real* dv[3];
  real* we[3];
  real* ddv = nullptr;   // damping force of ps (3 fields); default the step workspace
  real* hv = nullptr;    // divergences of dv (3 fields); default the step workspace
};

// the fields a stage computes that its small steps (and the adjoint) read:
// the step workspace, or a tape slot when the step is taped
struct StageFields {
  real* hda;       // 6  hyperdiffusion pass 1
  real* hdb;       // 6  hyperdiffusion pass 2
  real* gT;        // 6  large-step tendencies
  real* ethtot;    // 1
  double* M;       // 3  FP64 tridiagonal matrix Mc, Mu, Ml
};

inline StageFields ws_stage(const Context& ctx) {
  return {ctx.field(W_HDA), ctx.field(W_HDB), ctx.field(W_GT), ctx.field(W_ETHTOT), ctx.dfield(D_MC)};
}

// hyperdiffusion (numfilter_hdiffusion, mod_numfilter.py:845-1469, lap order 2):
// the iterated Laplacians of (vx, vy, vz, w, T, rho) in two passes, pass 1
// into sf.hda with a halo fill, pass 2 into sf.hdb (6 fields each).  The halo
// fill of pass 2 (fill_hdb) is for callers that read hdb off the compute
// points; the step reads it only at its own compute points (path 0).
template <class X>
void hdiff_passes(X& x, const Context& ctx, const real* prog, const real* diag, const StageFields& sf,
                  bool fill_hdb = true) {
  const Dims& d = ctx.d;
  const long NF = d.NF;
  HDiffPassP<0> f;
  f.d = d; f.c = ctx.c; f.G = ctx.G;
  f.in[0] = diag + I_VX * NF; f.in[1] = diag + I_VY * NF; f.in[2] = diag + I_VZ * NF;
  f.in[3] = diag + I_W * NF; f.in[4] = diag + I_TEM * NF; f.in[5] = prog;
  for (int q = 0; q < 6; ++q) f.out[q] = sf.hda + q * NF;
  f.rhog = prog;
  if (ctx.G.bs_on) {
    f.sub[4] = ctx.G.tem_bs;
    f.sub[5] = ctx.G.rho_bs;
  }
  const ColGeom cg = col_geom(d, kColChunks);
  const StGeom sg = st_geom(d, kColChunks);
  {
    ColStencil<LapC<6, HDiffPassP<0>>> cs;
    cs.op.pk = f;
    cs.op.d = d;
    for (int q = 0; q < 6; ++q) { cs.op.in[q] = f.in[q]; cs.op.out[q] = f.out[q]; }
    cs.op.pole_off = false;
    cs.cg = sg;
    x.run_cols(cs, sg.nblocks, decltype(cs)::lds_elems(sg));
  }
  halo_block(x, ctx, sf.hda, 6);
  // second pass: the Laplacians of fields 0..3, then the Kh-weighted
  // diffusion of fields 4, 5 (whose pole blocks do all six fields)
  HDiffPassP<1> f2;
  f2.d = d; f2.c = ctx.c; f2.G = ctx.G;
  for (int q = 0; q < 6; ++q) {
    f2.in[q] = sf.hda + q * NF;
    f2.out[q] = sf.hdb + q * NF;
  }
  f2.rhog = prog;
  {
    ColStencil<LapC<4, HDiffPassP<1>>> cs;
    cs.op.pk = f2;
    cs.op.d = d;
    for (int q = 0; q < 4; ++q) { cs.op.in[q] = f2.in[q]; cs.op.out[q] = f2.out[q]; }
    cs.op.pole_off = true;
    cs.cg = sg;
    x.run_cols(cs, sg.nblocks, decltype(cs)::lds_elems(sg));
  }
  {
    DiffKhK k2;
    k2.d = d; k2.c = ctx.c; k2.G = ctx.G; k2.cg = cg;
    for (int q = 0; q < 6; ++q) { k2.in[q] = f2.in[q]; k2.out[q] = f2.out[q]; }
    k2.rhog = prog;
    x.run_cols(k2, cg.nblocks, DiffKhK::lds_elems(cg));
  }
  if (fill_hdb) halo_block(x, ctx, sf.hdb, 6);
}

// large-step part of RK stage nl: diagnostics and boundary values (in place
// on prog's ghost rows), momentum advection, hyperdiffusion, divergence
// damping of prog, the sources and matrix of the small steps, ethtot and the
// start of the mean fluxes
template <class X>
void stage_large(X& x, const Context& ctx, real* prog, const real* rhogq, double dts, const StageFields& sf,
                 bool means = true) {
  const Dims& d = ctx.d;
  const long NF = d.NF;
  real* diag = ctx.field(W_DIAG);
  real* eth = ctx.field(W_ETH);
  real* gadv = ctx.field(W_GADV);
  {
    PrePostP f{d, ctx.c, ctx.G, prog, rhogq, diag, eth};
    x.run_pts(f, d);
  }
  // momentum advection: column tiles with 4-byte fields (gl07: 292 -> 189 us);
  // with 8-byte fields their registers (236) cost more than the loads they save
  // (466 -> 596 us), so the FP64 form keeps the level-staged point kernel
  if (sizeof(real) == sizeof(float)) {
    ColStencil<AdvMomC> cs;
    cs.op.pk = AdvMomP{d, ctx.c, ctx.G, prog, diag, gadv};
    cs.op.d = d;
    cs.cg = st_geom(d, kColChunks);
    x.run_cols(cs, cs.cg.nblocks, decltype(cs)::lds_elems(cs.cg));
  } else {
    run_staged(x, AdvMomP{d, ctx.c, ctx.G, prog, diag, gadv});
  }
  hdiff_passes(x, ctx, prog, diag, sf, false);
  // divergence damping of the large-step momentum
  divdamp(x, ctx, prog + 1 * NF, prog + 2 * NF, prog + 3 * NF, prog + 4 * NF,
          ctx.field(W_DD), ctx.field(W_DD + 1), ctx.field(W_DD + 2));
  {
    ViPath0P f;
    f.d = d; f.c = ctx.c; f.G = ctx.G;
    f.dt = dts;
    f.prog = prog; f.diag = diag; f.eth = eth; f.gadv = gadv;
    for (int q = 0; q < 6; ++q) f.vt[q] = sf.hdb + q * NF;
    for (int q = 0; q < 3; ++q) f.dd[q] = ctx.field(W_DD + q);
    f.gT = sf.gT;
    f.Mc = sf.M; f.Mu = sf.M + NF; f.Ml = sf.M + 2 * NF;
    f.getot = ctx.field(W_GETOT); f.frhog = ctx.field(W_FRHOG);
    run_staged(x, f);
  }
  {
    EthTotPMP f{d, ctx.c, ctx.G, prog, eth, sf.ethtot, means ? ctx.field(W_PM) : nullptr};
    x.run_pts(f, d);
  }
}

// the horizontal divergences of the small-step momentum dv (ViMainH) into hv (3 fields)
template <class X>
void small_hv(X& x, const Context& ctx, const real* prog, real* const dv[3], real* hv, const StageFields& sf) {
  ViMainH f;
  f.d = ctx.d; f.c = ctx.c; f.G = ctx.G;
  f.prog = prog;
  for (int q = 0; q < 3; ++q) f.dv[q] = dv[q];
  f.eth = ctx.field(W_ETH); f.ethtot = sf.ethtot; f.gT = sf.gT;
  for (int q = 0; q < 3; ++q) f.hv[q] = hv + q * ctx.d.NF;
  ColStencil<ViMainHC> cs;
  cs.op.pk = f;
  cs.op.d = ctx.d;
  cs.cg = st_geom(ctx.d, kColChunks);
  x.run_cols(cs, cs.cg.nblocks, ColStencil<ViMainHC>::lds_elems(cs.cg));
}

// one small step (mod_vi.py:1283-1450): divergence damping of the split
// momentum, path1, vi_main (horizontal, tridiagonal, mass/energy), halos
template <class X>
void small_step(X& x, const Context& ctx, const real* prog, const SmallBufs& s, double dts, double rweight,
                const StageFields& sf, bool means = true) {
  const Dims& d = ctx.d;
  const long NF = d.NF;
  real* gT = sf.gT;
  real* pm = means ? ctx.field(W_PM) : nullptr;   // the tracer's mean mass fluxes
  real* ddv = s.ddv ? s.ddv : ctx.field(W_DDV);
  real* hvb = s.hv ? s.hv : ctx.field(W_HV);
  divdamp(x, ctx, s.ps[1], s.ps[2], s.ps[3], s.ps[4], ddv, ddv + NF, ddv + 2 * NF);
  {
    Path1P f;
    f.d = d; f.c = ctx.c; f.G = ctx.G;
    f.dt = (real)dts; f.rweight = (real)rweight;
    f.prog = prog;
    for (int q = 0; q < 6; ++q) f.ps[q] = s.ps[q];
    f.gT = gT;
    for (int q = 0; q < 3; ++q) f.dd[q] = ddv + q * NF;
    for (int q = 0; q < 3; ++q) f.dv[q] = s.dv[q];
    f.pm = pm;
    run_staged(x, f);
  }
  halo(x, ctx, s.dv, 3);
  small_hv(x, ctx, prog, s.dv, hvb, sf);
  real* hv[3] = {hvb, hvb + NF, hvb + 2 * NF};
  {
    // the z40 column (39 unknowns) keeps the sweep in registers
    auto solve = [&](auto f) {
      f.d = d; f.c = ctx.c; f.G = ctx.G;
      f.dt = dts;
      for (int q = 0; q < 6; ++q) f.ps[q] = s.ps[q];
      f.gT = gT;
      f.Mc = sf.M; f.Mu = sf.M + NF; f.Ml = sf.M + 2 * NF;
      for (int q = 0; q < 3; ++q) f.hv[q] = hv[q];
      f.U = ctx.dfield(D_U); f.Gm = ctx.dfield(D_GM);
      f.out = s.we[1];
      x.run(f, ncols(d));
    };
    if (d.kmax - d.kmin == 39) solve(ViMainV<39>{});
    else solve(ViMainV<0>{});
  }
  {
    ViMainPost f;
    f.d = d; f.c = ctx.c; f.G = ctx.G;
    f.dt = (real)dts; f.rweight = (real)rweight;
    f.prog = prog;
    for (int q = 0; q < 6; ++q) f.ps[q] = s.ps[q];
    for (int q = 0; q < 3; ++q) f.dv[q] = s.dv[q];
    f.ethtot = sf.ethtot; f.gT = gT; f.getot = ctx.field(W_GETOT);
    for (int q = 0; q < 3; ++q) f.hv[q] = hv[q];
    for (int q = 0; q < 3; ++q) f.we[q] = s.we[q];
    f.pm = pm;
    x.run_pts(f, d);
  }
  halo(x, ctx, s.we, 3);
}

// PROG += PROG_split with the horizontal projection
template <class X>
void prog_update(X& x, const Context& ctx, real* prog, real* const ps[6]) {
  ProgUpdateP f;
  f.d = ctx.d; f.c = ctx.c; f.G = ctx.G;
  f.prog = prog;
  for (int q = 0; q < 6; ++q) f.ps[q] = ps[q];
  x.run_pts(f, ctx.d);
}

// The tape of one step (optional, for the adjoint): per RK stage its input
// state and the stage fields, per small step its split state and vi_main
// divergences, in field-precision fields of NF values, then the three
// stages' FP64 matrices.  The step writes them in place of its workspace
// fields, so taping adds no traffic beyond the stage-input copies.
struct TapeLayout {
  static constexpr int kStage = 25;   // input 6, hda 6, hdb 6, gT 6, ethtot 1
  static constexpr int kSlot = 9;     // split state 6, divergences 3
  int nsub[3];
  long real_fields() const { return 3L * kStage + (long)kSlot * (nsub[0] + nsub[1] + nsub[2] + 3); }
  long slot0(int nl) const {
    long s = 3L * kStage;
    for (int m = 0; m < nl; ++m) s += (long)kSlot * (nsub[m] + 1);
    return s;
  }
  static size_t bytes(long NF, int sstep_max) {
    TapeLayout t{{sstep_max / 3, sstep_max / 2, sstep_max}};
    return (size_t)t.real_fields() * NF * sizeof(real) + (size_t)9 * NF * sizeof(double);
  }
};

struct Tape {
  TapeLayout L;
  real* r;       // real part
  double* m;     // FP64 matrices
  long NF;
  Tape(void* base, long NF_, int sstep_max)
      : L{{sstep_max / 3, sstep_max / 2, sstep_max}}, r((real*)base), m(nullptr), NF(NF_) {
    m = (double*)(r + L.real_fields() * NF);
  }
  real* input(int nl) const { return r + (long)nl * TapeLayout::kStage * NF; }
  StageFields stage(int nl) const {
    real* b = input(nl);
    return {b + 6 * NF, b + 12 * NF, b + 18 * NF, b + 24 * NF, m + (long)nl * 3 * NF};
  }
  real* slot(int nl, int it) const { return r + (L.slot0(nl) + (long)TapeLayout::kSlot * it) * NF; }
};

// record (optional): the input state of each RK stage, 3 x 6 fields;
// tape (optional): the full tape of the step (record is then unused)
template <class X>
int dyn_step(X& x, const Context& ctx, real* prog, real* progq, int nq, double dtl, int sstep_max,
             real* record = nullptr, void* tape = nullptr) {
  const Dims& d = ctx.d;
  const long NF = d.NF;
  const double dts = dtl / (double)sstep_max;
  const int nsub[3] = {sstep_max / 3, sstep_max / 2, sstep_max};
  if (nq < 1 || !prog || !progq) return NH_ERR_ARG;   // tracer 0 is qv (zero in a dry run)
  real* prog0 = ctx.field(W_PROG0);
  x.d2d(prog0, prog, 6 * NF * sizeof(real));
  const Tape tp(tape, NF, sstep_max);
  for (int nl = 0; nl < 3; ++nl) {
    if (tape) x.d2d(tp.input(nl), prog, 6 * NF * sizeof(real));
    else if (record) x.d2d(record + (long)nl * 6 * NF, prog, 6 * NF * sizeof(real));
    const StageFields sf = tape ? tp.stage(nl) : ws_stage(ctx);
    stage_large(x, ctx, prog, progq, dts, sf);
    // split state ps[0..5], path1 momentum dv[0..2], vi_main output we[0..2]
    SmallBufs s;
    if (tape) {
      real* c0 = tp.slot(nl, 0);
      for (int q = 0; q < 6; ++q) s.ps[q] = c0 + q * NF;
    } else {
      for (int q = 0; q < 6; ++q) s.ps[q] = ctx.field(W_SPLIT + q);
      for (int q = 0; q < 3; ++q) s.dv[q] = ctx.field(W_SPLIT + 6 + q);
      for (int q = 0; q < 3; ++q) s.we[q] = ctx.field(W_SPLIT + 9 + q);
    }
    {
      SplitInit f{NF, prog, prog0, {s.ps[0], s.ps[1], s.ps[2], s.ps[3], s.ps[4], s.ps[5]}, nl == 0 ? 1 : 0};
      x.run(f, NF);
    }
    const double rweight = 1.0 / (double)nsub[nl];
    for (int it = 0; it < nsub[nl]; ++it) {
      if (tape) {
        // outputs straight into the next slot: (we0, dv0, dv1, dv2, we1, we2)
        real* c1 = tp.slot(nl, it + 1);
        s.we[0] = c1; s.dv[0] = c1 + NF; s.dv[1] = c1 + 2 * NF; s.dv[2] = c1 + 3 * NF;
        s.we[1] = c1 + 4 * NF; s.we[2] = c1 + 5 * NF;
        s.hv = tp.slot(nl, it) + 6 * NF;
        small_step(x, ctx, prog, s, dts, rweight, sf);
        for (int q = 0; q < 6; ++q) s.ps[q] = c1 + q * NF;
        continue;
      }
      small_step(x, ctx, prog, s, dts, rweight, sf);
      // rotate buffers: the new split state is (we0, dv0, dv1, dv2, we1, we2)
      real* old[6] = {s.ps[0], s.ps[1], s.ps[2], s.ps[3], s.ps[4], s.ps[5]};
      s.ps[0] = s.we[0]; s.ps[1] = s.dv[0]; s.ps[2] = s.dv[1]; s.ps[3] = s.dv[2]; s.ps[4] = s.we[1];
      s.ps[5] = s.we[2];
      s.we[0] = old[0]; s.dv[0] = old[1]; s.dv[1] = old[2]; s.dv[2] = old[3]; s.we[1] = old[4]; s.we[2] = old[5];
    }
    prog_update(x, ctx, prog, s.ps);
    // the mean fluxes feed the tracer transport after the last stage; each
    // stage restarts them at every point (EthTotPMP)
    if (nl == 2) halo_block(x, ctx, ctx.field(W_PM), 5);
    if (nl < 2) halo_block(x, ctx, prog, 6);
    int rc = x.check("dyn_step stage");
    if (rc) return rc;
  }
  tracer_transport(x, ctx, progq, nq, dtl, ctx.field(W_PM), ctx.field(W_PROG0), ctx.field(W_FRHOG));
  halo_block(x, ctx, prog, 6);
  halo_block(x, ctx, progq, nq);
  return x.check("dyn_step");
}

NH_END
