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
//                                                         (f.phase1, barrier, f.phase2 per block)
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
#include "nh_budget.h"
#include "nh_trcadv.h"
#include "../../include/nicamhip.h"

namespace nh {

// fields of the step workspace, each NF doubles
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
  W_MC = 44, W_MU = 45, W_ML = 46,
  W_GETOT = 47,           // 1
  W_FRHOG = 48,           // 1
  W_ETHTOT = 49,          // 1
  W_PM = 50,              // 5  time-mean mass fluxes
  W_SPLIT = 55,           // 12 split state (6) + path1 momentum (3) + vi_main out (3)
  // tracer transport
  W_FLXV = 67,            // 1  vertical mass flux of the half steps
  W_RHOG1 = 68,           // 1  density after the first vertical half step
  W_RHOG2 = 69,           // 1  density after the horizontal step
  W_FA = 70,              // 3  arc mass fluxes (AI, AIJ, AJ)
  W_XC = 73,              // 9  arc mass centroids
  W_GQ = 82,              // 3  gradient of the mixing ratio
  W_QA = 85,              // 3  upwind face values (faces 0..2)
  W_QOUT = 88,            // 2  limiter outflow bounds
  W_RQ = 90,              // 1  tracer after the horizontal step
  W_HV = 91,              // 3  horizontal divergences of the small-step momentum
  W_NFIELDS = 94
};

struct Context {
  Dims d;
  Consts c;
  Geo G;
  // owned device buffers
  double* geo_buf = nullptr;
  int* int_buf = nullptr;
  int* halo_dst = nullptr;
  int* halo_src = nullptr;
  long halo_n = 0;
  double* ws = nullptr;
  bool geo_ready = false;
  int device = 0;
  void* prof = nullptr;   // executor-specific kernel timer state
  double* field(int f) const { return ws + (long)f * d.NF; }
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

inline void set_consts(Consts& c, const nh_consts* in) {
  c.RADIUS = in->radius;
  c.OHM = in->ohm;
  c.GRAV = in->grav;
  c.Rdry = in->rdry;
  c.CVdry = in->cvdry;
  c.Rvap = in->rvap;
  c.CVW_qv = in->cvw_qv;
  c.ALPHA = in->alpha;
  c.GAMMA_H = in->gamma_h;
  c.HDIFF_FACT_RHO = in->hdiff_fact_rho;
  c.ALPHA_D = in->alpha_d;
  c.THUBURN_LIM = in->thuburn_lim;
}

// geometry arrays in upload order: regional [comp][nreg], pole [v][lp][...],
// vertical profiles [kall]
constexpr int kNReg = 13, kNPl = 12, kNVert = 14;

inline long geo_total(const Dims& d) {
  static const long rc[kNReg] = {21, 21, 7, 18, 18, 3, 6, 9, 1, 9, 1, 1, 1};
  static const long pc[kNPl] = {36, 36, 12, 108, 36, 36, 36, 36, 2, 36, 12, 12};
  long t = (long)kNVert * d.kall;
  for (int i = 0; i < kNReg; ++i) t += rc[i] * d.nreg;
  for (int i = 0; i < kNPl; ++i) t += pc[i];
  return t;
}

template <class X>
int upload_geometry(X& x, Context& ctx, const nh_geometry* g) {
  const Dims& d = ctx.d;
  static const long rc[kNReg] = {21, 21, 7, 18, 18, 3, 6, 9, 1, 9, 1, 1, 1};
  static const long pc[kNPl] = {36, 36, 12, 108, 36, 36, 36, 36, 2, 36, 12, 12};
  if (!ctx.geo_buf) {
    ctx.geo_buf = (double*)x.alloc(geo_total(d) * sizeof(double));
    ctx.int_buf = (int*)x.alloc(d.lall * sizeof(int));
    if (!ctx.geo_buf || !ctx.int_buf) return NH_ERR_ALLOC;
  }
  const double* regsrc[kNReg] = {g->coef_div, g->coef_grad, g->coef_lap, g->coef_intp, g->coef_diff,
                                 g->grd_x,    g->tri_w,     g->arc_hn,   g->rarea,      g->arc_xr,
                                 g->area,     g->lat,       g->lon};
  const double* plsrc[kNPl] = {g->coef_div_pl, g->coef_grad_pl, g->coef_lap_pl, g->coef_intp_pl,
                               g->coef_diff_pl, g->grd_x_pl,    g->tri_w_pl,    g->arc_hn_pl,
                               g->rarea_pl,     g->arc_xr_pl,   g->lat_pl,      g->lon_pl};
  const double* vsrc[kNVert] = {g->afact, g->bfact, g->cfact, g->dfact, g->rdgz, g->rdgzh, g->c2w0,
                                g->c2w1,  g->w2c0,  g->w2c1,  g->phi,   g->dgz,  g->gz,   g->gzh};
  for (int i = 0; i < kNReg; ++i)
    if (!regsrc[i]) return NH_ERR_ARG;
  for (int i = 0; i < kNPl; ++i)
    if (!plsrc[i]) return NH_ERR_ARG;
  for (int i = 0; i < kNVert; ++i)
    if (!vsrc[i]) return NH_ERR_ARG;
  if (!g->sgp) return NH_ERR_ARG;
  double* p = ctx.geo_buf;
  Geo& G = ctx.G;
  const double** regdst[kNReg] = {&G.cdiv, &G.cgrad, &G.clap, &G.cintp, &G.cdiff, &G.gx,  &G.twt,
                                  &G.ahn,  &G.rarea, &G.axr,  &G.area,  &G.lat,   &G.lon};
  const double** pldst[kNPl] = {&G.cdiv_pl, &G.cgrad_pl, &G.clap_pl,  &G.cintp_pl, &G.cdiff_pl, &G.gx_pl,
                                &G.twt_pl,  &G.ahn_pl,   &G.rarea_pl, &G.axr_pl,   &G.lat_pl,   &G.lon_pl};
  const double** vdst[kNVert] = {&G.afact, &G.bfact, &G.cfact, &G.dfact, &G.rdgz, &G.rdgzh, &G.c2w0,
                                 &G.c2w1,  &G.w2c0,  &G.w2c1,  &G.phi,   &G.dgz,  &G.gz,    &G.gzh};
  for (int i = 0; i < kNReg; ++i) {
    x.h2d(p, regsrc[i], rc[i] * d.nreg * sizeof(double));
    *regdst[i] = p;
    p += rc[i] * d.nreg;
  }
  for (int i = 0; i < kNPl; ++i) {
    x.h2d(p, plsrc[i], pc[i] * sizeof(double));
    *pldst[i] = p;
    p += pc[i];
  }
  for (int i = 0; i < kNVert; ++i) {
    x.h2d(p, vsrc[i], d.kall * sizeof(double));
    *vdst[i] = p;
    p += d.kall;
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
  if (n == 0) return 0;
  ctx.halo_dst = (int*)x.alloc(n * sizeof(int));
  ctx.halo_src = (int*)x.alloc(n * sizeof(int));
  if (!ctx.halo_dst || !ctx.halo_src) return NH_ERR_ALLOC;
  x.h2d(ctx.halo_dst, dst, n * sizeof(int));
  x.h2d(ctx.halo_src, src, n * sizeof(int));
  return x.check("upload_halo");
}

// in-place halo fill of up to 16 fields
template <class X>
void halo(X& x, const Context& ctx, double* const* f, int nf) {
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
void halo_block(X& x, const Context& ctx, double* base, int nf) {
  double* f[16];
  for (int q0 = 0; q0 < nf; q0 += 16) {
    int m = (nf - q0 < 16) ? nf - q0 : 16;
    for (int q = 0; q < m; ++q) f[q] = base + (long)(q0 + q) * ctx.d.NF;
    halo(x, ctx, f, m);
  }
}

inline long ncols(const Dims& d) { return d.nreg + 12; }

// ----------------------------------------------------------------------------
// 3-D divergence damping of (vx, vy, vz, w) into out (3 fields)
// ----------------------------------------------------------------------------
template <class X>
void divdamp(X& x, const Context& ctx, const double* vx, const double* vy, const double* vz, const double* w,
             double* out0, double* out1, double* out2) {
  const TileGeom tg = tile_geom(ctx.d, 2);
  DivDamp1T a;
  a.d = ctx.d; a.c = ctx.c; a.G = ctx.G; a.tg = tg;
  a.rvx = vx; a.rvy = vy; a.rvz = vz; a.rw = w;
  for (int q = 0; q < 3; ++q) a.out[q] = ctx.field(W_DDT + q);
  x.run_tiles(a, tg.nblocks, tg.lds);
  double* tmp[3] = {ctx.field(W_DDT), ctx.field(W_DDT + 1), ctx.field(W_DDT + 2)};
  halo(x, ctx, tmp, 3);
  DivDamp2T b;
  b.d = ctx.d; b.c = ctx.c; b.G = ctx.G; b.tg = tg;
  for (int q = 0; q < 3; ++q) b.in[q] = tmp[q];
  b.out[0] = out0; b.out[1] = out1; b.out[2] = out2;
  x.run_tiles(b, tg.nblocks, tg.lds);
}

// ----------------------------------------------------------------------------
// tracer transport with the time-mean mass fluxes of the last stage
// (mod_dynamics.py:2346-2461 -> mod_src_tracer.py:83-1228)
// ----------------------------------------------------------------------------
// pm: mass fluxes (rhog, rhogvx, rhogvy, rhogvz, rhogw) with field stride NF;
// rhog_in: density at the start of the step; frhog: density tendency
template <class X>
void tracer_transport(X& x, const Context& ctx, double* progq, int nq, double dt, const double* pm,
                      const double* rhog_in, const double* frhog) {
  const Dims& d = ctx.d;
  const long NF = d.NF, nc = ncols(d);
  double* rhog1 = ctx.field(W_RHOG1);
  double* rhog2 = ctx.field(W_RHOG2);
  double* flxv = ctx.field(W_FLXV);
  {
    TrcVert1 f;
    f.d = d; f.c = ctx.c; f.G = ctx.G;
    f.dt = dt;
    f.rw_mean = pm + 4 * NF;
    f.rhog_in = rhog_in;
    f.frhog = frhog;
    f.flxv = flxv; f.rhog1 = rhog1;
    f.rhogq = progq; f.nq = nq;
    f.limiter = ctx.c.THUBURN_LIM;
    x.run(f, nc);
  }
  double* fa[3] = {ctx.field(W_FA), ctx.field(W_FA + 1), ctx.field(W_FA + 2)};
  double* xc[9];
  for (int q = 0; q < 9; ++q) xc[q] = ctx.field(W_XC + q);
  {
    TrcHFlux f;
    f.d = d; f.c = ctx.c; f.G = ctx.G;
    f.dt = dt; f.pm = pm;
    for (int q = 0; q < 3; ++q) f.fa[q] = fa[q];
    for (int q = 0; q < 9; ++q) f.xc[q] = xc[q];
    x.run(f, nc);
  }
  {
    TrcRhog2 f;
    f.d = d; f.c = ctx.c; f.G = ctx.G;
    f.dt = dt;
    for (int q = 0; q < 3; ++q) f.fa[q] = fa[q];
    f.rhog1 = rhog1; f.frhog = frhog; f.rhog2 = rhog2;
    x.run(f, nc);
  }
  double* gq[3] = {ctx.field(W_GQ), ctx.field(W_GQ + 1), ctx.field(W_GQ + 2)};
  double* qa[3] = {ctx.field(W_QA), ctx.field(W_QA + 1), ctx.field(W_QA + 2)};
  double* qout[2] = {ctx.field(W_QOUT), ctx.field(W_QOUT + 1)};
  double* rqh = ctx.field(W_RQ);
  for (int iq = 0; iq < nq; ++iq) {
    double* rq = progq + (long)iq * NF;
    {
      TrcGradQ f;
      f.d = d; f.G = ctx.G;
      f.rq = rq; f.rhog1 = rhog1;
      for (int q = 0; q < 3; ++q) f.gq[q] = gq[q];
      x.run(f, nc);
    }
    halo(x, ctx, gq, 3);
    {
      TrcRemap f;
      f.d = d; f.c = ctx.c; f.G = ctx.G;
      f.dt = dt; f.rq = rq; f.rhog1 = rhog1;
      for (int q = 0; q < 3; ++q) { f.gq[q] = gq[q]; f.fa[q] = fa[q]; f.qa[q] = qa[q]; }
      for (int q = 0; q < 9; ++q) f.xc[q] = xc[q];
      x.run(f, nc);
    }
    if (ctx.c.THUBURN_LIM) {
      TrcQout f;
      f.d = d; f.c = ctx.c; f.G = ctx.G;
      f.dt = dt; f.rq = rq; f.rhog1 = rhog1; f.frhog = frhog;
      for (int q = 0; q < 3; ++q) f.fa[q] = fa[q];
      f.qout[0] = qout[0]; f.qout[1] = qout[1];
      x.run(f, nc);
      halo(x, ctx, qout, 2);
    }
    {
      TrcApply f;
      f.d = d; f.c = ctx.c; f.G = ctx.G;
      f.dt = dt; f.rq = rq; f.rhog1 = rhog1;
      for (int q = 0; q < 3; ++q) { f.fa[q] = fa[q]; f.qa[q] = qa[q]; }
      f.qout[0] = qout[0]; f.qout[1] = qout[1];
      f.out = rqh;
      x.run(f, nc);
    }
    {
      TrcVert2 f;
      f.d = d; f.c = ctx.c; f.G = ctx.G;
      f.dt = dt; f.flxv = flxv; f.rhog2 = rhog2; f.frhog = frhog;
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
int tracer_step(X& x, const Context& ctx, const double* prog, double* progq, int nq, double dtl) {
  const Dims& d = ctx.d;
  double* frhog = ctx.field(W_FRHOG);
  x.run(FillZero{frhog}, d.NF);
  tracer_transport(x, ctx, progq, nq, dtl, prog, prog, frhog);
  halo_block(x, ctx, progq, nq);
  return x.check("tracer_step");
}

template <class X>
int trcadv_winds(X& x, const Context& ctx, double* prog, double time, int kind, const double* par) {
  TrcadvWindsP f;
  f.d = ctx.d; f.G = ctx.G; f.prog = prog; f.time = time; f.kind = kind;
  for (int i = 0; i < 4; ++i) f.par[i] = par ? par[i] : 0.0;
  x.run_pts(f, ctx.d);
  return x.check("trcadv_winds");
}

// ----------------------------------------------------------------------------
// one large step (mod_dynamics.py:1725-2741 with the RK3/HEVI schedule)
// ----------------------------------------------------------------------------
// global budgets (Embudget._store_var, mod_embudget.py:82-139): block partials
// on the device into the W_GT scratch field, compensated host sum of them
template <class X>
int budget(X& x, const Context& ctx, const double* prog, const double* progq, double lhv, double out[6]) {
  const Dims& d = ctx.d;
  const int W = d.g - 2;
  const long npts = (long)d.lall * (d.kmax - d.kmin + 1) * W * W;
  const int nblocks = (int)((npts + 255) / 256);
  if ((long)nblocks * B_NTERMS * 2 > 6 * d.NF) return NH_ERR_ARG;
  double* part = ctx.field(W_GT);
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

template <class X>
int dyn_step(X& x, const Context& ctx, double* prog, double* progq, int nq, double dtl, int sstep_max) {
  const Dims& d = ctx.d;
  const long NF = d.NF;
  const long nc = ncols(d);
  const double dts = dtl / (double)sstep_max;
  const int nsub[3] = {sstep_max / 3, sstep_max / 2, sstep_max};
  if (nq < 1 || !prog || !progq) return NH_ERR_ARG;   // tracer 0 is qv (zero in a dry run)
  const double* rhogq = progq;
  double* prog0 = ctx.field(W_PROG0);
  double* diag = ctx.field(W_DIAG);
  double* eth = ctx.field(W_ETH);
  double* gadv = ctx.field(W_GADV);
  double* gT = ctx.field(W_GT);
  double* pm = ctx.field(W_PM);
  double* ethtot = ctx.field(W_ETHTOT);
  double* getot = ctx.field(W_GETOT);
  double* frhog = ctx.field(W_FRHOG);
  x.d2d(prog0, prog, 6 * NF * sizeof(double));
  for (int nl = 0; nl < 3; ++nl) {
    // diagnostics and boundary values
    {
      PrePostP f{d, ctx.c, ctx.G, prog, rhogq, diag, eth};
      x.run_pts(f, d);
    }
    {
      AdvMomP f{d, ctx.c, ctx.G, prog, diag, gadv};
      x.run_pts(f, d);
    }
    // hyperdiffusion: two passes with a halo fill after each
    {
      HDiffPassP<0> f;
      f.d = d; f.c = ctx.c; f.G = ctx.G;
      f.in[0] = diag + I_VX * NF; f.in[1] = diag + I_VY * NF; f.in[2] = diag + I_VZ * NF;
      f.in[3] = diag + I_W * NF; f.in[4] = diag + I_TEM * NF; f.in[5] = prog;
      for (int q = 0; q < 6; ++q) f.out[q] = ctx.field(W_HDA + q);
      f.rhog = prog;
      x.run_pts(f, d);
      halo_block(x, ctx, ctx.field(W_HDA), 6);
      HDiffPass2T f2;
      f2.d = d; f2.c = ctx.c; f2.G = ctx.G; f2.tg = tile_geom(d, 12);
      for (int q = 0; q < 6; ++q) {
        f2.in[q] = ctx.field(W_HDA + q);
        f2.out[q] = ctx.field(W_HDB + q);
      }
      f2.rhog = prog;
      x.run_tiles(f2, f2.tg.nblocks, f2.tg.lds);
      halo_block(x, ctx, ctx.field(W_HDB), 6);
    }
    // divergence damping of the large-step momentum
    divdamp(x, ctx, prog + 1 * NF, prog + 2 * NF, prog + 3 * NF, prog + 4 * NF,
            ctx.field(W_DD), ctx.field(W_DD + 1), ctx.field(W_DD + 2));
    {
      ViPath0P f;
      f.d = d; f.c = ctx.c; f.G = ctx.G;
      f.dt = dts;
      f.prog = prog; f.diag = diag; f.eth = eth; f.gadv = gadv;
      for (int q = 0; q < 6; ++q) f.vt[q] = ctx.field(W_HDB + q);
      for (int q = 0; q < 3; ++q) f.dd[q] = ctx.field(W_DD + q);
      f.gT = gT;
      f.Mc = ctx.field(W_MC); f.Mu = ctx.field(W_MU); f.Ml = ctx.field(W_ML);
      f.getot = getot; f.frhog = frhog;
      x.run_pts(f, d);
    }
    {
      EthTotPMP f{d, ctx.c, ctx.G, prog, eth, ethtot, pm};
      x.run_pts(f, d);
    }
    // split state: ps[0..5], then path1 momentum dv[0..2] and vi_main output we[0..2]
    double* ps[6];
    double* dv[3];
    double* we[3];
    for (int q = 0; q < 6; ++q) ps[q] = ctx.field(W_SPLIT + q);
    for (int q = 0; q < 3; ++q) dv[q] = ctx.field(W_SPLIT + 6 + q);
    for (int q = 0; q < 3; ++q) we[q] = ctx.field(W_SPLIT + 9 + q);
    {
      SplitInit f{NF, prog, prog0, {ps[0], ps[1], ps[2], ps[3], ps[4], ps[5]}, nl == 0 ? 1 : 0};
      x.run(f, NF);
    }
    const double rweight = 1.0 / (double)nsub[nl];
    for (int it = 0; it < nsub[nl]; ++it) {
      double* ddv = ctx.field(W_DDV);
      divdamp(x, ctx, ps[1], ps[2], ps[3], ps[4], ddv, ddv + NF, ddv + 2 * NF);
      {
        Path1P f;
        f.d = d; f.c = ctx.c; f.G = ctx.G;
        f.dt = dts; f.rweight = rweight;
        f.prog = prog;
        for (int q = 0; q < 6; ++q) f.ps[q] = ps[q];
        f.gT = gT;
        for (int q = 0; q < 3; ++q) f.dd[q] = ddv + q * NF;
        for (int q = 0; q < 3; ++q) f.dv[q] = dv[q];
        f.pm = pm;
        x.run_pts(f, d);
      }
      halo(x, ctx, dv, 3);
      double* hv[3] = {ctx.field(W_HV), ctx.field(W_HV + 1), ctx.field(W_HV + 2)};
      {
        ViMainH f;
        f.d = d; f.c = ctx.c; f.G = ctx.G;
        f.prog = prog;
        for (int q = 0; q < 3; ++q) f.dv[q] = dv[q];
        f.eth = eth; f.ethtot = ethtot; f.gT = gT;
        for (int q = 0; q < 3; ++q) f.hv[q] = hv[q];
        x.run_pts(f, d);
      }
      {
        ViMainV f;
        f.d = d; f.c = ctx.c; f.G = ctx.G;
        f.dt = dts;
        for (int q = 0; q < 6; ++q) f.ps[q] = ps[q];
        f.gT = gT;
        f.Mc = ctx.field(W_MC); f.Mu = ctx.field(W_MU); f.Ml = ctx.field(W_ML);
        for (int q = 0; q < 3; ++q) f.hv[q] = hv[q];
        for (int q = 0; q < 3; ++q) f.we[q] = we[q];
        x.run(f, nc);
      }
      {
        ViMainPost f;
        f.d = d; f.c = ctx.c; f.G = ctx.G;
        f.dt = dts; f.rweight = rweight;
        f.prog = prog;
        for (int q = 0; q < 6; ++q) f.ps[q] = ps[q];
        for (int q = 0; q < 3; ++q) f.dv[q] = dv[q];
        f.ethtot = ethtot; f.gT = gT; f.getot = getot;
        for (int q = 0; q < 3; ++q) f.hv[q] = hv[q];
        for (int q = 0; q < 3; ++q) f.we[q] = we[q];
        f.pm = pm;
        x.run_pts(f, d);
      }
      halo(x, ctx, we, 3);
      // rotate buffers: the new split state is (we0, dv0, dv1, dv2, we1, we2)
      double* old[6] = {ps[0], ps[1], ps[2], ps[3], ps[4], ps[5]};
      ps[0] = we[0]; ps[1] = dv[0]; ps[2] = dv[1]; ps[3] = dv[2]; ps[4] = we[1]; ps[5] = we[2];
      we[0] = old[0]; dv[0] = old[1]; dv[1] = old[2]; dv[2] = old[3]; we[1] = old[4]; we[2] = old[5];
    }
    {
      ProgUpdateP f;
      f.d = d; f.c = ctx.c; f.G = ctx.G;
      f.prog = prog;
      for (int q = 0; q < 6; ++q) f.ps[q] = ps[q];
      x.run_pts(f, d);
    }
    halo_block(x, ctx, pm, 5);
    if (nl < 2) halo_block(x, ctx, prog, 6);
    int rc = x.check("dyn_step stage");
    if (rc) return rc;
  }
  tracer_transport(x, ctx, progq, nq, dtl, ctx.field(W_PM), ctx.field(W_PROG0), ctx.field(W_FRHOG));
  halo_block(x, ctx, prog, 6);
  halo_block(x, ctx, progq, nq);
  return x.check("dyn_step");
}

}  // namespace nh
