// Adjoint of the dynamics step and the surface-pressure diagnostic.
//
// dyn_step_adj(prog_in, progq, prog_b): prog_b holds the bar of the state
// after one dyn_step from prog_in on entry and the bar of prog_in on exit.
// Checkpointing inside the step: a forward sweep records the input of each
// RK stage; the reverse sweep then, per stage, recomputes the stage's
// large-step fields and its small steps (each small step's split state kept
// in a checkpoint slot), and runs the transposed kernels of nh_adjoint.h in
// reverse order.  The tracer masses are held at progq (see nh_adjoint.h).
#pragma once

#include "nh_adjoint.h"
#include "nh_engine.h"
#include "nh_tracer_adj.h"

NH_BEGIN

// adjoint workspace, fields of NF values (field precision)
enum AwsField {
  A_PROG = 0,      // 6  state of the recomputed stage
  A_SC = 6,        // 18 stage inputs
  A_P0B = 24,      // 6  bar of the step's initial state (PROG0)
  A_DIAGB = 30,    // 6  stage bars from here to A_CK
  A_ETHB = 36,     // 1
  A_GADVB = 37,    // 4
  A_HDAB = 41,     // 6
  A_HDBB = 47,     // 6
  A_DDB = 53,      // 3
  A_DDTB = 56,     // 3
  A_GTB = 59,      // 6
  A_GETOTB = 65,   // 1
  A_ETHTOTB = 66,  // 1
  A_PSB = 67,      // 6  bar of the split state after the small step being reversed
  A_PSB2 = 73,     // 6  bar of the split state before it
  A_HVB = 79,      // 3
  A_DDVB = 82,     // 3
  A_TRIW = 85,     // 6  triangle weights of the damping transposes
  A_CK = 91        // per small-step checkpoint slot: its split state (6), then the
                   // divergences (3) of the step from it
};
constexpr int kSlot = 9;
// FP64: bars of the tridiagonal matrix, the recomputed solution
enum AdwsField { AD_MCB = 0, AD_MUB = 1, AD_MLB = 2, AD_X = 3, AD_NFIELDS = 4 };

inline long adj_fields(int nck) { return (long)A_CK + (long)kSlot * nck; }

inline real* afield(const Context& ctx, int f) { return ctx.aws + (long)f * ctx.d.NF; }
inline double* adfield(const Context& ctx, int f) { return ctx.adws + (long)f * ctx.d.NF; }

template <class X>
void zero_bytes(X& x, void* p, long nbytes) {
  const ZeroWords z = zero_words(p, nbytes);
  x.run(z, z.threads());
}
template <class X>
void zero_fields(X& x, real* p, long nfields, long NF) {
  zero_bytes(x, p, nfields * NF * (long)sizeof(real));
}

// transpose of halo(): destination bars onto their sources, up to 16 fields
template <class X>
void halo_adj(X& x, const Context& ctx, real* const* f, int nf) {
  if (ctx.halo_n == 0) return;
  for (int q0 = 0; q0 < nf; q0 += 16) {
    HaloScatterA h;
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
void halo_block_adj(X& x, const Context& ctx, real* base, int nf) {
  real* f[16];
  for (int q0 = 0; q0 < nf; q0 += 16) {
    const int n = (nf - q0 < 16) ? nf - q0 : 16;
    for (int q = 0; q < n; ++q) f[q] = base + (long)(q0 + q) * ctx.d.NF;
    halo_adj(x, ctx, f, n);
  }
}

// transpose of divdamp(): bars of the damping force ob -> bars of
// rhogvx, rhogvy, rhogvz, rhogw (vb); A_DDTB is the intermediate bar
template <class X>
void divdamp_adj(X& x, const Context& ctx, real* const ob[3], real* const vb[4]) {
  real* tb[3] = {afield(ctx, A_DDTB), afield(ctx, A_DDTB + 1), afield(ctx, A_DDTB + 2)};
  real* S[2] = {afield(ctx, A_TRIW), afield(ctx, A_TRIW + 1)};
  DivDampSA<2> s2{ctx.d, ctx.c, ctx.G, {ob[0], ob[1], ob[2]}, {S[0], S[1]}};
  x.run_pts(s2, ctx.d);
  DivDampGA<2> g2{ctx.d, ctx.c, ctx.G, {S[0], S[1]}, {tb[0], tb[1], tb[2], nullptr}};
  x.run_pts(g2, ctx.d);
  halo_adj(x, ctx, tb, 3);
  DivDampSA<1> s1{ctx.d, ctx.c, ctx.G, {tb[0], tb[1], tb[2]}, {S[0], S[1]}};
  x.run_pts(s1, ctx.d);
  DivDampGA<1> g1{ctx.d, ctx.c, ctx.G, {S[0], S[1]}, {vb[0], vb[1], vb[2], vb[3]}};
  x.run_pts(g1, ctx.d);
}

// the transposed Laplacian of NegLapGA: interior points as column tiles,
// then the border and pole points
template <int N, class X>
void lap_adj(X& x, const Context& ctx, const NegLapGA& l) {
  const StGeom sg = st_geom(ctx.d, kColChunks);
  ColStencil<LapTC<N>> cs;
  cs.op.d = ctx.d;
  cs.op.G = ctx.G;
  for (int q = 0; q < N; ++q) { cs.op.ob[q] = l.ob[q]; cs.op.ib[q] = l.ib[q]; }
  cs.cg = sg;
  x.run_cols(cs, sg.nblocks, decltype(cs)::lds_elems(sg));
  x.run(l, l.border_threads());
}

// reverse of one small step: ck = its input split state, nx = its output
// (we0, dv0, dv1, dv2, we1, we2); psb (bar of nx, consumed) -> psb2 (bar of
// ck, added) and the stage bars
// (hvf: the divergences the recomputed step kept)
template <class X>
void small_step_adj(X& x, const Context& ctx, const real* prog, real* const ck[6], real* const nx[6],
                    real* hvf, double dts, real* prog_b, real* const psb[6], real* const psb2[6],
                    const StageFields& sf) {
  const Dims& d = ctx.d;
  const long NF = d.NF;
  real* gT = sf.gT;
  real* gTb = afield(ctx, A_GTB);
  real* dvn[3] = {nx[1], nx[2], nx[3]};
  real* hv[3] = {hvf, hvf + NF, hvf + 2 * NF};
  real* hvb[3] = {afield(ctx, A_HVB), afield(ctx, A_HVB + 1), afield(ctx, A_HVB + 2)};
  real* web[3] = {psb[0], psb[4], psb[5]};
  real* dvb[3] = {psb[1], psb[2], psb[3]};
  halo_adj(x, ctx, web, 3);
  // Path1A stores the damping-force bars whole; ViMainPostG stores psb2 and
  // the divergence bars with 4-byte fields, and adds to them (cleared here
  // and by the caller) with 8-byte fields
  if (!ViMainPostG::kStore) zero_fields(x, hvb[0], 3, NF);
  {
    ViMainPostA f;
    f.d = d; f.c = ctx.c; f.G = ctx.G;
    f.dt = (real)dts;
    f.prog = prog;
    for (int q = 0; q < 6; ++q) { f.ps[q] = ck[q]; f.ps_b[q] = psb2[q]; }
    for (int q = 0; q < 3; ++q) { f.dv[q] = dvn[q]; f.dv_b[q] = dvb[q]; f.hv[q] = hv[q]; f.hv_b[q] = hvb[q]; }
    f.ethtot = sf.ethtot; f.gT = gT;
    f.we1 = nx[4];
    for (int q = 0; q < 3; ++q) f.we_b[q] = web[q];
    f.prog_b = prog_b; f.ethtot_b = afield(ctx, A_ETHTOTB); f.gT_b = gTb; f.getot_b = afield(ctx, A_GETOTB);
    x.run_pts(ViMainPostG{f}, d);
  }
  {
    // transposed solve: y per column, then the bars per point and level
    auto sweep = [&](auto f) {
      f.d = d;
      f.out_b = web[1];
      f.Mc = sf.M; f.Mu = sf.M + NF; f.Ml = sf.M + 2 * NF;
      f.Y = ctx.dfield(D_U); f.Gm = ctx.dfield(D_GM);
      x.run(f, ncols(d));
    };
    if (d.kmax - d.kmin == 39) sweep(ViMainVAy<39>{});
    else sweep(ViMainVAy<0>{});
    ViMainVAb f;
    f.d = d; f.c = ctx.c; f.G = ctx.G;
    f.dt = dts;
    f.x = nx[4];
    f.Y = ctx.dfield(D_U);
    for (int q = 0; q < 6; ++q) f.ps_b[q] = psb2[q];
    f.gT_b = gTb;
    f.Mc_b = adfield(ctx, AD_MCB); f.Mu_b = adfield(ctx, AD_MUB); f.Ml_b = adfield(ctx, AD_MLB);
    for (int q = 0; q < 3; ++q) f.hv_b[q] = hvb[q];
    x.run_pts(f, d);
  }
  {
    ViMainHA f;
    f.d = d; f.c = ctx.c; f.G = ctx.G;
    f.prog = prog;
    for (int q = 0; q < 3; ++q) { f.dv[q] = dvn[q]; f.hv_b[q] = hvb[q]; f.dv_b[q] = dvb[q]; }
    f.eth = ctx.field(W_ETH); f.ethtot = sf.ethtot;
    f.eth_b = afield(ctx, A_ETHB); f.ethtot_b = afield(ctx, A_ETHTOTB);
    f.prog_b = prog_b; f.gT_b = gTb;
    x.run_pts(f, d);
  }
  halo_adj(x, ctx, dvb, 3);
  real* ddvb[3] = {afield(ctx, A_DDVB), afield(ctx, A_DDVB + 1), afield(ctx, A_DDVB + 2)};
  {
    Path1A f;
    f.d = d; f.c = ctx.c; f.G = ctx.G;
    f.dt = (real)dts;
    f.prog = prog;
    for (int q = 0; q < 6; ++q) f.ps_b[q] = psb2[q];
    for (int q = 0; q < 3; ++q) { f.dv[q] = dvn[q]; f.dv_b[q] = dvb[q]; f.dd_b[q] = ddvb[q]; }
    f.prog_b = prog_b; f.gT_b = gTb;
    x.run_pts(f, d);
  }
  real* vb[4] = {psb2[1], psb2[2], psb2[3], psb2[4]};
  divdamp_adj(x, ctx, ddvb, vb);
}

// reverse of the large-step part of a stage (stage_large); prog is the
// recomputed boundary-conditioned state
template <class X>
void stage_large_adj(X& x, const Context& ctx, const real* prog, const real* rhogq, double dts, real* prog_b,
                     const StageFields& sf) {
  const Dims& d = ctx.d;
  const long NF = d.NF;
  real* diag = ctx.field(W_DIAG);
  real* diagb = afield(ctx, A_DIAGB);
  real* ethb = afield(ctx, A_ETHB);
  {
    EthTotA f;
    f.d = d; f.c = ctx.c; f.G = ctx.G;
    f.prog = prog;
    f.ethtot_b = afield(ctx, A_ETHTOTB); f.eth_b = ethb; f.prog_b = prog_b;
    x.run_pts(f, d);
  }
  {
    ViPath0A f;
    f.d = d; f.c = ctx.c; f.G = ctx.G;
    f.dt = dts;
    f.prog = prog; f.diag = diag; f.eth = ctx.field(W_ETH);
    for (int q = 0; q < 6; ++q) { f.vt[q] = sf.hdb + q * NF; f.vt_b[q] = afield(ctx, A_HDBB + q); }
    f.gT_b = afield(ctx, A_GTB);
    f.Mc_b = adfield(ctx, AD_MCB); f.Mu_b = adfield(ctx, AD_MUB); f.Ml_b = adfield(ctx, AD_MLB);
    f.getot_b = afield(ctx, A_GETOTB);
    f.prog_b = prog_b; f.diag_b = diagb; f.eth_b = ethb;
    f.gadv_b = afield(ctx, A_GADVB);
    for (int q = 0; q < 3; ++q) f.dd_b[q] = afield(ctx, A_DDB + q);
    // the stencil bars go through the damping transposes' triangle weights,
    // which are free until divdamp_adj below
    for (int q = 0; q < 5; ++q) f.S[q] = afield(ctx, A_TRIW + q);
    x.run_pts(f, d);
    ViPath0G h;
    h.d = d; h.G = ctx.G;
    h.prog = prog; h.eth = ctx.field(W_ETH);
    for (int q = 0; q < 5; ++q) h.S[q] = afield(ctx, A_TRIW + q);
    h.prog_b = prog_b; h.diag_b = diagb; h.eth_b = ethb;
    x.run_pts(h, d);
  }
  {
    real* ob[3] = {afield(ctx, A_DDB), afield(ctx, A_DDB + 1), afield(ctx, A_DDB + 2)};
    real* vb[4] = {prog_b + NF, prog_b + 2 * NF, prog_b + 3 * NF, prog_b + 4 * NF};
    divdamp_adj(x, ctx, ob, vb);
  }
  // the forward step fills no halo of pass 2's output (path 0 reads it at its
  // compute points only), so its bars stay zero off the compute points
  {
    // the triangles' vertex bars go through the split-state bars A_PSB ..
    // A_DDVB, which the small steps of the stage no longer need
    static_assert(A_TRIW - A_PSB >= 18, "18 scratch fields");
    HDiffKhTriA t;
    t.d = d; t.c = ctx.c; t.G = ctx.G;
    for (int q = 0; q < 2; ++q) {
      t.in[q] = sf.hda + (4 + q) * NF;
      t.ob[q] = afield(ctx, A_HDBB + 4 + q);
    }
    t.rhog = prog;
    for (int q = 0; q < 18; ++q) t.S[q] = afield(ctx, A_PSB + q);
    x.run_pts(t, d);
    HDiffKhG h;
    h.d = d; h.G = ctx.G;
    for (int q = 0; q < 18; ++q) h.S[q] = afield(ctx, A_PSB + q);
    for (int q = 0; q < 2; ++q) {
      h.ob[q] = afield(ctx, A_HDBB + 4 + q);
      h.ib[q] = afield(ctx, A_HDAB + 4 + q);
    }
    h.rhog_b = prog_b;
    x.run_pts(h, d);
    HDiffPass2A f;
    f.d = d; f.c = ctx.c; f.G = ctx.G;
    for (int q = 0; q < 6; ++q) {
      f.in[q] = sf.hda + q * NF;
      f.ob[q] = afield(ctx, A_HDBB + q);
      f.ib[q] = afield(ctx, A_HDAB + q);
    }
    f.rhog = prog;
    f.rhog_b = prog_b;
    x.run(f, f.threads());
    NegLapGA l;
    l.d = d; l.G = ctx.G; l.nq = 4;
    for (int q = 0; q < 4; ++q) {
      l.ob[q] = afield(ctx, A_HDBB + q);
      l.ib[q] = afield(ctx, A_HDAB + q);
    }
    lap_adj<4>(x, ctx, l);
  }
  halo_block_adj(x, ctx, afield(ctx, A_HDAB), 6);
  {
    NegLapGA l;
    l.d = d; l.G = ctx.G; l.nq = 6;
    for (int q = 0; q < 6; ++q) l.ob[q] = afield(ctx, A_HDAB + q);
    l.ib[0] = diagb + I_VX * NF; l.ib[1] = diagb + I_VY * NF; l.ib[2] = diagb + I_VZ * NF;
    l.ib[3] = diagb + I_W * NF; l.ib[4] = diagb + I_TEM * NF; l.ib[5] = prog_b;
    lap_adj<6>(x, ctx, l);
  }
  {
    // momentum advection transposed as gathers; the damping transposes'
    // triangle weights are dead here and hold the dvv bars and W
    real* S[3] = {afield(ctx, A_TRIW), afield(ctx, A_TRIW + 1), afield(ctx, A_TRIW + 2)};
    real* W = afield(ctx, A_TRIW + 3);
    AdvMomG1 f1;
    f1.d = d; f1.c = ctx.c; f1.G = ctx.G;
    f1.prog = prog; f1.diag = diag;
    f1.ob = afield(ctx, A_GADVB); f1.prog_b = prog_b;
    for (int q = 0; q < 3; ++q) f1.S[q] = S[q];
    x.run_pts(f1, d);
    AdvMomG2 f2;
    f2.d = d; f2.c = ctx.c; f2.G = ctx.G;
    f2.prog = prog; f2.diag = diag;
    for (int q = 0; q < 3; ++q) f2.S[q] = S[q];
    f2.prog_b = prog_b; f2.diag_b = diagb; f2.W = W;
    x.run_pts(f2, d);
    AdvMomG3 f3;
    f3.d = d; f3.G = ctx.G; f3.W = W; f3.wdb = diagb + I_W * NF;
    x.run_pts(f3, d);
  }
  {
    PrePostA f;
    f.d = d; f.c = ctx.c; f.G = ctx.G;
    f.prog = prog; f.rhogq = rhogq;
    f.prog_b = prog_b; f.diag_b = diagb; f.eth_b = ethb;
    x.run_pts(f, d);
  }
}

// stage_in (optional): the stage inputs recorded by the forward step (3 x 6
// fields, the first being prog_in); the forward sweep is then skipped
template <class X>
int dyn_step_adj(X& x, const Context& ctx, const real* prog_in, const real* progq, int nq, double dtl,
                 int sstep_max, real* prog_b, const real* stage_in = nullptr) {
  const Dims& d = ctx.d;
  const long NF = d.NF;
  if (stage_in) prog_in = stage_in;
  if (nq < 1 || !prog_in || !progq || !prog_b) return NH_ERR_ARG;
  if (!ctx.aws || ctx.adj_ck < sstep_max + 1) return NH_ERR_STATE;
  const double dts = dtl / (double)sstep_max;
  const int nsub[3] = {sstep_max / 3, sstep_max / 2, sstep_max};
  real* P = afield(ctx, A_PROG);
  real* prog0 = ctx.field(W_PROG0);
  auto ck_slot = [&](int i, real* f[6]) {
    for (int q = 0; q < 6; ++q) f[q] = afield(ctx, A_CK + kSlot * i + q);
  };
  auto slot_hv = [&](int i) { return afield(ctx, A_CK + kSlot * i + 6); };
  auto stage_input = [&](int nl) -> const real* {
    return stage_in ? stage_in + (long)nl * 6 * NF : afield(ctx, A_SC + 6 * nl);
  };
  // forward sweep: the inputs of stages 0..2 (unless recorded)
  x.d2d(prog0, prog_in, 6 * NF * sizeof(real));
  x.d2d(P, prog_in, 6 * NF * sizeof(real));
  for (int nl = 0; nl < 3 && !stage_in; ++nl) {
    x.d2d(afield(ctx, A_SC + 6 * nl), P, 6 * NF * sizeof(real));
    if (nl == 2) break;
    stage_large(x, ctx, P, progq, dts, ws_stage(ctx), false);
    SmallBufs s;
    for (int q = 0; q < 6; ++q) s.ps[q] = ctx.field(W_SPLIT + q);
    for (int q = 0; q < 3; ++q) s.dv[q] = ctx.field(W_SPLIT + 6 + q);
    for (int q = 0; q < 3; ++q) s.we[q] = ctx.field(W_SPLIT + 9 + q);
    x.run(SplitInit{NF, P, prog0, {s.ps[0], s.ps[1], s.ps[2], s.ps[3], s.ps[4], s.ps[5]}, nl == 0 ? 1 : 0}, NF);
    for (int it = 0; it < nsub[nl]; ++it) {
      small_step(x, ctx, P, s, dts, 1.0 / (double)nsub[nl], ws_stage(ctx), false);
      real* old[6] = {s.ps[0], s.ps[1], s.ps[2], s.ps[3], s.ps[4], s.ps[5]};
      s.ps[0] = s.we[0]; s.ps[1] = s.dv[0]; s.ps[2] = s.dv[1]; s.ps[3] = s.dv[2]; s.ps[4] = s.we[1];
      s.ps[5] = s.we[2];
      s.we[0] = old[0]; s.dv[0] = old[1]; s.dv[1] = old[2]; s.dv[2] = old[3]; s.we[1] = old[4]; s.we[2] = old[5];
    }
    prog_update(x, ctx, P, s.ps);
    halo_block(x, ctx, P, 6);
  }
  // reverse sweep
  halo_block_adj(x, ctx, prog_b, 6);   // the final halo fill of PROG
  zero_fields(x, afield(ctx, A_P0B), 6, NF);
  for (int nl = 2; nl >= 0; --nl) {
    if (nl < 2) halo_block_adj(x, ctx, prog_b, 6);
    // recompute the stage, keeping every small step's split state
    x.d2d(P, stage_input(nl), 6 * NF * sizeof(real));
    stage_large(x, ctx, P, progq, dts, ws_stage(ctx), false);
    real* ck[6];
    ck_slot(0, ck);
    x.run(SplitInit{NF, P, prog0, {ck[0], ck[1], ck[2], ck[3], ck[4], ck[5]}, nl == 0 ? 1 : 0}, NF);
    const double rweight = 1.0 / (double)nsub[nl];
    for (int it = 0; it < nsub[nl]; ++it) {
      real *c0[6], *c1[6];
      ck_slot(it, c0);
      ck_slot(it + 1, c1);
      SmallBufs s;
      for (int q = 0; q < 6; ++q) s.ps[q] = c0[q];
      s.we[0] = c1[0]; s.dv[0] = c1[1]; s.dv[1] = c1[2]; s.dv[2] = c1[3]; s.we[1] = c1[4]; s.we[2] = c1[5];
      s.hv = slot_hv(it);
      small_step(x, ctx, P, s, dts, rweight, ws_stage(ctx), false);
    }
    // adjoint of the stage
    // the stage bars the small steps accumulate into; A_PSB is stored whole by
    // ProgUpdateA, A_PSB2 .. A_DDVB by each small step (cleared first with
    // 8-byte fields)
    zero_fields(x, afield(ctx, A_DIAGB), A_PSB - A_DIAGB, NF);
    zero_bytes(x, adfield(ctx, AD_MCB), 3 * NF * (long)sizeof(double));
    real *psb[6], *psb2[6];
    for (int q = 0; q < 6; ++q) {
      psb[q] = afield(ctx, A_PSB + q);
      psb2[q] = afield(ctx, A_PSB2 + q);
    }
    {
      ProgUpdateA f;
      f.d = d; f.c = ctx.c; f.G = ctx.G;
      f.prog_b = prog_b;
      for (int q = 0; q < 6; ++q) f.ps_b[q] = psb[q];
      x.run_pts(f, d);
    }
    for (int it = nsub[nl] - 1; it >= 0; --it) {
      real *c0[6], *c1[6];
      ck_slot(it, c0);
      ck_slot(it + 1, c1);
      if (!ViMainPostG::kStore) zero_fields(x, psb2[0], 6, NF);
      small_step_adj(x, ctx, P, c0, c1, slot_hv(it), dts, prog_b, psb, psb2, ws_stage(ctx));
      for (int q = 0; q < 6; ++q) {
        real* t = psb[q];
        psb[q] = psb2[q];
        psb2[q] = t;
      }
    }
    x.run(SplitInitA{NF, prog_b, afield(ctx, A_P0B), {psb[0], psb[1], psb[2], psb[3], psb[4], psb[5]},
                     nl == 0 ? 1 : 0},
          NF);
    stage_large_adj(x, ctx, P, progq, dts, prog_b, ws_stage(ctx));
    int rc = x.check("dyn_step_adj stage");
    if (rc) return rc;
  }
  x.run(AxpyP{prog_b, afield(ctx, A_P0B)}, 6 * NF);
  return x.check("dyn_step_adj");
}

// the adjoint from the tape of the forward step (dyn_step with a tape): only
// the diagnostics of each stage are recomputed (Pre_Post, which also sets
// the ghost rows of the stage state), then the transposes run
template <class X>
int dyn_step_adj_tape(X& x, const Context& ctx, const void* tape, const real* progq, int nq, double dtl,
                      int sstep_max, real* prog_b) {
  const Dims& d = ctx.d;
  const long NF = d.NF;
  if (nq < 1 || !tape || !progq || !prog_b) return NH_ERR_ARG;
  if (!ctx.aws) return NH_ERR_STATE;
  const double dts = dtl / (double)sstep_max;
  const int nsub[3] = {sstep_max / 3, sstep_max / 2, sstep_max};
  const Tape tp(const_cast<void*>(tape), NF, sstep_max);
  real* P = afield(ctx, A_PROG);
  halo_block_adj(x, ctx, prog_b, 6);   // the final halo fill of PROG
  zero_fields(x, afield(ctx, A_P0B), 6, NF);
  for (int nl = 2; nl >= 0; --nl) {
    if (nl < 2) halo_block_adj(x, ctx, prog_b, 6);
    const StageFields sf = tp.stage(nl);
    x.d2d(P, tp.input(nl), 6 * NF * sizeof(real));
    {
      PrePostP f{d, ctx.c, ctx.G, P, progq, ctx.field(W_DIAG), ctx.field(W_ETH)};
      x.run_pts(f, d);
    }
    // the stage bars the small steps accumulate into; A_PSB is stored whole by
    // ProgUpdateA, A_PSB2 .. A_DDVB by each small step (cleared first with
    // 8-byte fields)
    zero_fields(x, afield(ctx, A_DIAGB), A_PSB - A_DIAGB, NF);
    zero_bytes(x, adfield(ctx, AD_MCB), 3 * NF * (long)sizeof(double));
    real *psb[6], *psb2[6];
    for (int q = 0; q < 6; ++q) {
      psb[q] = afield(ctx, A_PSB + q);
      psb2[q] = afield(ctx, A_PSB2 + q);
    }
    {
      ProgUpdateA f;
      f.d = d; f.c = ctx.c; f.G = ctx.G;
      f.prog_b = prog_b;
      for (int q = 0; q < 6; ++q) f.ps_b[q] = psb[q];
      x.run_pts(f, d);
    }
    for (int it = nsub[nl] - 1; it >= 0; --it) {
      real *c0[6], *c1[6];
      for (int q = 0; q < 6; ++q) {
        c0[q] = tp.slot(nl, it) + q * NF;
        c1[q] = tp.slot(nl, it + 1) + q * NF;
      }
      if (!ViMainPostG::kStore) zero_fields(x, psb2[0], 6, NF);
      small_step_adj(x, ctx, P, c0, c1, tp.slot(nl, it) + 6 * NF, dts, prog_b, psb, psb2, sf);
      for (int q = 0; q < 6; ++q) {
        real* t = psb[q];
        psb[q] = psb2[q];
        psb2[q] = t;
      }
    }
    x.run(SplitInitA{NF, prog_b, afield(ctx, A_P0B), {psb[0], psb[1], psb[2], psb[3], psb[4], psb[5]},
                     nl == 0 ? 1 : 0},
          NF);
    stage_large_adj(x, ctx, P, progq, dts, prog_b, sf);
    int rc = x.check("dyn_step_adj_tape stage");
    if (rc) return rc;
  }
  x.run(AxpyP{prog_b, afield(ctx, A_P0B)}, 6 * NF);
  return x.check("dyn_step_adj_tape");
}

// surface pressure per column (forward) or its transpose added into prog_b
template <class X>
int sfc_pressure(X& x, const Context& ctx, const real* prog, const real* rhogq, double* ps, const double* ps_b,
                 real* prog_b) {
  if (!prog || !rhogq || (!ps_b && !ps) || (ps_b && !prog_b)) return NH_ERR_ARG;
  if (ctx.d.kmin + 2 > ctx.d.kmax) return NH_ERR_ARG;
  SfcPresP f{ctx.d, ctx.c, ctx.G, prog, rhogq, ps, ps_b, prog_b};
  x.run(f, ncols(ctx.d));
  return x.check(ps_b ? "nh_sfc_pressure_adj" : "nh_sfc_pressure");
}

// ----------------------------------------------------------------------------
// adjoint of the offline tracer step (tracer_step) with respect to the tracer
// masses; the mass fluxes of prog are held fixed, the limiters' switches are
// those of the forward pass from progq.  progq_b: bar of the tracers after
// the step on entry, bar of progq on exit (every point, halo points
// included: the caller's halo values are inputs of the step too).
// Each tracer is independent: its transport is recomputed up to the
// flux-form update, then the transposes run in reverse order.
// ----------------------------------------------------------------------------
enum TrcAdjField {
  TA_RQ1 = 0,      // 1  tracer after the first vertical half step
  TA_RQHB = 1,     // 1  bar of the horizontal step's output
  TA_QAB = 2,      // 3  bars of the upwind face values
  TA_QOB = 5,      // 2  bars of the outflow bounds
  TA_GQB = 7,      // 3  bars of the gradient
  TA_IB = 10,      // 1  bar of the tracer before the first half step
  TA_QHS = 11,     // 5  face values of the vertical half step as linear forms
  TA_NFIELDS = 16
};

template <class X>
int tracer_step_adj(X& x, const Context& ctx, const real* prog, const real* progq, int nq, double dtl,
                    real* progq_b) {
  const Dims& d = ctx.d;
  const long NF = d.NF, nc = ncols(d);
  if (nq < 1 || !prog || !progq || !progq_b) return NH_ERR_ARG;
  if (!ctx.aws || adj_fields(ctx.adj_ck) < TA_NFIELDS) return NH_ERR_STATE;
  const real dt = (real)dtl;
  const int lim = ctx.c.THUBURN_LIM;
  real* frhog = ctx.field(W_FRHOG);
  real* rhog1 = ctx.field(W_RHOG1);
  real* rhog2 = ctx.field(W_RHOG2);
  real* flxv = ctx.field(W_FLXV);
  real* fa[3] = {ctx.field(W_FA), ctx.field(W_FA + 1), ctx.field(W_FA + 2)};
  real* xc[9];
  for (int q = 0; q < 9; ++q) xc[q] = ctx.field(W_XC + q);
  real* gq[3] = {ctx.field(W_GQ), ctx.field(W_GQ + 1), ctx.field(W_GQ + 2)};
  real* qa[3] = {ctx.field(W_QA), ctx.field(W_QA + 1), ctx.field(W_QA + 2)};
  real* qout[2] = {ctx.field(W_QOUT), ctx.field(W_QOUT + 1)};
  real* rqh = ctx.field(W_RQ);
  real* rq1 = afield(ctx, TA_RQ1);
  real* rqh_b = afield(ctx, TA_RQHB);
  real* qa_b[3] = {afield(ctx, TA_QAB), afield(ctx, TA_QAB + 1), afield(ctx, TA_QAB + 2)};
  real* qo_b[2] = {afield(ctx, TA_QOB), afield(ctx, TA_QOB + 1)};
  real* gq_b[3] = {afield(ctx, TA_GQB), afield(ctx, TA_GQB + 1), afield(ctx, TA_GQB + 2)};
  real* ib = afield(ctx, TA_IB);
  real* qhs = afield(ctx, TA_QHS);
  x.run(FillZero{frhog}, NF);
  {
    TrcHFlux f;
    f.d = d; f.c = ctx.c; f.G = ctx.G;
    f.dt = dt; f.pm = prog;
    for (int q = 0; q < 3; ++q) f.fa[q] = fa[q];
    for (int q = 0; q < 9; ++q) f.xc[q] = xc[q];
    x.run_pts(f, d);
  }
  halo_block_adj(x, ctx, progq_b, nq);   // the final halo fill of progq
  for (int iq = 0; iq < nq; ++iq) {
    const real* q0 = progq + (long)iq * NF;
    real* rq_b = progq_b + (long)iq * NF;
    // forward: first half step (also the vertical flux and rhog1), then the
    // horizontal step up to its output rqh
    x.d2d(rq1, q0, NF * sizeof(real));
    {
      TrcVert1 f;
      f.d = d; f.c = ctx.c; f.G = ctx.G;
      f.dt = dt; f.rw_mean = prog + 4 * NF; f.rhog_in = prog; f.frhog = frhog;
      f.flxv = flxv; f.rhog1 = rhog1; f.rhogq = rq1; f.nq = 1; f.limiter = lim;
      x.run(f, nc);
    }
    if (iq == 0) {
      TrcRhog2 f;
      f.d = d; f.c = ctx.c; f.G = ctx.G;
      f.dt = dt;
      for (int q = 0; q < 3; ++q) f.fa[q] = fa[q];
      f.rhog1 = rhog1; f.frhog = frhog; f.rhog2 = rhog2;
      x.run_pts(f, d);
    }
    {
      TrcGradQ f;
      f.d = d; f.G = ctx.G; f.rq = rq1; f.rhog1 = rhog1;
      for (int q = 0; q < 3; ++q) f.gq[q] = gq[q];
      x.run_pts(f, d);
    }
    halo(x, ctx, gq, 3);
    {
      TrcRemap f;
      f.d = d; f.c = ctx.c; f.G = ctx.G;
      f.dt = dt; f.rq = rq1; f.rhog1 = rhog1;
      for (int q = 0; q < 3; ++q) { f.gq[q] = gq[q]; f.fa[q] = fa[q]; f.qa[q] = qa[q]; }
      for (int q = 0; q < 9; ++q) f.xc[q] = xc[q];
      x.run_pts(f, d);
    }
    if (lim) {
      TrcQout f;
      f.d = d; f.c = ctx.c; f.G = ctx.G;
      f.dt = dt; f.rq = rq1; f.rhog1 = rhog1; f.frhog = frhog;
      for (int q = 0; q < 3; ++q) f.fa[q] = fa[q];
      f.qout[0] = qout[0]; f.qout[1] = qout[1];
      x.run_pts(f, d);
      halo(x, ctx, qout, 2);
    }
    {
      TrcApply f;
      f.d = d; f.c = ctx.c; f.G = ctx.G;
      f.dt = dt; f.rq = rq1; f.rhog1 = rhog1;
      for (int q = 0; q < 3; ++q) { f.fa[q] = fa[q]; f.qa[q] = qa[q]; }
      f.qout[0] = qout[0]; f.qout[1] = qout[1];
      f.out = rqh; f.limiter = lim;
      for (int q = 0; q < 3; ++q) f.gq[q] = gq[q];
      for (int q = 0; q < 9; ++q) f.xc[q] = xc[q];
      x.run_pts(f, d);
    }
    // reverse; rq_b becomes the bar of rq1 once the compute cells' bars have
    // gone through the second half step (elsewhere rq1 passes unchanged)
    zero_fields(x, rqh_b, TA_QHS - TA_RQHB, NF);
    {
      TrcVert2A f;
      f.d = d; f.c = ctx.c; f.G = ctx.G;
      f.dt = dt; f.flxv = flxv; f.rhog2 = rhog2; f.frhog = frhog;
      f.in = rqh; f.ob = rq_b; f.ib = rqh_b; f.qhs = qhs; f.limiter = lim;
      x.run(f, nc);
    }
    TgSink sk;
    sk.rhog1 = rhog1; sk.rq_b = rq_b;
    for (int q = 0; q < 3; ++q) sk.qa_b[q] = qa_b[q];
    for (int q = 0; q < 2; ++q) sk.qo_b[q] = qo_b[q];
    {
      TrcApplyA f;
      f.d = d; f.c = ctx.c; f.G = ctx.G;
      f.dt = dt; f.rq = rq1; f.rhog1 = rhog1;
      for (int q = 0; q < 3; ++q) { f.fa[q] = fa[q]; f.qa[q] = qa[q]; }
      f.qout[0] = qout[0]; f.qout[1] = qout[1];
      f.ob = rqh_b; f.sk = sk; f.limiter = lim;
      for (int q = 0; q < 3; ++q) f.gq_b[q] = gq_b[q];
      for (int q = 0; q < 9; ++q) f.xc[q] = xc[q];
      x.run_pts(f, d);
    }
    if (lim) {
      halo_adj(x, ctx, qo_b, 2);
      TrcQoutA f;
      f.d = d; f.c = ctx.c; f.G = ctx.G;
      f.dt = dt; f.rq = rq1; f.rhog1 = rhog1; f.frhog = frhog;
      for (int q = 0; q < 3; ++q) f.fa[q] = fa[q];
      f.qo_b[0] = qo_b[0]; f.qo_b[1] = qo_b[1];
      f.sk = sk;
      x.run_pts(f, d);
    }
    {
      TrcRemapA f;
      f.d = d; f.c = ctx.c; f.G = ctx.G;
      f.dt = dt; f.rhog1 = rhog1;
      for (int q = 0; q < 3; ++q) { f.fa[q] = fa[q]; f.qa_b[q] = qa_b[q]; f.gq_b[q] = gq_b[q]; }
      for (int q = 0; q < 9; ++q) f.xc[q] = xc[q];
      f.rq_b = rq_b;
      x.run_pts(f, d);
    }
    halo_adj(x, ctx, gq_b, 3);
    {
      TrcGradQA f;
      f.d = d; f.G = ctx.G; f.rhog1 = rhog1;
      for (int q = 0; q < 3; ++q) f.gq_b[q] = gq_b[q];
      f.rq_b = rq_b;
      x.run_pts(f, d);
    }
    {
      TrcVert1A f;
      f.d = d; f.c = ctx.c; f.G = ctx.G;
      f.dt = dt; f.rw_mean = prog + 4 * NF; f.rhog_in = prog; f.frhog = frhog;
      f.in = q0; f.ob = rq_b; f.ib = ib; f.qhs = qhs; f.limiter = lim;
      x.run(f, nc);
    }
    x.d2d(rq_b, ib, NF * sizeof(real));
    int rc = x.check("tracer_step_adj");
    if (rc) return rc;
  }
  return 0;
}

NH_END
