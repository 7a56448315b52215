// Transpose of the tracer transport with respect to the tracer masses, with
// the mass fluxes and densities held fixed and the limiter switches frozen
// from the forward pass (the offline tracer step, TIME_integ_type 'TRCADV').
//
// With the fluxes fixed every kernel of the transport is piecewise linear in
// rhogq: the Thuburn limiters' minima, maxima and clips select one of their
// inputs, the upwind masks are exactly 0 or 1.  Each transposed kernel re-runs
// the forward arithmetic on "tagged" values that carry which input they came
// from (a selection keeps the winner's tag), then sends the output bar to the
// selected inputs.  The vertical limiter's face values are linear forms over
// the four levels k-2 .. k+1 of their column.
#pragma once

#include "nh_tracer.h"
#include "nh_adjoint.h"

NH_BEGIN

// ----------------------------------------------------------------------------
// tagged values: v and where it came from
// ----------------------------------------------------------------------------
enum { TG_CONST = 0, TG_Q = 1, TG_QA = 2, TG_QO = 5 };   // TG_QA + arc, TG_QO + 0/1
struct Tg {
  real v;
  int kind;
  long addr;
};
NH_HD inline Tg tg_c(real v) { return {v, TG_CONST, 0}; }
NH_HD inline Tg tmin(const Tg& a, const Tg& b) { return (a.v < b.v) ? a : b; }   // as dmin
NH_HD inline Tg tmax(const Tg& a, const Tg& b) { return (a.v > b.v) ? a : b; }   // as dmax
NH_HD inline Tg tclip(const Tg& x, const Tg& lo, const Tg& hi) { return tmin(tmax(x, lo), hi); }
// cm * x + (1 - cm) * y with cm exactly 0 or 1
NH_HD inline Tg tsel(real cm, const Tg& x, const Tg& y) { return (cm == real(1.0)) ? x : y; }

// the bar w of a tagged value onto its source
struct TgSink {
  const real* rhog1;
  real* rq_b;
  real* qa_b[3];
  real* qo_b[2];
  NH_HD void add(const Tg& t, real w) const {
    if (t.kind == TG_Q) aadd(&rq_b[t.addr], w / rhog1[t.addr]);
    else if (t.kind >= TG_QA && t.kind < TG_QO) aadd(&qa_b[t.kind - TG_QA][t.addr], w);
    else if (t.kind >= TG_QO) aadd(&qo_b[t.kind - TG_QO][t.addr], w);
  }
};

// the limiter's face bounds with tags (face_bounds / qin_of_cell / qin_of_owner)
struct FaceBoundsT {
  Tg mn[6], mx[6];
  real cm[6];
  bool pent;
};

NH_HD inline void face_bounds_t(const Tg q[7], const real cmo[6], bool pentagon, FaceBoundsT& fb) {
  const int sets[6][4] = {{0, 6, 1, 2}, {0, 1, 2, 3}, {0, 2, 3, 4}, {4, 5, 0, 3}, {5, 6, 0, 4}, {6, 1, 0, 5}};
  for (int m = 0; m < 6; ++m) {
    const int* s = sets[m];
    int s1 = s[1], s2 = s[2];
    if (m == 4 && pentagon) { s1 = 0; s2 = 1; }
    const Tg a = q[s[0]], b = q[s1], c = q[s2], e = q[s[3]];
    fb.mn[m] = tmin(tmin(tmin(a, b), c), e);
    fb.mx[m] = tmax(tmax(tmax(a, b), c), e);
    fb.cm[m] = cmo[m];
  }
  fb.pent = pentagon;
}

NH_HD inline void qin_of_cell_t(const FaceBoundsT& fb, int m, Tg& lo, Tg& hi) {
  const real c = fb.cm[m];
  if (m < 3) {
    lo = tsel(c, fb.mn[m], tg_c(NH_BIG));
    hi = tsel(c, fb.mx[m], tg_c(-NH_BIG));
  } else if (m == 4 && fb.pent) {
    lo = (c == real(1.0)) ? tg_c(NH_BIG) : fb.mn[m];
    hi = (c == real(1.0)) ? tg_c(-NH_BIG) : fb.mx[m];
  } else {
    lo = tsel(c, tg_c(NH_BIG), fb.mn[m]);
    hi = tsel(c, tg_c(-NH_BIG), fb.mx[m]);
  }
}

NH_HD inline void qin_of_owner_t(const FaceBoundsT& fb, int m, Tg& lo, Tg& hi) {
  const real c = fb.cm[m];
  if (m == 4 && fb.pent) {
    lo = (c == real(1.0)) ? fb.mn[m] : tg_c(NH_BIG);
    hi = (c == real(1.0)) ? fb.mx[m] : tg_c(-NH_BIG);
  } else {
    lo = tsel(c, fb.mn[m], tg_c(NH_BIG));
    hi = tsel(c, fb.mx[m], tg_c(-NH_BIG));
  }
}

// ----------------------------------------------------------------------------
// flux-form update with the limited face values (TrcApply); ob: bar of its
// output (consumed)
// ----------------------------------------------------------------------------
struct TrcApplyA {
  static constexpr const char* kName = "trc_apply_adj";
  Dims d; Consts c; Geo G;
  real dt;
  const real* rq;
  const real* rhog1;
  const real* fa[3];
  const real* qa[3];
  const real* qout[2];
  real* ob;
  TgSink sk;
  int limiter;
  real* gq_b[3];       // limiter off: bars of the gradient (faces 3..5)
  const real* xc[9];
  // per level on the point grid (levels independent; the adds to other
  // points are atomic), or every level of a column
  NH_HD void operator()(long t) const { levels(cell_of(d, t), 0, d.kall); }
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (pt_of(d, plane, x, p)) levels(cell_of_pt(p), p.k, p.k + 1);
  }
  NH_HD void levels(const Cell& cl, int k0, int k1) const {
    if (cl.pole ? cl.v != 0 : !cl.inner) return;
    const int g = d.g;
    for (int k = k0; k < k1; ++k) {
      const long a = cl.base + k * cl.ks;
      const real w = ob[a];
      ob[a] = real(0.0);
      if (w == real(0.0)) continue;
      aadd(&sk.rq_b[a], w);
      if (!limiter) {
        if (cl.pole) {
          for (int v = 1; v <= 5; ++v) aadd(&sk.qa_b[0][a + v], -fa[0][a + v] * G.rarea_pl[cl.lp] * dt * w);
        } else {
          real f[6];
          for (int m = 0; m < 6; ++m) f[m] = face_flux(G, d, fa, a, cl.ci, cl.i, cl.j, cl.l, m, dt);
          for (int m = 0; m < 3; ++m) aadd(&sk.qa_b[m][a], -f[m] * w);
          // faces 3..5: cm * (extrapolation from here) + (1 - cm) * (from the owner)
          for (int m = 3; m < 6; ++m) {
            const int arc = m - 3;
            const long o = (m == 3) ? a - 1 : (m == 4) ? a - 1 - g : a - g;
            const long co = (m == 3) ? cl.ci - 1 : (m == 4) ? cl.ci - 1 - g : cl.ci - g;
            const real cm = upwind_mask(f[m] / rhog1[a]);
            const long src = (cm == real(1.0)) ? a : o;
            const long cs = (cm == real(1.0)) ? cl.ci : co;
            const real wb = -f[m] * w;
            aadd(&sk.rq_b[src], wb / rhog1[src]);
            for (int dd = 0; dd < 3; ++dd) aadd(&gq_b[dd][src], wb * (xc[arc * 3 + dd][o] - rc(G.gx, d.nreg, dd, cs)));
          }
        }
        continue;
      }
      auto qt = [&](long x) { return Tg{rq[x] / rhog1[x], TG_Q, x}; };
      auto qot = [&](int s, long x) { return Tg{qout[s][x], TG_QO + s, x}; };
      if (cl.pole) {
        const real ra = G.rarea_pl[cl.lp];
        const Tg q0 = qt(a);
        Tg qv[6];
        for (int v = 1; v <= 5; ++v) qv[v] = qt(a + v);
        for (int v = 1; v <= 5; ++v) {
          const int vp = (v == 5) ? 1 : v + 1, vm = (v == 1) ? 5 : v - 1;
          const Tg mn = tmin(tmin(tmin(q0, qv[v]), qv[vm]), qv[vp]);
          const Tg mx = tmax(tmax(tmax(q0, qv[v]), qv[vm]), qv[vp]);
          const real fl = fa[0][a + v] * ra * dt;
          const real cm = upwind_mask(fl / rhog1[a]);
          const Tg lo0 = (cm == real(1.0)) ? mn : tg_c(NH_BIG), hi0 = (cm == real(1.0)) ? mx : tg_c(-NH_BIG);
          const Tg lo1 = (cm == real(1.0)) ? tg_c(NH_BIG) : mn, hi1 = (cm == real(1.0)) ? tg_c(-NH_BIG) : mx;
          Tg s{qa[0][a + v], TG_QA + 0, a + v};
          s = tsel(cm, tclip(s, lo0, hi0), tclip(s, lo1, hi1));
          const Tg s2 = tmax(tmin(s, qot(1, a + v)), qot(0, a + v));
          const Tg s3 = tmax(tmin(s, qot(1, a)), qot(0, a));
          s = tsel(cm, s2, s3);
          sk.add(s, -fl * w);
        }
        continue;
      }
      TrcCellCtx x;
      trc_cell(G, d, cl, a, dt, rq, rhog1, fa, x);
      long off[7];
      stencil_offsets(g, off);
      Tg q[7];
      for (int m = 0; m < 7; ++m) q[m] = qt(a + off[m]);
      FaceBoundsT fb;
      face_bounds_t(q, x.cmo, cl.i == 1 && cl.j == 1 && G.sgp[cl.l], fb);
      const long own[6] = {a, a, a, a - 1, a - 1 - g, a - g};
      const long dst[6] = {a + 1, a + 1 + g, a + g, a, a, a};
      const int arc[6] = {0, 1, 2, 0, 1, 2};
      for (int m = 0; m < 6; ++m) {
        const real cm = fb.cm[m];
        Tg loA, hiA, loB, hiB;
        if (m < 3) {
          qin_of_cell_t(fb, m, loA, hiA);
          loB = tsel(cm, tg_c(NH_BIG), fb.mn[m]);
          hiB = tsel(cm, tg_c(-NH_BIG), fb.mx[m]);
        } else {
          qin_of_owner_t(fb, m, loA, hiA);
          qin_of_cell_t(fb, m, loB, hiB);
        }
        const Tg s{qa[arc[m]][own[m]], TG_QA + arc[m], own[m]};
        const Tg v1 = tsel(cm, tclip(s, loA, hiA), tclip(s, loB, hiB));
        const long o = own[m], dn = dst[m];
        const Tg sf = tsel(cm, tmax(tmin(v1, qot(1, dn)), qot(0, dn)), tmax(tmin(v1, qot(1, o)), qot(0, o)));
        sk.add(sf, -x.f[m] * w);
      }
    }
  }
};

// ----------------------------------------------------------------------------
// outflow bounds of the limiter (TrcQout); qo_b: bars of qout (consumed)
// ----------------------------------------------------------------------------
struct TrcQoutA {
  static constexpr const char* kName = "trc_qout_adj";
  Dims d; Consts c; Geo G;
  real dt;
  const real* rq;
  const real* rhog1;
  const real* frhog;
  const real* fa[3];
  real* qo_b[2];
  TgSink sk;
  // per level on the point grid (levels independent; the adds to other
  // points are atomic), or every level of a column
  NH_HD void operator()(long t) const { levels(cell_of(d, t), 0, d.kall); }
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (pt_of(d, plane, x, p)) levels(cell_of_pt(p), p.k, p.k + 1);
  }
  NH_HD void levels(const Cell& cl, int k0, int k1) const {
    if (cl.pole ? cl.v != 0 : !cl.inner) return;
    for (int k = k0; k < k1; ++k) {
      const long a = cl.base + k * cl.ks;
      const real b0 = qo_b[0][a], b1 = qo_b[1][a];
      qo_b[0][a] = real(0.0);
      qo_b[1][a] = real(0.0);
      if (b0 == real(0.0) && b1 == real(0.0)) continue;
      const real dk = real(1.0) * frhog[a] / rhog1[a] * dt;
      auto qt = [&](long x) { return Tg{rq[x] / rhog1[x], TG_Q, x}; };
      const Tg q0 = qt(a);
      Tg nmin = q0, nmax = q0;
      Tg lo[6], hi[6];
      real chm[6];
      real Cin, Cout;
      int nf = 0;
      if (cl.pole) {
        const real ra = G.rarea_pl[cl.lp];
        Tg qv[6];
        for (int v = 1; v <= 5; ++v) qv[v] = qt(a + v);
        Cin = real(0.0); Cout = real(0.0);
        for (int v = 1; v <= 5; ++v) {
          const int vp = (v == 5) ? 1 : v + 1, vm = (v == 1) ? 5 : v - 1;
          const Tg mn = tmin(tmin(tmin(q0, qv[v]), qv[vm]), qv[vp]);
          const Tg mx = tmax(tmax(tmax(q0, qv[v]), qv[vm]), qv[vp]);
          const real ch = fa[0][a + v] * ra * dt / rhog1[a];
          const real cm = upwind_mask(ch);
          lo[nf] = (cm == real(1.0)) ? mn : tg_c(NH_BIG);
          hi[nf] = (cm == real(1.0)) ? mx : tg_c(-NH_BIG);
          nmin = tmin(nmin, lo[nf]);
          nmax = tmax(nmax, hi[nf]);
          chm[nf] = cm * ch;
          Cin = Cin + chm[nf];
          Cout = Cout + (ch - chm[nf]);
          ++nf;
        }
      } else {
        TrcCellCtx x;
        trc_cell(G, d, cl, a, dt, rq, rhog1, fa, x);
        long off[7];
        stencil_offsets(d.g, off);
        Tg q[7];
        for (int m = 0; m < 7; ++m) q[m] = qt(a + off[m]);
        FaceBoundsT fb;
        face_bounds_t(q, x.cmo, cl.i == 1 && cl.j == 1 && G.sgp[cl.l], fb);
        for (int m = 0; m < 6; ++m) {
          qin_of_cell_t(fb, m, lo[m], hi[m]);
          nmin = tmin(nmin, lo[m]);
          nmax = tmax(nmax, hi[m]);
        }
        Cin = -real(0.0); Cout = -real(0.0);
        for (int m = 0; m < 6; ++m) {
          chm[m] = (x.ch[m] < real(0.0)) ? x.ch[m] : real(0.0);
          Cin = Cin + chm[m];
          Cout = Cout + (x.ch[m] - chm[m]);
        }
        nf = 6;
      }
      const real zsw = real(0.5) - nsign(real(0.5), nabs(Cout) - NH_EPS);
      const real den = Cout + zsw, fac = real(1.0) - zsw;
      const real X = real(1.0) - Cin - Cout + dk;
      // qout0 = (q0 - sum chm hi - nmax X) / den fac + q0 zsw; qout1 with lo, nmin
      sk.add(q0, (b0 + b1) * (fac / den + zsw));
      sk.add(nmax, -b0 * X * fac / den);
      sk.add(nmin, -b1 * X * fac / den);
      for (int m = 0; m < nf; ++m) {
        if (chm[m] == real(0.0)) continue;
        sk.add(hi[m], -b0 * chm[m] * fac / den);
        sk.add(lo[m], -b1 * chm[m] * fac / den);
      }
    }
  }
};

// ----------------------------------------------------------------------------
// upwind face values (TrcRemap); qa_b consumed
// ----------------------------------------------------------------------------
struct TrcRemapA {
  static constexpr const char* kName = "trc_remap_adj";
  Dims d; Consts c; Geo G;
  real dt;
  const real* rhog1;
  const real* fa[3];
  const real* xc[9];
  real* qa_b[3];
  real* rq_b;
  real* gq_b[3];
  // per level on the point grid (levels independent; the adds to other
  // points are atomic), or every level of a column
  NH_HD void operator()(long t) const { levels(cell_of(d, t), 0, d.kall); }
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (pt_of(d, plane, x, p)) levels(cell_of_pt(p), p.k, p.k + 1);
  }
  NH_HD void levels(const Cell& cl, int k0, int k1) const {
    if (cl.pole) {
      if (cl.v == 0) return;
      const int lp = cl.lp, v = cl.v;
      const long R0 = d.R + (long)lp * d.kall * 6;
      const real* xp0 = G.gx_pl + (0 * 2 + lp) * 3;
      const real* xpv = G.gx_pl + (v * 2 + lp) * 3;
      for (int k = k0; k < k1; ++k) {
        const long a0 = R0 + k * 6, av = a0 + v;
        const real w = qa_b[0][av];
        qa_b[0][av] = real(0.0);
        const real ch = fa[0][av] * G.rarea_pl[lp] * dt / rhog1[a0];
        const real cm = upwind_mask(ch);
        const long src = (cm == real(1.0)) ? av : a0;
        const real* xs = (cm == real(1.0)) ? xpv : xp0;
        aadd(&rq_b[src], w / rhog1[src]);
        for (int dd = 0; dd < 3; ++dd) aadd(&gq_b[dd][src], w * (xc[dd][av] - xs[dd]));
      }
      return;
    }
    const int g = d.g, i = cl.i, j = cl.j;
    if (i > g - 2 || j > g - 2) return;
    const long nreg = d.nreg, ci = cl.ci, ks = cl.ks;
    const long noff[3] = {1, 1 + (long)g, (long)g};
    const bool valid[3] = {j >= 1, true, i >= 1};
    for (int k = k0; k < k1; ++k) {
      const long a = cl.base + k * ks;
      for (int arc = 0; arc < 3; ++arc) {
        const real w = qa_b[arc][a];
        qa_b[arc][a] = real(0.0);
        if (!valid[arc] || w == real(0.0)) continue;
        const long an = a + noff[arc], cn = ci + noff[arc];
        const real ch = fa[arc][a] * rc(G.rarea, nreg, 0, ci) * dt / rhog1[a];
        const real cm = upwind_mask(ch);
        const long src = (cm == real(1.0)) ? an : a;
        const long cs = (cm == real(1.0)) ? cn : ci;
        aadd(&rq_b[src], w / rhog1[src]);
        for (int dd = 0; dd < 3; ++dd)
          aadd(&gq_b[dd][src], w * (xc[arc * 3 + dd][a] - rc(G.gx, nreg, dd, cs)));
      }
    }
  }
};

// ----------------------------------------------------------------------------
// gradient of q = rhogq / rhog1 (TrcGradQ); gq_b consumed
// ----------------------------------------------------------------------------
struct TrcGradQA {
  static constexpr const char* kName = "trc_gradq_adj";
  Dims d; Geo G;
  const real* rhog1;
  real* gq_b[3];
  real* rq_b;
  // per level on the point grid (levels independent; the adds to other
  // points are atomic), or every level of a column
  NH_HD void operator()(long t) const { levels(cell_of(d, t), 0, d.kall); }
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (pt_of(d, plane, x, p)) levels(cell_of_pt(p), p.k, p.k + 1);
  }
  NH_HD void levels(const Cell& cl, int k0, int k1) const {
    if (cl.pole ? cl.v != 0 : !cl.inner) return;
    long nb[7];
    const int nn = neighbours(d, cl, nb);
    for (int k = k0; k < k1; ++k) {
      const long kk = k * cl.ks;
      real gb[3];
      for (int dd = 0; dd < 3; ++dd) {
        gb[dd] = gq_b[dd][cl.base + kk];
        gq_b[dd][cl.base + kk] = real(0.0);
      }
      for (int m = 0; m < nn; ++m) {
        real s = real(0.0);
        for (int dd = 0; dd < 3; ++dd)
          s += (cl.pole ? G.cgrad_pl[(m * 2 + cl.lp) * 3 + dd] : rc(G.cgrad, d.nreg, dd * 7 + m, cl.ci)) * gb[dd];
        const long x = nb[m] + kk;
        aadd(&rq_b[x], s / rhog1[x]);
      }
    }
  }
};

// ----------------------------------------------------------------------------
// vertical half step (vert_sweep) transposed, one thread per column.  The
// face value qh of level k is a linear form over q at levels k-2 .. k+1
// (LF, coefficient 0 = level k-2); the forward sweep is re-run on LFs.
// ----------------------------------------------------------------------------
struct LF {
  real v;
  real c[4];
};
NH_HD inline LF lf_unit(real v, int i) {
  LF r{v, {real(0.0), real(0.0), real(0.0), real(0.0)}};
  r.c[i] = real(1.0);
  return r;
}
NH_HD inline LF lf_const(real v) { return LF{v, {real(0.0), real(0.0), real(0.0), real(0.0)}}; }
NH_HD inline LF lf_min(const LF& a, const LF& b) { return (a.v < b.v) ? a : b; }
NH_HD inline LF lf_max(const LF& a, const LF& b) { return (a.v > b.v) ? a : b; }
NH_HD inline LF lf_clip(const LF& x, const LF& lo, const LF& hi) { return lf_min(lf_max(x, lo), hi); }
// one level down the frame (the form of level k-1 seen from level k)
NH_HD inline LF lf_shift(const LF& a) { return LF{a.v, {a.c[1], a.c[2], a.c[3], real(0.0)}}; }
// the form of w0 a0 + w1 a1 + w2 a2 + w3 a3 with the value v of the forward expression
NH_HD inline LF lf_comb(real v, real w0, const LF& a0, real w1, const LF& a1, real w2, const LF& a2, real w3,
                        const LF& a3) {
  LF r = lf_const(v);
  for (int i = 0; i < 4; ++i) r.c[i] = ((w0 * a0.c[i] + w1 * a1.c[i]) + w2 * a2.c[i]) + w3 * a3.c[i];
  return r;
}

// column transpose of vert_sweep: ob (bars of out, consumed) -> ib (bars of
// in, added); qhs: 5 fields of column scratch (value + 4 coefficients of each
// level's face value)
template <class DEN, class FLX, class DV>
NH_HD inline void vert_sweep_t(const Geo& G, int kmin, int kmax, const real* in, real* ob, real* ib, real* qhs,
                               long NF, long base, long ks, DEN den, FLX flx, DV dv, bool limiter, bool clipneg) {
  auto ck0 = [&](int k) -> real { return (k >= kmin && k <= kmax) ? -flx(k) / den(k) * G.rdgz[k] : real(0.0); };
  auto at = [&](int k) { return base + (long)k * ks; };
  // forward on linear forms (frame of level k: c[0..3] = levels k-2 .. k+1)
  real q_km1 = in[at(kmin - 1)] / den(kmin - 1);
  real q_k = in[at(kmin)] / den(kmin);
  LF omin_prev = lf_const(real(0.0)), omax_prev = lf_const(real(0.0));
  real c0_k = ck0(kmin);
  for (int k = kmin; k <= kmax; ++k) {
    const real q_kp1 = in[at(k + 1)] / den(k + 1);
    const real c0_kp1 = ck0(k + 1);
    const real c1_k = flx(k + 1) / den(k) * G.rdgz[k];
    const LF Qm = lf_unit(q_km1, 1), Qk = lf_unit(q_k, 2), Qp = lf_unit(q_kp1, 3);
    const LF zero = lf_const(real(0.0));
    LF qh = lf_comb(G.afact[k] * q_k + G.bfact[k] * q_km1, G.afact[k], Qk, G.bfact[k], Qm, real(0.0), zero,
                    real(0.0), zero);
    LF omin = lf_const(real(0.0)), omax = lf_const(real(0.0));
    if (limiter) {
      const real inL = real(0.5) - nsign(real(0.5), c0_k);
      const real inU = real(0.5) + nsign(real(0.5), c0_kp1);
      const LF big = lf_const(NH_BIG), nbig = lf_const(-NH_BIG);
      const LF minL = (inL == real(1.0)) ? lf_min(Qk, Qm) : big;
      const LF minU = (inU == real(1.0)) ? lf_min(Qk, Qp) : big;
      const LF maxL = (inL == real(1.0)) ? lf_max(Qk, Qm) : nbig;
      const LF maxU = (inU == real(1.0)) ? lf_max(Qk, Qp) : nbig;
      const LF nmin = lf_min(lf_min(minL, minU), Qk);
      const LF nmax = lf_max(lf_max(maxL, maxU), Qk);
      const real Cin = inL * c0_k + inU * c1_k;
      const real Cout = (real(1.0) - inL) * c0_k + (real(1.0) - inU) * c1_k;
      const real zsw = real(0.5) - nsign(real(0.5), nabs(Cout) - NH_EPS);
      const real dk = dv(k);
      const real den2 = Cout + zsw, fac = real(1.0) - zsw, X = Cin + Cout - dk;
      // omin = ((q_k - nmax) + nmax X - CQmax) / den2 fac + q_k zsw, CQmax = inL c0 maxL + inU c1 maxU
      const real CQmax = inL * c0_k * maxL.v + inU * c1_k * maxU.v;
      const real CQmin = inL * c0_k * minL.v + inU * c1_k * minU.v;
      const real vmin = ((q_k - nmax.v) + nmax.v * X - CQmax) / den2 * fac + q_k * zsw;
      const real vmax = ((q_k - nmin.v) + nmin.v * X - CQmin) / den2 * fac + q_k * zsw;
      const real wq = fac / den2 + zsw, wn = (X - real(1.0)) * fac / den2;
      const real wl = (inL == real(1.0)) ? -c0_k * fac / den2 : real(0.0);
      const real wu = (inU == real(1.0)) ? -c1_k * fac / den2 : real(0.0);
      omin = lf_comb(vmin, wq, Qk, wn, nmax, wl, maxL, wu, maxU);
      omax = lf_comb(vmax, wq, Qk, wn, nmin, wl, minL, wu, minU);
      if (k > kmin) {
        const LF op = lf_shift(omin_prev), xp = lf_shift(omax_prev);
        qh = (inL == real(1.0)) ? lf_clip(qh, op, xp) : lf_clip(qh, omin, omax);
      }
    }
    if (k == kmin) qh = lf_const(real(0.0));
    const long a = at(k);
    qhs[a] = qh.v;
    for (int i = 0; i < 4; ++i) qhs[(long)(1 + i) * NF + a] = qh.c[i];
    omin_prev = omin;
    omax_prev = omax;
    q_km1 = q_k;
    q_k = q_kp1;
    c0_k = c0_kp1;
  }
  // the outputs: out(k) = r(k) - (flx(k+1) qh(k+1) - flx(k) qh(k)) rdgz(k), qh(kmax+1) = 0;
  // a clipped small negative output has zero derivative; ghost levels are constants
  auto eff = [&](int k) -> real {
    if (k < kmin || k > kmax) return real(0.0);
    real w = ob[at(k)];
    if (clipneg) {
      const real qk1 = (k + 1 <= kmax) ? qhs[at(k + 1)] : real(0.0);
      const real o = in[at(k)] - (flx(k + 1) * qk1 - flx(k) * qhs[at(k)]) * G.rdgz[k];
      if (o > -real(1.0e-10) && o < real(0.0)) w = real(0.0);
    }
    return w;
  };
  real qbar[NH_KMAX];
  for (int k = 0; k < NH_KMAX; ++k) qbar[k] = real(0.0);
  real e_k = eff(kmin);
  for (int k = kmin; k <= kmax; ++k) {
    const real e_kp1 = eff(k + 1);
    ib[at(k)] += e_k;
    // qh(k) enters out(k-1) with -flx(k) rdgz(k-1) and out(k) with +flx(k) rdgz(k)
    real hb = e_k * flx(k) * G.rdgz[k];
    if (k > kmin) hb -= eff(k - 1) * flx(k) * G.rdgz[k - 1];
    if (hb != real(0.0))
      for (int i = 0; i < 4; ++i) {
        const int lv = k - 2 + i;
        if (lv >= 0 && lv < NH_KMAX) qbar[lv] += hb * qhs[(long)(1 + i) * NF + at(k)];
      }
    e_k = e_kp1;
  }
  for (int k = kmin; k <= kmax; ++k) ob[at(k)] = real(0.0);
  ob[at(kmin - 1)] = real(0.0);
  ob[at(kmax + 1)] = real(0.0);
  for (int k = kmin - 1; k <= kmax + 1; ++k)
    if (qbar[k] != real(0.0)) ib[at(k)] += qbar[k] / den(k);
}

struct TrcVert1A {
  static constexpr const char* kName = "trc_vert1_adj";
  Dims d; Consts c; Geo G;
  real dt;
  const real* rw_mean;
  const real* rhog_in;
  const real* frhog;
  const real* in;     // the tracer before the half step
  real* ob;           // bars after (consumed)
  real* ib;           // bars before (added)
  real* qhs;          // 5 fields of scratch
  int limiter;
  NH_HD void operator()(long t) const {
    Cell cl = cell_of(d, t);
    const long ks = cl.ks, b = cl.base;
    const int kmin = d.kmin, kmax = d.kmax;
    auto flx = [&](int k) -> real {
      return (k >= kmin + 1 && k <= kmax) ? (real(0.0) + rw_mean[b + k * ks] * real(1.0)) * real(0.5) * dt : real(0.0);
    };
    auto den = [&](int k) { return rhog_in[b + k * ks]; };
    auto dv = [&](int k) { return real(0.0) * frhog[b + k * ks] / rhog_in[b + k * ks] * dt; };
    vert_sweep_t(G, kmin, kmax, in, ob, ib, qhs, d.NF, b, ks, den, flx, dv, limiter != 0, false);
  }
};

struct TrcVert2A {
  static constexpr const char* kName = "trc_vert2_adj";
  Dims d; Consts c; Geo G;
  real dt;
  const real* flxv;
  const real* rhog2;
  const real* frhog;
  const real* in;
  real* ob;
  real* ib;
  real* qhs;
  int limiter;
  NH_HD void operator()(long t) const {
    Cell cl = cell_of(d, t);
    if (cl.pole ? cl.v != 0 : !cl.inner) return;
    const long ks = cl.ks, b = cl.base;
    auto flx = [&](int k) { return flxv[b + k * ks]; };
    auto den = [&](int k) { return rhog2[b + k * ks]; };
    auto dv = [&](int k) { return real(0.0) * frhog[b + k * ks] / rhog2[b + k * ks] * dt; };
    vert_sweep_t(G, d.kmin, d.kmax, in, ob, ib, qhs, d.NF, b, ks, den, flx, dv, limiter != 0, true);
  }
};

NH_END
