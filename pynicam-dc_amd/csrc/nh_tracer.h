// Tracer transport of the last RK stage: MIURA2004 flux form with the
// Thuburn (1996) limiters (nhm/dynamics/mod_src_tracer.py:83-1228).
//
//   vertical half step   TrcVert (phase 1, every column, all tracers in place)
//   horizontal step      TrcHFlux (arc mass fluxes, mass centroids), TrcRhog2,
//                        then per tracer TrcGradQ -> halo -> TrcRemap -> TrcQout
//                        -> halo -> TrcApply
//   vertical half step   TrcVert (phase 2, compute cells, per tracer)
//
// The face mass fluxes of the reference's flx_h(i, j, face) are formed on the
// fly from one value per hexagon arc: face 0/1/2 of (i, j) is the AI/AIJ/AJ arc
// of (i, j); faces 3/4/5 are the same arcs of (i-1, j), (i-1, j-1), (i, j-1)
// with the sign flipped and this cell's area.  The limiter's inflow bounds
// (Qin) are recomputed from q on the 7-point stencil instead of being stored.
#pragma once

#include "nh_step.h"

NH_BEGIN

// CONST_EPS / CONST_HUGE of the field precision (np.finfo(rdtype).eps / .max,
// mod_const.py:170-172)
template <class T> struct Lim;
template <> struct Lim<double> {
  static constexpr double eps = 2.220446049250313e-16, big = 1.7976931348623157e308;   // NH_KEEP64
};
template <> struct Lim<float> {
  static constexpr float eps = 1.1920928955078125e-07f, big = 3.4028234663852886e+38f;
};
#define NH_EPS (Lim<real>::eps)
#define NH_BIG (Lim<real>::big)

NH_HD inline real dmin(real a, real b) { return (a < b) ? a : b; }
NH_HD inline real dmax(real a, real b) { return (a > b) ? a : b; }
NH_HD inline real clip(real x, real lo, real hi) { return dmin(dmax(x, lo), hi); }
NH_HD inline real upwind_mask(real ch) { return real(0.5) - nsign(real(0.5), ch - NH_EPS); }

// ----------------------------------------------------------------------------
// vertical half step of one tracer column (mod_src_tracer.py:420-533 and the
// limiter mod_src_tracer.py:1909-2074), streamed in k
// ----------------------------------------------------------------------------
template <class DEN, class FLX, class DV>
NH_HD inline void vert_sweep(const Geo& G, int kmin, int kmax, const real* in, real* out, long base, long ks,
                             DEN den, FLX flx, DV dv, bool limiter, bool clipneg) {
  auto ck0 = [&](int k) -> real { return (k >= kmin && k <= kmax) ? -flx(k) / den(k) * G.rdgz[k] : real(0.0); };
  real r_km1 = in[base + (kmin - 1) * ks];
  real r_k = in[base + kmin * ks];
  real q_km1 = r_km1 / den(kmin - 1);
  real q_k = r_k / den(kmin);
  real omin_prev = real(0.0), omax_prev = real(0.0);
  real qh_prev = real(0.0);      // limited face value at k-1 (zero at kmin)
  real r_prev = r_km1;
  real c0_k = ck0(kmin);
  for (int k = kmin; k <= kmax; ++k) {
    const real r_kp1 = in[base + (k + 1) * ks];
    const real q_kp1 = r_kp1 / den(k + 1);
    const real c0_kp1 = ck0(k + 1);
    const real c1_k = flx(k + 1) / den(k) * G.rdgz[k];
    real qh = G.afact[k] * q_k + G.bfact[k] * q_km1;
    real omin = real(0.0), omax = real(0.0);
    if (limiter) {
      const real inL = real(0.5) - nsign(real(0.5), c0_k);
      const real inU = real(0.5) + nsign(real(0.5), c0_kp1);
      const real minL = (inL == real(1.0)) ? dmin(q_k, q_km1) : NH_BIG;
      const real minU = (inU == real(1.0)) ? dmin(q_k, q_kp1) : NH_BIG;
      const real maxL = (inL == real(1.0)) ? dmax(q_k, q_km1) : -NH_BIG;
      const real maxU = (inU == real(1.0)) ? dmax(q_k, q_kp1) : -NH_BIG;
      const real nmin = dmin(dmin(minL, minU), q_k);
      const real nmax = dmax(dmax(maxL, maxU), q_k);
      const real Cin = inL * c0_k + inU * c1_k;
      const real Cout = (real(1.0) - inL) * c0_k + (real(1.0) - inU) * c1_k;
      const real CQmin = inL * c0_k * minL + inU * c1_k * minU;
      const real CQmax = inL * c0_k * maxL + inU * c1_k * maxU;
      const real zsw = real(0.5) - nsign(real(0.5), nabs(Cout) - NH_EPS);
      const real dk = dv(k);
      omin = ((q_k - nmax) + nmax * (Cin + Cout - dk) - CQmax) / (Cout + zsw) * (real(1.0) - zsw) + q_k * zsw;
      omax = ((q_k - nmin) + nmin * (Cin + Cout - dk) - CQmin) / (Cout + zsw) * (real(1.0) - zsw) + q_k * zsw;
      if (k > kmin) qh = inL * clip(qh, omin_prev, omax_prev) + (real(1.0) - inL) * clip(qh, omin, omax);
    }
    if (k == kmin) qh = real(0.0);
    if (k > kmin) {
      // update of level k-1 with the face values at k-1 and k
      real o = r_prev - (flx(k) * qh - flx(k - 1) * qh_prev) * G.rdgz[k - 1];
      if (clipneg && o > -real(1.0e-10) && o < real(0.0)) o = real(0.0);
      out[base + (k - 1) * ks] = o;
    }
    omin_prev = omin;
    omax_prev = omax;
    qh_prev = qh;
    r_prev = r_k;
    r_k = r_kp1;
    q_km1 = q_k;
    q_k = q_kp1;
    c0_k = c0_kp1;
  }
  {
    real o = r_prev - (flx(kmax + 1) * real(0.0) - flx(kmax) * qh_prev) * G.rdgz[kmax];
    if (clipneg && o > -real(1.0e-10) && o < real(0.0)) o = real(0.0);
    out[base + kmax * ks] = o;
  }
  out[base + (kmin - 1) * ks] = real(0.0);
  out[base + (kmax + 1) * ks] = real(0.0);
}

// Phase 1: vertical flux of the mean mass flux, half-step density and the
// first vertical half step of every tracer, on every column (the horizontal
// step reads halo values; computing them here equals filling them).
// (kernels/tracervertflux.py:44-122)
struct TrcVert1 {
  static constexpr const char* kName = "trc_vert1";
  Dims d; Consts c; Geo G;
  real dt;
  const real* rw_mean;    // time-mean rhogw
  const real* rhog_in;    // density at the start of the large step
  const real* frhog;      // hyperdiffusion density tendency
  real* flxv;             // out
  real* rhog1;            // out
  real* rhogq;            // nq tracer fields, in place
  int nq;
  int limiter;
  NH_HD void operator()(long t) const {
    Cell cl = cell_of(d, t);
    const long ks = cl.ks, b = cl.base;
    const int kmin = d.kmin, kmax = d.kmax;
    auto flx = [&](int k) -> real {
      return (k >= kmin + 1 && k <= kmax) ? (real(0.0) + rw_mean[b + k * ks] * real(1.0)) * real(0.5) * dt : real(0.0);
    };
    for (int k = 0; k < d.kall; ++k) flxv[b + k * ks] = flx(k);
    for (int k = kmin; k <= kmax; ++k)
      rhog1[b + k * ks] = rhog_in[b + k * ks] - (flx(k + 1) - flx(k)) * G.rdgz[k] + real(0.0) * frhog[b + k * ks] * dt;
    rhog1[b + (kmin - 1) * ks] = rhog_in[b + kmin * ks];
    rhog1[b + (kmax + 1) * ks] = rhog_in[b + kmax * ks];
    auto den = [&](int k) { return rhog_in[b + k * ks]; };
    auto dv = [&](int k) { return real(0.0) * frhog[b + k * ks] / rhog_in[b + k * ks] * dt; };
    for (int q = 0; q < nq; ++q) {
      real* f = rhogq + (long)q * d.NF;
      vert_sweep(G, kmin, kmax, f, f, b, ks, den, flx, dv, limiter != 0, false);
    }
  }
};

// Phase 2: second vertical half step of one tracer on the compute cells,
// with the density after the horizontal step, and the small-negative clip
// (mod_src_tracer.py:947-1204).
struct TrcVert2 {
  static constexpr const char* kName = "trc_vert2";
  Dims d; Consts c; Geo G;
  real dt;
  const real* flxv;
  const real* rhog2;
  const real* frhog;
  const real* in;     // tracer after the horizontal step
  real* out;          // tracer field
  int limiter;
  NH_HD void operator()(long t) const {
    Cell cl = cell_of(d, t);
    if (cl.pole ? cl.v != 0 : !cl.inner) return;
    const long ks = cl.ks, b = cl.base;
    auto flx = [&](int k) { return flxv[b + k * ks]; };
    auto den = [&](int k) { return rhog2[b + k * ks]; };
    auto dv = [&](int k) { return real(0.0) * frhog[b + k * ks] / rhog2[b + k * ks] * dt; };
    vert_sweep(G, d.kmin, d.kmax, in, out, b, ks, den, flx, dv, limiter != 0, true);
  }
};

// ----------------------------------------------------------------------------
// horizontal mass fluxes (mod_src_tracer.py:1376-1533)
// ----------------------------------------------------------------------------
// field layout of the arc quantities: fa[arc] (0.5 * rhov.HN of the arc), xc[arc*3+d]
struct TrcHFlux {
  static constexpr const char* kName = "trc_hflux";
  Dims d; Consts c; Geo G;
  real dt;
  const real* pm;     // 5 mean-flux fields: rhog, rhogvx, rhogvy, rhogvz, rhogw
  real* fa[3];
  real* xc[9];
  // the column launch walks every level; the point launch (more threads, the
  // levels being independent) does one
  NH_HD void operator()(long t) const { levels(cell_of(d, t), 0, d.kall); }
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (pt_of(d, plane, x, p)) levels(cell_of_pt(p), p.k, p.k + 1);
  }
  NH_HD void levels(const Cell& cl, int k0, int k1) const {
    const long NF = d.NF;
    if (cl.pole) {
      if (cl.v != 0) return;
      const int lp = cl.lp;
      for (int k = k0; k < k1; ++k) {
        const long a0 = cl.base + k * 6;
        real rt[6][4];
        for (int v = 1; v <= 5; ++v) {
          int vp = (v == 5) ? 1 : v + 1;
          const real* w = G.twt_pl + (v * 2 + lp) * 3;
          for (int q = 0; q < 4; ++q) {
            const real* f = pm + q * NF;
            rt[v][q] = f[a0] * w[0] + f[a0 + v] * w[1] + f[a0 + vp] * w[2];
          }
        }
        for (int v = 1; v <= 5; ++v) {
          int vm = (v == 1) ? 5 : v - 1;
          const real rr = real(1.0) / dmax(rt[vm][0] + rt[v][0], NH_EPS);
          const real v2x = rt[vm][1] + rt[v][1], v2y = rt[vm][2] + rt[v][2], v2z = rt[vm][3] + rt[v][3];
          const real* hn = G.ahn_pl + (v * 2 + lp) * 3;
          fa[0][a0 + v] = real(0.5) * (v2x * hn[0] + v2y * hn[1] + v2z * hn[2]);
          const real* xr = G.axr_pl + (v * 2 + lp) * 3;
          xc[0][a0 + v] = xr[0] - v2x * rr * dt * real(0.5);
          xc[1][a0 + v] = xr[1] - v2y * rr * dt * real(0.5);
          xc[2][a0 + v] = xr[2] - v2z * rr * dt * real(0.5);
        }
      }
      return;
    }
    const int g = d.g, i = cl.i, j = cl.j;
    if (i > g - 2 || j > g - 2) return;
    const long nreg = d.nreg, ci = cl.ci, ks = cl.ks;
    const real m0 = G.sgp[cl.l] ? real(0.0) : real(1.0), m1 = G.sgp[cl.l] ? real(1.0) : real(0.0);
    for (int k = k0; k < k1; ++k) {
      const long a = cl.base + k * ks;
      // triangle value of type T with corner at offset (di, dj) from this cell
      auto tri = [&](int T, int di, int dj, int q) -> real {
        const real* f = pm + q * NF;
        const long o = a + di + (long)dj * g;
        const long cc = ci + di + (long)dj * g;
        const real w1 = rcp(G.twt, nreg, T * 3 + 0, cc), w2 = rcp(G.twt, nreg, T * 3 + 1, cc),
                     w3 = rcp(G.twt, nreg, T * 3 + 2, cc);
        if (T == 0) {
          real s = f[o] * w1 + f[o + 1] * w2;
          return s + f[o + 1 + g] * w3;
        }
        real s = f[o] * w1;
        return s + (f[o + 1 + g] * w2 + f[o + g] * w3);
      };
      // the (0, 0) TI triangle is replaced by TJ(1, 0) on pentagon regions (arithmetic mask, all regions)
      auto triv = [&](int T, int di, int dj, int q) -> real {
        real v = tri(T, di, dj, q);
        if (T == 0 && i + di == 0 && j + dj == 0) v = v * m0 + tri(1, 1 - i, -j, q) * m1;
        return v;
      };
      // arcs: AI (TJ(i, j-1) + TI(i, j)), AIJ (TI(i, j) + TJ(i, j)), AJ (TJ(i, j) + TI(i-1, j))
      for (int arc = 0; arc < 3; ++arc) {
        int Ta, dia, dja, Tb, dib, djb;
        bool valid = true;
        if (arc == 0) { Ta = 1; dia = 0; dja = -1; Tb = 0; dib = 0; djb = 0; valid = (j >= 1); }
        else if (arc == 1) { Ta = 0; dia = 0; dja = 0; Tb = 1; dib = 0; djb = 0; }
        else { Ta = 1; dia = 0; dja = 0; Tb = 0; dib = -1; djb = 0; valid = (i >= 1); }
        if (!valid) {
          fa[arc][a] = real(0.0);
          for (int dd = 0; dd < 3; ++dd) xc[arc * 3 + dd][a] = real(0.0);
          continue;
        }
        const real rs = triv(Ta, dia, dja, 0) + triv(Tb, dib, djb, 0);
        const real rr = real(1.0) / dmax(rs, NH_EPS);
        real v2[3];
        for (int q = 0; q < 3; ++q) v2[q] = triv(Ta, dia, dja, q + 1) + triv(Tb, dib, djb, q + 1);
        const real hx = rcp(G.ahn, nreg, arc * 3 + 0, ci), hy = rcp(G.ahn, nreg, arc * 3 + 1, ci),
                     hz = rcp(G.ahn, nreg, arc * 3 + 2, ci);
        fa[arc][a] = real(0.5) * (v2[0] * hx + v2[1] * hy + v2[2] * hz);
        for (int dd = 0; dd < 3; ++dd)
          xc[arc * 3 + dd][a] = rcp(G.axr, nreg, arc * 3 + dd, ci) - v2[dd] * rr * dt * real(0.5);
      }
    }
  }
};

// face mass flux of face m (0..5) of regional cell at address a / column ci
NH_HD inline real face_flux(const Geo& G, const Dims& d, const real* const* fa, long a, long ci, int i, int j,
                              int l, int m, real dt) {
  const int g = d.g;
  const real ra = rcp(G.rarea, d.nreg, 0, ci);
  switch (m) {
    case 0: return fa[0][a] * ra * dt;
    case 1: return fa[1][a] * ra * dt;
    case 2: return fa[2][a] * ra * dt;
    case 3: return -fa[0][a - 1] * ra * dt;
    case 4: return -fa[1][a - 1 - g] * ra * dt;
    default: {
      real f = -fa[2][a - g] * ra * dt;
      if (i == 1 && j == 1) f *= (G.sgp[l] ? real(0.0) : real(1.0));
      return f;
    }
  }
}

// density after the horizontal step (mod_src_tracer.py:889-929)
struct TrcRhog2 {
  static constexpr const char* kName = "trc_rhog2";
  Dims d; Consts c; Geo G;
  real dt;
  const real* fa[3];
  const real* rhog1;
  const real* frhog;
  real* rhog2;
  // the column launch walks every level; the point launch (more threads, the
  // levels being independent) does one
  NH_HD void operator()(long t) const { levels(cell_of(d, t), 0, d.kall); }
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (pt_of(d, plane, x, p)) levels(cell_of_pt(p), p.k, p.k + 1);
  }
  NH_HD void levels(const Cell& cl, int k0, int k1) const {
    if (cl.pole ? cl.v != 0 : !cl.inner) return;
    for (int k = k0; k < k1; ++k) {
      const long a = cl.base + k * cl.ks;
      if (cl.pole) {
        real r = rhog1[a];
        for (int v = 1; v <= 5; ++v) r -= fa[0][a + v] * G.rarea_pl[cl.lp] * dt;
        r += real(1.0) * frhog[a] * dt;
        rhog2[a] = r;
      } else {
        real f[6];
        for (int m = 0; m < 6; ++m) f[m] = face_flux(G, d, fa, a, cl.ci, cl.i, cl.j, cl.l, m, dt);
        rhog2[a] = rhog1[a] - ((((f[0] + f[1]) + f[2]) + f[3]) + f[4] + f[5] - (real(1.0) * frhog[a] * dt));
      }
    }
  }
};

// gradient of q = rhogq / rhog1 (OPRT_gradient, kernels/oprtgradient.py:45-88)
struct TrcGradQ {
  static constexpr const char* kName = "trc_gradq";
  Dims d; Geo G;
  const real* rq;
  const real* rhog1;
  real* gq[3];
  // the column launch walks every level; the point launch (more threads, the
  // levels being independent) does one
  NH_HD void operator()(long t) const { levels(cell_of(d, t), 0, d.kall); }
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (pt_of(d, plane, x, p)) levels(cell_of_pt(p), p.k, p.k + 1);
  }
  NH_HD void levels(const Cell& cl, int k0, int k1) const {
    if (cl.pole ? cl.v != 0 : !cl.inner) return;
    long nb[7];
    int nn = neighbours(d, cl, nb);
    for (int k = k0; k < k1; ++k) {
      const long kk = k * cl.ks;
      real qv[7];
      for (int m = 0; m < nn; ++m) qv[m] = rq[nb[m] + kk] / rhog1[nb[m] + kk];
      for (int dd = 0; dd < 3; ++dd) {
        real acc;
        if (cl.pole) {
          acc = real(0.0);
          for (int v = 0; v < 6; ++v) {
            real pv = G.cgrad_pl[(v * 2 + cl.lp) * 3 + dd] * qv[v];
            acc = (v == 0) ? pv : acc + pv;
          }
        } else {
          acc = rcp(G.cgrad, d.nreg, dd * 7, cl.ci) * qv[0];
          for (int m = 1; m < 7; ++m) acc = acc + rcp(G.cgrad, d.nreg, dd * 7 + m, cl.ci) * qv[m];
        }
        gq[dd][cl.base + kk] = acc;
      }
    }
  }
};

// ----------------------------------------------------------------------------
// upwind face values (horizontal_remap, mod_src_tracer.py:1642-1843)
// ----------------------------------------------------------------------------
struct TrcRemap {
  static constexpr const char* kName = "trc_remap";
  Dims d; Consts c; Geo G;
  real dt;
  const real* rq;
  const real* rhog1;
  const real* gq[3];
  const real* fa[3];
  const real* xc[9];
  real* qa[3];
  // the column launch walks every level; the point launch (more threads, the
  // levels being independent) does one
  NH_HD void operator()(long t) const { levels(cell_of(d, t), 0, d.kall); }
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (pt_of(d, plane, x, p)) levels(cell_of_pt(p), p.k, p.k + 1);
  }
  NH_HD void levels(const Cell& cl, int k0, int k1) const {
    if (cl.pole) {
      if (cl.v == 0) return;
      // pole arc v: upwind value from the pole centre or spoke v
      const int lp = cl.lp, v = cl.v;
      const long R0 = d.R + (long)lp * d.kall * 6;
      const real* xp0 = G.gx_pl + (0 * 2 + lp) * 3;
      const real* xpv = G.gx_pl + (v * 2 + lp) * 3;
      for (int k = k0; k < k1; ++k) {
        const long a0 = R0 + k * 6, av = a0 + v;
        const real q0 = rq[a0] / rhog1[a0], qv = rq[av] / rhog1[av];
        const real xcx = xc[0][av], xcy = xc[1][av], xcz = xc[2][av];
        const real qap = q0 + gq[0][a0] * (xcx - xp0[0]) + gq[1][a0] * (xcy - xp0[1]) + gq[2][a0] * (xcz - xp0[2]);
        const real qam = qv + gq[0][av] * (xcx - xpv[0]) + gq[1][av] * (xcy - xpv[1]) + gq[2][av] * (xcz - xpv[2]);
        const real ch = fa[0][av] * G.rarea_pl[lp] * dt / rhog1[a0];
        const real cm = upwind_mask(ch);
        qa[0][av] = cm * qam + (real(1.0) - cm) * qap;
      }
      return;
    }
    const int g = d.g, i = cl.i, j = cl.j;
    if (i > g - 2 || j > g - 2) return;
    const long nreg = d.nreg, ci = cl.ci, ks = cl.ks;
    // neighbour across each arc: AI (i+1, j), AIJ (i+1, j+1), AJ (i, j+1)
    const long noff[3] = {1, 1 + (long)g, (long)g};
    const bool valid[3] = {j >= 1, true, i >= 1};
    for (int k = k0; k < k1; ++k) {
      const long a = cl.base + k * ks;
      for (int arc = 0; arc < 3; ++arc) {
        if (!valid[arc]) { qa[arc][a] = real(0.0); continue; }
        const long an = a + noff[arc];
        const long cn = ci + noff[arc];
        const real xcx = xc[arc * 3 + 0][a], xcy = xc[arc * 3 + 1][a], xcz = xc[arc * 3 + 2][a];
        const real qc = rq[a] / rhog1[a], qn = rq[an] / rhog1[an];
        const real qap = qc + gq[0][a] * (xcx - rcp(G.gx, nreg, 0, ci)) + gq[1][a] * (xcy - rcp(G.gx, nreg, 1, ci))
                         + gq[2][a] * (xcz - rcp(G.gx, nreg, 2, ci));
        const real qam = qn + gq[0][an] * (xcx - rcp(G.gx, nreg, 0, cn)) + gq[1][an] * (xcy - rcp(G.gx, nreg, 1, cn))
                         + gq[2][an] * (xcz - rcp(G.gx, nreg, 2, cn));
        const real ch = fa[arc][a] * rcp(G.rarea, nreg, 0, ci) * dt / rhog1[a];
        const real cm = upwind_mask(ch);
        qa[arc][a] = cm * qam + (real(1.0) - cm) * qap;
      }
    }
  }
};

// ----------------------------------------------------------------------------
// Thuburn horizontal limiter (horizontal_limiter_thuburn, mod_src_tracer.py:2190-2660)
// ----------------------------------------------------------------------------
// q on the 7-point stencil of one regional cell, and the inflow bounds of the
// six faces.  Stencil: 0 (i,j) 1 (i+1,j) 2 (i+1,j+1) 3 (i,j+1) 4 (i-1,j)
// 5 (i-1,j-1) 6 (i,j-1).
struct FaceBounds {
  real mn[6], mx[6];      // min/max of q over the four cells around each face
  real cm[6];             // upwind mask of the face's owner
  bool pent;                // face 4 uses the pentagon override
};

NH_HD inline void face_bounds(const real q[7], const real cmo[6], bool pentagon, FaceBounds& fb) {
  // owner sets: own AI {0,6,1,2}, own AIJ {0,1,2,3}, own AJ {0,2,3,4},
  // (i-1,j) AI {4,5,0,3}, (i-1,j-1) AIJ {5,6,0,4} (pentagon: {5,0,1,4}), (i,j-1) AJ {6,1,0,5}
  const int sets[6][4] = {{0, 6, 1, 2}, {0, 1, 2, 3}, {0, 2, 3, 4}, {4, 5, 0, 3}, {5, 6, 0, 4}, {6, 1, 0, 5}};
  for (int m = 0; m < 6; ++m) {
    const int* s = sets[m];
    int s1 = s[1], s2 = s[2];
    if (m == 4 && pentagon) { s1 = 0; s2 = 1; }
    const real a = q[s[0]], b = q[s1], c = q[s2], e = q[s[3]];
    fb.mn[m] = dmin(dmin(dmin(a, b), c), e);
    fb.mx[m] = dmax(dmax(dmax(a, b), c), e);
    fb.cm[m] = cmo[m];
  }
  fb.pent = pentagon;
}

// inflow bounds as seen from this cell (Qin(i, j, *, m)): faces 0..2 are owned by the
// cell, faces 3..5 by the neighbour and see the complementary mask
NH_HD inline void qin_of_cell(const FaceBounds& fb, int m, real& lo, real& hi) {
  const real c = fb.cm[m];
  if (m < 3) {
    lo = c * fb.mn[m] + (real(1.0) - c) * NH_BIG;
    hi = c * fb.mx[m] + (real(1.0) - c) * (-NH_BIG);
  } else if (m == 4 && fb.pent) {
    lo = (c == real(1.0)) ? NH_BIG : fb.mn[m];
    hi = (c == real(1.0)) ? -NH_BIG : fb.mx[m];
  } else {
    lo = c * NH_BIG + (real(1.0) - c) * fb.mn[m];
    hi = c * (-NH_BIG) + (real(1.0) - c) * fb.mx[m];
  }
}

// inflow bounds of the same face as seen from its owner (m >= 3 only)
NH_HD inline void qin_of_owner(const FaceBounds& fb, int m, real& lo, real& hi) {
  const real c = fb.cm[m];
  if (m == 4 && fb.pent) {
    lo = (c == real(1.0)) ? fb.mn[m] : NH_BIG;
    hi = (c == real(1.0)) ? fb.mx[m] : -NH_BIG;
  } else {
    lo = c * fb.mn[m] + (real(1.0) - c) * NH_BIG;
    hi = c * fb.mx[m] + (real(1.0) - c) * (-NH_BIG);
  }
}

// shared per-cell setup: q on the stencil, owner masks, own Courant numbers
struct TrcCellCtx {
  real q[7];
  real cmo[6];    // owner mask per face
  real ch[6];     // this cell's Courant numbers
  real f[6];      // this cell's face fluxes
};

NH_HD inline void trc_cell(const Geo& G, const Dims& d, const Cell& cl, long a, real dt, const real* rq,
                           const real* rhog1, const real* const* fa, TrcCellCtx& x) {
  long off[7];
  stencil_offsets(d.g, off);
  for (int m = 0; m < 7; ++m) x.q[m] = rq[a + off[m]] / rhog1[a + off[m]];
  for (int m = 0; m < 6; ++m) x.f[m] = face_flux(G, d, fa, a, cl.ci, cl.i, cl.j, cl.l, m, dt);
  for (int m = 0; m < 6; ++m) x.ch[m] = x.f[m] / rhog1[a];
  for (int m = 0; m < 3; ++m) x.cmo[m] = upwind_mask(x.ch[m]);
  // faces 3..5: the owner's face 0/1/2 mask with the owner's own flux and density
  const long nreg = d.nreg;
  const long o3 = a - 1, o4 = a - 1 - d.g, o5 = a - d.g;
  const long c3 = cl.ci - 1, c4 = cl.ci - 1 - d.g, c5 = cl.ci - d.g;
  x.cmo[3] = upwind_mask(fa[0][o3] * rcp(G.rarea, nreg, 0, c3) * dt / rhog1[o3]);
  x.cmo[4] = upwind_mask(fa[1][o4] * rcp(G.rarea, nreg, 0, c4) * dt / rhog1[o4]);
  x.cmo[5] = upwind_mask(fa[2][o5] * rcp(G.rarea, nreg, 0, c5) * dt / rhog1[o5]);
}

struct TrcQout {
  static constexpr const char* kName = "trc_qout";
  Dims d; Consts c; Geo G;
  real dt;
  const real* rq;
  const real* rhog1;
  const real* frhog;
  const real* fa[3];
  real* qout[2];
  // the column launch walks every level; the point launch (more threads, the
  // levels being independent) does one
  NH_HD void operator()(long t) const { levels(cell_of(d, t), 0, d.kall); }
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (pt_of(d, plane, x, p)) levels(cell_of_pt(p), p.k, p.k + 1);
  }
  NH_HD void levels(const Cell& cl, int k0, int k1) const {
    if (cl.pole ? cl.v != 0 : !cl.inner) return;
    for (int k = k0; k < k1; ++k) {
      const long a = cl.base + k * cl.ks;
      const real dk = real(1.0) * frhog[a] / rhog1[a] * dt;
      if (cl.pole) {
        const real ra = G.rarea_pl[cl.lp];
        const real q0 = rq[a] / rhog1[a];
        real qv[6];
        for (int v = 1; v <= 5; ++v) qv[v] = rq[a + v] / rhog1[a + v];
        real nmin = q0, nmax = q0, Cin = real(0.0), Cout = real(0.0), CQmin = real(0.0), CQmax = real(0.0);
        for (int v = 1; v <= 5; ++v) {
          const int vp = (v == 5) ? 1 : v + 1, vm = (v == 1) ? 5 : v - 1;
          const real mn = dmin(dmin(dmin(q0, qv[v]), qv[vm]), qv[vp]);
          const real mx = dmax(dmax(dmax(q0, qv[v]), qv[vm]), qv[vp]);
          const real ch = fa[0][a + v] * ra * dt / rhog1[a];
          const real cm = upwind_mask(ch);
          const real lo = (cm == real(1.0)) ? mn : NH_BIG;
          const real hi = (cm == real(1.0)) ? mx : -NH_BIG;
          nmin = dmin(nmin, lo);
          nmax = dmax(nmax, hi);
          const real chm = cm * ch;
          Cin = Cin + chm;
          Cout = Cout + (ch - chm);
          CQmin = CQmin + chm * lo;
          CQmax = CQmax + chm * hi;
        }
        const real zsw = real(0.5) - nsign(real(0.5), nabs(Cout) - NH_EPS);
        const real den = Cout + zsw, fac = real(1.0) - zsw;
        qout[0][a] = (q0 - CQmax - nmax * (real(1.0) - Cin - Cout + dk)) / den * fac + q0 * zsw;
        qout[1][a] = (q0 - CQmin - nmin * (real(1.0) - Cin - Cout + dk)) / den * fac + q0 * zsw;
        continue;
      }
      TrcCellCtx x;
      trc_cell(G, d, cl, a, dt, rq, rhog1, fa, x);
      FaceBounds fb;
      face_bounds(x.q, x.cmo, cl.i == 1 && cl.j == 1 && G.sgp[cl.l], fb);
      real nmin = x.q[0], nmax = x.q[0];
      real lo[6], hi[6];
      for (int m = 0; m < 6; ++m) {
        qin_of_cell(fb, m, lo[m], hi[m]);
        nmin = dmin(nmin, lo[m]);
        nmax = dmax(nmax, hi[m]);
      }
      real Cin = -real(0.0), Cout = -real(0.0), CQmin = -real(0.0), CQmax = -real(0.0);
      for (int m = 0; m < 6; ++m) {
        const real chm = (x.ch[m] < real(0.0)) ? x.ch[m] : real(0.0);
        Cin = Cin + chm;
        Cout = Cout + (x.ch[m] - chm);
        CQmin = CQmin + chm * lo[m];
        CQmax = CQmax + chm * hi[m];
      }
      const real zsw = real(0.5) - nsign(real(0.5), nabs(Cout) - NH_EPS);
      const real qq = x.q[0];
      qout[0][a] = (qq - CQmax - nmax * (real(1.0) - Cin - Cout + dk)) / (Cout + zsw) * (real(1.0) - zsw) + qq * zsw;
      qout[1][a] = (qq - CQmin - nmin * (real(1.0) - Cin - Cout + dk)) / (Cout + zsw) * (real(1.0) - zsw) + qq * zsw;
    }
  }
};

// limited face values and the flux-form update of rhogq (mod_src_tracer.py:2536-2660, :838-874)
struct TrcApply {
  static constexpr const char* kName = "trc_apply";
  Dims d; Consts c; Geo G;
  real dt;
  const real* rq;
  const real* rhog1;
  const real* fa[3];
  const real* qa[3];
  const real* qout[2];
  real* out;
  // limiter off (mod_src_tracer.py:789): the face values go in unlimited; faces
  // 3..5 then take the remap's own values of this cell (:1785), upwind by this
  // cell's mask between the extrapolations from here and from the face's owner
  int limiter;
  const real* gq[3];
  const real* xc[9];
  NH_HD real face_unlimited(const Cell& cl, long a, int m, real fm) const {
    if (m < 3) return qa[m][a];
    const int g = d.g, arc = m - 3;
    const long o = (m == 3) ? a - 1 : (m == 4) ? a - 1 - g : a - g;
    const long co = (m == 3) ? cl.ci - 1 : (m == 4) ? cl.ci - 1 - g : cl.ci - g;
    const long nreg = d.nreg;
    const real* x0 = xc[arc * 3];
    const real* x1 = xc[arc * 3 + 1];
    const real* x2 = xc[arc * 3 + 2];
    const real am = rq[a] / rhog1[a] + gq[0][a] * (x0[o] - rcp(G.gx, nreg, 0, cl.ci))
                    + gq[1][a] * (x1[o] - rcp(G.gx, nreg, 1, cl.ci)) + gq[2][a] * (x2[o] - rcp(G.gx, nreg, 2, cl.ci));
    const real ap = rq[o] / rhog1[o] + gq[0][o] * (x0[o] - rcp(G.gx, nreg, 0, co))
                    + gq[1][o] * (x1[o] - rcp(G.gx, nreg, 1, co)) + gq[2][o] * (x2[o] - rcp(G.gx, nreg, 2, co));
    const real cm = upwind_mask(fm / rhog1[a]);
    return cm * am + (real(1.0) - cm) * ap;
  }
  // the column launch walks every level; the point launch (more threads, the
  // levels being independent) does one
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
      if (!limiter) {
        if (cl.pole) {
          real r = rq[a];
          for (int v = 1; v <= 5; ++v) r -= fa[0][a + v] * G.rarea_pl[cl.lp] * dt * qa[0][a + v];
          out[a] = r;
        } else {
          real f[6], s[6];
          for (int m = 0; m < 6; ++m) f[m] = face_flux(G, d, fa, a, cl.ci, cl.i, cl.j, cl.l, m, dt);
          for (int m = 0; m < 6; ++m) s[m] = face_unlimited(cl, a, m, f[m]);
          out[a] = rq[a] - (f[0] * s[0] + f[1] * s[1] + f[2] * s[2] + f[3] * s[3] + f[4] * s[4] + f[5] * s[5]);
        }
        continue;
      }
      if (cl.pole) {
        const real ra = G.rarea_pl[cl.lp];
        const real q0 = rq[a] / rhog1[a];
        real qv[6];
        for (int v = 1; v <= 5; ++v) qv[v] = rq[a + v] / rhog1[a + v];
        real r = rq[a];
        for (int v = 1; v <= 5; ++v) {
          const int vp = (v == 5) ? 1 : v + 1, vm = (v == 1) ? 5 : v - 1;
          const real mn = dmin(dmin(dmin(q0, qv[v]), qv[vm]), qv[vp]);
          const real mx = dmax(dmax(dmax(q0, qv[v]), qv[vm]), qv[vp]);
          const real fl = fa[0][a + v] * ra * dt;
          const real cm = upwind_mask(fl / rhog1[a]);
          const real lo0 = (cm == real(1.0)) ? mn : NH_BIG, hi0 = (cm == real(1.0)) ? mx : -NH_BIG;
          const real lo1 = (cm == real(1.0)) ? NH_BIG : mn, hi1 = (cm == real(1.0)) ? -NH_BIG : mx;
          real s = qa[0][a + v];
          const real s0 = clip(s, lo0, hi0), s1 = clip(s, lo1, hi1);
          s = cm * s0 + (real(1.0) - cm) * s1;
          const real s2 = dmax(dmin(s, qout[1][a + v]), qout[0][a + v]);
          const real s3 = dmax(dmin(s, qout[1][a]), qout[0][a]);
          s = cm * s2 + (real(1.0) - cm) * s3;
          r -= fl * s;
        }
        out[a] = r;
        continue;
      }
      TrcCellCtx x;
      trc_cell(G, d, cl, a, dt, rq, rhog1, fa, x);
      FaceBounds fb;
      face_bounds(x.q, x.cmo, cl.i == 1 && cl.j == 1 && G.sgp[cl.l], fb);
      // owner cell, downstream cell of the owner's face, and the face's unlimited value
      const long own[6] = {a, a, a, a - 1, a - 1 - g, a - g};
      const long dst[6] = {a + 1, a + 1 + g, a + g, a, a, a};
      const int arc[6] = {0, 1, 2, 0, 1, 2};
      real s_face[6];
      for (int m = 0; m < 6; ++m) {
        const real cm = fb.cm[m];
        real loA, hiA, loB, hiB;     // A: owner side, B: downstream side
        if (m < 3) {
          qin_of_cell(fb, m, loA, hiA);
          // the downstream cell sees the complementary mask
          loB = cm * NH_BIG + (real(1.0) - cm) * fb.mn[m];
          hiB = cm * (-NH_BIG) + (real(1.0) - cm) * fb.mx[m];
        } else {
          qin_of_owner(fb, m, loA, hiA);
          qin_of_cell(fb, m, loB, hiB);
        }
        const real s = qa[arc[m]][own[m]];
        const real v1 = cm * clip(s, loA, hiA) + (real(1.0) - cm) * clip(s, loB, hiB);
        const long o = own[m], dn = dst[m];
        s_face[m] = cm * dmax(dmin(v1, qout[1][dn]), qout[0][dn]) + (real(1.0) - cm) * dmax(dmin(v1, qout[1][o]), qout[0][o]);
      }
      out[a] = rq[a] - (x.f[0] * s_face[0] + x.f[1] * s_face[1] + x.f[2] * s_face[2] + x.f[3] * s_face[3]
                        + x.f[4] * s_face[4] + x.f[5] * s_face[5]);
    }
  }
};

NH_END
