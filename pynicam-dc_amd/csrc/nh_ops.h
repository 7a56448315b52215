// Standalone operators of the reference's kernel API (the pure compute_*
// functions of nhm/dynamics/kernels/ and the numfilter / limiter routines
// built from them), for callers that use the operators one by one instead of
// the whole step.  Each takes and returns fields in the device layout; the
// step engine (nh_engine.h) fuses the same arithmetic into fewer kernels.
//
// Vertical metrics are the flat-topography k-profiles the context holds
// (GSQRT = GAM = 1), so the metric factors of the reference expressions are
// exactly 1 and are left out: x * 1 == x and 1 * 1 == 1 in IEEE arithmetic,
// which keeps every result bit-identical to the expression with the factor.
#pragma once

#include "nh_engine.h"

NH_BEGIN

// ----------------------------------------------------------------------------
// tridiagonal matrix of the rhogw solve, FP64 in every form
// (compute_rhow_matrix_reg/_pl, kernels/vimatrix.py:111-165, with the
// pre-combined metric products of precombine_matrix_reg, :67-86)
// ----------------------------------------------------------------------------
// one thread per column (every regional column and every pole point); rows
// kmin+1..kmax of Mc, Mu, Ml; eth and g_tilde at half levels
struct ViMatrixOp {
  static constexpr const char* kName = "op_vi_matrix";
  Dims d; Consts c; Geo G;
  double dt;
  const real* eth;
  const real* gt;
  double* Mc; double* Mu; double* Ml;
  NH_HD void operator()(long t) const {
    const Cell cl = cell_of(d, t);
    const VProf<double>& V = G.v64;
    const ConstsT<double>& C = c.d;
    const double GCVovR = C.GRAV * C.CVdry / C.Rdry;
    const double ACVovRt2 = C.ALPHA * C.CVdry / C.Rdry / (dt * dt);
    const long ks = cl.ks;
    for (int k = d.kmin + 1; k <= d.kmax; ++k) {
      const long a = cl.base + (long)k * ks;
      const double e0 = eth[a], ep = eth[a + ks], em = eth[a - ks];
      const double g0 = gt[a], gp = gt[a + ks], gm = gt[a - ks];
      const double Ca = V.rdgz[k] + V.rdgz[k - 1];      // (RGSGAM2 rdgz + RGSGAM2(k-1) rdgz(k-1)) GAM2H
      const double Cb = V.dfact[k] - V.cfact[k - 1];
      const double Cu1 = V.rdgz[k], Cl1 = V.rdgz[k];    // RGSGAM2 rdgz GAM2H(k+-1)
      Mc[a] = ACVovRt2 + V.rdgzh[k] * (Ca * e0 - Cb * (g0 + GCVovR));
      Mu[a] = -V.rdgzh[k] * (Cu1 * ep + V.cfact[k] * (gp + GCVovR));
      Ml[a] = -V.rdgzh[k] * (Cl1 * em - V.dfact[k - 1] * (gm + GCVovR));
    }
  }
};

// ----------------------------------------------------------------------------
// rhogw column solve, FP64 in every form (compute_rhow_solver_reg/_pl,
// kernels/virhowsolver.py:67-172: source assembly, boundary rows, Thomas
// sweeps of _thomas :42-64)
// ----------------------------------------------------------------------------
// one thread per column; the forward sweep keeps u and gamma in registers
// (NK = kmax - kmin unknowns, or a 78-entry private array for NK = 0).  out
// may be rhogw itself: each column reads its boundary rows before it stores.
template <int NK>
struct ViSolveOp {
  static constexpr const char* kName = "op_vi_solve";
  static constexpr int kMax = NK > 0 ? NK : 78;
  Dims d; Consts c; Geo G;
  double dt;
  const real *rhogw, *rhogw0, *preg0, *rhog0, *Srho, *Sw, *Spre;
  const double *Mc, *Mu, *Ml;
  real* out;
  NH_HD void operator()(long t) const {
    const Cell cl = cell_of(d, t);
    const VProf<double>& V = G.v64;
    const ConstsT<double>& C = c.d;
    const long b = cl.base, ks = cl.ks;
    const int kmin = d.kmin, kmax = d.kmax;
    const int N = NK > 0 ? NK : kmax - kmin;
    const double CVovRt2 = C.CVdry / C.Rdry / (dt * dt);
    const double rw_bot = (double)rhogw[b + (long)kmin * ks];          // rhogw RGSGAM2H
    const double rw_top = (double)rhogw[b + (long)(kmax + 1) * ks];
    double u[kMax], gm[kMax];
    double beta = 0.0;
    for (int n = 0; n < N; ++n) {
      const int k = kmin + 1 + n;
      const long a = b + (long)k * ks, am = a - ks;
      const double A = (double)rhogw0[a] * C.ALPHA + dt * (double)Sw[a];
      const double B = ((double)preg0[a] + dt * (double)Spre[a]) - ((double)preg0[am] + dt * (double)Spre[am]);
      const double Cc = ((double)rhog0[a] + dt * (double)Srho[a]) * V.afact[k]
                        + ((double)rhog0[am] + dt * (double)Srho[am]) * V.bfact[k];
      double S = (A - B * dt * V.rdgzh[k] - Cc * dt * C.GRAV) * CVovRt2;
      if (n == 0) S = S - Ml[a] * rw_bot;
      if (n == N - 1) S = S - Mu[a] * rw_top;
      if (n == 0) {
        beta = Mc[a];
        u[0] = S / beta;
      } else {
        const double g = Mu[am] / beta;
        beta = Mc[a] - Ml[a] * g;
        u[n] = (S - Ml[a] * u[n - 1]) / beta;
        gm[n] = g;
      }
    }
    // the rows below kmin and above kmax+1 are copied, the boundary rows scaled by GSGAM2H
    if (out != rhogw) {
      for (int k = 0; k < kmin; ++k) out[b + (long)k * ks] = rhogw[b + (long)k * ks];
      for (int k = kmax + 2; k < d.kall; ++k) out[b + (long)k * ks] = rhogw[b + (long)k * ks];
    }
    double x = u[N - 1];
    out[b + (long)kmax * ks] = (real)x;
    for (int n = N - 2; n >= 0; --n) {
      x = u[n] - gm[n + 1] * x;
      out[b + (long)(kmin + 1 + n) * ks] = (real)x;
    }
    out[b + (long)kmin * ks] = (real)rw_bot;
    out[b + (long)(kmax + 1) * ks] = (real)rw_top;
  }
};

// ----------------------------------------------------------------------------
// hyperdiffusion tendency (Numf._hdiff_tendency_host, mod_numfilter.py:1519-1578,
// with the Laplacian-1 filter off, then OPRT_horizontalize_vec of the momentum,
// mod_oprt.py:1404-1427)
// ----------------------------------------------------------------------------
// one thread per point; vt the six iterated-Laplacian fields of the second
// pass (vx, vy, vz, w, T, rho), out the six tendencies in PROG order
// (rhog, rhogvx, rhogvy, rhogvz, rhogw, rhoge)
struct HDiffTendOp {
  static constexpr const char* kName = "op_hdiff_tend";
  Dims d; Consts c; Geo G;
  const real* rhog;
  const real* vt[6];
  real* out[6];
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (!pt_of(d, plane, x, p)) return;
    const long a = p.a, ks = p.ks;
    const int k = p.k;
    const real Kh = c.GAMMA_H;
    const real rg = rhog[a];
    real fvx = -(vt[0][a] * Kh + real(0.0)) * rg;
    real fvy = -(vt[1][a] * Kh + real(0.0)) * rg;
    real fvz = -(vt[2][a] * Kh + real(0.0)) * rg;
    const real rgh = (k >= d.kmin && k <= d.kmax + 1) ? G.c2w0[k] * rg + G.c2w1[k] * rhog[a - ks] : real(0.0);
    const real R = c.RADIUS;
    if (p.pole) {
      const real gx = G.gx_pl[(p.v * 2 + p.lp) * 3 + 0], gy = G.gx_pl[(p.v * 2 + p.lp) * 3 + 1],
                 gz = G.gx_pl[(p.v * 2 + p.lp) * 3 + 2];
      const real prd = fvx * gx / R + fvy * gy / R + fvz * gz / R;
      fvx = fvx - prd * gx / R; fvy = fvy - prd * gy / R; fvz = fvz - prd * gz / R;
    } else if (p.inner) {
      const real gx = G.gx[p.ci], gy = G.gx[d.nreg + p.ci], gz = G.gx[2 * d.nreg + p.ci];
      const real prd = (fvx * gx + fvy * gy + fvz * gz) / R;
      fvx = fvx - prd * gx / R; fvy = fvy - prd * gy / R; fvz = fvz - prd * gz / R;
    }
    out[0][a] = -(vt[5][a] + real(0.0));
    out[1][a] = fvx;
    out[2][a] = fvy;
    out[3][a] = fvz;
    out[4][a] = -(vt[3][a] * Kh + real(0.0)) * rgh;
    out[5][a] = -(vt[4][a] + real(0.0));
  }
};

// ----------------------------------------------------------------------------
// vertical Thuburn limiter of the face values (compute_vertical_limiter,
// kernels/verticallimiter.py:41-92)
// ----------------------------------------------------------------------------
// one thread per column; q_h rows kmin+1..kmax are limited in place, ck0/ck1
// are the two Courant numbers of each cell
struct LimVOp {
  static constexpr const char* kName = "op_limiter_v";
  Dims d;
  real* qh;
  const real* q;
  const real* dq;
  const real* ck0;
  const real* ck1;
  NH_HD void operator()(long t) const {
    const Cell cl = cell_of(d, t);
    const long b = cl.base, ks = cl.ks;
    const int kmin = d.kmin, kmax = d.kmax;
    real omin_prev = real(0.0), omax_prev = real(0.0);
    for (int k = kmin; k <= kmax; ++k) {
      const long a = b + (long)k * ks;
      const real inL = real(0.5) - nsign(real(0.5), ck0[a]);
      const real inU = real(0.5) + nsign(real(0.5), ck0[a + ks]);
      const real qc = q[a], qb = q[a - ks], qa = q[a + ks];
      const real c1 = ck0[a], c2 = ck1[a];
      const real minL = (inL == real(1.0)) ? dmin(qc, qb) : NH_BIG;
      const real minU = (inU == real(1.0)) ? dmin(qc, qa) : NH_BIG;
      const real maxL = (inL == real(1.0)) ? dmax(qc, qb) : -NH_BIG;
      const real maxU = (inU == real(1.0)) ? dmax(qc, qa) : -NH_BIG;
      const real nmin = dmin(dmin(minL, minU), qc);
      const real nmax = dmax(dmax(maxL, maxU), qc);
      const real Cin = inL * c1 + inU * c2;
      const real Cout = (real(1.0) - inL) * c1 + (real(1.0) - inU) * c2;
      const real CQmin = inL * c1 * minL + inU * c2 * minU;
      const real CQmax = inL * c1 * maxL + inU * c2 * maxU;
      const real zsw = real(0.5) - nsign(real(0.5), nabs(Cout) - NH_EPS);
      const real dk = dq[a];
      const real omin = ((qc - nmax) + nmax * (Cin + Cout - dk) - CQmax) / (Cout + zsw) * (real(1.0) - zsw) + qc * zsw;
      const real omax = ((qc - nmin) + nmin * (Cin + Cout - dk) - CQmin) / (Cout + zsw) * (real(1.0) - zsw) + qc * zsw;
      if (k > kmin) {
        const real v = qh[a];
        qh[a] = inL * clip(v, omin_prev, omax_prev) + (real(1.0) - inL) * clip(v, omin, omax);
      }
      omin_prev = omin;
      omax_prev = omax;
    }
  }
};

// ----------------------------------------------------------------------------
// horizontal Thuburn limiter of the face values (compute_horizontal_limiter_qout
// and _apply, kernels/horizontallimiter.py:132-255, and the pole kernels :51-129;
// the Qout halo exchange between them as in mod_src_tracer.py)
// ----------------------------------------------------------------------------
// Regional fields: q, d (one value per cell), ch and q_a (six faces per cell:
// 0/1/2 the AI/AIJ/AJ faces owned by the cell, 3/4/5 the faces owned by
// (i-1, j), (i-1, j-1), (i, j-1)), cmask (its faces 0..2 are read: the upwind
// mask of each owned face).  Pole points use component 0 of ch, cmask and q_a.
// The inflow bounds Qin are recomputed from q and cmask wherever they are read.
struct LimHGeom {
  Dims d;
  const int* sgp;
  const real* q;
  const real* cmask;
  // q at regional cell (i, j) of region l
  NH_HD real qat(int l, int i, int j, int k) const {
    return q[(((long)l * d.kall + k) * d.g + j) * d.g + i];
  }
  NH_HD real cm(int l, int i, int j, int k, int m) const {
    return cmask[(long)m * d.NF + (((long)l * d.kall + k) * d.g + j) * d.g + i];
  }
  // min and max of q over the four cells around face m (0..2) of owner (i, j),
  // with the reference's substitutions at i = 0 / j = 0 and the pentagon's AIJ face
  NH_HD bool face_minmax(int l, int i, int j, int k, int m, real& mn, real& mx) const {
    real s[4];
    bool pent = false;
    s[0] = qat(l, i, j, k);
    if (m == 0) {
      s[1] = (j > 0) ? qat(l, i, j - 1, k) : qat(l, 0, 0, k);
      s[2] = qat(l, i + 1, j, k);
      s[3] = qat(l, i + 1, j + 1, k);
    } else if (m == 1) {
      pent = (i == 0 && j == 0 && sgp[l]);
      if (pent) {
        s[1] = qat(l, 1, 1, k);
        s[2] = qat(l, 2, 1, k);
        s[3] = qat(l, 0, 1, k);
      } else {
        s[1] = qat(l, i + 1, j, k);
        s[2] = qat(l, i + 1, j + 1, k);
        s[3] = qat(l, i, j + 1, k);
      }
    } else {
      s[1] = qat(l, i + 1, j + 1, k);
      s[2] = qat(l, i, j + 1, k);
      s[3] = (i > 0) ? qat(l, i - 1, j, k) : qat(l, 0, 0, k);
    }
    mn = dmin(dmin(dmin(s[0], s[1]), s[2]), s[3]);
    mx = dmax(dmax(dmax(s[0], s[1]), s[2]), s[3]);
    return pent;
  }
  // Qin of face m of owner (i, j) on the owner's side (down = false) or on the
  // downstream cell's side (its face m + 3)
  NH_HD void qin(int l, int i, int j, int k, int m, bool down, real& lo, real& hi) const {
    real mn, mx;
    const bool pent = face_minmax(l, i, j, k, m, mn, mx);
    const real c = cm(l, i, j, k, m);
    if (pent) {
      lo = ((c == real(1.0)) != down) ? mn : NH_BIG;
      hi = ((c == real(1.0)) != down) ? mx : -NH_BIG;
    } else if (!down) {
      lo = c * mn + (real(1.0) - c) * NH_BIG;
      hi = c * mx + (real(1.0) - c) * (-NH_BIG);
    } else {
      lo = c * NH_BIG + (real(1.0) - c) * mn;
      hi = c * (-NH_BIG) + (real(1.0) - c) * mx;
    }
  }
};

// Qout of every regional cell (interior: the Thuburn bound; border: q) and of
// the pole centres; the pole ring points are left for the halo fill
struct LimHQoutOp {
  static constexpr const char* kName = "op_limiter_h_qout";
  LimHGeom H;
  const real* dq;
  const real* ch;
  real* qout[2];
  NH_HD void operator()(int plane, int x) const {
    const Dims& d = H.d;
    Pt p;
    if (!pt_of(d, plane, x, p)) return;
    const long a = p.a, NF = d.NF;
    if (p.pole) {
      if (p.v != 0) return;
      const real q0 = H.q[a];
      real nmin = q0, nmax = q0;
      real lo[6], hi[6];
      for (int v = 1; v <= 5; ++v) {
        const int vp = (v == 5) ? 1 : v + 1, vm = (v == 1) ? 5 : v - 1;
        const real qv = H.q[a + v], qm = H.q[a + vm], qp = H.q[a + vp];
        const real mn = dmin(dmin(dmin(q0, qv), qm), qp);
        const real mx = dmax(dmax(dmax(q0, qv), qm), qp);
        const bool in = H.cmask[a + v] == real(1.0);
        lo[v] = in ? mn : NH_BIG;
        hi[v] = in ? mx : -NH_BIG;
        nmin = dmin(nmin, lo[v]);
        nmax = dmax(nmax, hi[v]);
      }
      real Cin = real(0.0), Cout = real(0.0), CQmin = real(0.0), CQmax = real(0.0);
      for (int v = 1; v <= 5; ++v) {
        const real chm = H.cmask[a + v] * ch[a + v];
        Cin = Cin + chm;
        Cout = Cout + (ch[a + v] - chm);
        CQmin = CQmin + chm * lo[v];
        CQmax = CQmax + chm * hi[v];
      }
      const real zsw = real(0.5) - nsign(real(0.5), nabs(Cout) - NH_EPS);
      const real den = Cout + zsw, fac = real(1.0) - zsw;
      const real dd = dq[a];
      qout[0][a] = (q0 - CQmax - nmax * (real(1.0) - Cin - Cout + dd)) / den * fac + q0 * zsw;
      qout[1][a] = (q0 - CQmin - nmin * (real(1.0) - Cin - Cout + dd)) / den * fac + q0 * zsw;
      return;
    }
    const real qq = H.q[a];
    if (!p.inner) {
      qout[0][a] = qq;
      qout[1][a] = qq;
      return;
    }
    const int l = p.l, i = p.i, j = p.j, k = p.k;
    real lo[6], hi[6];
    for (int m = 0; m < 3; ++m) H.qin(l, i, j, k, m, false, lo[m], hi[m]);
    H.qin(l, i - 1, j, k, 0, true, lo[3], hi[3]);
    H.qin(l, i - 1, j - 1, k, 1, true, lo[4], hi[4]);
    H.qin(l, i, j - 1, k, 2, true, lo[5], hi[5]);
    real nmin = qq, nmax = qq;
    for (int m = 0; m < 6; ++m) {
      nmin = dmin(nmin, lo[m]);
      nmax = dmax(nmax, hi[m]);
    }
    real Cin = -real(0.0), Cout = -real(0.0), CQmin = -real(0.0), CQmax = -real(0.0);
    for (int m = 0; m < 6; ++m) {
      const real c6 = ch[(long)m * NF + a];
      const real chm = (c6 < real(0.0)) ? c6 : real(0.0);
      Cin = Cin + chm;
      Cout = Cout + (c6 - chm);
      CQmin = CQmin + chm * lo[m];
      CQmax = CQmax + chm * hi[m];
    }
    const real zsw = real(0.5) - nsign(real(0.5), nabs(Cout) - NH_EPS);
    const real dd = dq[a];
    qout[0][a] = (qq - CQmax - nmax * (real(1.0) - Cin - Cout + dd)) / (Cout + zsw) * (real(1.0) - zsw) + qq * zsw;
    qout[1][a] = (qq - CQmin - nmin * (real(1.0) - Cin - Cout + dd)) / (Cout + zsw) * (real(1.0) - zsw) + qq * zsw;
  }
};

// limited face values: every regional cell computes the faces it holds (its
// own faces 0..2 when it owns them, faces 3..5 from their owners), so that
// no thread writes another's values; pole ring points limit component 0
struct LimHApplyOp {
  static constexpr const char* kName = "op_limiter_h_apply";
  LimHGeom H;
  const real* qin_a;     // q_a on entry (6 faces)
  const real* qout[2];
  real* qa;              // q_a on exit (may be qin_a)
  NH_HD real face(int l, int oi, int oj, int k, int m) const {
    const Dims& d = H.d;
    const long NF = d.NF;
    const long o = (((long)l * d.kall + k) * d.g + oj) * d.g + oi;
    const long dn = o + ((m == 0) ? 1 : (m == 1) ? 1 + d.g : d.g);
    const real c = H.cm(l, oi, oj, k, m);
    real loA, hiA, loB, hiB;
    H.qin(l, oi, oj, k, m, false, loA, hiA);
    H.qin(l, oi, oj, k, m, true, loB, hiB);
    const real s = qin_a[(long)m * NF + o];
    const real v1 = c * dmin(dmax(s, loA), hiA) + (real(1.0) - c) * dmin(dmax(s, loB), hiB);
    return c * dmax(dmin(v1, qout[1][dn]), qout[0][dn]) + (real(1.0) - c) * dmax(dmin(v1, qout[1][o]), qout[0][o]);
  }
  NH_HD void operator()(int plane, int x) const {
    const Dims& d = H.d;
    Pt p;
    if (!pt_of(d, plane, x, p)) return;
    const long a = p.a, NF = d.NF;
    if (p.pole) {
      if (p.v == 0) return;
      const long c0 = a - p.v;
      const int v = p.v;
      const int vp = (v == 5) ? 1 : v + 1, vm = (v == 1) ? 5 : v - 1;
      const real q0 = H.q[c0], qv = H.q[a], qm = H.q[c0 + vm], qp = H.q[c0 + vp];
      const real mn = dmin(dmin(dmin(q0, qv), qm), qp);
      const real mx = dmax(dmax(dmax(q0, qv), qm), qp);
      const real cmv = H.cmask[a];
      const bool in = cmv == real(1.0);
      const real s = qin_a[a];
      const real s0 = dmin(dmax(s, in ? mn : NH_BIG), in ? mx : -NH_BIG);
      const real s1 = dmin(dmax(s, in ? NH_BIG : mn), in ? -NH_BIG : mx);
      const real sv = cmv * s0 + (real(1.0) - cmv) * s1;
      const real s2 = dmax(dmin(sv, qout[1][a]), qout[0][a]);
      const real s3 = dmax(dmin(sv, qout[1][c0]), qout[0][c0]);
      qa[a] = cmv * s2 + (real(1.0) - cmv) * s3;
      return;
    }
    const int l = p.l, i = p.i, j = p.j, k = p.k, G1 = d.g - 1;
    real f[6];
    bool has[6];
    has[0] = has[1] = has[2] = (i < G1 && j < G1);
    has[3] = (i >= 1 && j < G1);
    has[4] = (i >= 1 && j >= 1);
    has[5] = (j >= 1 && i < G1);
    for (int m = 0; m < 3; ++m)
      if (has[m]) f[m] = face(l, i, j, k, m);
    if (has[3]) f[3] = face(l, i - 1, j, k, 0);
    if (has[4]) f[4] = face(l, i - 1, j - 1, k, 1);
    if (has[5]) f[5] = face(l, i, j - 1, k, 2);
    for (int m = 0; m < 6; ++m) {
      const long am = (long)m * NF + a;
      qa[am] = has[m] ? f[m] : qin_a[am];
    }
  }
};

NH_END
