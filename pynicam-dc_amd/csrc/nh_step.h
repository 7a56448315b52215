// Column functors of the HEVI large/small step.
//
// Vertical metrics enter as k-profiles (flat topography: GSQRT = GAM = 1, the
// terrain-following gradient terms vanish); the host verifies this before it
// creates a context.  Operation order follows the reference kernels so that
// the FP64 results track the numpy path to round-off.
#pragma once

#include "nh_columns.h"

namespace nh {

// ----------------------------------------------------------------------------
// Pre_Post: diagnostics + boundary conditions + eth
// (kernels/diag.py:53-143, mod_bndcnd.py:153-238, kernels/thrmdyn.py:56-71)
// ----------------------------------------------------------------------------
struct PrePost {
  static constexpr const char* kName = "prepost";
  Dims d; Consts c; Geo G;
  double* prog;        // 6 fields, in/out (ghost rows written)
  const double* rhogq; // qv * rhog
  double* diag;        // 6 fields out
  double* eth;         // out
  NH_HD void operator()(long t) const {
    Cell cl = cell_of(d, t);
    const long NF = d.NF;
    const int kmin = d.kmin, kmax = d.kmax, kall = d.kall;
    double* rg = prog + 0 * NF;
    double* rvx = prog + 1 * NF;
    double* rvy = prog + 2 * NF;
    double* rvz = prog + 3 * NF;
    double* rw = prog + 4 * NF;
    double* re = prog + 5 * NF;
    double* pre = diag + I_PRE * NF;
    double* tem = diag + I_TEM * NF;
    double* vx = diag + I_VX * NF;
    double* vy = diag + I_VY * NF;
    double* vz = diag + I_VZ * NF;
    double* w = diag + I_W * NF;
    double rho[NH_KMAX], ein[NH_KMAX];
    for (int k = 0; k < kall; ++k) {
      long a = cl.base + k * cl.ks;
      double rhog = rg[a];
      rho[k] = rhog / 1.0;
      vx[a] = rvx[a] / rhog;
      vy[a] = rvy[a] / rhog;
      vz[a] = rvz[a] / rhog;
      ein[k] = re[a] / rhog;
      double q = rhogq[a] / rhog;
      double qd = 1.0 - q;
      double cv = q * c.CVW_qv + qd * c.CVdry;
      double tm = ein[k] / cv;
      tem[a] = tm;
      pre[a] = rho[k] * tm * (qd * c.Rdry + q * c.Rvap);
    }
    for (int k = kmin + 1; k <= kmax; ++k) {
      long a = cl.base + k * cl.ks;
      w[a] = rw[a] / (G.c2w0[k] * rg[a] + G.c2w1[k] * rg[a - cl.ks]);
    }
    // thermodynamic boundary values (TEM type at top and bottom)
    const long at = cl.base + (kmax + 1) * cl.ks, ab = cl.base + (kmin - 1) * cl.ks;
    const long ak = cl.base + kmax * cl.ks, ak1 = cl.base + (kmax - 1) * cl.ks;
    const long an = cl.base + kmin * cl.ks, an1 = cl.base + (kmin + 1) * cl.ks;
    double tem_t = tem[ak], tem_b = tem[an];
    double pre_t = pre[ak1] - rho[kmax] * (G.phi[kmax + 1] - G.phi[kmax - 1]);
    double pre_b = pre[an1] - rho[kmin] * (G.phi[kmin - 1] - G.phi[kmin + 1]);
    tem[at] = tem_t; tem[ab] = tem_b;
    pre[at] = pre_t; pre[ab] = pre_b;
    rho[kmax + 1] = pre_t / (c.Rdry * tem_t);
    rho[kmin - 1] = pre_b / (c.Rdry * tem_b);
    rg[at] = rho[kmax + 1] * 1.0;
    rg[ab] = rho[kmin - 1] * 1.0;
    ein[kmax + 1] = c.CVdry * tem[at];
    ein[kmin - 1] = c.CVdry * tem[ab];
    re[at] = rg[at] * ein[kmax + 1];
    re[ab] = rg[ab] * ein[kmin - 1];
    // momentum: free-slip top, rigid bottom
    double* rv[3] = {rvx, rvy, rvz};
    double* vv[3] = {vx, vy, vz};
    for (int q = 0; q < 3; ++q) {
      double top = 1.0 * (rv[q][ak] / rg[ak]) * rg[at];
      double btm = -1.0 * (rv[q][an] / rg[an]) * rg[ab];
      rv[q][at] = top;
      rv[q][ab] = btm;
    }
    for (int q = 0; q < 3; ++q) {
      vv[q][at] = rv[q][at] / rg[at];
      vv[q][ab] = rv[q][ab] / rg[ab];
    }
    // vertical mass flux at the boundaries (no terrain: zero normal flux)
    rw[at] = -0.0;
    rw[an] = 0.0;
    rw[ab] = 0.0;
    w[at] = rw[at] / (G.c2w0[kmax + 1] * rg[at] + G.c2w1[kmax + 1] * rg[ak]);
    w[an] = rw[an] / (G.c2w0[kmin] * rg[an] + G.c2w1[kmin] * rg[ab]);
    w[ab] = 0.0;
    for (int k = 0; k < kall; ++k) {
      long a = cl.base + k * cl.ks;
      eth[a] = ein[k] + pre[a] / rho[k];
    }
  }
};

// ----------------------------------------------------------------------------
// Momentum advection, Coriolis force and horizontal projection
// (mod_src.py:96-497, kernels/advconvmom.py:54-194, advconv.py, fluxconv.py)
// ----------------------------------------------------------------------------
struct AdvMom {
  static constexpr const char* kName = "adv_mom";
  Dims d; Consts c; Geo G;
  const double* prog;
  const double* diag;
  double* out;   // 4 fields: grhogvx, grhogvy, grhogvz, grhogw
  NH_HD void operator()(long t) const {
    Cell cl = cell_of(d, t);
    const long NF = d.NF;
    const int kmin = d.kmin, kmax = d.kmax, g = d.g;
    const double R = c.RADIUS;
    if (cl.pole && cl.v != 0) return;
    if (!cl.pole && !cl.inner) {
      for (int k = 0; k < d.kall; ++k)
        for (int q = 0; q < 4; ++q) out[q * NF + cl.base + k * cl.ks] = 0.0;
      return;
    }
    const double* rvx = prog + 1 * NF;
    const double* rvy = prog + 2 * NF;
    const double* rvz = prog + 3 * NF;
    const double* rw = prog + 4 * NF;
    const double* rg = prog + 0 * NF;
    const double* vxd = diag + I_VX * NF;
    const double* vyd = diag + I_VY * NF;
    const double* vzd = diag + I_VZ * NF;
    const double* wd = diag + I_W * NF;
    // neighbour columns and their position vectors
    long nb[7];
    double px[7], py[7], pz[7];
    int nn = cl.pole ? 6 : 7;
    if (cl.pole) {
      for (int v = 0; v < 6; ++v) {
        nb[v] = d.R + (long)cl.lp * d.kall * 6 + v;
        px[v] = G.gx_pl[(v * 2 + cl.lp) * 3 + 0];
        py[v] = G.gx_pl[(v * 2 + cl.lp) * 3 + 1];
        pz[v] = G.gx_pl[(v * 2 + cl.lp) * 3 + 2];
      }
    } else {
      long off[7];
      stencil_offsets(g, off);
      for (int m = 0; m < 7; ++m) {
        nb[m] = cl.base + off[m];
        long cim = cl.ci + off[m];
        px[m] = G.gx[0 * d.nreg + cim];
        py[m] = G.gx[1 * d.nreg + cim];
        pz[m] = G.gx[2 * d.nreg + cim];
      }
    }
    const long ks = cl.ks;
    const double gxc = px[0] / R, gyc = py[0] / R, gzc = pz[0] / R;
    // merged velocity (v + w_full * r/|r|) at level k of neighbour m
    auto vv3 = [&](int m, int k, double* o) {
      long a = nb[m] + k * ks;
      double wc = G.cfact[k] * wd[a + ks] + G.dfact[k] * wd[a];
      o[0] = vxd[a] + wc * px[m] / R;
      o[1] = vyd[a] + wc * py[m] / R;
      o[2] = vzd[a] + wc * pz[m] / R;
    };
    double gwc_prev = 0.0;
    // vertical flux of each advected component at the lower half level
    double vmh_lo[3] = {0.0, 0.0, 0.0};
    for (int k = kmin; k <= kmax; ++k) {
      const long a = cl.base + k * ks;
      // horizontal flux divergence for the three advected components
      double vvm[7][3];
      for (int m = 0; m < nn; ++m) vv3(m, k, vvm[m]);
      double div[3];
      for (int q = 0; q < 3; ++q) {
        if (cl.pole) {
          div[q] = div_pl(G, cl.lp,
                          [&](int v) { return rvx[nb[v] + k * ks] * vvm[v][q] * 1.0; },
                          [&](int v) { return rvy[nb[v] + k * ks] * vvm[v][q] * 1.0; },
                          [&](int v) { return rvz[nb[v] + k * ks] * vvm[v][q] * 1.0; });
        } else {
          div[q] = div7(G, d.nreg, cl.ci,
                        [&](int m) { return rvx[nb[m] + k * ks] * vvm[m][q] * 1.0; },
                        [&](int m) { return rvy[nb[m] + k * ks] * vvm[m][q] * 1.0; },
                        [&](int m) { return rvz[nb[m] + k * ks] * vvm[m][q] * 1.0; });
        }
      }
      // vertical flux at the upper half level k+1 (zero at kmax+1)
      double vmh_hi[3] = {0.0, 0.0, 0.0};
      if (k + 1 <= kmax) {
        double vv_up[3];
        vv3(0, k + 1, vv_up);
        double wup = rw[a + ks];
        for (int q = 0; q < 3; ++q)
          vmh_hi[q] = 0.0 * 1.0 + wup * (G.afact[k + 1] * vv_up[q] + G.bfact[k + 1] * vvm[0][q]) * 1.0;
      }
      double dvv[3];
      for (int q = 0; q < 3; ++q) dvv[q] = -div[q] - (vmh_hi[q] - vmh_lo[q]) * G.rdgz[k];
      for (int q = 0; q < 3; ++q) vmh_lo[q] = vmh_hi[q];
      // Coriolis and projection onto the sphere tangent plane
      double rh = rg[a];
      double dx = dvv[0] - (-2.0 * rh * (c.OHM * vvm[0][1]));
      double dy = dvv[1] - (2.0 * rh * (c.OHM * vvm[0][0]));
      double dz = dvv[2];
      double prd = dx * gxc + dy * gyc + dz * gzc;
      out[0 * NF + a] = dx - prd * gxc;
      out[1 * NF + a] = dy - prd * gyc;
      out[2 * NF + a] = dz - prd * gzc;
      double gwc = prd * c.ALPHA;
      out[3 * NF + a] = (k > kmin) ? (G.c2w0[k] * gwc + G.c2w1[k] * gwc_prev) : 0.0;
      gwc_prev = gwc;
    }
    for (int q = 0; q < 4; ++q) {
      out[q * NF + cl.base + (kmin - 1) * ks] = 0.0;
      out[q * NF + cl.base + (kmax + 1) * ks] = 0.0;
    }
  }
};

// ----------------------------------------------------------------------------
// Horizontal operators used by the hyperdiffusion and divergence damping
// ----------------------------------------------------------------------------
// Laplacian of one field at one level of a regional cell (oprtlaplacian.py:41-55)
NH_HD inline double lap7(const Geo& G, long nreg, long ci, const double* f, long a, int g) {
  const long off[7] = {0, 1, 1 + (long)g, (long)g, -1, -1 - (long)g, -(long)g};
  double s = rc(G.clap, nreg, 0, ci) * f[a + off[0]];
  for (int m = 1; m < 7; ++m) s = s + rc(G.clap, nreg, m, ci) * f[a + off[m]];
  return s;
}

NH_HD inline double lap_pl(const Geo& G, int lp, const double* f, long a0) {
  // a0: address of pole point v=0 at this level; neighbours at a0 + v
  double s = 0.0;
  for (int v = 0; v < 6; ++v) {
    double pv = G.clap_pl[v * 2 + lp] * f[a0 + v];
    s = (v == 0) ? pv : s + pv;
  }
  return s;
}


// ----------------------------------------------------------------------------
// Kh-weighted diffusion (kernels/oprtdiffusion.py:59-177)
// ----------------------------------------------------------------------------
// Triangle value of a diffusion-type operator: triangle of type T whose (i, j)
// corner has cell index cj and address a.  TI spans (i,j)(i+1,j)(i+1,j+1),
// TJ spans (i,j)(i+1,j+1)(i,j+1).
NH_HD inline void tri_grad(const Geo& G, long nreg, long cj, int T, const double* s, long a, int g,
                           double out[3]) {
  double s0 = s[a];
  double s1 = (T == 0) ? s[a + 1] : s[a + 1 + g];
  double s2 = (T == 0) ? s[a + 1 + g] : s[a + g];
  for (int dd = 0; dd < 3; ++dd) {
    double c1 = rc(G.cintp, nreg, (dd * 3 + 0) * 2 + T, cj);
    double c2 = rc(G.cintp, nreg, (dd * 3 + 1) * 2 + T, cj);
    double c3 = rc(G.cintp, nreg, (dd * 3 + 2) * 2 + T, cj);
    out[dd] = ((2.0 * c1 - c2 - c3) * s0 + (-c1 + 2.0 * c2 - c3) * s1 + (-c1 - c2 + 2.0 * c3) * s2) / 3.0;
  }
}

template <class KH>
NH_HD inline double diffusion7(const Geo& G, const Dims& d, const Cell& cl, const double* s, KH kh, long a) {
  const int g = d.g;
  const long nreg = d.nreg, ci = cl.ci;
  double t0i[3], t0j[3], t1i[3], t2j[3], t2i[3], t3j[3];
  tri_grad(G, nreg, ci, 0, s, a, g, t0i);                   // TI(i, j)
  tri_grad(G, nreg, ci, 1, s, a, g, t0j);                   // TJ(i, j)
  tri_grad(G, nreg, ci - 1, 0, s, a - 1, g, t1i);           // TI(i-1, j)
  tri_grad(G, nreg, ci - 1 - g, 1, s, a - 1 - g, g, t2j);   // TJ(i-1, j-1)
  tri_grad(G, nreg, ci - 1 - g, 0, s, a - 1 - g, g, t2i);   // TI(i-1, j-1)
  tri_grad(G, nreg, ci - g, 1, s, a - g, g, t3j);           // TJ(i, j-1)
  if (cl.i == 1 && cl.j == 1) {
    double m0 = G.sgp[cl.l] ? 0.0 : 1.0, m1 = G.sgp[cl.l] ? 1.0 : 0.0;
    double tjx[3];
    tri_grad(G, nreg, ci - 1 - g + 1, 1, s, a - 1 - g + 1, g, tjx);   // TJ(i, j-1) seen from (0,0)
    for (int q = 0; q < 3; ++q) t2i[q] = t2i[q] * m0 + tjx[q] * m1;
  }
  double k0 = kh(a);
  double kf1 = 0.5 * (k0 + kh(a + 1 + g));
  double kf2 = 0.5 * (k0 + kh(a + g));
  double kf3 = 0.5 * (k0 + kh(a - 1));
  double kf4 = 0.5 * (k0 + kh(a - 1 - g));
  double kf5 = 0.5 * (k0 + kh(a - g));
  double kf6 = 0.5 * (k0 + kh(a + 1));
  double inner = 0.0;
  for (int dd = 0; dd < 3; ++dd) {
    const double* cd = G.cdiff + (long)(dd * 6) * nreg + ci;
    double blk = kf1 * cd[0] * (t0i[dd] + t0j[dd]) + kf2 * cd[nreg] * (t0j[dd] + t1i[dd])
               + kf3 * cd[2 * nreg] * (t1i[dd] + t2j[dd]) + kf4 * cd[3 * nreg] * (t2j[dd] + t2i[dd])
               + kf5 * cd[4 * nreg] * (t2i[dd] + t3j[dd]) + kf6 * cd[5 * nreg] * (t3j[dd] + t0i[dd]);
    inner = (dd == 0) ? blk : inner + blk;
  }
  return inner;
}

template <class KH>
NH_HD inline double diffusion_pl(const Geo& G, int lp, const double* s, KH kh, long a0) {
  double vt[5][3];
  for (int v = 1; v <= 5; ++v) {
    int vp = (v == 5) ? 1 : v + 1;
    for (int dd = 0; dd < 3; ++dd) {
      const double* cI = G.cintp_pl + ((v * 2 + lp) * 3 + dd) * 3;
      double c0 = cI[0], c1 = cI[1], c2 = cI[2];
      vt[v - 1][dd] = ((2.0 * c0 - c1 - c2) * s[a0] + (-c0 + 2.0 * c1 - c2) * s[a0 + v]
                       + (-c0 - c1 + 2.0 * c2) * s[a0 + vp]) / 3.0;
    }
  }
  double row = 0.0;
  for (int v = 1; v <= 5; ++v) {
    int vm = (v == 1) ? 5 : v - 1;
    double kavg = 0.5 * (kh(a0) + kh(a0 + v));
    const double* cd = G.cdiff_pl + (v * 2 + lp) * 3;
    double sum = cd[0] * (vt[vm - 1][0] + vt[v - 1][0]);
    sum = sum + cd[1] * (vt[vm - 1][1] + vt[v - 1][1]);
    sum = sum + cd[2] * (vt[vm - 1][2] + vt[v - 1][2]);
    double cb = kavg * sum;
    row = (v == 1) ? cb : row + cb;
  }
  return row;
}

// One hyperdiffusion pass: out_q = -L(in_q) for six fields (mod_numfilter.py:1125-1264).
// In the last pass components 4 (temperature) and 5 (density) use the
// Kh-weighted diffusion with Kh*rhog*CVdry and Kh*rhog*hdiff_fact_rho.
// LAST = 0: first pass (six Laplacians); LAST = 1: second pass.
template <int LAST>
struct HDiffPass {
  static constexpr const char* kName = LAST ? "hdiff_pass2" : "hdiff_pass1";
  Dims d; Consts c; Geo G;
  const double* in[6];
  double* out[6];
  const double* rhog;
  NH_HD void operator()(long t) const {
    Cell cl = cell_of(d, t);
    if (cl.pole && cl.v != 0) return;
    if (!cl.pole && !cl.inner) return;
    for (int k = 0; k < d.kall; ++k) {
      long a = cl.base + k * cl.ks;
      for (int q = 0; q < 6; ++q) {
        double r;
        if (LAST && q >= 4) {
          double fac = (q == 4) ? c.CVdry : c.HDIFF_FACT_RHO;
          const double* rg = rhog;
          const double gam = c.GAMMA_H;
          auto kh = [rg, fac, gam](long x) { return rg[x] * fac * gam; };
          r = cl.pole ? diffusion_pl(G, cl.lp, in[q], kh, a) : diffusion7(G, d, cl, in[q], kh, a);
        } else {
          r = cl.pole ? lap_pl(G, cl.lp, in[q], a) : lap7(G, d.nreg, cl.ci, in[q], a, d.g);
        }
        out[q][a] = -r;
      }
    }
  }
};

// ----------------------------------------------------------------------------
// 3-D divergence damping (mod_numfilter.py:2032-2268, kernels/oprt3ddivdamp.py,
// kernels/oprtdivdamp.py, kernels/divdamppostcomm.py)
// ----------------------------------------------------------------------------
// triangle value sum_d sum_vertex c * v_d(vertex) for the divdamp operators
NH_HD inline double tri_div(const Geo& G, long nreg, long cj, int T, const double* vx, const double* vy,
                            const double* vz, long a, int g, double sgn) {
  long o1 = (T == 0) ? 1 : 1 + g;
  long o2 = (T == 0) ? 1 + g : g;
  const double* v3[3] = {vx, vy, vz};
  double s = 0.0;
  bool first = true;
  for (int dd = 0; dd < 3; ++dd) {
    double c0 = rc(G.cintp, nreg, (dd * 3 + 0) * 2 + T, cj);
    double c1 = rc(G.cintp, nreg, (dd * 3 + 1) * 2 + T, cj);
    double c2 = rc(G.cintp, nreg, (dd * 3 + 2) * 2 + T, cj);
    double t0 = c0 * (sgn * v3[dd][a]);
    double t1 = c1 * (sgn * v3[dd][a + o1]);
    double t2 = c2 * (sgn * v3[dd][a + o2]);
    if (first) { s = t0 + t1 + t2; first = false; }
    else s = s + t0 + t1 + t2;
  }
  return s;
}

NH_HD inline double hex_sum(const Geo& G, long nreg, long ci, int dd, const double T[6]) {
  // T = {TI(i,j), TJ(i,j), TI(i-1,j), TJ(i-1,j-1), TI(i-1,j-1), TJ(i,j-1)}
  const double* cd = G.cdiff + (long)(dd * 6) * nreg + ci;
  return cd[0] * (T[0] + T[1]) + cd[nreg] * (T[1] + T[2]) + cd[2 * nreg] * (T[2] + T[3])
       + cd[3 * nreg] * (T[3] + T[4]) + cd[4 * nreg] * (T[4] + T[5]) + cd[5 * nreg] * (T[5] + T[0]);
}

// First operator: triangle-mean 3-D divergence, gradient back to cells.
struct DivDamp1 {
  static constexpr const char* kName = "divdamp_3d";
  Dims d; Consts c; Geo G;
  const double *rvx, *rvy, *rvz, *rw;
  double* out[3];
  NH_HD void operator()(long t) const {
    Cell cl = cell_of(d, t);
    if (cl.pole && cl.v != 0) return;
    if (!cl.pole && !cl.inner) return;
    const int kmin = d.kmin, kmax = d.kmax, g = d.g;
    const long ks = cl.ks;
    for (int q = 0; q < 3; ++q) {
      out[q][cl.base + (kmin - 1) * ks] = 0.0;
      out[q][cl.base + (kmax + 1) * ks] = 0.0;
    }
    auto rwv = [&](long a, int k) { return (k >= kmin + 1 && k <= kmax) ? (0.0 * 1.0 + rw[a] * 1.0) : 0.0; };
    for (int k = kmin; k <= kmax; ++k) {
      long a = cl.base + k * ks;
      if (cl.pole) {
        const long a0 = a;
        double sclt[5];
        for (int v = 1; v <= 5; ++v) {
          int vp = (v == 5) ? 1 : v + 1;
          double s = 0.0;
          bool first = true;
          const double* vv[3] = {rvx, rvy, rvz};
          for (int dd = 0; dd < 3; ++dd) {
            const double* cI = G.cintp_pl + ((v * 2 + cl.lp) * 3 + dd) * 3;
            double t0 = cI[0] * (vv[dd][a0] * 1.0), t1 = cI[1] * (vv[dd][a0 + v] * 1.0),
                   t2 = cI[2] * (vv[dd][a0 + vp] * 1.0);
            if (first) { s = t0 + t1 + t2; first = false; } else s = s + t0 + t1 + t2;
          }
          double up = (rwv(a0 + ks, k + 1) + rwv(a0 + v + ks, k + 1) + rwv(a0 + vp + ks, k + 1));
          double dn = (rwv(a0, k) + rwv(a0 + v, k) + rwv(a0 + vp, k));
          sclt[v - 1] = s + (up - dn) / 3.0 * G.rdgz[k];
        }
        for (int dd = 0; dd < 3; ++dd) {
          double row = 0.0;
          for (int v = 1; v <= 5; ++v) {
            int vm = (v == 1) ? 5 : v - 1;
            double cb = G.cdiff_pl[(v * 2 + cl.lp) * 3 + dd] * (sclt[vm - 1] + sclt[v - 1]);
            row = (v == 1) ? cb : row + cb;
          }
          out[dd][a] = row;
        }
        continue;
      }
      const long nreg = d.nreg, ci = cl.ci;
      // the six triangles around the cell: (corner offset, type)
      const long co[6] = {0, 0, -1, -1 - (long)g, -1 - (long)g, -(long)g};
      const int ty[6] = {0, 1, 0, 1, 0, 1};
      double T[6];
      for (int n = 0; n < 6; ++n) {
        long aa = a + co[n];
        long cj = ci + co[n];
        long o1 = (ty[n] == 0) ? 1 : 1 + g;
        long o2 = (ty[n] == 0) ? 1 + g : g;
        double s = tri_div(G, nreg, cj, ty[n], rvx, rvy, rvz, aa, g, 1.0);
        double up = (rwv(aa + ks, k + 1) + rwv(aa + o1 + ks, k + 1) + rwv(aa + o2 + ks, k + 1));
        double dn = (rwv(aa, k) + rwv(aa + o1, k) + rwv(aa + o2, k));
        T[n] = s + (up - dn) / 3.0 * G.rdgz[k];
      }
      if (cl.i == 1 && cl.j == 1) {
        // pentagon fix of the (i-1, j-1) TI triangle: TJ(i, j-1)
        double m0 = G.sgp[cl.l] ? 0.0 : 1.0, m1 = G.sgp[cl.l] ? 1.0 : 0.0;
        T[4] = T[4] * m0 + T[5] * m1;
      }
      for (int dd = 0; dd < 3; ++dd) out[dd][a] = hex_sum(G, nreg, ci, dd, T);
    }
  }
};

// Second operator on -(first result), scaled by the damping coefficient and
// projected onto the tangent plane.
struct DivDamp2 {
  static constexpr const char* kName = "divdamp_post";
  Dims d; Consts c; Geo G;
  const double* in[3];
  double* out[3];
  NH_HD void operator()(long t) const {
    Cell cl = cell_of(d, t);
    const int kmin = d.kmin, kmax = d.kmax, g = d.g;
    const long ks = cl.ks;
    const double R = c.RADIUS;
    if (cl.pole && cl.v != 0) {
      for (int k = 0; k < d.kall; ++k)
        for (int q = 0; q < 3; ++q) out[q][cl.base + k * ks] = 0.0;
      return;
    }
    if (!cl.pole && !cl.inner) {
      for (int k = 0; k < d.kall; ++k)
        for (int q = 0; q < 3; ++q) out[q][cl.base + k * ks] = 0.0;
      return;
    }
    for (int k = 0; k < d.kall; ++k) {
      long a = cl.base + k * ks;
      double o[3];
      if (k < kmin || k > kmax) {
        o[0] = o[1] = o[2] = 0.0;
      } else if (cl.pole) {
        double sclt[5];
        for (int v = 1; v <= 5; ++v) {
          int vp = (v == 5) ? 1 : v + 1;
          double s = 0.0;
          bool first = true;
          for (int dd = 0; dd < 3; ++dd) {
            const double* cI = G.cintp_pl + ((v * 2 + cl.lp) * 3 + dd) * 3;
            double t0 = cI[0] * (-in[dd][a]), t1 = cI[1] * (-in[dd][a + v]), t2 = cI[2] * (-in[dd][a + vp]);
            if (first) { s = t0 + t1 + t2; first = false; } else s = s + t0 + t1 + t2;
          }
          sclt[v - 1] = s;
        }
        for (int dd = 0; dd < 3; ++dd) {
          double row = 0.0;
          for (int v = 1; v <= 5; ++v) {
            int vm = (v == 1) ? 5 : v - 1;
            double cb = G.cdiff_pl[(v * 2 + cl.lp) * 3 + dd] * (sclt[vm - 1] + sclt[v - 1]);
            row = (v == 1) ? cb : row + cb;
          }
          o[dd] = row;
        }
      } else {
        const long nreg = d.nreg, ci = cl.ci;
        const long co[6] = {0, 0, -1, -1 - (long)g, -1 - (long)g, -(long)g};
        const int ty[6] = {0, 1, 0, 1, 0, 1};
        double T[6];
        for (int n = 0; n < 6; ++n)
          T[n] = tri_div(G, nreg, ci + co[n], ty[n], in[0], in[1], in[2], a + co[n], g, -1.0);
        for (int dd = 0; dd < 3; ++dd) o[dd] = hex_sum(G, nreg, ci, dd, T);
      }
      double gd[3] = {c.ALPHA_D * o[0], c.ALPHA_D * o[1], c.ALPHA_D * o[2]};
      double gx, gy, gz;
      if (cl.pole) {
        gx = G.gx_pl[(0 * 2 + cl.lp) * 3 + 0];
        gy = G.gx_pl[(0 * 2 + cl.lp) * 3 + 1];
        gz = G.gx_pl[(0 * 2 + cl.lp) * 3 + 2];
        double prd = gd[0] * gx / R + gd[1] * gy / R + gd[2] * gz / R;
        out[0][a] = gd[0] - prd * gx / R;
        out[1][a] = gd[1] - prd * gy / R;
        out[2][a] = gd[2] - prd * gz / R;
      } else {
        gx = G.gx[cl.ci]; gy = G.gx[d.nreg + cl.ci]; gz = G.gx[2 * d.nreg + cl.ci];
        double prd = (gd[0] * gx + gd[1] * gy + gd[2] * gz) / R;
        out[0][a] = gd[0] - prd * gx / R;
        out[1][a] = gd[1] - prd * gy / R;
        out[2][a] = gd[2] - prd * gz / R;
      }
    }
  }
};

}  // namespace nh

namespace nh {

// 7-point gradient of s at one level (oprtgradient / presgrad _gradient_interior)
NH_HD inline void grad7(const Geo& G, long nreg, long ci, const double* s, long a, int g, double o[3]) {
  const long off[7] = {0, 1, 1 + (long)g, (long)g, -1, -1 - (long)g, -(long)g};
  for (int dd = 0; dd < 3; ++dd) {
    double acc = rc(G.cgrad, nreg, dd * 7 + 0, ci) * (s[a] * 1.0);
    for (int m = 1; m < 7; ++m) acc = acc + rc(G.cgrad, nreg, dd * 7 + m, ci) * (s[a + off[m]] * 1.0);
    o[dd] = acc;
  }
}

NH_HD inline void grad_pl(const Geo& G, int lp, const double* s, long a0, double o[3]) {
  for (int dd = 0; dd < 3; ++dd) {
    double acc = 0.0;
    for (int v = 0; v < 6; ++v) {
      double pv = G.cgrad_pl[(v * 2 + lp) * 3 + dd] * (s[a0 + v] * 1.0 * 1.0);
      acc = (v == 0) ? pv : acc + pv;
    }
    o[dd] = acc;
  }
}

// horizontal pressure-gradient force at level k with the first-layer remedy
// (kernels/presgrad.py:126-265 with flat terrain: the vertical term is zero)
NH_HD inline void presgrad_h(const Geo& G, const Dims& d, const Cell& cl, const double* P, int k,
                             double R, double o[3]) {
  int kk = (k == d.kmin) ? d.kmin + 1 : k;
  long a = cl.base + kk * cl.ks;
  double gr[3];
  if (cl.pole) grad_pl(G, cl.lp, P, a, gr);
  else grad7(G, d.nreg, cl.ci, P, a, d.g, gr);
  for (int q = 0; q < 3; ++q) gr[q] = gr[q] + 0.0;
  double gx, gy, gz;
  if (cl.pole) {
    gx = G.gx_pl[(0 * 2 + cl.lp) * 3 + 0];
    gy = G.gx_pl[(0 * 2 + cl.lp) * 3 + 1];
    gz = G.gx_pl[(0 * 2 + cl.lp) * 3 + 2];
    double prd = gr[0] * gx / R + gr[1] * gy / R + gr[2] * gz / R;
    o[0] = gr[0] - prd * gx / R;
    o[1] = gr[1] - prd * gy / R;
    o[2] = gr[2] - prd * gz / R;
  } else {
    gx = G.gx[cl.ci]; gy = G.gx[d.nreg + cl.ci]; gz = G.gx[2 * d.nreg + cl.ci];
    double prd = (gr[0] * gx + gr[1] * gy + gr[2] * gz) / R;
    o[0] = gr[0] - prd * gx / R;
    o[1] = gr[1] - prd * gy / R;
    o[2] = gr[2] - prd * gz / R;
  }
}

// neighbour column bases of a compute cell (7 regional, 6 pole)
NH_HD inline int neighbours(const Dims& d, const Cell& cl, long nb[7]) {
  if (cl.pole) {
    for (int v = 0; v < 6; ++v) nb[v] = d.R + (long)cl.lp * d.kall * 6 + v;
    return 6;
  }
  long off[7];
  stencil_offsets(d.g, off);
  for (int m = 0; m < 7; ++m) nb[m] = cl.base + off[m];
  return 7;
}

// horizontal divergence of (fx, fy, fz) given per neighbour index at one level
template <class FX, class FY, class FZ>
NH_HD inline double hdiv(const Geo& G, const Dims& d, const Cell& cl, FX fx, FY fy, FZ fz) {
  if (cl.pole) return div_pl(G, cl.lp, fx, fy, fz);
  return div7(G, d.nreg, cl.ci, fx, fy, fz);
}

// ----------------------------------------------------------------------------
// vi path0: large-step source terms of the small steps + tridiagonal matrix
// (mod_vi.py:132-881, kernels/vimatrix.py:67-136)
// ----------------------------------------------------------------------------
struct ViPath0 {
  static constexpr const char* kName = "vi_path0";
  Dims d; Consts c; Geo G;
  double dt;
  const double* prog;
  const double* diag;
  const double* eth;
  const double* gadv;   // 4 fields from AdvMom
  const double* vt[6];  // hyperdiffusion result after the last pass (haloed)
  const double* dd[3];  // divergence damping of PROG
  double* gT;           // 6 fields out
  double* Mc; double* Mu; double* Ml;
  double* getot;        // large-step rhoge tendency (grhogetot0)
  double* frhog;        // hyperdiffusion rhog tendency
  NH_HD void operator()(long t) const {
    Cell cl = cell_of(d, t);
    if (cl.pole && cl.v != 0) return;
    if (!cl.pole && !cl.inner) return;
    const long NF = d.NF, ks = cl.ks;
    const int kmin = d.kmin, kmax = d.kmax, kall = d.kall;
    const double R = c.RADIUS;
    const double* rg = prog;
    const double* rvx = prog + NF;
    const double* rvy = prog + 2 * NF;
    const double* rvz = prog + 3 * NF;
    const double* rw = prog + 4 * NF;
    const double* pre = diag + I_PRE * NF;
    long nb[7];
    neighbours(d, cl, nb);
    double gx, gy, gz;
    if (cl.pole) {
      gx = G.gx_pl[cl.lp * 3 + 0]; gy = G.gx_pl[cl.lp * 3 + 1]; gz = G.gx_pl[cl.lp * 3 + 2];
    } else {
      gx = G.gx[cl.ci]; gy = G.gx[d.nreg + cl.ci]; gz = G.gx[2 * d.nreg + cl.ci];
    }
    const double Kh = c.GAMMA_H;
    // half-level quantities
    double rhog_h[NH_KMAX], eth_h[NH_KMAX], gzt[NH_KMAX], pwh[NH_KMAX], dpgw[NH_KMAX], dbuo[NH_KMAX];
    for (int k = 0; k < kall; ++k) {
      long a = cl.base + k * ks;
      if (k >= kmin) {
        rhog_h[k] = G.c2w0[k] * rg[a] + G.c2w1[k] * rg[a - ks];
        eth_h[k] = G.afact[k] * eth[a] + G.bfact[k] * eth[a - ks];
      }
      if (k >= kmin + 1 && k <= kmax) {
        dpgw[k] = 1.0 * (pre[a] * 1.0 - pre[a - ks] * 1.0) * G.rdgzh[k];
        dbuo[k] = -c.GRAV * (G.c2w0[k] * (rg[a] / 1.0 * 1.0) + G.c2w1[k] * (rg[a - ks] / 1.0 * 1.0));
      } else {
        dpgw[k] = 0.0;
        dbuo[k] = 0.0;
      }
    }
    rhog_h[kmin - 1] = rhog_h[kmin];
    eth_h[kmin - 1] = eth_h[kmin];
    for (int k = 0; k < kall; ++k) {
      gzt[k] = c.GRAV - (dpgw[k] - dbuo[k]) / rhog_h[k];
      pwh[k] = -gzt[k] * rw[cl.base + k * ks];
    }
    // hyperdiffusion tendency + momentum advection = large-step tendency g0
    for (int k = 0; k < kall; ++k) {
      long a = cl.base + k * ks;
      double rgk = rg[a];
      double fvx = -(vt[0][a] * Kh + 0.0) * rgk;
      double fvy = -(vt[1][a] * Kh + 0.0) * rgk;
      double fvz = -(vt[2][a] * Kh + 0.0) * rgk;
      double rgh = (k >= kmin) ? G.c2w0[k] * rgk + G.c2w1[k] * rg[a - ks] : 0.0;
      double fw = -(vt[3][a] * Kh + 0.0) * rgh;
      double fe = -(vt[4][a] + 0.0);
      double fr = -(vt[5][a] + 0.0);
      if (cl.pole) {
        double prd = fvx * gx / R + fvy * gy / R + fvz * gz / R;
        fvx = fvx - prd * gx / R; fvy = fvy - prd * gy / R; fvz = fvz - prd * gz / R;
      } else {
        double prd = (fvx * gx + fvy * gy + fvz * gz) / R;
        fvx = fvx - prd * gx / R; fvy = fvy - prd * gy / R; fvz = fvz - prd * gz / R;
      }
      double g0[6];
      g0[0] = 0.0 + fr;
      g0[1] = gadv[0 * NF + a] + fvx;
      g0[2] = gadv[1 * NF + a] + fvy;
      g0[3] = gadv[2 * NF + a] + fvz;
      g0[4] = gadv[3 * NF + a] + fw;
      g0[5] = 0.0 + fe;
      getot[a] = g0[5];
      frhog[a] = g0[0];
      // source terms
      double drhog = 0.0, drhoge = 0.0, pw = 0.0;
      double dpg[3] = {0.0, 0.0, 0.0};
      if (k >= kmin && k <= kmax) {
        double hd = hdiv(G, d, cl,
                         [&](int m) { return rvx[nb[m] + k * ks] * 1.0; },
                         [&](int m) { return rvy[nb[m] + k * ks] * 1.0; },
                         [&](int m) { return rvz[nb[m] + k * ks] * 1.0; });
        double vmh_hi = (k + 1 <= kmax) ? 0.0 * 1.0 + 1.0 * rw[a + ks] * 1.0 : 0.0;
        double vmh_lo = (k >= kmin + 1) ? 0.0 * 1.0 + 1.0 * rw[a] * 1.0 : 0.0;
        drhog = -hd - (vmh_hi - vmh_lo) * G.rdgz[k];
        double he = hdiv(G, d, cl,
                         [&](int m) { return rvx[nb[m] + k * ks] * eth[nb[m] + k * ks] * 1.0; },
                         [&](int m) { return rvy[nb[m] + k * ks] * eth[nb[m] + k * ks] * 1.0; },
                         [&](int m) { return rvz[nb[m] + k * ks] * eth[nb[m] + k * ks] * 1.0; });
        double ws_hi = (k + 1 <= kmax)
            ? 0.0 * 1.0 + 1.0 * (rw[a + ks] * (G.afact[k + 1] * eth[a + ks] + G.bfact[k + 1] * eth[a])) * 1.0 : 0.0;
        double ws_lo = (k >= kmin + 1)
            ? 0.0 * 1.0 + 1.0 * (rw[a] * (G.afact[k] * eth[a] + G.bfact[k] * eth[a - ks])) * 1.0 : 0.0;
        drhoge = -he - (ws_hi - ws_lo) * G.rdgz[k];
        presgrad_h(G, d, cl, pre, k, R, dpg);
        const double* vx = diag + I_VX * NF;
        const double* vy = diag + I_VY * NF;
        const double* vz = diag + I_VZ * NF;
        pw = vx[a] * dpg[0] + vy[a] * dpg[1] + vz[a] * dpg[2] + G.w2c0[k] * pwh[k + 1] + G.w2c1[k] * pwh[k];
      }
      gT[0 * NF + a] = g0[0] + drhog;
      gT[1 * NF + a] = g0[1] - dpg[0] + dd[0][a] + 0.0;
      gT[2 * NF + a] = g0[2] - dpg[1] + dd[1][a] + 0.0;
      gT[3 * NF + a] = g0[3] - dpg[2] + dd[2][a] + 0.0;
      gT[4 * NF + a] = g0[4] + 0.0 * c.ALPHA - dpgw[k] + dbuo[k];
      gT[5 * NF + a] = g0[5] + drhoge + pw;
    }
    // tridiagonal coefficients of the vertical implicit rhogw equation
    const double GCVovR = c.GRAV * c.CVdry / c.Rdry;
    const double ACVovRt2 = c.ALPHA * c.CVdry / c.Rdry / (dt * dt);
    for (int k = kmin + 1; k <= kmax; ++k) {
      long a = cl.base + k * ks;
      double Ca = (1.0 * G.rdgz[k] + 1.0 * G.rdgz[k - 1]) * 1.0;
      double Cb = G.dfact[k] - G.cfact[k - 1];
      double Cu1 = 1.0 * G.rdgz[k] * 1.0;
      double Cu2 = 1.0 * 1.0 * GCVovR;
      double Cl1 = 1.0 * G.rdgz[k] * 1.0;
      double Cl2 = 1.0 * 1.0 * GCVovR;
      Mc[a] = ACVovRt2 / 1.0 + G.rdgzh[k] * (Ca * eth_h[k] - Cb * (gzt[k] + GCVovR));
      Mu[a] = -G.rdgzh[k] * (Cu1 * eth_h[k + 1] + G.cfact[k] * (gzt[k + 1] + Cu2));
      Ml[a] = -G.rdgzh[k] * (Cl1 * eth_h[k - 1] - G.dfact[k - 1] * (gzt[k - 1] + Cl2));
    }
  }
};

// kinetic energy at level k of a column given accessors (kernels/rhogkin.py:38-79)
template <class F>
NH_HD inline double rhogkin_k(const Geo& G, int kmin, int kmax, int k, F get) {
  // get(k, q): q = 0 rhog, 1..3 rhogv, 4 rhogw
  if (k < kmin || k > kmax) return 0.0;
  double rgk = get(k, 0);
  double vx = get(k, 1), vy = get(k, 2), vz = get(k, 3);
  double kh = 0.5 * (vx * vx + vy * vy + vz * vz) / rgk;
  double vup = 0.0, vdn = 0.0;
  if (k + 1 <= kmax) {
    double w = get(k + 1, 4);
    vup = 0.5 * (w * w) / (G.c2w0[k + 1] * get(k + 1, 0) + G.c2w1[k + 1] * rgk);
  }
  if (k >= kmin + 1) {
    double w = get(k, 4);
    vdn = 0.5 * (w * w) / (G.c2w0[k] * rgk + G.c2w1[k] * get(k - 1, 0));
  }
  return kh + G.w2c0[k] * vup + G.w2c1[k] * vdn;
}

// total specific energy of the large-step state (for the small-step energy
// flux) and the running mean of the small-step mass fluxes, all columns
struct EthTotPM {
  static constexpr const char* kName = "ethtot_mean";
  Dims d; Consts c; Geo G;
  const double* prog;
  const double* eth;
  double* ethtot;
  double* pm;   // 5 fields
  NH_HD void operator()(long t) const {
    Cell cl = cell_of(d, t);
    const long NF = d.NF, ks = cl.ks;
    auto get = [&](int k, int q) { return prog[q * NF + cl.base + k * ks]; };
    for (int k = 0; k < d.kall; ++k) {
      long a = cl.base + k * ks;
      double kin0 = rhogkin_k(G, d.kmin, d.kmax, k, get);
      ethtot[a] = eth[a] + kin0 / prog[a] + G.phi[k];
      for (int q = 0; q < 5; ++q) pm[q * NF + a] = prog[q * NF + a];
    }
  }
};

// split state at the start of an RK stage: 0 (first stage) or PROG0 - PROG
struct SplitInit {
  static constexpr const char* kName = "split_init";
  long NF;
  const double* prog;
  const double* prog0;
  double* ps[6];
  int zero;
  NH_HD void operator()(long t) const {
    for (int q = 0; q < 6; ++q) ps[q][t] = zero ? 0.0 : prog0[q * NF + t] - prog[q * NF + t];
  }
};

// ----------------------------------------------------------------------------
// small step path1: horizontally explicit momentum update (kernels/vipath1.py:68-160)
// ----------------------------------------------------------------------------
struct Path1 {
  static constexpr const char* kName = "vi_path1";
  Dims d; Consts c; Geo G;
  double dt, rweight;
  const double* prog;
  const double* ps[6];
  const double* gT;
  const double* dd[3];
  double* dv[3];
  double* pm;          // accumulate momentum means (fields 1..3)
  NH_HD void operator()(long t) const {
    Cell cl = cell_of(d, t);
    if (cl.pole && cl.v != 0) return;
    if (!cl.pole && !cl.inner) return;
    const long NF = d.NF, ks = cl.ks;
    const int kmin = d.kmin, kmax = d.kmax;
    const double RovCV = c.Rdry / c.CVdry;
    // pressure of this small step from the split energy: PS_e * R/Cv, formed at the gradient points
    const double* pse = ps[I_RHOGE];
    for (int k = kmin; k <= kmax; ++k) {
      long a = cl.base + k * ks;
      double Pg[3];
      {
        int kk = (k == kmin) ? kmin + 1 : k;
        long ak = cl.base + kk * ks;
        double gr[3];
        if (cl.pole) {
          for (int dd2 = 0; dd2 < 3; ++dd2) {
            double acc = 0.0;
            for (int v = 0; v < 6; ++v) {
              double pv = G.cgrad_pl[(v * 2 + cl.lp) * 3 + dd2] * ((pse[ak + v] * RovCV) * 1.0 * 1.0);
              acc = (v == 0) ? pv : acc + pv;
            }
            gr[dd2] = acc;
          }
        } else {
          const long off[7] = {0, 1, 1 + (long)d.g, (long)d.g, -1, -1 - (long)d.g, -(long)d.g};
          for (int dd2 = 0; dd2 < 3; ++dd2) {
            double acc = rc(G.cgrad, d.nreg, dd2 * 7, cl.ci) * ((pse[ak] * RovCV) * 1.0);
            for (int m = 1; m < 7; ++m)
              acc = acc + rc(G.cgrad, d.nreg, dd2 * 7 + m, cl.ci) * ((pse[ak + off[m]] * RovCV) * 1.0);
            gr[dd2] = acc;
          }
        }
        for (int q = 0; q < 3; ++q) gr[q] = gr[q] + 0.0;
        const double R = c.RADIUS;
        if (cl.pole) {
          double gx = G.gx_pl[cl.lp * 3 + 0], gy = G.gx_pl[cl.lp * 3 + 1], gz = G.gx_pl[cl.lp * 3 + 2];
          double prd = gr[0] * gx / R + gr[1] * gy / R + gr[2] * gz / R;
          Pg[0] = gr[0] - prd * gx / R; Pg[1] = gr[1] - prd * gy / R; Pg[2] = gr[2] - prd * gz / R;
        } else {
          double gx = G.gx[cl.ci], gy = G.gx[d.nreg + cl.ci], gz = G.gx[2 * d.nreg + cl.ci];
          double prd = (gr[0] * gx + gr[1] * gy + gr[2] * gz) / R;
          Pg[0] = gr[0] - prd * gx / R; Pg[1] = gr[1] - prd * gy / R; Pg[2] = gr[2] - prd * gz / R;
        }
      }
      for (int q = 0; q < 3; ++q) {
        double drv = gT[(1 + q) * NF + a] - Pg[q] + dd[q][a] + 0.0;
        dv[q][a] = ps[1 + q][a] + drv * dt;
      }
    }
    // ghost levels: free-slip top, rigid bottom, scaled by the large-step density
    const long at = cl.base + (kmax + 1) * ks, ab = cl.base + (kmin - 1) * ks;
    const long ak = cl.base + kmax * ks, an = cl.base + kmin * ks;
    const double* rg = prog;
    for (int q = 0; q < 3; ++q) {
      double top = 1.0 * (dv[q][ak] / rg[ak]) * rg[at];
      double btm = -1.0 * (dv[q][an] / rg[an]) * rg[ab];
      dv[q][at] = top;
      dv[q][ab] = btm;
    }
    for (int k = 0; k < d.kall; ++k) {
      long a = cl.base + k * ks;
      for (int q = 0; q < 3; ++q) pm[(1 + q) * NF + a] = pm[(1 + q) * NF + a] + dv[q][a] * rweight;
    }
  }
};

// ----------------------------------------------------------------------------
// small step vi_main: vertical implicit solve and mass/energy update
// (mod_vi.py:1925-2432, kernels/vimain.py:127-262, kernels/virhowsolver.py:42-120)
// ----------------------------------------------------------------------------
struct ViMain {
  static constexpr const char* kName = "vi_main";
  Dims d; Consts c; Geo G;
  double dt, rweight;
  const double* prog;
  const double* ps[6];
  const double* dv[3];     // momentum after path1 (haloed)
  const double* eth;
  const double* ethtot;
  const double* gT;
  const double* getot;
  const double* Mc; const double* Mu; const double* Ml;
  double* we[3];           // rhog, rhogw, rhoge split out
  double* pm;              // accumulate mass-flux means (fields 0 and 4)
  NH_HD void operator()(long t) const {
    Cell cl = cell_of(d, t);
    if (cl.pole && cl.v != 0) return;
    if (!cl.pole && !cl.inner) return;
    const long NF = d.NF, ks = cl.ks;
    const int kmin = d.kmin, kmax = d.kmax, kall = d.kall;
    const double RovCV = c.Rdry / c.CVdry;
    long nb[7];
    neighbours(d, cl, nb);
    const double* pg = ps[I_RHOG];
    const double* pw = ps[I_RHOGW];
    const double* pe = ps[I_RHOGE];
    double pe_k[NH_KMAX];
    for (int k = 0; k < kall; ++k) pe_k[k] = pe[cl.base + k * ks];
    pe_k[kmin - 1] = pe_k[kmin];
    pe_k[kmax + 1] = pe_k[kmax];
    double divs1[NH_KMAX], grhog1[NH_KMAX], gpre[NH_KMAX];
    for (int k = kmin; k <= kmax; ++k) {
      long kk = k * ks;
      divs1[k] = hdiv(G, d, cl,
                      [&](int m) { return dv[0][nb[m] + kk] * 1.0; },
                      [&](int m) { return dv[1][nb[m] + kk] * 1.0; },
                      [&](int m) { return dv[2][nb[m] + kk] * 1.0; });
      double drhog = -divs1[k] - (0.0 - 0.0) * G.rdgz[k];
      double he = hdiv(G, d, cl,
                       [&](int m) { return dv[0][nb[m] + kk] * eth[nb[m] + kk] * 1.0; },
                       [&](int m) { return dv[1][nb[m] + kk] * eth[nb[m] + kk] * 1.0; },
                       [&](int m) { return dv[2][nb[m] + kk] * eth[nb[m] + kk] * 1.0; });
      double drhoge = -he - (0.0 - 0.0) * G.rdgz[k];
      long a = cl.base + kk;
      grhog1[k] = gT[0 * NF + a] + drhog;
      double grhoge1 = gT[5 * NF + a] + drhoge;
      gpre[k] = grhoge1 * c.Rdry / c.CVdry;
    }
    // source vector of the tridiagonal system (rows kmin+1 .. kmax)
    const int N = kmax - kmin;
    const double CVovRt2 = c.CVdry / c.Rdry / (dt * dt);
    double S[NH_KMAX], u[NH_KMAX], gam[NH_KMAX];
    for (int n = 0; n < N; ++n) {
      int k = kmin + 1 + n;
      long a = cl.base + k * ks;
      double sw = gT[4 * NF + a] + 0.0 * c.ALPHA;
      double p1 = pe_k[k] * RovCV, p0 = pe_k[k - 1] * RovCV;
      double A = (pw[a] * c.ALPHA + dt * sw) * 1.0;
      double B = (p1 + dt * gpre[k]) * 1.0 - (p0 + dt * gpre[k - 1]) * 1.0;
      double Cc = (pg[a] + dt * grhog1[k]) * 1.0 * G.afact[k] + (pg[a - ks] + dt * grhog1[k - 1]) * 1.0 * G.bfact[k];
      S[n] = (A - B * dt * G.rdgzh[k] - Cc * dt * c.GRAV) * CVovRt2;
    }
    const double rw_kmin = 0.0 * 1.0;
    const double rw_top = -0.0 * 1.0;
    S[0] = S[0] - Ml[cl.base + (kmin + 1) * ks] * rw_kmin;
    S[N - 1] = S[N - 1] - Mu[cl.base + kmax * ks] * rw_top;
    // Thomas sweep
    double beta = Mc[cl.base + (kmin + 1) * ks];
    u[0] = S[0] / beta;
    for (int n = 1; n < N; ++n) {
      long a = cl.base + (kmin + 1 + n) * ks;
      double gm = Mu[a - ks] / beta;
      beta = Mc[a] - Ml[a] * gm;
      u[n] = (S[n] - Ml[a] * u[n - 1]) / beta;
      gam[n] = gm;
    }
    double rw1[NH_KMAX];
    for (int k = 0; k < kall; ++k) rw1[k] = 0.0;
    double x = u[N - 1];
    rw1[kmax] = x * 1.0;
    for (int n = N - 2; n >= 0; --n) {
      x = u[n] - gam[n + 1] * x;
      rw1[kmin + 1 + n] = x * 1.0;
    }
    rw1[kmin] = rw_kmin * 1.0;
    rw1[kmax + 1] = rw_top * 1.0;
    // new density split and kinetic-energy correction terms
    const double* rg = prog;
    double rs1[NH_KMAX];
    for (int k = 0; k < kall; ++k) {
      long a = cl.base + k * ks;
      double drhog2 = 0.0;
      if (k >= kmin && k <= kmax) {
        double hi = (k + 1 <= kmax) ? 0.0 * 1.0 + 1.0 * rw1[k + 1] * 1.0 : 0.0;
        double lo = (k >= kmin + 1) ? 0.0 * 1.0 + 1.0 * rw1[k] * 1.0 : 0.0;
        drhog2 = -divs1[k] - (hi - lo) * G.rdgz[k];
      }
      rs1[k] = pg[a] + (gT[0 * NF + a] + drhog2) * dt;
    }
    auto g10 = [&](int k, int q) {
      long a = cl.base + k * ks;
      return prog[q * NF + a] + ps[q][a];
    };
    auto g11 = [&](int k, int q) {
      long a = cl.base + k * ks;
      if (q == 0) return rg[a] + rs1[k];
      if (q == 4) return prog[4 * NF + a] + rw1[k];
      return prog[q * NF + a] + dv[q - 1][a];
    };
    for (int k = 0; k < kall; ++k) {
      long a = cl.base + k * ks;
      double kin10 = rhogkin_k(G, kmin, kmax, k, g10);
      double kin11 = rhogkin_k(G, kmin, kmax, k, g11);
      double dtot = 0.0;
      if (k >= kmin && k <= kmax) {
        long kk = k * ks;
        double hd = hdiv(G, d, cl,
                         [&](int m) { return (prog[1 * NF + nb[m] + kk] + dv[0][nb[m] + kk]) * ethtot[nb[m] + kk] * 1.0; },
                         [&](int m) { return (prog[2 * NF + nb[m] + kk] + dv[1][nb[m] + kk]) * ethtot[nb[m] + kk] * 1.0; },
                         [&](int m) { return (prog[3 * NF + nb[m] + kk] + dv[2][nb[m] + kk]) * ethtot[nb[m] + kk] * 1.0; });
        double hi = 0.0, lo = 0.0;
        if (k + 1 <= kmax)
          hi = 0.0 * 1.0 + 1.0 * ((prog[4 * NF + a + ks] + rw1[k + 1]) *
                                  (G.afact[k + 1] * ethtot[a + ks] + G.bfact[k + 1] * ethtot[a])) * 1.0;
        if (k >= kmin + 1)
          lo = 0.0 * 1.0 + 1.0 * ((prog[4 * NF + a] + rw1[k]) *
                                  (G.afact[k] * ethtot[a] + G.bfact[k] * ethtot[a - ks])) * 1.0;
        dtot = -hd - (hi - lo) * G.rdgz[k];
      }
      double re1 = pe_k[k] + (getot[a] + dtot) * dt + (kin10 - kin11) + (pg[a] - rs1[k]) * G.phi[k];
      we[0][a] = rs1[k];
      we[1][a] = rw1[k];
      we[2][a] = re1;
      pm[0 * NF + a] = pm[0 * NF + a] + rs1[k] * rweight;
      pm[4 * NF + a] = pm[4 * NF + a] + rw1[k] * rweight;
    }
  }
};

// ----------------------------------------------------------------------------
// end of the small steps: PROG += PROG_split and horizontal projection (mod_vi.py:1572-1586)
// ----------------------------------------------------------------------------
struct ProgUpdate {
  static constexpr const char* kName = "prog_update";
  Dims d; Consts c; Geo G;
  double* prog;
  const double* ps[6];
  NH_HD void operator()(long t) const {
    Cell cl = cell_of(d, t);
    const long NF = d.NF, ks = cl.ks;
    const double R = c.RADIUS;
    double gx = 0, gy = 0, gz = 0;
    bool proj = cl.pole || cl.inner;
    if (cl.pole) {
      gx = G.gx_pl[(cl.v * 2 + cl.lp) * 3 + 0];
      gy = G.gx_pl[(cl.v * 2 + cl.lp) * 3 + 1];
      gz = G.gx_pl[(cl.v * 2 + cl.lp) * 3 + 2];
    } else if (cl.inner) {
      gx = G.gx[cl.ci]; gy = G.gx[d.nreg + cl.ci]; gz = G.gx[2 * d.nreg + cl.ci];
    }
    for (int k = 0; k < d.kall; ++k) {
      long a = cl.base + k * ks;
      double v[6];
      for (int q = 0; q < 6; ++q) v[q] = prog[q * NF + a] + ps[q][a];
      if (proj) {
        if (cl.pole) {
          double prd = v[1] * gx / R + v[2] * gy / R + v[3] * gz / R;
          v[1] = v[1] - prd * gx / R; v[2] = v[2] - prd * gy / R; v[3] = v[3] - prd * gz / R;
        } else {
          double prd = (v[1] * gx + v[2] * gy + v[3] * gz) / R;
          v[1] = v[1] - prd * gx / R; v[2] = v[2] - prd * gy / R; v[3] = v[3] - prd * gz / R;
        }
      }
      for (int q = 0; q < 6; ++q) prog[q * NF + a] = v[q];
    }
  }
};

}  // namespace nh
