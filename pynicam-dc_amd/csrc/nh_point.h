// Point kernels of the HEVI step: one thread per (column, level).
//
// Horizontal stencils have no vertical recurrence, so they run one thread per
// point on the 2-D grid of nh_columns.h (plane = region level, x = point in
// the plane).  Vertical couplings that reach only k-2..k+1 (half-level
// interpolation, boundary rows, the tridiagonal coefficients) are recomputed
// by the thread that needs them, which keeps every kernel free of per-thread
// column arrays.  Only the tridiagonal solve (ViMainV) stays a column kernel.
// The arithmetic of every value follows the reference kernels cited per
// functor, in the reference's operation order.
#pragma once

#include "nh_step.h"

NH_BEGIN

// ----------------------------------------------------------------------------
// single operators (tests and callers of the C-ABI)
// ----------------------------------------------------------------------------
struct OpLapP {
  static constexpr const char* kName = "oprt_laplacian";
  Dims d; Geo G;
  const real* s;
  real* o;
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (!pt_of(d, plane, x, p) || !pt_compute(p)) return;
    o[p.a] = p.pole ? lap_pl(G, p.lp, s, p.a) : lap7(G, d.nreg, p.ci, s, p.a, d.g);
  }
};

struct OpDivP {
  static constexpr const char* kName = "oprt_divergence";
  Dims d; Geo G;
  const real *vx, *vy, *vz;
  real* o;
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (!pt_of(d, plane, x, p) || !pt_compute(p)) return;
    const long a = p.a;
    if (p.pole) {
      o[a] = div_pl(G, p.lp, [&](int v) { return vx[a + v]; }, [&](int v) { return vy[a + v]; },
                    [&](int v) { return vz[a + v]; });
    } else {
      long off[7];
      stencil_offsets(d.g, off);
      o[a] = div7(G, d.nreg, p.ci, [&](int m) { return vx[a + off[m]]; }, [&](int m) { return vy[a + off[m]]; },
                  [&](int m) { return vz[a + off[m]]; });
    }
  }
};

struct OpGradP {
  static constexpr const char* kName = "oprt_gradient";
  Dims d; Geo G;
  const real* s;
  real *ox, *oy, *oz;
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (!pt_of(d, plane, x, p) || !pt_compute(p)) return;
    real o[3];
    if (p.pole) grad_pl(G, p.lp, s, p.a, o);
    else grad7(G, d.nreg, p.ci, s, p.a, d.g, o);
    ox[p.a] = o[0];
    oy[p.a] = o[1];
    oz[p.a] = o[2];
  }
};

// ----------------------------------------------------------------------------
// Pre_Post (kernels/diag.py:53-143, mod_bndcnd.py:153-238, kernels/thrmdyn.py:56-71)
// ----------------------------------------------------------------------------
struct PrePostP {
  static constexpr const char* kName = "prepost";
  Dims d; Consts c; Geo G;
  real* prog;
  const real* rhogq;
  real* diag;
  real* eth;
  struct Th { real rho, ein, tem, pre; };
  NH_HD Th thermo_v(real rhog, real rhoge, real rhogq0) const {
    Th t;
    t.rho = rhog / real(1.0);
    t.ein = rhoge / rhog;
    const real q = rhogq0 / rhog;
    const real qd = real(1.0) - q;
    const real cv = q * c.CVW_qv + qd * c.CVdry;
    t.tem = t.ein / cv;
    t.pre = t.rho * t.tem * (qd * c.Rdry + q * c.Rvap);
    return t;
  }
  NH_HD Th thermo(long a) const { return thermo_v(prog[a], prog[5 * d.NF + a], rhogq[a]); }
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (!pt_of(d, plane, x, p)) return;
    const long NF = d.NF, a = p.a, ks = p.ks;
    const int kmin = d.kmin, kmax = d.kmax, k = p.k;
    real* rg = prog;
    real* rw = prog + 4 * NF;
    real* pre = diag + I_PRE * NF;
    real* tem = diag + I_TEM * NF;
    real* w = diag + I_W * NF;
    if (k >= kmin && k <= kmax) {
      // every value is read before the first store (the stores could alias
      // the reads for all the compiler knows, which would serialize them)
      const long al = (k > kmin) ? a - ks : a + ks;   // level below, or above at kmin
      const real rhog = rg[a], rhoge = prog[5 * NF + a], q0 = rhogq[a];
      const real v1 = prog[1 * NF + a], v2 = prog[2 * NF + a], v3 = prog[3 * NF + a];
      const real rwa = rw[a], rgl = rg[al], rgel = prog[5 * NF + al], ql = rhogq[al];
      Th t = thermo_v(rhog, rhoge, q0);
      diag[I_VX * NF + a] = v1 / rhog;
      diag[I_VY * NF + a] = v2 / rhog;
      diag[I_VZ * NF + a] = v3 / rhog;
      tem[a] = t.tem;
      pre[a] = t.pre;
      if (k > kmin) {
        w[a] = rwa / (G.c2w0[k] * rhog + G.c2w1[k] * rgl);
      } else {
        // bottom: rigid lid, with the density of the ghost row after the boundary condition
        const Th t1 = thermo_v(rgl, rgel, ql);
        const real pre_b = t1.pre - t.rho * (G.phi[kmin - 1] - G.phi[kmin + 1]);
        const real rho_b = pre_b / (c.Rdry * t.tem);
        rw[a] = real(0.0);
        w[a] = real(0.0) / (G.c2w0[kmin] * rhog + G.c2w1[kmin] * (rho_b * real(1.0)));
      }
      eth[a] = t.ein + t.pre / t.rho;
      return;
    }
    // ghost rows: TEM thermodynamic condition, free-slip top, rigid bottom
    const bool top = (k == kmax + 1);
    const long an = top ? a - ks : a + ks;            // nearest physical level
    const long an2 = top ? a - 2 * ks : a + 2 * ks;    // next one
    const Th tn = thermo(an);
    const Th tn2 = thermo(an2);
    const real tem_g = tn.tem;
    const real pre_g = top ? tn2.pre - tn.rho * (G.phi[kmax + 1] - G.phi[kmax - 1])
                             : tn2.pre - tn.rho * (G.phi[kmin - 1] - G.phi[kmin + 1]);
    const real rho_g = pre_g / (c.Rdry * tem_g);
    const real rg_g = rho_g * real(1.0);
    const real ein_g = c.CVdry * tem_g;
    tem[a] = tem_g;
    pre[a] = pre_g;
    rg[a] = rg_g;
    prog[5 * NF + a] = rg_g * ein_g;
    const real sgn = top ? real(1.0) : -real(1.0);
    for (int q = 0; q < 3; ++q) {
      const real v = sgn * (prog[(1 + q) * NF + an] / rg[an]) * rg_g;
      prog[(1 + q) * NF + a] = v;
      diag[(I_VX + q) * NF + a] = v / rg_g;
    }
    if (top) {
      rw[a] = -real(0.0);
      w[a] = rw[a] / (G.c2w0[kmax + 1] * rg_g + G.c2w1[kmax + 1] * rg[an]);
    } else {
      rw[a] = real(0.0);
      w[a] = real(0.0);
    }
    eth[a] = ein_g + pre_g / rho_g;
  }
};

// ----------------------------------------------------------------------------
// momentum advection, Coriolis, projection (mod_src.py:96-497, kernels/advconvmom.py)
// out[0..2]: momentum tendencies; out[3]: the radial part gwc at full levels
// (the half-level w tendency c2w0*gwc(k) + c2w1*gwc(k-1) is formed by its user)
// ----------------------------------------------------------------------------
struct AdvMomP {
  static constexpr const char* kName = "adv_mom";
  // fields read on the stencil: vx, vy, vz, w(k), w(k+1), rhogvx, rhogvy, rhogvz
  static constexpr int kNStage = 8;
  static constexpr bool kStageF32 = false;
  Dims d; Consts c; Geo G;
  const real* prog;
  const real* diag;
  real* out;
  // field f of the stencil set, positioned so that the point's own level reads it
  NH_HD const real* stage_ptr(int f, int) const {
    const long NF = d.NF;
    switch (f) {
      case 0: return diag + I_VX * NF;
      case 1: return diag + I_VY * NF;
      case 2: return diag + I_VZ * NF;
      case 3: return diag + I_W * NF;
      case 4: return diag + I_W * NF + (long)d.g * d.g;   // next level of the region
      default: return prog + (long)(f - 4) * NF;
    }
  }
  NH_HD bool stage_level(int k) const { return k >= d.kmin && k <= d.kmax; }
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (!pt_of(d, plane, x, p)) return;
    if (!pt_compute(p)) {
      for (int q = 0; q < 4; ++q) out[q * d.NF + p.a] = real(0.0);
      return;
    }
    const long ks = p.ks;
    at(p, [&](int f, long b) { return f == 4 ? diag[I_W * d.NF + b + ks] : stage_ptr(f, p.k)[b]; });
  }
  template <class LD>
  NH_HD void at(const Pt& p, LD ld) const {
    const long NF = d.NF, a = p.a;
    const int kmin = d.kmin, kmax = d.kmax, k = p.k;
    if (k < kmin || k > kmax) {
      for (int q = 0; q < 4; ++q) out[q * NF + a] = real(0.0);
      return;
    }
    // separate instances for the 6-point pole and 7-point regional stencils:
    // fixed trip counts keep the per-neighbour arrays in registers
    if (p.pole) point<true>(p, ld);
    else point<false>(p, ld);
  }
  template <bool POLE, class LD>
  NH_HD void point(const Pt& p, LD ld) const {
    constexpr int NN = POLE ? 6 : 7;
    const long NF = d.NF, a = p.a, ks = p.ks;
    const int kmin = d.kmin, kmax = d.kmax, k = p.k;
    const real R = c.RADIUS;
    const real* rw = prog + 4 * NF;
    const real* rg = prog;
    const real* vxd = diag + I_VX * NF;
    const real* vyd = diag + I_VY * NF;
    const real* vzd = diag + I_VZ * NF;
    const real* wd = diag + I_W * NF;
    const long g = d.g;
    // stencil offset of neighbour m: (di, dj) + 1 packed two bits per point,
    // (0,0) (+1,0) (+1,+1) (0,+1) (-1,0) (-1,-1) (0,-1)
    auto off = [&](int m) -> long {
      const int di = ((4201 >> (2 * m)) & 3) - 1, dj = ((421 >> (2 * m)) & 3) - 1;
      return di + dj * g;
    };
    // address and position of neighbour m
    auto nbr = [&](int m) { return POLE ? a + m : a + off(m); };
    auto pos = [&](int m, real* x) {
      if (POLE) {
        for (int q = 0; q < 3; ++q) x[q] = G.gx_pl[(m * 2 + p.lp) * 3 + q];
      } else {
        for (int q = 0; q < 3; ++q) x[q] = G.gx[q * d.nreg + p.ci + off(m)];
      }
    };
    // merged velocity v + w_full r/|r| at address b (level kk) with position x
    auto vv3 = [&](long b, int kk, const real* x, real* o) {
      const real wc = G.cfact[kk] * wd[b + ks] + G.dfact[kk] * wd[b];
      o[0] = vxd[b] + wc * x[0] / R;
      o[1] = vyd[b] + wc * x[1] / R;
      o[2] = vzd[b] + wc * x[2] / R;
    };
    real x0[3];
    pos(0, x0);
    // divergence of rhogv * vv_q, one neighbour at a time: per direction and
    // component the sum over the neighbours runs in stencil order, as in
    // div7 / div_pl, without keeping every neighbour's values live
    real vvm0[3];
    real sd[3][3];   // [direction][q] (regional); the pole sums the directions per point
    real div[3];
#ifdef NH_ADV_UNROLL
#pragma unroll
#else
#pragma unroll 1
#endif
    for (int m = 0; m < NN; ++m) {
      real xm[3], vv[3];
      pos(m, xm);
      const long bm = nbr(m);
      {
        const real wc = G.cfact[k] * ld(4, bm) + G.dfact[k] * ld(3, bm);
        vv[0] = ld(0, bm) + wc * xm[0] / R;
        vv[1] = ld(1, bm) + wc * xm[1] / R;
        vv[2] = ld(2, bm) + wc * xm[2] / R;
      }
      if (m == 0)
        for (int q = 0; q < 3; ++q) vvm0[q] = vv[q];
      const real fx0 = ld(5, bm), fy0 = ld(6, bm), fz0 = ld(7, bm);
      if (POLE) {
        const real* cc = G.cdiv_pl + (m * 2 + p.lp) * 3;
        for (int q = 0; q < 3; ++q) {
          const real pv = cc[0] * (fx0 * vv[q] * real(1.0)) + cc[1] * (fy0 * vv[q] * real(1.0)) +
                          cc[2] * (fz0 * vv[q] * real(1.0));
          div[q] = (m == 0) ? pv : div[q] + pv;
        }
      } else {
        const real cx = rc(G.cdiv, d.nreg, m, p.ci), cy = rc(G.cdiv, d.nreg, 7 + m, p.ci),
                   cz = rc(G.cdiv, d.nreg, 14 + m, p.ci);
        for (int q = 0; q < 3; ++q) {
          const real tx = cx * (fx0 * vv[q] * real(1.0));
          const real ty = cy * (fy0 * vv[q] * real(1.0));
          const real tz = cz * (fz0 * vv[q] * real(1.0));
          sd[0][q] = (m == 0) ? tx : sd[0][q] + tx;
          sd[1][q] = (m == 0) ? ty : sd[1][q] + ty;
          sd[2][q] = (m == 0) ? tz : sd[2][q] + tz;
        }
      }
    }
    if (!POLE)
      for (int q = 0; q < 3; ++q) div[q] = sd[0][q] + sd[1][q] + sd[2][q];
    // vertical fluxes at the half levels k+1 and k (zero at kmin and kmax+1)
    real hi[3] = {real(0.0), real(0.0), real(0.0)}, lo[3] = {real(0.0), real(0.0), real(0.0)};
    if (k + 1 <= kmax) {
      real up[3];
      vv3(a + ks, k + 1, x0, up);
      const real wup = rw[a + ks];
      for (int q = 0; q < 3; ++q) hi[q] = real(0.0) * real(1.0) + wup * (G.afact[k + 1] * up[q] + G.bfact[k + 1] * vvm0[q]) * real(1.0);
    }
    if (k >= kmin + 1) {
      real dn[3];
      vv3(a - ks, k - 1, x0, dn);
      const real w0 = rw[a];
      for (int q = 0; q < 3; ++q) lo[q] = real(0.0) * real(1.0) + w0 * (G.afact[k] * vvm0[q] + G.bfact[k] * dn[q]) * real(1.0);
    }
    real dvv[3];
    for (int q = 0; q < 3; ++q) dvv[q] = -div[q] - (hi[q] - lo[q]) * G.rdgz[k];
    const real gxc = x0[0] / R, gyc = x0[1] / R, gzc = x0[2] / R;
    const real rh = rg[a];
    const real dx = dvv[0] - (-real(2.0) * rh * (c.OHM * vvm0[1]));
    const real dy = dvv[1] - (real(2.0) * rh * (c.OHM * vvm0[0]));
    const real dz = dvv[2];
    const real prd = dx * gxc + dy * gyc + dz * gzc;
    out[0 * NF + a] = dx - prd * gxc;
    out[1 * NF + a] = dy - prd * gyc;
    out[2 * NF + a] = dz - prd * gzc;
    out[3 * NF + a] = prd * c.ALPHA;
  }
};

// ----------------------------------------------------------------------------
// hyperdiffusion passes (mod_numfilter.py:1125-1264)
// ----------------------------------------------------------------------------
template <int LAST>
struct HDiffPassP {
  static constexpr const char* kName = LAST ? "hdiff_pass2" : "hdiff_pass1";
  static constexpr int kNStage = 6;    // the six filtered fields (first pass)
  static constexpr bool kStageF32 = false;
  Dims d; Consts c; Geo G;
  const real* in[6];
  real* out[6];
  const real* rhog;
  // first pass: FP64 k-profile subtracted from input f (tem - tem_bs,
  // rho - rho_bs, mod_numfilter.py:1068-1076), or none
  const double* sub[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  NH_HD const real* stage_ptr(int f, int) const { return in[f]; }
  NH_HD bool stage_level(int) const { return true; }
  NH_HD real input(int f, long b, int k) const {
    return sub[f] ? (real)((double)in[f][b] - sub[f][k]) : in[f][b];
  }
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (!pt_of(d, plane, x, p) || !pt_compute(p)) return;
    if (!LAST) {
      at(p, [&](int f, long b) { return input(f, b, p.k); });
      return;
    }
    const long a = p.a;
    const Cell cl = cell_of_pt(p);
    for (int q = 0; q < 6; ++q) {
      real r;
      if (q >= 4) {
        const real fac = (q == 4) ? c.CVdry : c.HDIFF_FACT_RHO;
        const real* rg = rhog;
        const real gam = c.GAMMA_H;
        auto kh = [rg, fac, gam](long xx) { return rg[xx] * fac * gam; };
        r = p.pole ? diffusion_pl(G, p.lp, in[q], kh, a) : diffusion7(G, d, cl, in[q], kh, a);
      } else {
        r = p.pole ? lap_pl(G, p.lp, in[q], a) : lap7(G, d.nreg, p.ci, in[q], a, d.g);
      }
      out[q][a] = -r;
    }
  }
  // first pass: minus the Laplacian of each field, fields read through ld(f, address)
  template <class LD>
  NH_HD void at(const Pt& p, LD ld) const {
    const long a = p.a;
    for (int q = 0; q < 6; ++q) {
      real r;
      if (p.pole) {
        r = real(0.0);
        for (int v = 0; v < 6; ++v) {
          const real pv = G.clap_pl[v * 2 + p.lp] * ld(q, a + v);
          r = (v == 0) ? pv : r + pv;
        }
      } else {
        const long g = d.g;
        r = rc(G.clap, d.nreg, 0, p.ci) * ld(q, a);
        for (int m = 1; m < 7; ++m) {
          const long off = (((4201 >> (2 * m)) & 3) - 1) + (((421 >> (2 * m)) & 3) - 1) * g;
          r = r + rc(G.clap, d.nreg, m, p.ci) * ld(q, a + off);
        }
      }
      out[q][a] = -r;
    }
  }
};

// ----------------------------------------------------------------------------
// 3-D divergence damping (kernels/oprt3ddivdamp.py, oprtdivdamp.py, divdamppostcomm.py)
// ----------------------------------------------------------------------------
struct DivDamp1P {
  static constexpr const char* kName = "divdamp_3d";
  Dims d; Consts c; Geo G;
  const real *rvx, *rvy, *rvz, *rw;
  real* out[3];
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (!pt_of(d, plane, x, p) || !pt_compute(p)) return;
    const int kmin = d.kmin, kmax = d.kmax, g = d.g, k = p.k;
    const long ks = p.ks, a = p.a;
    if (k < kmin || k > kmax) {
      for (int q = 0; q < 3; ++q) out[q][a] = real(0.0);
      return;
    }
    auto rwv = [&](long aa, int kk) { return (kk >= kmin + 1 && kk <= kmax) ? (real(0.0) * real(1.0) + rw[aa] * real(1.0)) : real(0.0); };
    if (p.pole) {
      real sclt[5];
      const real* vv[3] = {rvx, rvy, rvz};
      for (int v = 1; v <= 5; ++v) {
        const int vp = (v == 5) ? 1 : v + 1;
        real s = real(0.0);
        for (int dd = 0; dd < 3; ++dd) {
          const real* cI = G.cintp_pl + ((v * 2 + p.lp) * 3 + dd) * 3;
          const real t0 = cI[0] * (vv[dd][a] * real(1.0)), t1 = cI[1] * (vv[dd][a + v] * real(1.0)),
                       t2 = cI[2] * (vv[dd][a + vp] * real(1.0));
          s = (dd == 0) ? (t0 + t1 + t2) : (s + t0 + t1 + t2);
        }
        const real up = (rwv(a + ks, k + 1) + rwv(a + v + ks, k + 1) + rwv(a + vp + ks, k + 1));
        const real dn = (rwv(a, k) + rwv(a + v, k) + rwv(a + vp, k));
        sclt[v - 1] = s + (up - dn) / real(3.0) * G.rdgz[k];
      }
      for (int dd = 0; dd < 3; ++dd) {
        real row = real(0.0);
        for (int v = 1; v <= 5; ++v) {
          const int vm = (v == 1) ? 5 : v - 1;
          const real cb = G.cdiff_pl[(v * 2 + p.lp) * 3 + dd] * (sclt[vm - 1] + sclt[v - 1]);
          row = (v == 1) ? cb : row + cb;
        }
        out[dd][a] = row;
      }
      return;
    }
    const long nreg = d.nreg, ci = p.ci;
    const long co[6] = {0, 0, -1, -1 - (long)g, -1 - (long)g, -(long)g};
    const int ty[6] = {0, 1, 0, 1, 0, 1};
    real T[6];
    for (int n = 0; n < 6; ++n) {
      const long aa = a + co[n];
      const long cj = ci + co[n];
      const long o1 = (ty[n] == 0) ? 1 : 1 + g;
      const long o2 = (ty[n] == 0) ? 1 + g : g;
      const real s = tri_div(G, nreg, cj, ty[n], rvx, rvy, rvz, aa, g, real(1.0));
      const real up = (rwv(aa + ks, k + 1) + rwv(aa + o1 + ks, k + 1) + rwv(aa + o2 + ks, k + 1));
      const real dn = (rwv(aa, k) + rwv(aa + o1, k) + rwv(aa + o2, k));
      T[n] = s + (up - dn) / real(3.0) * G.rdgz[k];
    }
    if (p.i == 1 && p.j == 1) {
      const real m0 = G.sgp[p.l] ? real(0.0) : real(1.0), m1 = G.sgp[p.l] ? real(1.0) : real(0.0);
      T[4] = T[4] * m0 + T[5] * m1;
    }
    for (int dd = 0; dd < 3; ++dd) out[dd][a] = hex_sum(G, nreg, ci, dd, T);
  }
};

struct DivDamp2P {
  static constexpr const char* kName = "divdamp_post";
  Dims d; Consts c; Geo G;
  const real* in[3];
  real* out[3];
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (!pt_of(d, plane, x, p)) return;
    const long a = p.a;
    const int kmin = d.kmin, kmax = d.kmax, g = d.g, k = p.k;
    if (!pt_compute(p)) {
      for (int q = 0; q < 3; ++q) out[q][a] = real(0.0);
      return;
    }
    const real R = c.RADIUS;
    real o[3];
    if (k < kmin || k > kmax) {
      o[0] = o[1] = o[2] = real(0.0);
    } else if (p.pole) {
      real sclt[5];
      for (int v = 1; v <= 5; ++v) {
        const int vp = (v == 5) ? 1 : v + 1;
        real s = real(0.0);
        for (int dd = 0; dd < 3; ++dd) {
          const real* cI = G.cintp_pl + ((v * 2 + p.lp) * 3 + dd) * 3;
          const real t0 = cI[0] * (-in[dd][a]), t1 = cI[1] * (-in[dd][a + v]), t2 = cI[2] * (-in[dd][a + vp]);
          s = (dd == 0) ? (t0 + t1 + t2) : (s + t0 + t1 + t2);
        }
        sclt[v - 1] = s;
      }
      for (int dd = 0; dd < 3; ++dd) {
        real row = real(0.0);
        for (int v = 1; v <= 5; ++v) {
          const int vm = (v == 1) ? 5 : v - 1;
          const real cb = G.cdiff_pl[(v * 2 + p.lp) * 3 + dd] * (sclt[vm - 1] + sclt[v - 1]);
          row = (v == 1) ? cb : row + cb;
        }
        o[dd] = row;
      }
    } else {
      const long nreg = d.nreg, ci = p.ci;
      const long co[6] = {0, 0, -1, -1 - (long)g, -1 - (long)g, -(long)g};
      const int ty[6] = {0, 1, 0, 1, 0, 1};
      real T[6];
      for (int n = 0; n < 6; ++n) T[n] = tri_div(G, nreg, ci + co[n], ty[n], in[0], in[1], in[2], a + co[n], g, -real(1.0));
      for (int dd = 0; dd < 3; ++dd) o[dd] = hex_sum(G, nreg, ci, dd, T);
    }
    const real gd[3] = {c.ALPHA_D * o[0], c.ALPHA_D * o[1], c.ALPHA_D * o[2]};
    if (p.pole) {
      const real gx = G.gx_pl[p.lp * 3 + 0], gy = G.gx_pl[p.lp * 3 + 1], gz = G.gx_pl[p.lp * 3 + 2];
      const real prd = gd[0] * gx / R + gd[1] * gy / R + gd[2] * gz / R;
      out[0][a] = gd[0] - prd * gx / R;
      out[1][a] = gd[1] - prd * gy / R;
      out[2][a] = gd[2] - prd * gz / R;
    } else {
      const real gx = G.gx[p.ci], gy = G.gx[d.nreg + p.ci], gz = G.gx[2 * d.nreg + p.ci];
      const real prd = (gd[0] * gx + gd[1] * gy + gd[2] * gz) / R;
      out[0][a] = gd[0] - prd * gx / R;
      out[1][a] = gd[1] - prd * gy / R;
      out[2][a] = gd[2] - prd * gz / R;
    }
  }
};

// ----------------------------------------------------------------------------
// large-step source terms of the small steps and the tridiagonal coefficients
// (mod_vi.py:132-881, kernels/vimatrix.py:67-136); column quantities of levels
// k-1..k+1 are recomputed per point
// ----------------------------------------------------------------------------
struct ViPath0P {
  static constexpr const char* kName = "vi_path0";
  Dims d; Consts c; Geo G;
  double dt;
  const real* prog;
  const real* diag;
  const real* eth;
  const real* gadv;   // 3 momentum tendencies + gwc at full levels
  const real* vt[6];
  const real* dd[3];
  real* gT;
  double* Mc; double* Mu; double* Ml;   // FP64 in every form
  real* getot;
  real* frhog;
  // fields read on the stencil: rhogvx, rhogvy, rhogvz, eth at the point's
  // level, the pressure at the level of its gradient (kmin -> kmin+1)
  static constexpr int kNStage = 5;
  static constexpr bool kStageF32 = false;
  NH_HD const real* stage_ptr(int f, int k) const {
    const long NF = d.NF;
    if (f < 3) return prog + (long)(1 + f) * NF;
    if (f == 3) return eth;
    return diag + I_PRE * NF + (k == d.kmin ? (long)d.g * d.g : 0L);
  }
  NH_HD bool stage_level(int k) const { return k >= d.kmin && k <= d.kmax; }
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (!pt_of(d, plane, x, p) || !pt_compute(p)) return;
    const long sh = (p.k == d.kmin) ? p.ks : 0;   // the level stride differs at the poles
    at(p, [&](int f, long b) { return f < 4 ? stage_ptr(f, p.k)[b] : diag[I_PRE * d.NF + b + sh]; });
  }
  template <class LD>
  NH_HD void at(const Pt& p, LD ld) const {
    const long NF = d.NF, ks = p.ks, a = p.a;
    const int kmin = d.kmin, kmax = d.kmax, k = p.k;
    const real R = c.RADIUS;
    const real* rg = prog;
    const real* rw = prog + 4 * NF;
    const real* pre = diag + I_PRE * NF;
    const long b = a - (long)k * ks;     // column base
    auto lev = [&](int kk) { return b + (long)kk * ks; };
    // the column window k-2 .. k+1 of rhog, pressure, eth and rhogw, loaded
    // first at levels clamped into range (a clamped value is never used: the
    // level tests below are the reference's), read through selects
    real Wrg[4], Wpr[4], Wet[4], Wrw[4];
    for (int jw = 0; jw < 4; ++jw) {
      int kw = k - 2 + jw;
      kw = (kw < 0) ? 0 : ((kw > d.kall - 1) ? d.kall - 1 : kw);
      const long aw = lev(kw);
      Wrg[jw] = rg[aw];
      Wpr[jw] = pre[aw];
      Wet[jw] = eth[aw];
      Wrw[jw] = rw[aw];
    }
    auto win = [&](const real* W, int kk) {
      const int jw = kk - k + 2;
      return (jw <= 0) ? W[0] : (jw == 1) ? W[1] : (jw == 2) ? W[2] : W[3];
    };
    auto RG = [&](int kk) { return win(Wrg, kk); };
    auto PR = [&](int kk) { return win(Wpr, kk); };
    // pregd = (pre - pre_bs) GSGAM2, rhogd = (rho - rho_bs) GSGAM2 with the basic
    // state in FP64 (mod_dynamics.py:1056-1057); GSGAM2 = 1
    auto PRD = [&](int kk) { return (real)((double)PR(kk) - G.pre_bs[kk]) * real(1.0); };
    auto RGD = [&](int kk) { return (real)((double)(RG(kk) / real(1.0)) - G.rho_bs[kk]) * real(1.0); };
    auto ET = [&](int kk) { return win(Wet, kk); };
    auto RW = [&](int kk) { return win(Wrw, kk); };
    auto rhog_h = [&](int kk) {
      const int k2 = (kk < kmin) ? kmin : kk;
      return G.c2w0[k2] * RG(k2) + G.c2w1[k2] * RG(k2 - 1);
    };
    auto dpgw = [&](int kk) {
      return (kk >= kmin + 1 && kk <= kmax) ? real(1.0) * (PRD(kk) * real(1.0) - PRD(kk - 1) * real(1.0)) * G.rdgzh[kk] : real(0.0);
    };
    auto dbuo = [&](int kk) {
      return (kk >= kmin + 1 && kk <= kmax)
                 ? -c.GRAV * (G.c2w0[kk] * RGD(kk) + G.c2w1[kk] * RGD(kk - 1))
                 : real(0.0);
    };
    auto gzt = [&](int kk) { return c.GRAV - (dpgw(kk) - dbuo(kk)) / rhog_h(kk); };
    auto pwh = [&](int kk) { return -gzt(kk) * RW(kk); };
    real gx, gy, gz;
    if (p.pole) {
      gx = G.gx_pl[p.lp * 3 + 0]; gy = G.gx_pl[p.lp * 3 + 1]; gz = G.gx_pl[p.lp * 3 + 2];
    } else {
      gx = G.gx[p.ci]; gy = G.gx[d.nreg + p.ci]; gz = G.gx[2 * d.nreg + p.ci];
    }
    const real Kh = c.GAMMA_H;
    // hyperdiffusion tendency + momentum advection = large-step tendency g0
    const real rgk = RG(k);
    real fvx = -(vt[0][a] * Kh + real(0.0)) * rgk;
    real fvy = -(vt[1][a] * Kh + real(0.0)) * rgk;
    real fvz = -(vt[2][a] * Kh + real(0.0)) * rgk;
    const real rgh = (k >= kmin) ? G.c2w0[k] * rgk + G.c2w1[k] * RG(k - 1) : real(0.0);
    const real fw = -(vt[3][a] * Kh + real(0.0)) * rgh;
    const real fe = -(vt[4][a] + real(0.0));
    const real fr = -(vt[5][a] + real(0.0));
    if (p.pole) {
      const real prd = fvx * gx / R + fvy * gy / R + fvz * gz / R;
      fvx = fvx - prd * gx / R; fvy = fvy - prd * gy / R; fvz = fvz - prd * gz / R;
    } else {
      const real prd = (fvx * gx + fvy * gy + fvz * gz) / R;
      fvx = fvx - prd * gx / R; fvy = fvy - prd * gy / R; fvz = fvz - prd * gz / R;
    }
    const real gadv_w = (k > kmin && k <= kmax) ? (G.c2w0[k] * gadv[3 * NF + a] + G.c2w1[k] * gadv[3 * NF + a - ks]) : real(0.0);
    real g0[6];
    g0[0] = real(0.0) + fr;
    g0[1] = gadv[0 * NF + a] + fvx;
    g0[2] = gadv[1 * NF + a] + fvy;
    g0[3] = gadv[2 * NF + a] + fvz;
    g0[4] = gadv_w + fw;
    g0[5] = real(0.0) + fe;
    getot[a] = g0[5];
    frhog[a] = g0[0];
    real drhog = real(0.0), drhoge = real(0.0), pw = real(0.0);
    real dpg[3] = {real(0.0), real(0.0), real(0.0)};
    if (k >= kmin && k <= kmax) {
      // div(rhogv), div(rhogv eth)
      real dvs[2];
      hdiv_multi<2>(G, d, p, [&](int, long bn, real (*v)[3]) {
        const real e = ld(3, bn);
        const real r3[3] = {ld(0, bn), ld(1, bn), ld(2, bn)};
        for (int q = 0; q < 3; ++q) {
          v[0][q] = r3[q] * real(1.0);
          v[1][q] = r3[q] * e * real(1.0);
        }
      }, dvs);
      const real hd = dvs[0];
      const real vmh_hi = (k + 1 <= kmax) ? real(0.0) * real(1.0) + real(1.0) * RW(k + 1) * real(1.0) : real(0.0);
      const real vmh_lo = (k >= kmin + 1) ? real(0.0) * real(1.0) + real(1.0) * RW(k) * real(1.0) : real(0.0);
      drhog = -hd - (vmh_hi - vmh_lo) * G.rdgz[k];
      const real he = dvs[1];
      const real ws_hi = (k + 1 <= kmax)
          ? real(0.0) * real(1.0) + real(1.0) * (RW(k + 1) * (G.afact[k + 1] * ET(k + 1) + G.bfact[k + 1] * ET(k))) * real(1.0) : real(0.0);
      const real ws_lo = (k >= kmin + 1)
          ? real(0.0) * real(1.0) + real(1.0) * (RW(k) * (G.afact[k] * ET(k) + G.bfact[k] * ET(k - 1))) * real(1.0) : real(0.0);
      drhoge = -he - (ws_hi - ws_lo) * G.rdgz[k];
      // the gradient of pregd at the level the pressure is read from (kmin -> kmin+1)
      const int kg = (k == kmin) ? kmin + 1 : k;
      presgrad_ld(p, [&](int f, long b) {
        return f == 4 ? (real)((double)ld(4, b) - G.pre_bs[kg]) * real(1.0) : ld(f, b);
      }, R, dpg);
      pw = diag[I_VX * NF + a] * dpg[0] + diag[I_VY * NF + a] * dpg[1] + diag[I_VZ * NF + a] * dpg[2]
         + G.w2c0[k] * pwh(k + 1) + G.w2c1[k] * pwh(k);
    }
    gT[0 * NF + a] = g0[0] + drhog;
    gT[1 * NF + a] = g0[1] - dpg[0] + dd[0][a] + real(0.0);
    gT[2 * NF + a] = g0[2] - dpg[1] + dd[1][a] + real(0.0);
    gT[3 * NF + a] = g0[3] - dpg[2] + dd[2][a] + real(0.0);
    gT[4 * NF + a] = g0[4] + real(0.0) * c.ALPHA - dpgw(k) + dbuo(k);
    gT[5 * NF + a] = g0[5] + drhoge + pw;
    if (k >= kmin + 1 && k <= kmax) {
      // tridiagonal coefficients, FP64 in every form: the column quantities of
      // levels k-1..k+1 again, from the stored fields, in double
      const VProf<double>& V = G.v64;
      const ConstsT<double>& C = c.d;
      auto rhog_h64 = [&](int kk) {
        const int k2 = (kk < kmin) ? kmin : kk;
        return V.c2w0[k2] * (double)RG(k2) + V.c2w1[k2] * (double)RG(k2 - 1);
      };
      auto eth_h64 = [&](int kk) {
        const int k2 = (kk < kmin) ? kmin : kk;
        return V.afact[k2] * (double)ET(k2) + V.bfact[k2] * (double)ET(k2 - 1);
      };
      auto dpgw64 = [&](int kk) {
        return (kk >= kmin + 1 && kk <= kmax)
                   ? 1.0 * (((double)PR(kk) - G.pre_bs[kk]) * 1.0 - ((double)PR(kk - 1) - G.pre_bs[kk - 1]) * 1.0)
                         * V.rdgzh[kk]
                   : 0.0;
      };
      auto dbuo64 = [&](int kk) {
        return (kk >= kmin + 1 && kk <= kmax)
                   ? -C.GRAV * (V.c2w0[kk] * (((double)RG(kk) / 1.0 - G.rho_bs[kk]) * 1.0) +
                                V.c2w1[kk] * (((double)RG(kk - 1) / 1.0 - G.rho_bs[kk - 1]) * 1.0))
                   : 0.0;
      };
      auto gzt64 = [&](int kk) { return C.GRAV - (dpgw64(kk) - dbuo64(kk)) / rhog_h64(kk); };
      const double GCVovR = C.GRAV * C.CVdry / C.Rdry;
      const double ACVovRt2 = C.ALPHA * C.CVdry / C.Rdry / (dt * dt);
      const double Ca = (1.0 * V.rdgz[k] + 1.0 * V.rdgz[k - 1]) * 1.0;
      const double Cb = V.dfact[k] - V.cfact[k - 1];
      const double Cu1 = 1.0 * V.rdgz[k] * 1.0;
      const double Cu2 = 1.0 * 1.0 * GCVovR;
      const double Cl1 = 1.0 * V.rdgz[k] * 1.0;
      const double Cl2 = 1.0 * 1.0 * GCVovR;
      Mc[a] = ACVovRt2 / 1.0 + V.rdgzh[k] * (Ca * eth_h64(k) - Cb * (gzt64(k) + GCVovR));
      Mu[a] = -V.rdgzh[k] * (Cu1 * eth_h64(k + 1) + V.cfact[k] * (gzt64(k + 1) + Cu2));
      Ml[a] = -V.rdgzh[k] * (Cl1 * eth_h64(k - 1) - V.dfact[k - 1] * (gzt64(k - 1) + Cl2));
    }
  }
  // horizontal pressure-gradient force at the point with the first-layer
  // remedy (presgrad_h), the pressure of the gradient's level read as ld(4, .)
  template <class LD>
  NH_HD void presgrad_ld(const Pt& p, LD ld, real R, real o[3]) const {
    const long a = p.a;
    real gr[3];
    if (p.pole) {
      for (int dd2 = 0; dd2 < 3; ++dd2) {
        real acc = real(0.0);
        for (int v = 0; v < 6; ++v) {
          const real pv = G.cgrad_pl[(v * 2 + p.lp) * 3 + dd2] * (ld(4, a + v) * real(1.0) * real(1.0));
          acc = (v == 0) ? pv : acc + pv;
        }
        gr[dd2] = acc;
      }
    } else {
      const long g = d.g;
      for (int dd2 = 0; dd2 < 3; ++dd2) {
        real acc = rc(G.cgrad, d.nreg, dd2 * 7 + 0, p.ci) * (ld(4, a) * real(1.0));
        for (int m = 1; m < 7; ++m) {
          const long off = (((4201 >> (2 * m)) & 3) - 1) + (((421 >> (2 * m)) & 3) - 1) * g;
          acc = acc + rc(G.cgrad, d.nreg, dd2 * 7 + m, p.ci) * (ld(4, a + off) * real(1.0));
        }
        gr[dd2] = acc;
      }
    }
    for (int q = 0; q < 3; ++q) gr[q] = gr[q] + real(0.0);
    real gx, gy, gz;
    if (p.pole) {
      gx = G.gx_pl[(0 * 2 + p.lp) * 3 + 0];
      gy = G.gx_pl[(0 * 2 + p.lp) * 3 + 1];
      gz = G.gx_pl[(0 * 2 + p.lp) * 3 + 2];
      const real prd = gr[0] * gx / R + gr[1] * gy / R + gr[2] * gz / R;
      o[0] = gr[0] - prd * gx / R;
      o[1] = gr[1] - prd * gy / R;
      o[2] = gr[2] - prd * gz / R;
    } else {
      gx = G.gx[p.ci]; gy = G.gx[d.nreg + p.ci]; gz = G.gx[2 * d.nreg + p.ci];
      const real prd = (gr[0] * gx + gr[1] * gy + gr[2] * gz) / R;
      o[0] = gr[0] - prd * gx / R;
      o[1] = gr[1] - prd * gy / R;
      o[2] = gr[2] - prd * gz / R;
    }
  }
};

// total specific energy of the large-step state and the start of the time-mean fluxes
struct EthTotPMP {
  static constexpr const char* kName = "ethtot_mean";
  Dims d; Consts c; Geo G;
  const real* prog;
  const real* eth;
  real* ethtot;
  real* pm;
  // the column values are read first (levels k-1, k+1 clamped into range;
  // rhogkin_k uses them only where they exist), then the stores
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (!pt_of(d, plane, x, p)) return;
    const long NF = d.NF, a = p.a, ks = p.ks;
    const int k = p.k;
    const long am = (k > 0) ? a - ks : a, ap = (k < d.kall - 1) ? a + ks : a;
    const real rg[3] = {prog[am], prog[a], prog[ap]};
    const real rv[3] = {prog[NF + a], prog[2 * NF + a], prog[3 * NF + a]};
    const real rw[2] = {prog[4 * NF + a], prog[4 * NF + ap]};
    const real e = eth[a];
    auto get = [&](int kk, int q) {
      if (q == 0) return rg[kk - k + 1];
      if (q == 4) return rw[kk - k];
      return rv[q - 1];
    };
    const real kin0 = rhogkin_k(G, d.kmin, d.kmax, k, get);
    ethtot[a] = e + kin0 / rg[1] + G.phi[k];
    if (pm) {   // null: the mean fluxes are not wanted (adjoint recompute)
      pm[a] = rg[1];
      for (int q = 0; q < 3; ++q) pm[(1 + q) * NF + a] = rv[q];
      pm[4 * NF + a] = rw[0];
    }
  }
};

// ----------------------------------------------------------------------------
// small step path1 (kernels/vipath1.py:68-160); the ghost rows recompute the
// nearest physical level
// ----------------------------------------------------------------------------
struct Path1P {
  static constexpr const char* kName = "vi_path1";
  // the split rhoge is read on the stencil, at the level whose pressure
  // gradient the point needs (stage_shift levels above the point's own)
  static constexpr int kNStage = 1;
  static constexpr bool kStageF32 = true;
  Dims d; Consts c; Geo G;
  real dt, rweight;
  const real* prog;
  const real* ps[6];
  const real* gT;
  const real* dd[3];
  real* dv[3];
  real* pm;
  // level of the gradient for a point at level k: kmin -> kmin+1 (first-layer
  // remedy); the ghost rows use the nearest physical level's value, kmin-1 ->
  // kmin (gradient at kmin+1) and kmax+1 -> kmax
  NH_HD int stage_shift(int k) const {
    const int kmin = d.kmin, kmax = d.kmax;
    if (k == kmin || k == kmin - 1) return kmin + 1 - k;
    if (k == kmax + 1) return -1;
    return 0;
  }
  NH_HD const real* stage_ptr(int, int k) const { return ps[I_RHOGE] + (long)stage_shift(k) * d.g * d.g; }
  NH_HD bool stage_level(int) const { return true; }
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (!pt_of(d, plane, x, p) || !pt_compute(p)) return;
    const long sh = (long)stage_shift(p.k) * p.ks;
    at(p, [&](int, long b) { return ps[I_RHOGE][b + sh]; });
  }
  // momentum of the small step at level k (column address a); the gradient of
  // the split pressure comes through grad(m, dir)
  template <class GR>
  NH_HD void level(const Pt& p, long a, GR grad, real out[3]) const {
    const long NF = d.NF;
    const real R = c.RADIUS;
    real gr[3];
    for (int dd2 = 0; dd2 < 3; ++dd2) gr[dd2] = grad(dd2);
    for (int q = 0; q < 3; ++q) gr[q] = gr[q] + real(0.0);
    real Pg[3];
    if (p.pole) {
      const real gx = G.gx_pl[p.lp * 3 + 0], gy = G.gx_pl[p.lp * 3 + 1], gz = G.gx_pl[p.lp * 3 + 2];
      const real prd = gr[0] * gx / R + gr[1] * gy / R + gr[2] * gz / R;
      Pg[0] = gr[0] - prd * gx / R; Pg[1] = gr[1] - prd * gy / R; Pg[2] = gr[2] - prd * gz / R;
    } else {
      const real gx = G.gx[p.ci], gy = G.gx[d.nreg + p.ci], gz = G.gx[2 * d.nreg + p.ci];
      const real prd = (gr[0] * gx + gr[1] * gy + gr[2] * gz) / R;
      Pg[0] = gr[0] - prd * gx / R; Pg[1] = gr[1] - prd * gy / R; Pg[2] = gr[2] - prd * gz / R;
    }
    for (int q = 0; q < 3; ++q) {
      const real drv = gT[(1 + q) * NF + a] - Pg[q] + dd[q][a] + real(0.0);
      out[q] = ps[1 + q][a] + drv * dt;
    }
  }
  // ld(0, address on the point's level): split rhoge at the gradient's level
  template <class LD>
  NH_HD void at(const Pt& p, LD ld) const {
    const long NF = d.NF, a = p.a, ks = p.ks;
    const int kmin = d.kmin, kmax = d.kmax, k = p.k;
    const real RovCV = c.Rdry / c.CVdry;
    auto grad = [&](int dd2) {
      real acc;
      if (p.pole) {
        acc = real(0.0);
        for (int v = 0; v < 6; ++v) {
          const real pv = G.cgrad_pl[(v * 2 + p.lp) * 3 + dd2] * ((ld(0, a + v) * RovCV) * real(1.0) * real(1.0));
          acc = (v == 0) ? pv : acc + pv;
        }
      } else {
        const long g = d.g;
        acc = rc(G.cgrad, d.nreg, dd2 * 7, p.ci) * ((ld(0, a) * RovCV) * real(1.0));
        for (int m = 1; m < 7; ++m) {
          const long off = (((4201 >> (2 * m)) & 3) - 1) + (((421 >> (2 * m)) & 3) - 1) * g;
          acc = acc + rc(G.cgrad, d.nreg, dd2 * 7 + m, p.ci) * ((ld(0, a + off) * RovCV) * real(1.0));
        }
      }
      return acc;
    };
    // the mass-flux sums are read before any store (a store to dv could alias
    // them for all the compiler knows, which would put each read behind it)
    real pmq[3] = {real(0.0), real(0.0), real(0.0)};
    if (pm)
      for (int q = 0; q < 3; ++q) pmq[q] = pm[(1 + q) * NF + a];
    real v[3];
    if (k >= kmin && k <= kmax) {
      level(p, a, grad, v);
    } else {
      // free-slip top / rigid bottom, scaled by the large-step density
      const bool top = (k == kmax + 1);
      const long an = top ? a - ks : a + ks;
      real vn[3];
      level(p, an, grad, vn);
      const real* rg = prog;
      for (int q = 0; q < 3; ++q) v[q] = top ? real(1.0) * (vn[q] / rg[an]) * rg[a] : -real(1.0) * (vn[q] / rg[an]) * rg[a];
    }
    for (int q = 0; q < 3; ++q) {
      dv[q][a] = v[q];
      if (pm) pm[(1 + q) * NF + a] = pmq[q] + v[q] * rweight;
    }
  }
};

// ----------------------------------------------------------------------------
// small step vi_main, horizontal part: the three flux divergences of the new
// momentum (mod_vi.py:1925-2432, kernels/vimain.py:127-262)
//   hv[0] = div(dv), hv[1] = vertical-pressure source gpre, hv[2] = div((prog + dv) * ethtot)
// ----------------------------------------------------------------------------
struct ViMainH {
  static constexpr const char* kName = "vi_main_h";
  static constexpr int kNStage = 8;    // fields read on the stencil: dv 0..2, eth, ethtot, prog 1..3
  static constexpr bool kStageF32 = false;
  Dims d; Consts c; Geo G;
  const real* prog;
  const real* dv[3];
  const real* eth;
  const real* ethtot;
  const real* gT;
  real* hv[3];
  NH_HD const real* stage_ptr(int f, int) const {
    return f < 3 ? dv[f] : (f == 3 ? eth : (f == 4 ? ethtot : prog + (long)(f - 4) * d.NF));
  }
  NH_HD bool stage_level(int k) const { return k >= d.kmin && k <= d.kmax; }
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (!pt_of(d, plane, x, p) || !pt_compute(p)) return;
    at(p, [&](int f, long b) { return stage_ptr(f, p.k)[b]; });
  }
  // ld(f, address): field f of the stencil set at an address of the point's stencil
  template <class LD>
  NH_HD void at(const Pt& p, LD ld) const {
    const int kmin = d.kmin, kmax = d.kmax, k = p.k;
    if (k < kmin || k > kmax) return;
    const long NF = d.NF, a = p.a;
    // div(dv), div(dv eth), div((prog + dv) ethtot)
    real dvs[3];
    hdiv_multi<3>(G, d, p, [&](int, long b, real (*v)[3]) {
      const real e = ld(3, b), et = ld(4, b);
      for (int q = 0; q < 3; ++q) {
        const real dq = ld(q, b);
        v[0][q] = dq * real(1.0);
        v[1][q] = dq * e * real(1.0);
        v[2][q] = (ld(5 + q, b) + dq) * et * real(1.0);
      }
    }, dvs);
    const real divs1 = dvs[0], he = dvs[1], hd = dvs[2];
    const real drhoge = -he - (real(0.0) - real(0.0)) * G.rdgz[k];
    const real grhoge1 = gT[5 * NF + a] + drhoge;
    hv[0][a] = divs1;
    hv[1][a] = grhoge1 * c.Rdry / c.CVdry;
    hv[2][a] = hd;
  }
};

// vertical part, one thread per compute column: source vector of the
// tridiagonal rhogw system and the Thomas sweeps, streamed in k with the
// previous level kept in registers (kernels/virhowsolver.py:42-120).  FP64 in
// every form: the field values are widened on load.  NK > 0 is the number of
// unknowns (kmax - kmin) fixed at compile time: u and gamma of the forward
// sweep then stay in registers; NK = 0 keeps them in the FP64 scratch U, Gm.
// The backward sweep writes the new rhogw split to out on every level.
template <int NK>
struct ViMainV {
  static constexpr const char* kName = "vi_main_v";
  static constexpr int kUnroll = NK > 0 ? NK : 4;
  Dims d; Consts c; Geo G;
  double dt;
  const real* ps[6];
  const real* gT;
  const double* Mc; const double* Mu; const double* Ml;
  const real* hv[3];
  double* U;
  double* Gm;
  real* out;
  NH_HD void operator()(long t) const {
    Cell cl = cell_of(d, t);
    if (cl.pole ? cl.v != 0 : !cl.inner) return;
    const long NF = d.NF, ks = cl.ks, b = cl.base;
    const int kmin = d.kmin, kmax = d.kmax;
    const VProf<double>& V = G.v64;
    const ConstsT<double>& C = c.d;
    const double RovCV = C.Rdry / C.CVdry;
    const double CVovRt2 = C.CVdry / C.Rdry / (dt * dt);
    const real* __restrict__ pg = ps[I_RHOG];
    const real* __restrict__ pw = ps[I_RHOGW];
    const real* __restrict__ pe = ps[I_RHOGE];
    const real* __restrict__ gT0 = gT;
    const real* __restrict__ gT4 = gT + 4 * NF;
    const real* __restrict__ hv0 = hv[0];
    const real* __restrict__ hv1 = hv[1];
    const double* __restrict__ mc = Mc;
    const double* __restrict__ mu = Mu;
    const double* __restrict__ ml = Ml;
    double* __restrict__ Uo = U;
    double* __restrict__ Go = Gm;
    real* __restrict__ Xo = out;
    const int N = (NK > 0) ? NK : kmax - kmin;
    double ub[NK > 0 ? NK : 1], gb[NK > 0 ? NK : 1];
    // level k-1 values, starting at kmin
    long am = b + (long)kmin * ks;
    double pe_m = pe[am], hv1_m = hv1[am], pg_m = pg[am];
    double gr_m = (double)gT0[am] + (-(double)hv0[am] - (0.0 - 0.0) * V.rdgz[kmin]);
    double beta = 0.0, u_prev = 0.0, mu_m = 0.0;
#pragma unroll kUnroll
    for (int n = 0; n < N; ++n) {
      const int k = kmin + 1 + n;
      const long a = am + ks;
      const double pe_k = pe[a], hv1_k = hv1[a], pg_k = pg[a];
      const double gr_k = (double)gT0[a] + (-(double)hv0[a] - (0.0 - 0.0) * V.rdgz[k]);
      const double sw = (double)gT4[a] + 0.0 * C.ALPHA;
      const double A = ((double)pw[a] * C.ALPHA + dt * sw) * 1.0;
      const double B = (pe_k * RovCV + dt * hv1_k) * 1.0 - (pe_m * RovCV + dt * hv1_m) * 1.0;
      const double Cc = (pg_k + dt * gr_k) * 1.0 * V.afact[k] + (pg_m + dt * gr_m) * 1.0 * V.bfact[k];
      double S = (A - B * dt * V.rdgzh[k] - Cc * dt * C.GRAV) * CVovRt2;
      const double mc_k = mc[a], ml_k = ml[a], mu_k = mu[a];
      if (n == 0) S = S - ml_k * (0.0 * 1.0);
      if (n == N - 1) S = S - mu_k * (-0.0 * 1.0);
      double u;
      if (n == 0) {
        beta = mc_k;
        u = S / beta;
      } else {
        const double gm = mu_m / beta;
        beta = mc_k - ml_k * gm;
        u = (S - ml_k * u_prev) / beta;
        if (NK > 0) gb[NK > 0 ? n : 0] = gm;
        else Go[a] = gm;
      }
      if (NK > 0) ub[NK > 0 ? n : 0] = u;
      else Uo[a] = u;
      u_prev = u;
      mu_m = mu_k;
      pe_m = pe_k; hv1_m = hv1_k; pg_m = pg_k; gr_m = gr_k;
      am = a;
    }
    // backward sweep
    long a = b + (long)kmax * ks;
    double xv = u_prev;
    Xo[a] = (real)(xv * 1.0);
    if (NK > 0) {
#pragma unroll
      for (int n = N - 2; n >= 0; --n) {
        a -= ks;
        xv = ub[NK > 0 ? n : 0] - gb[NK > 0 ? n + 1 : 0] * xv;
        Xo[a] = (real)(xv * 1.0);
      }
    } else {
      double gm_next = Go[a];
#pragma unroll 4
      for (int n = N - 2; n >= 0; --n) {
        a -= ks;
        xv = Uo[a] - gm_next * xv;
        gm_next = Go[a];
        Xo[a] = (real)(xv * 1.0);
      }
    }
    Xo[b + (long)kmin * ks] = (real)(0.0 * 1.0);
    Xo[b + (long)(kmax + 1) * ks] = (real)(-0.0 * 1.0);
    Xo[b + (long)(kmin - 1) * ks] = real(0.0);
  }
};

// mass and energy update of the small step from the new rhogw split, one
// thread per point (kernels/vimain.py:190-262, rhogkin.py:38-79)
struct ViMainPost {
  static constexpr const char* kName = "vi_main_post";
  Dims d; Consts c; Geo G;
  real dt, rweight;
  const real* prog;
  const real* ps[6];
  const real* dv[3];
  const real* ethtot;
  const real* gT;
  const real* getot;
  const real* hv[3];
  real* we[3];       // we[1] (rhogw split) is read
  real* pm;
  // Every value the point needs is loaded first, at levels clamped into
  // 0..kall-1 (a clamped level's value is never used: the level tests below
  // are the reference's), so that the loads issue together instead of one
  // memory round trip per level test.
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (!pt_of(d, plane, x, p) || !pt_compute(p)) return;
    const long NF = d.NF, ks = p.ks, a = p.a;
    const int kmin = d.kmin, kmax = d.kmax, k = p.k;
    const long b = a - (long)k * ks;
    // window levels k-1 .. k+2 -> index 0..3
    long A[4];
    for (int j = 0; j < 4; ++j) {
      int kk = k - 1 + j;
      kk = (kk < 0) ? 0 : ((kk > d.kall - 1) ? d.kall - 1 : kk);
      A[j] = b + (long)kk * ks;
    }
    const real* pg = ps[I_RHOG];
    const real* rw1 = we[1];
    real P0[3], S0[3], H0[3], T0[3], ET[3], RW[4], P4[2], S4[2], PQ[3], SQ[3], DQ[3];
    for (int j = 0; j < 3; ++j) {
      P0[j] = prog[A[j]];
      S0[j] = pg[A[j]];
      H0[j] = hv[0][A[j]];
      T0[j] = gT[A[j]];
      ET[j] = ethtot[A[j]];
    }
    for (int j = 0; j < 4; ++j) RW[j] = rw1[A[j]];
    for (int j = 0; j < 2; ++j) {
      P4[j] = prog[4 * NF + A[1 + j]];
      S4[j] = ps[4][A[1 + j]];
    }
    for (int q = 0; q < 3; ++q) {
      PQ[q] = prog[(1 + q) * NF + a];
      SQ[q] = ps[1 + q][a];
      DQ[q] = dv[q][a];
    }
    const real h2 = hv[2][a], ge = getot[a];
    const int ke = (k < kmin) ? kmin : ((k > kmax) ? kmax : k);
    const real pe_k = ps[I_RHOGE][b + (long)ke * ks];
    const real pm0 = pm ? pm[0 * NF + a] : real(0.0), pm4 = pm ? pm[4 * NF + a] : real(0.0);
    // j: window index of level kk (k - 1 + j)
    auto rs1 = [&](int kk, int j) {
      real drhog2 = real(0.0);
      if (kk >= kmin && kk <= kmax) {
        const real hi = (kk + 1 <= kmax) ? real(0.0) * real(1.0) + real(1.0) * RW[j + 1] * real(1.0) : real(0.0);
        const real lo = (kk >= kmin + 1) ? real(0.0) * real(1.0) + real(1.0) * RW[j] * real(1.0) : real(0.0);
        drhog2 = -H0[j] - (hi - lo) * G.rdgz[kk];
      }
      return S0[j] + (T0[j] + drhog2) * dt;
    };
    auto g10 = [&](int kk, int q) {
      const int j = kk - k + 1;
      if (q == 0) return P0[j] + S0[j];
      if (q == 4) return P4[j - 1] + S4[j - 1];
      return PQ[q - 1] + SQ[q - 1];
    };
    auto g11 = [&](int kk, int q) {
      const int j = kk - k + 1;
      if (q == 0) return P0[j] + rs1(kk, j);
      if (q == 4) return P4[j - 1] + RW[j];
      return PQ[q - 1] + DQ[q - 1];
    };
    const real kin10 = rhogkin_k(G, kmin, kmax, k, g10);
    const real kin11 = rhogkin_k(G, kmin, kmax, k, g11);
    real dtot = real(0.0);
    if (k >= kmin && k <= kmax) {
      real hi = real(0.0), lo = real(0.0);
      if (k + 1 <= kmax)
        hi = real(0.0) * real(1.0) + real(1.0) * ((P4[1] + RW[2]) * (G.afact[k + 1] * ET[2] + G.bfact[k + 1] * ET[1]))
             * real(1.0);
      if (k >= kmin + 1)
        lo = real(0.0) * real(1.0) + real(1.0) * ((P4[0] + RW[1]) * (G.afact[k] * ET[1] + G.bfact[k] * ET[0]))
             * real(1.0);
      dtot = -h2 - (hi - lo) * G.rdgz[k];
    }
    const real r1 = rs1(k, 1);
    const real re1 = pe_k + (ge + dtot) * dt + (kin10 - kin11) + (S0[1] - r1) * G.phi[k];
    we[0][a] = r1;
    we[2][a] = re1;
    if (pm) {
      pm[0 * NF + a] = pm0 + r1 * rweight;
      pm[4 * NF + a] = pm4 + RW[1] * rweight;
    }
  }
};

// PROG += PROG_split and horizontal projection (mod_vi.py:1572-1586)
struct ProgUpdateP {
  static constexpr const char* kName = "prog_update";
  Dims d; Consts c; Geo G;
  real* prog;
  const real* ps[6];
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (!pt_of(d, plane, x, p)) return;
    const long NF = d.NF, a = p.a;
    const real R = c.RADIUS;
    real v[6];
    for (int q = 0; q < 6; ++q) v[q] = prog[q * NF + a] + ps[q][a];
    if (p.pole) {
      const real gx = G.gx_pl[(p.v * 2 + p.lp) * 3 + 0];
      const real gy = G.gx_pl[(p.v * 2 + p.lp) * 3 + 1];
      const real gz = G.gx_pl[(p.v * 2 + p.lp) * 3 + 2];
      const real prd = v[1] * gx / R + v[2] * gy / R + v[3] * gz / R;
      v[1] = v[1] - prd * gx / R; v[2] = v[2] - prd * gy / R; v[3] = v[3] - prd * gz / R;
    } else if (p.inner) {
      const real gx = G.gx[p.ci], gy = G.gx[d.nreg + p.ci], gz = G.gx[2 * d.nreg + p.ci];
      const real prd = (v[1] * gx + v[2] * gy + v[3] * gz) / R;
      v[1] = v[1] - prd * gx / R; v[2] = v[2] - prd * gy / R; v[3] = v[3] - prd * gz / R;
    }
    for (int q = 0; q < 6; ++q) prog[q * NF + a] = v[q];
  }
};

NH_END
