// Stencil and column helpers of the HEVI large/small step (used by the
// point kernels of nh_point.h and the tracer kernels of nh_tracer.h).
//
// Vertical metrics enter as k-profiles (flat topography: GSQRT = GAM = 1, the
// terrain-following gradient terms vanish); the host verifies this before it
// creates a context.  Operation order follows the reference kernels so that
// the FP64 results track the numpy path to round-off.
#pragma once

#include "nh_columns.h"

NH_BEGIN



// ----------------------------------------------------------------------------
// Horizontal operators used by the hyperdiffusion and divergence damping
// ----------------------------------------------------------------------------
// Laplacian of one field at one level of a regional cell (oprtlaplacian.py:41-55)
NH_HD inline real lap7(const Geo& G, long nreg, long ci, const real* f, long a, int g) {
  const long off[7] = {0, 1, 1 + (long)g, (long)g, -1, -1 - (long)g, -(long)g};
  real s = rc(G.clap, nreg, 0, ci) * f[a + off[0]];
  for (int m = 1; m < 7; ++m) s = s + rc(G.clap, nreg, m, ci) * f[a + off[m]];
  return s;
}

NH_HD inline real lap_pl(const Geo& G, int lp, const real* f, long a0) {
  // a0: address of pole point v=0 at this level; neighbours at a0 + v
  real s = real(0.0);
  for (int v = 0; v < 6; ++v) {
    real pv = G.clap_pl[v * 2 + lp] * f[a0 + v];
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
NH_HD inline void tri_grad(const Geo& G, long nreg, long cj, int T, const real* s, long a, int g,
                           real out[3]) {
  real s0 = s[a];
  real s1 = (T == 0) ? s[a + 1] : s[a + 1 + g];
  real s2 = (T == 0) ? s[a + 1 + g] : s[a + g];
  for (int dd = 0; dd < 3; ++dd) {
    real c1 = rc(G.cintp, nreg, (dd * 3 + 0) * 2 + T, cj);
    real c2 = rc(G.cintp, nreg, (dd * 3 + 1) * 2 + T, cj);
    real c3 = rc(G.cintp, nreg, (dd * 3 + 2) * 2 + T, cj);
    out[dd] = ((real(2.0) * c1 - c2 - c3) * s0 + (-c1 + real(2.0) * c2 - c3) * s1 + (-c1 - c2 + real(2.0) * c3) * s2) / real(3.0);
  }
}

template <class KH>
NH_HD inline real diffusion7(const Geo& G, const Dims& d, const Cell& cl, const real* s, KH kh, long a) {
  const int g = d.g;
  const long nreg = d.nreg, ci = cl.ci;
  real t0i[3], t0j[3], t1i[3], t2j[3], t2i[3], t3j[3];
  tri_grad(G, nreg, ci, 0, s, a, g, t0i);                   // TI(i, j)
  tri_grad(G, nreg, ci, 1, s, a, g, t0j);                   // TJ(i, j)
  tri_grad(G, nreg, ci - 1, 0, s, a - 1, g, t1i);           // TI(i-1, j)
  tri_grad(G, nreg, ci - 1 - g, 1, s, a - 1 - g, g, t2j);   // TJ(i-1, j-1)
  tri_grad(G, nreg, ci - 1 - g, 0, s, a - 1 - g, g, t2i);   // TI(i-1, j-1)
  tri_grad(G, nreg, ci - g, 1, s, a - g, g, t3j);           // TJ(i, j-1)
  if (cl.i == 1 && cl.j == 1) {
    real m0 = G.sgp[cl.l] ? real(0.0) : real(1.0), m1 = G.sgp[cl.l] ? real(1.0) : real(0.0);
    real tjx[3];
    tri_grad(G, nreg, ci - 1 - g + 1, 1, s, a - 1 - g + 1, g, tjx);   // TJ(i, j-1) seen from (0,0)
    for (int q = 0; q < 3; ++q) t2i[q] = t2i[q] * m0 + tjx[q] * m1;
  }
  real k0 = kh(a);
  real kf1 = real(0.5) * (k0 + kh(a + 1 + g));
  real kf2 = real(0.5) * (k0 + kh(a + g));
  real kf3 = real(0.5) * (k0 + kh(a - 1));
  real kf4 = real(0.5) * (k0 + kh(a - 1 - g));
  real kf5 = real(0.5) * (k0 + kh(a - g));
  real kf6 = real(0.5) * (k0 + kh(a + 1));
  real inner = real(0.0);
  for (int dd = 0; dd < 3; ++dd) {
    const real* cd = G.cdiff + (long)(dd * 6) * nreg + ci;
    real blk = kf1 * cd[0] * (t0i[dd] + t0j[dd]) + kf2 * cd[nreg] * (t0j[dd] + t1i[dd])
               + kf3 * cd[2 * nreg] * (t1i[dd] + t2j[dd]) + kf4 * cd[3 * nreg] * (t2j[dd] + t2i[dd])
               + kf5 * cd[4 * nreg] * (t2i[dd] + t3j[dd]) + kf6 * cd[5 * nreg] * (t3j[dd] + t0i[dd]);
    inner = (dd == 0) ? blk : inner + blk;
  }
  return inner;
}

template <class KH>
NH_HD inline real diffusion_pl(const Geo& G, int lp, const real* s, KH kh, long a0) {
  real vt[5][3];
  for (int v = 1; v <= 5; ++v) {
    int vp = (v == 5) ? 1 : v + 1;
    for (int dd = 0; dd < 3; ++dd) {
      const real* cI = G.cintp_pl + ((v * 2 + lp) * 3 + dd) * 3;
      real c0 = cI[0], c1 = cI[1], c2 = cI[2];
      vt[v - 1][dd] = ((real(2.0) * c0 - c1 - c2) * s[a0] + (-c0 + real(2.0) * c1 - c2) * s[a0 + v]
                       + (-c0 - c1 + real(2.0) * c2) * s[a0 + vp]) / real(3.0);
    }
  }
  real row = real(0.0);
  for (int v = 1; v <= 5; ++v) {
    int vm = (v == 1) ? 5 : v - 1;
    real kavg = real(0.5) * (kh(a0) + kh(a0 + v));
    const real* cd = G.cdiff_pl + (v * 2 + lp) * 3;
    real sum = cd[0] * (vt[vm - 1][0] + vt[v - 1][0]);
    sum = sum + cd[1] * (vt[vm - 1][1] + vt[v - 1][1]);
    sum = sum + cd[2] * (vt[vm - 1][2] + vt[v - 1][2]);
    real cb = kavg * sum;
    row = (v == 1) ? cb : row + cb;
  }
  return row;
}


// ----------------------------------------------------------------------------
// 3-D divergence damping (mod_numfilter.py:2032-2268, kernels/oprt3ddivdamp.py,
// kernels/oprtdivdamp.py, kernels/divdamppostcomm.py)
// ----------------------------------------------------------------------------
// triangle value sum_d sum_vertex c * v_d(vertex) for the divdamp operators
NH_HD inline real tri_div(const Geo& G, long nreg, long cj, int T, const real* vx, const real* vy,
                            const real* vz, long a, int g, real sgn) {
  long o1 = (T == 0) ? 1 : 1 + g;
  long o2 = (T == 0) ? 1 + g : g;
  const real* v3[3] = {vx, vy, vz};
  real s = real(0.0);
  bool first = true;
  for (int dd = 0; dd < 3; ++dd) {
    real c0 = rc(G.cintp, nreg, (dd * 3 + 0) * 2 + T, cj);
    real c1 = rc(G.cintp, nreg, (dd * 3 + 1) * 2 + T, cj);
    real c2 = rc(G.cintp, nreg, (dd * 3 + 2) * 2 + T, cj);
    real t0 = c0 * (sgn * v3[dd][a]);
    real t1 = c1 * (sgn * v3[dd][a + o1]);
    real t2 = c2 * (sgn * v3[dd][a + o2]);
    if (first) { s = t0 + t1 + t2; first = false; }
    else s = s + t0 + t1 + t2;
  }
  return s;
}

NH_HD inline real hex_sum(const Geo& G, long nreg, long ci, int dd, const real T[6]) {
  // T = {TI(i,j), TJ(i,j), TI(i-1,j), TJ(i-1,j-1), TI(i-1,j-1), TJ(i,j-1)}
  const real* cd = G.cdiff + (long)(dd * 6) * nreg + ci;
  return cd[0] * (T[0] + T[1]) + cd[nreg] * (T[1] + T[2]) + cd[2 * nreg] * (T[2] + T[3])
       + cd[3 * nreg] * (T[3] + T[4]) + cd[4 * nreg] * (T[4] + T[5]) + cd[5 * nreg] * (T[5] + T[0]);
}



NH_END

NH_BEGIN

// 7-point gradient of s at one level (oprtgradient / presgrad _gradient_interior)
NH_HD inline void grad7(const Geo& G, long nreg, long ci, const real* s, long a, int g, real o[3]) {
  const long off[7] = {0, 1, 1 + (long)g, (long)g, -1, -1 - (long)g, -(long)g};
  for (int dd = 0; dd < 3; ++dd) {
    real acc = rc(G.cgrad, nreg, dd * 7 + 0, ci) * (s[a] * real(1.0));
    for (int m = 1; m < 7; ++m) acc = acc + rc(G.cgrad, nreg, dd * 7 + m, ci) * (s[a + off[m]] * real(1.0));
    o[dd] = acc;
  }
}

NH_HD inline void grad_pl(const Geo& G, int lp, const real* s, long a0, real o[3]) {
  for (int dd = 0; dd < 3; ++dd) {
    real acc = real(0.0);
    for (int v = 0; v < 6; ++v) {
      real pv = G.cgrad_pl[(v * 2 + lp) * 3 + dd] * (s[a0 + v] * real(1.0) * real(1.0));
      acc = (v == 0) ? pv : acc + pv;
    }
    o[dd] = acc;
  }
}

// horizontal pressure-gradient force at level k with the first-layer remedy
// (kernels/presgrad.py:126-265 with flat terrain: the vertical term is zero)
NH_HD inline void presgrad_h(const Geo& G, const Dims& d, const Cell& cl, const real* P, int k,
                             real R, real o[3]) {
  int kk = (k == d.kmin) ? d.kmin + 1 : k;
  long a = cl.base + kk * cl.ks;
  real gr[3];
  if (cl.pole) grad_pl(G, cl.lp, P, a, gr);
  else grad7(G, d.nreg, cl.ci, P, a, d.g, gr);
  for (int q = 0; q < 3; ++q) gr[q] = gr[q] + real(0.0);
  real gx, gy, gz;
  if (cl.pole) {
    gx = G.gx_pl[(0 * 2 + cl.lp) * 3 + 0];
    gy = G.gx_pl[(0 * 2 + cl.lp) * 3 + 1];
    gz = G.gx_pl[(0 * 2 + cl.lp) * 3 + 2];
    real prd = gr[0] * gx / R + gr[1] * gy / R + gr[2] * gz / R;
    o[0] = gr[0] - prd * gx / R;
    o[1] = gr[1] - prd * gy / R;
    o[2] = gr[2] - prd * gz / R;
  } else {
    gx = G.gx[cl.ci]; gy = G.gx[d.nreg + cl.ci]; gz = G.gx[2 * d.nreg + cl.ci];
    real prd = (gr[0] * gx + gr[1] * gy + gr[2] * gz) / R;
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
NH_HD inline real hdiv(const Geo& G, const Dims& d, const Cell& cl, FX fx, FY fy, FZ fz) {
  if (cl.pole) return div_pl(G, cl.lp, fx, fy, fz);
  return div7(G, d.nreg, cl.ci, fx, fy, fz);
}


// NDIV horizontal divergences at one compute point, one neighbour at a time:
// f(m, addr, v) fills v[i][dir] for divergence i at neighbour m (address
// addr).  Per divergence and direction the sum over the neighbours runs in
// stencil order, then x + y + z, as div7 does (the pole: per point
// cx*fx + cy*fy + cz*fz, summed over v, as div_pl).  The neighbour loop is not
// unrolled so that only one neighbour's values are live.
template <int NDIV, class F>
NH_HD inline void hdiv_multi(const Geo& G, const Dims& d, const Pt& p, F f, real out[NDIV]) {
  real v[NDIV][3];
  if (p.pole) {
#pragma unroll 1
    for (int m = 0; m < 6; ++m) {
      f(m, p.a + m, v);
      const real* c = G.cdiv_pl + (m * 2 + p.lp) * 3;
      for (int i = 0; i < NDIV; ++i) {
        const real pv = c[0] * v[i][0] + c[1] * v[i][1] + c[2] * v[i][2];
        out[i] = (m == 0) ? pv : out[i] + pv;
      }
    }
    return;
  }
  real s[NDIV][3];
  const long g = d.g;
#pragma unroll 1
  for (int m = 0; m < 7; ++m) {
    // (di, dj) + 1 of neighbour m packed two bits per point:
    // (0,0) (+1,0) (+1,+1) (0,+1) (-1,0) (-1,-1) (0,-1)
    const long off = (((4201 >> (2 * m)) & 3) - 1) + (((421 >> (2 * m)) & 3) - 1) * g;
    f(m, p.a + off, v);
    const real cx = rc(G.cdiv, d.nreg, m, p.ci), cy = rc(G.cdiv, d.nreg, 7 + m, p.ci),
               cz = rc(G.cdiv, d.nreg, 14 + m, p.ci);
    for (int i = 0; i < NDIV; ++i) {
      const real tx = cx * v[i][0], ty = cy * v[i][1], tz = cz * v[i][2];
      s[i][0] = (m == 0) ? tx : s[i][0] + tx;
      s[i][1] = (m == 0) ? ty : s[i][1] + ty;
      s[i][2] = (m == 0) ? tz : s[i][2] + tz;
    }
  }
  for (int i = 0; i < NDIV; ++i) out[i] = s[i][0] + s[i][1] + s[i][2];
}

// kinetic energy at level k of a column given accessors (kernels/rhogkin.py:38-79),
// in the precision T of the vertical profiles V (the field precision, or FP64
// for the global sums)
template <class T, class F>
NH_HD inline T rhogkin_k(const VProf<T>& V, int kmin, int kmax, int k, F get) {
  // get(k, q): q = 0 rhog, 1..3 rhogv, 4 rhogw
  if (k < kmin || k > kmax) return T(0.0);
  T rgk = get(k, 0);
  T vx = get(k, 1), vy = get(k, 2), vz = get(k, 3);
  T kh = T(0.5) * (vx * vx + vy * vy + vz * vz) / rgk;
  T vup = T(0.0), vdn = T(0.0);
  if (k + 1 <= kmax) {
    T w = get(k + 1, 4);
    vup = T(0.5) * (w * w) / (V.c2w0[k + 1] * get(k + 1, 0) + V.c2w1[k + 1] * rgk);
  }
  if (k >= kmin + 1) {
    T w = get(k, 4);
    vdn = T(0.5) * (w * w) / (V.c2w0[k] * rgk + V.c2w1[k] * get(k - 1, 0));
  }
  return kh + V.w2c0[k] * vup + V.w2c1[k] * vdn;
}


// split state at the start of an RK stage: 0 (first stage) or PROG0 - PROG
struct SplitInit {
  static constexpr const char* kName = "split_init";
  long NF;
  const real* prog;
  const real* prog0;
  real* ps[6];
  int zero;
  NH_HD void operator()(long t) const {
    for (int q = 0; q < 6; ++q) ps[q][t] = zero ? real(0.0) : prog0[q * NF + t] - prog[q * NF + t];
  }
};




NH_END
