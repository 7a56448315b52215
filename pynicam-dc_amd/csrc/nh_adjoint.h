// Adjoint (transposed) kernels of the dynamics step.
//
// Each functor XxxA is the transpose of the Jacobian of the forward functor
// Xxx at the forward state it is given: it reads the adjoint of Xxx's outputs
// ("bars"), consumes them (sets them to zero, so that a buffer reused by the
// forward step starts clean for its previous use), and adds the adjoint of
// Xxx's inputs into the input bars.  The forward values a kernel needs (for
// products, quotients) are read from the same fields the forward kernel read;
// the step driver recomputes them before each adjoint kernel.
//
// Stencil kernels are written in scatter form: every compute point adds its
// contribution to the bars of the inputs it read, with atomic adds, so that a
// kernel's adjoint follows the forward code line by line.  Column and
// point-local kernels (the tridiagonal solve, the pointwise updates) write
// without atomics.  The linearisation is exact for every smooth term;
// sign/branch switches (ghost-row selection, k ranges) follow the forward.
//
// Tracer masses enter the dynamics only through q = rhogq / rhog in the
// thermodynamics; they are held at their trajectory values here, so the
// adjoint is the full derivative when the tracers are zero (the dry runs).
#pragma once

#include "nh_point.h"
#include "nh_tile.h"

NH_BEGIN

NH_HD inline void aadd(float* p, float v) {
#if defined(__HIP_DEVICE_COMPILE__)
  unsafeAtomicAdd(p, v);
#else
#pragma omp atomic
  *p += v;
#endif
}
NH_HD inline void aadd(double* p, double v) {
#if defined(__HIP_DEVICE_COMPILE__)
  unsafeAtomicAdd(p, v);
#else
#pragma omp atomic
  *p += v;
#endif
}

// regional stencil offset of neighbour m: (0,0) (+1,0) (+1,+1) (0,+1) (-1,0) (-1,-1) (0,-1)
NH_HD inline long nb_off(int m, long g) {
  return (((4201 >> (2 * m)) & 3) - 1) + (((421 >> (2 * m)) & 3) - 1) * g;
}

// position of a compute point for the horizontal projections: the pole
// centre's GRD_x (pole) or the cell's (regional)
NH_HD inline void point_pos(const Geo& G, const Dims& d, const Pt& p, real x[3]) {
  if (p.pole) {
    for (int q = 0; q < 3; ++q) x[q] = G.gx_pl[p.lp * 3 + q];
  } else {
    for (int q = 0; q < 3; ++q) x[q] = G.gx[q * d.nreg + p.ci];
  }
}

// transpose of o = v - ((v.x)/R) x/R (symmetric): in place on v
NH_HD inline void proj_t(const real x[3], real R, real v[3]) {
  const real prd = (v[0] * x[0] + v[1] * x[1] + v[2] * x[2]) / R;
  for (int q = 0; q < 3; ++q) v[q] = v[q] - prd * x[q] / R;
}

// ----------------------------------------------------------------------------
// halo gather: f[dst] = f[src]  ->  fb[src] += fb[dst]; fb[dst] = 0
// (the host checks that destinations are unique and are never sources)
// ----------------------------------------------------------------------------
struct HaloScatterA {
  static constexpr const char* kName = "halo_adj";
  Dims d;
  const int* dst;   // level-0 field addresses, as in HaloGather
  const int* src;
  long n;
  real* f[16];
  int nf;
  NH_HD void operator()(long t) const {
    HaloGather h;
    h.d = d;
    const unsigned tt = (unsigned)t, nn = (unsigned)n;
    const unsigned k = tt / nn, hh = tt - k * nn;
    const long bd0 = dst[hh], bs0 = src[hh];
    const long bd = bd0 + (long)k * h.stride(bd0);
    const long bs = bs0 + (long)k * h.stride(bs0);
    // every destination bar read before any store or add (one memory round
    // trip rather than one per field)
    real v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q)
      if (q < nf) v[q] = f[q][bd];
#pragma unroll
    for (int q = 0; q < 16; ++q)
      if (q < nf) {
        f[q][bd] = real(0.0);
        if (v[q] != real(0.0)) aadd(&f[q][bs], v[q]);
      }
  }
};

// zero a range of 4-byte words: thread t < nvec clears the 16-byte chunk t
// (from the first 16-byte boundary, one dwordx4 store), the next threads the
// head words before that boundary and the tail words after the last chunk
struct alignas(16) Zero4 {
  unsigned v[4];
};
struct ZeroWords {
  static constexpr const char* kName = "fill_zero";
  unsigned* w;   // first word
  long head;     // words before the first 16-byte boundary (0..3)
  long nvec;     // 16-byte chunks
  long nw;       // words in all
  NH_HD void operator()(long t) const {
    if (t < nvec) {
      const Zero4 z = {{0u, 0u, 0u, 0u}};
      reinterpret_cast<Zero4*>(w + head)[t] = z;
      return;
    }
    t -= nvec;
    const long i = t < head ? t : head + 4 * nvec + (t - head);
    if (i < nw) w[i] = 0u;
  }
  // threads to launch
  long threads() const { return nvec + (nw - 4 * nvec); }
};
inline ZeroWords zero_words(void* p, long nbytes) {
  ZeroWords z;
  z.w = static_cast<unsigned*>(p);
  z.nw = nbytes / 4;
  const long mis = (long)((reinterpret_cast<unsigned long>(p) >> 2) & 3);
  z.head = mis ? 4 - mis : 0;
  if (z.head > z.nw) z.head = z.nw;
  z.nvec = (z.nw - z.head) / 4;
  return z;
}

// y += x over a flat range
struct AxpyP {
  static constexpr const char* kName = "adj_accumulate";
  real* y;
  const real* x;
  NH_HD void operator()(long t) const { y[t] += x[t]; }
};

// ----------------------------------------------------------------------------
// split state at a stage start: ps = prog0 - prog (nl > 0) or 0
// ----------------------------------------------------------------------------
struct SplitInitA {
  static constexpr const char* kName = "split_init_adj";
  long NF;
  real* prog_b;
  real* prog0_b;
  real* ps_b[6];
  int zero;
  NH_HD void operator()(long t) const {
    for (int q = 0; q < 6; ++q) {
      const real v = ps_b[q][t];
      ps_b[q][t] = real(0.0);
      if (!zero) {
        prog0_b[q * NF + t] += v;
        prog_b[q * NF + t] -= v;
      }
    }
  }
};

// ----------------------------------------------------------------------------
// PROG = P(PROG + PROG_split): the bar of the updated PROG becomes the bar of
// the stage's PROG and of the final split state (set, not added)
// ----------------------------------------------------------------------------
struct ProgUpdateA {
  static constexpr const char* kName = "prog_update_adj";
  Dims d; Consts c; Geo G;
  real* prog_b;
  real* ps_b[6];
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (!pt_of(d, plane, x, p)) return;
    const long NF = d.NF, a = p.a;
    const real R = c.RADIUS;
    real v[6];
    for (int q = 0; q < 6; ++q) v[q] = prog_b[q * NF + a];
    if (p.pole || p.inner) {
      real xx[3];
      if (p.pole) {
        for (int q = 0; q < 3; ++q) xx[q] = G.gx_pl[(p.v * 2 + p.lp) * 3 + q];
      } else {
        for (int q = 0; q < 3; ++q) xx[q] = G.gx[q * d.nreg + p.ci];
      }
      proj_t(xx, R, v + 1);
    }
    for (int q = 0; q < 6; ++q) {
      prog_b[q * NF + a] = v[q];
      ps_b[q][a] = v[q];
    }
  }
};

// ----------------------------------------------------------------------------
// divergence damping: helpers of the transposes (the kernels are below)
// ----------------------------------------------------------------------------
// transpose of hex_sum over the three directions: o[dd] -> Tb[6]
NH_HD inline void hex_sum_t(const Geo& G, long nreg, long ci, const real o[3], real Tb[6]) {
  for (int n = 0; n < 6; ++n) Tb[n] = real(0.0);
  for (int dd = 0; dd < 3; ++dd) {
    const real* cd = G.cdiff + (long)(dd * 6) * nreg + ci;
    const real c0 = cd[0], c1 = cd[nreg], c2 = cd[2 * nreg], c3 = cd[3 * nreg], c4 = cd[4 * nreg],
               c5 = cd[5 * nreg];
    Tb[0] += (c0 + c5) * o[dd];
    Tb[1] += (c0 + c1) * o[dd];
    Tb[2] += (c1 + c2) * o[dd];
    Tb[3] += (c2 + c3) * o[dd];
    Tb[4] += (c3 + c4) * o[dd];
    Tb[5] += (c4 + c5) * o[dd];
  }
}

// transpose of the pole ring sum row[dd] = sum_v cdiff_pl(v, dd) (sclt[v-1] + sclt[v-2])
NH_HD inline void pole_ring_t(const Geo& G, int lp, const real o[3], real sb[5]) {
  for (int v = 0; v < 5; ++v) sb[v] = real(0.0);
  for (int v = 1; v <= 5; ++v) {
    const int vm = (v == 1) ? 5 : v - 1;
    real s = real(0.0);
    for (int dd = 0; dd < 3; ++dd) s += G.cdiff_pl[(v * 2 + lp) * 3 + dd] * o[dd];
    sb[vm - 1] += s;
    sb[v - 1] += s;
  }
}

// ----------------------------------------------------------------------------
// hyperdiffusion passes
// ----------------------------------------------------------------------------

// transpose of tri_grad: ob3[dd] -> sb at the triangle's corners
NH_HD inline void tri_grad_t(const Geo& G, long nreg, long cj, int T, real* sb, long a, int g, const real ob3[3]) {
  real b0 = real(0.0), b1 = real(0.0), b2 = real(0.0);
  for (int dd = 0; dd < 3; ++dd) {
    const real c1 = rc(G.cintp, nreg, (dd * 3 + 0) * 2 + T, cj);
    const real c2 = rc(G.cintp, nreg, (dd * 3 + 1) * 2 + T, cj);
    const real c3 = rc(G.cintp, nreg, (dd * 3 + 2) * 2 + T, cj);
    const real o = ob3[dd] / real(3.0);
    b0 += (real(2.0) * c1 - c2 - c3) * o;
    b1 += (-c1 + real(2.0) * c2 - c3) * o;
    b2 += (-c1 - c2 + real(2.0) * c3) * o;
  }
  aadd(&sb[a], b0);
  aadd(&sb[(T == 0) ? a + 1 : a + 1 + g], b1);
  aadd(&sb[(T == 0) ? a + 1 + g : a + g], b2);
}

// ----------------------------------------------------------------------------
// gather-form helpers: the output points whose stencil reads point x
// ----------------------------------------------------------------------------
// regional point x = (i, j) of a region is neighbour m of the output point
// (i - di_m, j - dj_m); that point computes when it is interior
NH_HD inline bool src_of(const Dims& d, const Pt& x, int m, long* a, long* ci) {
  const int di = ((4201 >> (2 * m)) & 3) - 1, dj = ((421 >> (2 * m)) & 3) - 1;
  const int i = x.i - di, j = x.j - dj;
  if (i < 1 || i > d.g - 2 || j < 1 || j > d.g - 2) return false;
  const long o = (long)di + (long)dj * d.g;
  *a = x.a - o;
  *ci = x.ci - o;
  return true;
}

// -lap transposed, gather form: ib[q](x) += -sum over the compute points a
// whose stencil holds x of clap(a, m) ob[q](a) (both hyperdiffusion passes)
// (the interior points run as column tiles, LapTC; this functor's 1-D form
// does the rest: the region borders, one thread per border point and level,
// and the pole points)
struct NegLapGA {
  static constexpr const char* kName = "hdiff_lap_adj";
  Dims d; Geo G;
  int nq;
  const real* ob[6];
  real* ib[6];
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (!pt_of(d, plane, x, p)) return;
    point(p);
  }
  // t < lall * kall * 4 (g - 1): border point e of plane t / (4 (g - 1)),
  // going round the region (j = 0, i = g - 1, j = g - 1, i = 0); then the
  // pole points
  NH_HD long border_threads() const { return (long)d.lall * d.kall * 4 * (d.g - 1) + 2L * d.kall * 6; }
  NH_HD void operator()(long t) const {
    const int g = d.g, per = 4 * (g - 1);
    const long nb = (long)d.lall * d.kall * per;
    Pt p;
    if (t < nb) {
      const int plane = (int)(t / per), e = (int)(t - (long)plane * per);
      int i, j;
      if (e < g - 1) { i = e; j = 0; }
      else if (e < 2 * (g - 1)) { i = g - 1; j = e - (g - 1); }
      else if (e < 3 * (g - 1)) { i = g - 1 - (e - 2 * (g - 1)); j = g - 1; }
      else { i = 0; j = g - 1 - (e - 3 * (g - 1)); }
      if (!pt_of(d, plane, j * g + i, p)) return;
    } else {
      const long u = t - nb;
      if (!pt_of(d, d.lall * d.kall + (int)(u / 6), (int)(u % 6), p)) return;
    }
    point(p);
  }
  NH_HD void point(const Pt& p) const {
    const long a = p.a;
    real acc[6] = {real(0.0), real(0.0), real(0.0), real(0.0), real(0.0), real(0.0)};
    if (p.pole) {
      const real cl = G.clap_pl[p.v * 2 + p.lp];
      for (int q = 0; q < nq; ++q) acc[q] = cl * ob[q][a - p.v];
    } else {
      for (int m = 0; m < 7; ++m) {
        long ao, co;
        if (!src_of(d, p, m, &ao, &co)) continue;
        const real cl = rc(G.clap, d.nreg, m, co);
        for (int q = 0; q < nq; ++q) acc[q] += cl * ob[q][ao];
      }
    }
    for (int q = 0; q < nq; ++q) ib[q][a] += -acc[q];
  }
};

// second pass, diffused fields: -Kh-weighted diffusion of the temperature and
// density fields with Kh = rhog * fac * GAMMA_H (the four Laplacian fields
// go through NegLapGA).  Three launches: HDiffKhTriA stores, per triangle,
// the bars of its three vertices; HDiffKhG gathers them per point; this
// functor then scatters, with atomic adds, the compute points the two leave
// out: the pole centres and the pentagon point (1, 1) of the regions whose
// stencil there drops a triangle (sgp), one thread each.
struct HDiffPass2A {
  static constexpr const char* kName = "hdiff_pass2_adj";
  Dims d; Consts c; Geo G;
  const real* in[6];   // forward inputs (the first pass, halo-filled)
  const real* rhog;
  real* ob[6];
  real* ib[6];
  real* rhog_b;
  // t < lall * kall: the point (1, 1) of region t / kall at level t % kall
  // (sgp regions only); then the pole centres, one per pole plane
  NH_HD long threads() const { return (long)(d.lall + 2) * d.kall; }
  NH_HD void operator()(long t) const {
    const long nrp = (long)d.lall * d.kall;
    Pt p;
    if (t < nrp) {
      if (!G.sgp[t / d.kall]) return;
      if (!pt_of(d, (int)t, d.g + 1, p)) return;
    } else if (!pt_of(d, (int)t, 0, p)) {
      return;
    }
    if (!pt_compute(p)) return;
    point(p);
  }
  NH_HD void point(const Pt& p) const {
    const long a = p.a;
    // regional points: the bars of rhog and of each field summed per stencil
    // point first, then one atomic add per stencil point
    real rb[7] = {real(0.0), real(0.0), real(0.0), real(0.0), real(0.0), real(0.0), real(0.0)};
    for (int f = 0; f < 2; ++f) {
      const real o = -ob[4 + f][a];
      ob[4 + f][a] = real(0.0);
      const real kfac = ((f == 0) ? c.CVdry : c.HDIFF_FACT_RHO) * c.GAMMA_H;
      if (p.pole) diff_pl_t(p, in[4 + f], ib[4 + f], kfac, o);
      else diff7_t(p, in[4 + f], ib[4 + f], kfac, o, rb);
    }
    if (!p.pole)
      for (int m = 0; m < 7; ++m) aadd(&rhog_b[a + nb_off(m, d.g)], rb[m]);
  }
  // tri_grad_t's three vertex bars without the adds
  NH_HD void tri_grad_b(long cj, int T, const real ob3[3], real& b0, real& b1, real& b2) const {
    const long nreg = d.nreg;
    b0 = real(0.0); b1 = real(0.0); b2 = real(0.0);
    for (int dd = 0; dd < 3; ++dd) {
      const real c1 = rc(G.cintp, nreg, (dd * 3 + 0) * 2 + T, cj);
      const real c2 = rc(G.cintp, nreg, (dd * 3 + 1) * 2 + T, cj);
      const real c3 = rc(G.cintp, nreg, (dd * 3 + 2) * 2 + T, cj);
      const real o = ob3[dd] / real(3.0);
      b0 += (real(2.0) * c1 - c2 - c3) * o;
      b1 += (-c1 + real(2.0) * c2 - c3) * o;
      b2 += (-c1 - c2 + real(2.0) * c3) * o;
    }
  }
  NH_HD void diff7_t(const Pt& p, const real* s, real* sb, real kfac, real o, real rb[7]) const {
    const int g = d.g;
    const long nreg = d.nreg, ci = p.ci, a = p.a;
    // the six triangles: TI(i,j) TJ(i,j) TI(i-1,j) TJ(i-1,j-1) TI(i-1,j-1) TJ(i,j-1)
    const long tc[6] = {0, 0, -1, -1 - (long)g, -1 - (long)g, -(long)g};
    const int tT[6] = {0, 1, 0, 1, 0, 1};
    real tv[6][3];
    for (int n = 0; n < 6; ++n) tri_grad(G, nreg, ci + tc[n], tT[n], s, a + tc[n], g, tv[n]);
    const bool pent = (p.i == 1 && p.j == 1);
    const real m0 = G.sgp[p.l] ? real(0.0) : real(1.0), m1 = G.sgp[p.l] ? real(1.0) : real(0.0);
    if (pent)
      for (int dd = 0; dd < 3; ++dd) tv[4][dd] = tv[4][dd] * m0 + tv[5][dd] * m1;
    // kf of the six arcs: neighbours (+1,+1) (0,+1) (-1,0) (-1,-1) (0,-1) (+1,0)
    const long kn[6] = {1 + (long)g, (long)g, -1, -1 - (long)g, -(long)g, 1};
    const real k0 = rhog[a] * kfac;
    real kf[6];
    for (int e = 0; e < 6; ++e) kf[e] = real(0.5) * (k0 + rhog[a + kn[e]] * kfac);
    real tb[6][3];
    for (int n = 0; n < 6; ++n)
      for (int dd = 0; dd < 3; ++dd) tb[n][dd] = real(0.0);
    real kfb[6] = {real(0.0), real(0.0), real(0.0), real(0.0), real(0.0), real(0.0)};
    for (int dd = 0; dd < 3; ++dd) {
      const real* cd = G.cdiff + (long)(dd * 6) * nreg + ci;
      for (int e = 0; e < 6; ++e) {
        const int na = e, nb = (e + 1) % 6;   // arc e joins triangles e and e+1
        const real cc = cd[e * nreg] * o;
        kfb[e] += cc * (tv[na][dd] + tv[nb][dd]);
        tb[na][dd] += cc * kf[e];
        tb[nb][dd] += cc * kf[e];
      }
    }
    // stencil index of the arcs' neighbours kn[e]: (+1,+1) 2, (0,+1) 3, (-1,0) 4, (-1,-1) 5, (0,-1) 6, (+1,0) 1
    const int em[6] = {2, 3, 4, 5, 6, 1};
    real k0b = real(0.0);
    for (int e = 0; e < 6; ++e) {
      k0b += real(0.5) * kfb[e];
      rb[em[e]] += real(0.5) * kfb[e] * kfac;
    }
    rb[0] += k0b * kfac;
    if (pent)
      for (int dd = 0; dd < 3; ++dd) {
        tb[5][dd] += tb[4][dd] * m1;
        tb[4][dd] = tb[4][dd] * m0;
      }
    // the triangles' vertices as stencil indices
    const int vm[6][3] = {{0, 1, 2}, {0, 2, 3}, {4, 0, 3}, {5, 0, 4}, {5, 6, 0}, {6, 1, 0}};
    real acc[7] = {real(0.0), real(0.0), real(0.0), real(0.0), real(0.0), real(0.0), real(0.0)};
    for (int n = 0; n < 6; ++n) {
      real b0, b1, b2;
      tri_grad_b(ci + tc[n], tT[n], tb[n], b0, b1, b2);
      acc[vm[n][0]] += b0;
      acc[vm[n][1]] += b1;
      acc[vm[n][2]] += b2;
    }
    for (int m = 0; m < 7; ++m) aadd(&sb[a + nb_off(m, g)], acc[m]);
  }
  NH_HD void diff_pl_t(const Pt& p, const real* s, real* sb, real kfac, real o) const {
    const long a0 = p.a;
    const int lp = p.lp;
    real vt[5][3];
    for (int v = 1; v <= 5; ++v) {
      const int vp = (v == 5) ? 1 : v + 1;
      for (int dd = 0; dd < 3; ++dd) {
        const real* cI = G.cintp_pl + ((v * 2 + lp) * 3 + dd) * 3;
        const real c0 = cI[0], c1 = cI[1], c2 = cI[2];
        vt[v - 1][dd] = ((real(2.0) * c0 - c1 - c2) * s[a0] + (-c0 + real(2.0) * c1 - c2) * s[a0 + v]
                         + (-c0 - c1 + real(2.0) * c2) * s[a0 + vp]) / real(3.0);
      }
    }
    real vb[5][3];
    for (int v = 0; v < 5; ++v)
      for (int dd = 0; dd < 3; ++dd) vb[v][dd] = real(0.0);
    for (int v = 1; v <= 5; ++v) {
      const int vm = (v == 1) ? 5 : v - 1;
      const real kavg = real(0.5) * (rhog[a0] * kfac + rhog[a0 + v] * kfac);
      const real* cd = G.cdiff_pl + (v * 2 + lp) * 3;
      const real sum = cd[0] * (vt[vm - 1][0] + vt[v - 1][0]) + cd[1] * (vt[vm - 1][1] + vt[v - 1][1])
                       + cd[2] * (vt[vm - 1][2] + vt[v - 1][2]);
      const real kb = o * sum;
      aadd(&rhog_b[a0], real(0.5) * kb * kfac);
      aadd(&rhog_b[a0 + v], real(0.5) * kb * kfac);
      const real smb = o * kavg;
      for (int dd = 0; dd < 3; ++dd) {
        vb[vm - 1][dd] += smb * cd[dd];
        vb[v - 1][dd] += smb * cd[dd];
      }
    }
    for (int v = 1; v <= 5; ++v) {
      const int vp = (v == 5) ? 1 : v + 1;
      real b0 = real(0.0), b1 = real(0.0), b2 = real(0.0);
      for (int dd = 0; dd < 3; ++dd) {
        const real* cI = G.cintp_pl + ((v * 2 + lp) * 3 + dd) * 3;
        const real c0 = cI[0], c1 = cI[1], c2 = cI[2];
        const real w = vb[v - 1][dd] / real(3.0);
        b0 += (real(2.0) * c0 - c1 - c2) * w;
        b1 += (-c0 + real(2.0) * c1 - c2) * w;
        b2 += (-c0 - c1 + real(2.0) * c2) * w;
      }
      aadd(&sb[a0], b0);
      aadd(&sb[a0 + v], b1);
      aadd(&sb[a0 + vp], b2);
    }
  }
};

// the Kh-weighted diffusion transposed per triangle.  Triangle (c, T) of
// corner c has the vertices c, c + 1, c + 1 + g (T = 0) or c, c + 1 + g, c + g
// (T = 1) and enters the stencils of those three points, each through the two
// arcs from that point to the other two vertices.  One thread per corner and
// level sums, for both fields and both triangles, the contributions of the
// triangle's users (interior points other than the sgp pentagon point) to the
// bar of its gradient, transposes the gradient, and stores the bars of its
// vertices: S[f * 6 + T * 3 + v] at the corner's address (f = 0, 1 the two
// fields, f = 2 rhog, whose bar from arc (u, w) lands on u and w).  Corners
// 0..g-2 in i and j, every level.
struct HDiffKhTriA {
  static constexpr const char* kName = "hdiff_pass2_adj";
  Dims d; Consts c; Geo G;
  const real* in[2];   // fields 4, 5 of the first pass (halo-filled)
  const real* ob[2];   // bars of pass 2's outputs 4, 5
  const real* rhog;
  real* S[18];
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (!pt_of(d, plane, x, p) || p.pole || p.i > d.g - 2 || p.j > d.g - 2) return;
    const long g = d.g, nreg = d.nreg, ac = p.a, cc = p.ci;
    const bool sgp = G.sgp[p.l] != 0;
    const real kfac[2] = {c.CVdry * c.GAMMA_H, c.HDIFF_FACT_RHO * c.GAMMA_H};
    for (int T = 0; T < 2; ++T) {
      // vertex offsets from the corner
      const long vo[3] = {0, (T == 0) ? 1 : 1 + g, (T == 0) ? 1 + g : g};
      const int vi[3] = {0, 1, (T == 0) ? 1 : 0}, vj[3] = {0, (T == 0) ? 0 : 1, 1};
      real tv[2][3];
      for (int f = 0; f < 2; ++f) tri_grad(G, nreg, cc, T, in[f], ac, (int)g, tv[f]);
      real tb[2][3] = {{real(0.0), real(0.0), real(0.0)}, {real(0.0), real(0.0), real(0.0)}};
      real rb[3] = {real(0.0), real(0.0), real(0.0)};
      real rv[3];
      for (int v = 0; v < 3; ++v) rv[v] = rhog[ac + vo[v]];
      // user u (vertex u) reaches the other two vertices through its arcs
      // eA (to vertex vA) and eB (to vertex vB)
      for (int u = 0; u < 3; ++u) {
        const int ui = p.i + vi[u], uj = p.j + vj[u];
        if (ui < 1 || ui > d.g - 2 || uj < 1 || uj > d.g - 2) continue;
        if (sgp && ui == 1 && uj == 1) continue;
        int eA, eB, vA, vB;
        if (T == 0) {
          if (u == 0) { eA = 0; vA = 2; eB = 5; vB = 1; }
          else if (u == 1) { eA = 2; vA = 0; eB = 1; vB = 2; }
          else { eA = 4; vA = 1; eB = 3; vB = 0; }
        } else {
          if (u == 0) { eA = 1; vA = 2; eB = 0; vB = 1; }
          else if (u == 1) { eA = 3; vA = 0; eB = 2; vB = 2; }
          else { eA = 5; vA = 1; eB = 4; vB = 0; }
        }
        const long au = ac + vo[u], cu = cc + vo[u];
        real cA[3], cB[3];
        for (int dd = 0; dd < 3; ++dd) {
          cA[dd] = G.cdiff[(long)(dd * 6 + eA) * nreg + cu];
          cB[dd] = G.cdiff[(long)(dd * 6 + eB) * nreg + cu];
        }
        for (int f = 0; f < 2; ++f) {
          const real o = -ob[f][au];
          const real kfA = real(0.5) * (rv[u] * kfac[f] + rv[vA] * kfac[f]);
          const real kfB = real(0.5) * (rv[u] * kfac[f] + rv[vB] * kfac[f]);
          real kbA = real(0.0), kbB = real(0.0);
          for (int dd = 0; dd < 3; ++dd) {
            const real ccA = cA[dd] * o, ccB = cB[dd] * o;
            tb[f][dd] += ccA * kfA + ccB * kfB;
            kbA += ccA * tv[f][dd];
            kbB += ccB * tv[f][dd];
          }
          const real hA = real(0.5) * kbA * kfac[f], hB = real(0.5) * kbB * kfac[f];
          rb[u] += hA + hB;
          rb[vA] += hA;
          rb[vB] += hB;
        }
      }
      for (int f = 0; f < 2; ++f) {
        real b0 = real(0.0), b1 = real(0.0), b2 = real(0.0);
        for (int dd = 0; dd < 3; ++dd) {
          const real c1 = rc(G.cintp, nreg, (dd * 3 + 0) * 2 + T, cc);
          const real c2 = rc(G.cintp, nreg, (dd * 3 + 1) * 2 + T, cc);
          const real c3 = rc(G.cintp, nreg, (dd * 3 + 2) * 2 + T, cc);
          const real o = tb[f][dd] / real(3.0);
          b0 += (real(2.0) * c1 - c2 - c3) * o;
          b1 += (-c1 + real(2.0) * c2 - c3) * o;
          b2 += (-c1 - c2 + real(2.0) * c3) * o;
        }
        S[f * 6 + T * 3 + 0][ac] = b0;
        S[f * 6 + T * 3 + 1][ac] = b1;
        S[f * 6 + T * 3 + 2][ac] = b2;
      }
      for (int v = 0; v < 3; ++v) S[12 + T * 3 + v][ac] = rb[v];
    }
  }
};

// the gather of HDiffKhTriA's vertex bars: point x is vertex 0 of triangles
// (x, 0) and (x, 1), vertex 1 of (x - 1, 0) and (x - 1 - g, 1), vertex 2 of
// (x - 1 - g, 0) and (x - g, 1); clears the output bars it has consumed
struct HDiffKhG {
  static constexpr const char* kName = "hdiff_pass2_adj";
  Dims d; Geo G;
  const real* S[18];
  real* ob[2];
  real* ib[2];
  real* rhog_b;
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (!pt_of(d, plane, x, p) || p.pole) return;
    const int gm = d.g - 2;
    const long g = d.g, a = p.a;
    // (corner offset, T, vertex, valid)
    const long co[6] = {0, 0, -1, -1 - g, -1 - g, -g};
    const int tT[6] = {0, 1, 0, 1, 0, 1}, tv[6] = {0, 0, 1, 1, 2, 2};
    const bool ok[6] = {p.i <= gm && p.j <= gm, p.i <= gm && p.j <= gm, p.i >= 1 && p.j <= gm,
                        p.i >= 1 && p.j >= 1, p.i >= 1 && p.j >= 1, p.i <= gm && p.j >= 1};
    real acc[3] = {real(0.0), real(0.0), real(0.0)};
    for (int n = 0; n < 6; ++n) {
      if (!ok[n]) continue;
      const int sidx = tT[n] * 3 + tv[n];
      for (int f = 0; f < 3; ++f) acc[f] += S[f * 6 + sidx][a + co[n]];
    }
    ib[0][a] += acc[0];
    ib[1][a] += acc[1];
    rhog_b[a] += acc[2];
    if (p.inner && !(p.i == 1 && p.j == 1 && G.sgp[p.l])) {
      ob[0][a] = real(0.0);
      ob[1][a] = real(0.0);
    }
  }
};

// ----------------------------------------------------------------------------
// kinetic energy at level k of a column (rhogkin_k): transpose for the
// accessor's values; add(kk, q, v) receives the bar of get(kk, q)
// ----------------------------------------------------------------------------
// (always inlined: callers that sum the bars in registers need the level
// offsets of add() to be constants)
template <class F, class ADD>
NH_HD inline __attribute__((always_inline)) void rhogkin_k_t(const VProf<real>& V, int kmin, int kmax, int k, F get,
                                                             ADD add, real kb) {
  if (k < kmin || k > kmax) return;
  const real rgk = get(k, 0);
  const real vx = get(k, 1), vy = get(k, 2), vz = get(k, 3);
  const real kh = real(0.5) * (vx * vx + vy * vy + vz * vz) / rgk;
  real rgk_b = -kb * kh / rgk;
  add(k, 1, kb * vx / rgk);
  add(k, 2, kb * vy / rgk);
  add(k, 3, kb * vz / rgk);
  if (k + 1 <= kmax) {
    const real w = get(k + 1, 4);
    const real D = V.c2w0[k + 1] * get(k + 1, 0) + V.c2w1[k + 1] * rgk;
    const real vup = real(0.5) * (w * w) / D;
    const real ub = kb * V.w2c0[k];
    add(k + 1, 4, ub * w / D);
    const real Db = -ub * vup / D;
    add(k + 1, 0, Db * V.c2w0[k + 1]);
    rgk_b += Db * V.c2w1[k + 1];
  }
  if (k >= kmin + 1) {
    const real w = get(k, 4);
    const real D = V.c2w0[k] * rgk + V.c2w1[k] * get(k - 1, 0);
    const real vdn = real(0.5) * (w * w) / D;
    const real db = kb * V.w2c1[k];
    add(k, 4, db * w / D);
    const real Db = -db * vdn / D;
    rgk_b += Db * V.c2w0[k];
    add(k - 1, 0, Db * V.c2w1[k]);
  }
  add(k, 0, rgk_b);
}

// ----------------------------------------------------------------------------
// momentum advection (AdvMomP): gadv bars -> bars of vx, vy, vz, w (diag),
// rhog, rhogvx..vz, rhogw (prog), as three gathers, without atomics:
//   AdvMomG1, every point: the bars ob are consumed; at a compute point of a
//     physical level the bar of the advection tendency dvv (the tail's
//     projection and ALPHA folded in) -> S, the Coriolis term's rhog bar into
//     prog_b; S = 0 elsewhere;
//   AdvMomG2, every point and level: the bars the point receives — from the
//     divergence terms of the compute points whose stencil holds it, from its
//     own tail (Coriolis, the vertical terms) and from the vertical terms of
//     its column's levels k - 1 and k + 1 — into prog_b (rhogv, rhogw) and
//     diag_b (v), and the w part of its merged-velocity bar -> W;
//   AdvMomG3, every point and level: the w bars of the merged velocity's
//     interpolation, from W at the level and the one below.
// Each sum is a forward point's transposed terms, grouped by the point they
// land on (a compute point's stencil neighbour m is the point it adds to with
// that point's coefficient m).
// ----------------------------------------------------------------------------
struct AdvMomG1 {
  static constexpr const char* kName = "adv_mom_adj1";
  Dims d; Consts c; Geo G;
  const real* prog;
  const real* diag;
  real* ob;        // 4 fields: bars of out[0..3] (consumed)
  real* prog_b;    // rhog bar (Coriolis)
  real* S[3];      // out: the dvv bars
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (!pt_of(d, plane, x, p)) return;
    const long NF = d.NF, a = p.a, ks = p.ks;
    const int k = p.k;
    real o[4];
    for (int q = 0; q < 4; ++q) {
      o[q] = ob[q * NF + a];
      ob[q * NF + a] = real(0.0);
    }
    real dv[3] = {real(0.0), real(0.0), real(0.0)};
    if (pt_compute(p) && k >= d.kmin && k <= d.kmax) {
      const real R = c.RADIUS;
      real x0[3];
      point_pos(G, d, p, x0);
      const real* wd = diag + I_W * NF;
      const real wc = G.cfact[k] * wd[a + ks] + G.dfact[k] * wd[a];
      real vvm0[3];
      for (int q = 0; q < 3; ++q) vvm0[q] = diag[(I_VX + q) * NF + a] + wc * x0[q] / R;
      const real gxc = x0[0] / R, gyc = x0[1] / R, gzc = x0[2] / R;
      const real prd_b = o[3] * c.ALPHA - (o[0] * gxc + o[1] * gyc + o[2] * gzc);
      dv[0] = o[0] + prd_b * gxc;
      dv[1] = o[1] + prd_b * gyc;
      dv[2] = o[2] + prd_b * gzc;
      prog_b[a] = prog_b[a] + real(2.0) * c.OHM * (dv[0] * vvm0[1] - dv[1] * vvm0[0]);
    }
    for (int q = 0; q < 3; ++q) S[q][a] = dv[q];
  }
};

struct AdvMomG2 {
  static constexpr const char* kName = "adv_mom_adj2";
  Dims d; Consts c; Geo G;
  const real* prog;
  const real* diag;
  const real* S[3];
  real* prog_b;
  real* diag_b;
  real* W;
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (!pt_of(d, plane, x, p)) return;
    const long NF = d.NF, a = p.a, ks = p.ks;
    const int kmin = d.kmin, kmax = d.kmax, k = p.k;
    if (k < kmin || k > kmax) {
      W[a] = real(0.0);
      return;
    }
    const real R = c.RADIUS;
    const real* rw = prog + 4 * NF;
    const real* rg = prog;
    const real* wd = diag + I_W * NF;
    // the point's position (a pole point's own) and merged velocity
    real xp[3];
    if (p.pole) {
      for (int q = 0; q < 3; ++q) xp[q] = G.gx_pl[(p.v * 2 + p.lp) * 3 + q];
    } else {
      for (int q = 0; q < 3; ++q) xp[q] = G.gx[q * d.nreg + p.ci];
    }
    auto vv3 = [&](long b, int kk, real* vo) {
      const real wc = G.cfact[kk] * wd[b + ks] + G.dfact[kk] * wd[b];
      for (int q = 0; q < 3; ++q) vo[q] = diag[(I_VX + q) * NF + b] + wc * xp[q] / R;
    };
    real vvP[3];
    vv3(a, k, vvP);
    real vvb[3] = {real(0.0), real(0.0), real(0.0)};
    real fb[3] = {real(0.0), real(0.0), real(0.0)};
    // divergence terms of the compute points Q whose stencil holds this point
    auto from = [&](long aQ, const real cdir[3]) {
      real divb[3];
      for (int q = 0; q < 3; ++q) divb[q] = -S[q][aQ];
      for (int q = 0; q < 3; ++q)
        for (int dd = 0; dd < 3; ++dd) {
          const real t = cdir[dd] * divb[q];
          fb[dd] += t * vvP[q];
          vvb[q] += t * prog[(1 + dd) * NF + a];
        }
    };
    if (p.pole) {
      real cdir[3];
      for (int dd = 0; dd < 3; ++dd) cdir[dd] = G.cdiv_pl[(p.v * 2 + p.lp) * 3 + dd];
      from(a - p.v, cdir);
    } else {
      const long g = d.g;
      for (int m = 0; m < 7; ++m) {
        const int iq = p.i - (((4201 >> (2 * m)) & 3) - 1), jq = p.j - (((421 >> (2 * m)) & 3) - 1);
        if (iq < 1 || iq > d.g - 2 || jq < 1 || jq > d.g - 2) continue;
        const long off = nb_off(m, g);
        real cdir[3];
        for (int dd = 0; dd < 3; ++dd) cdir[dd] = rc(G.cdiv, d.nreg, dd * 7 + m, p.ci - off);
        from(a - off, cdir);
      }
    }
    real rwb = real(0.0);
    if (pt_compute(p)) {
      // its own tail: Coriolis and the vertical terms' vvm0 parts
      const real dv[3] = {S[0][a], S[1][a], S[2][a]};
      const real rh = rg[a];
      vvb[1] += real(2.0) * rh * c.OHM * dv[0];
      vvb[0] -= real(2.0) * rh * c.OHM * dv[1];
      if (k + 1 <= kmax) {
        const real wup = rw[a + ks];
        for (int q = 0; q < 3; ++q) vvb[q] += (-dv[q] * G.rdgz[k]) * wup * G.bfact[k + 1];
      }
      if (k >= kmin + 1) {
        const real w0 = rw[a];
        real dn[3];
        vv3(a - ks, k - 1, dn);
        real w0b = real(0.0);
        for (int q = 0; q < 3; ++q) {
          const real lob = dv[q] * G.rdgz[k];
          w0b += lob * (G.afact[k] * vvP[q] + G.bfact[k] * dn[q]);
          vvb[q] += lob * w0 * G.afact[k];
        }
        rwb += w0b;
      }
      // the up term of level k - 1 and the down term of level k + 1
      if (k - 1 >= kmin) {
        const real wup = rw[a];
        real vm[3];
        vv3(a - ks, k - 1, vm);
        real wupb = real(0.0);
        for (int q = 0; q < 3; ++q) {
          const real hib = -S[q][a - ks] * G.rdgz[k - 1];
          wupb += hib * (G.afact[k] * vvP[q] + G.bfact[k] * vm[q]);
          vvb[q] += hib * wup * G.afact[k];
        }
        rwb += wupb;
      }
      if (k + 1 <= kmax) {
        const real w0 = rw[a + ks];
        for (int q = 0; q < 3; ++q) vvb[q] += (S[q][a + ks] * G.rdgz[k + 1]) * w0 * G.bfact[k + 1];
      }
    }
    for (int dd = 0; dd < 3; ++dd) prog_b[(1 + dd) * NF + a] += fb[dd];
    prog_b[4 * NF + a] += rwb;
    real wcb = real(0.0);
    for (int q = 0; q < 3; ++q) {
      diag_b[(I_VX + q) * NF + a] += vvb[q];
      wcb += vvb[q] * xp[q] / R;
    }
    W[a] = wcb;
  }
};

struct AdvMomG3 {
  static constexpr const char* kName = "adv_mom_adj3";
  Dims d; Geo G;
  const real* W;
  real* wdb;       // bar of diag w
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (!pt_of(d, plane, x, p)) return;
    const long a = p.a, ks = p.ks;
    const int k = p.k;
    real add = real(0.0);
    if (k - 1 >= d.kmin && k - 1 <= d.kmax) add += W[a - ks] * G.cfact[k - 1];
    if (k >= d.kmin && k <= d.kmax) add += W[a] * G.dfact[k];
    wdb[a] += add;
  }
};

// ----------------------------------------------------------------------------
// Pre_Post (PrePostP): in place on prog_b (bar of the boundary-conditioned
// PROG on entry, bar of the PROG before the kernel on exit), with the bars of
// DIAG and eth (consumed)
// ----------------------------------------------------------------------------
struct PrePostA {
  static constexpr const char* kName = "prepost_adj";
  Dims d; Consts c; Geo G;
  const real* prog;    // boundary-conditioned PROG (the forward kernel's result)
  const real* rhogq;
  real* prog_b;
  real* diag_b;
  real* eth_b;
  struct Th { real rho, ein, q, qd, cv, tem, Rm, pre; };
  NH_HD Th thermo(long a) const {
    const long NF = d.NF;
    Th t;
    const real rhog = prog[a];
    t.rho = rhog / real(1.0);
    t.ein = prog[5 * NF + a] / rhog;
    t.q = rhogq[a] / rhog;
    t.qd = real(1.0) - t.q;
    t.cv = t.q * c.CVW_qv + t.qd * c.CVdry;
    t.tem = t.ein / t.cv;
    t.Rm = t.qd * c.Rdry + t.q * c.Rvap;
    t.pre = t.rho * t.tem * t.Rm;
    return t;
  }
  // bars of (rho, tem, pre) at an interior point -> prog_b (rhog, rhoge)
  NH_HD void thermo_t(long a, const Th& t, real rhob, real temb, real preb) const {
    const long NF = d.NF;
    const real rhog = prog[a];
    // pre = rho * tem * Rm
    rhob += preb * t.tem * t.Rm;
    temb += preb * t.rho * t.Rm;
    const real Rmb = preb * t.rho * t.tem;
    // tem = ein / cv
    const real einb = temb / t.cv;
    const real cvb = -temb * t.tem / t.cv;
    // cv, Rm from q
    const real qb = cvb * (c.CVW_qv - c.CVdry) + Rmb * (c.Rvap - c.Rdry);
    // ein = rhoge / rhog, q = rhogq / rhog, rho = rhog
    aadd(&prog_b[5 * NF + a], einb / rhog);
    aadd(&prog_b[a], rhob - einb * t.ein / rhog - qb * t.q / rhog);
  }
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (!pt_of(d, plane, x, p)) return;
    const long NF = d.NF, a = p.a, ks = p.ks;
    const int kmin = d.kmin, kmax = d.kmax, k = p.k;
    real db[6];
    for (int q = 0; q < 6; ++q) {
      db[q] = diag_b[q * NF + a];
      diag_b[q * NF + a] = real(0.0);
    }
    const real eb = eth_b[a];
    eth_b[a] = real(0.0);
    if (k >= kmin && k <= kmax) {
      const Th t = thermo(a);
      const real rhog = prog[a];
      real rhob = real(0.0), temb = db[I_TEM], preb = db[I_PRE];
      // eth = ein + pre / rho
      real einb_x = eb;
      preb += eb / t.rho;
      rhob += -eb * t.pre / (t.rho * t.rho);
      aadd(&prog_b[5 * NF + a], einb_x / rhog);
      real rgb = -einb_x * t.ein / rhog;
      // velocities
      for (int q = 0; q < 3; ++q) {
        const real v = prog[(1 + q) * NF + a] / rhog;
        aadd(&prog_b[(1 + q) * NF + a], db[I_VX + q] / rhog);
        rgb += -db[I_VX + q] * v / rhog;
      }
      aadd(&prog_b[a], rgb);
      if (k > kmin) {
        const real D = G.c2w0[k] * prog[a] + G.c2w1[k] * prog[a - ks];
        const real rw = prog[4 * NF + a];
        const real wb = db[I_W];
        aadd(&prog_b[4 * NF + a], wb / D);
        const real Db = -wb * rw / (D * D);
        aadd(&prog_b[a], Db * G.c2w0[k]);
        aadd(&prog_b[a - ks], Db * G.c2w1[k]);
      } else {
        // rhogw at kmin is set to zero: its incoming bar is dropped, w = 0
        prog_b[4 * NF + a] = real(0.0);
      }
      thermo_t(a, t, rhob, temb, preb);
      return;
    }
    // ghost rows: every PROG value written here is a function of the nearest
    // two physical levels; the incoming bars of the ghost values are consumed
    const bool top = (k == kmax + 1);
    const long an = top ? a - ks : a + ks;
    const long an2 = top ? a - 2 * ks : a + 2 * ks;
    const Th tn = thermo(an);
    const Th tn2 = thermo(an2);
    const real dphi = top ? (G.phi[kmax + 1] - G.phi[kmax - 1]) : (G.phi[kmin - 1] - G.phi[kmin + 1]);
    const real tem_g = tn.tem;
    const real pre_g = tn2.pre - tn.rho * dphi;
    const real rho_g = pre_g / (c.Rdry * tem_g);
    const real rg_g = rho_g * real(1.0);
    const real ein_g = c.CVdry * tem_g;
    real pb[6];
    for (int q = 0; q < 6; ++q) {
      pb[q] = prog_b[q * NF + a];
      prog_b[q * NF + a] = real(0.0);
    }
    const real sgn = top ? real(1.0) : -real(1.0);
    real rggb = pb[0] + pb[5] * ein_g;
    real eingb = pb[5] * rg_g;
    // eth = ein_g + pre_g / rho_g
    eingb += eb;
    real pregb = db[I_PRE] + eb / rho_g;
    real rhogb = -eb * pre_g / (rho_g * rho_g);
    real temgb = db[I_TEM];
    // velocities: v = sgn * (rhogv(an) / rhog(an)) * rg_g; diag = v / rg_g
    real rgnb = real(0.0);
    const real rgn = prog[an];
    for (int q = 0; q < 3; ++q) {
      const real u = prog[(1 + q) * NF + an] / rgn;
      const real v = sgn * u * rg_g;
      real vb = pb[1 + q] + db[I_VX + q] / rg_g;
      rggb += -db[I_VX + q] * v / (rg_g * rg_g);
      const real ub = vb * sgn * rg_g;
      rggb += vb * sgn * u;
      aadd(&prog_b[(1 + q) * NF + an], ub / rgn);
      rgnb += -ub * u / rgn;
    }
    aadd(&prog_b[an], rgnb);
    // rg_g = rho_g = pre_g / (Rdry tem_g)
    rhogb += rggb;
    pregb += rhogb / (c.Rdry * tem_g);
    temgb += -rhogb * rho_g / tem_g;
    temgb += eingb * c.CVdry;
    // pre_g = tn2.pre - tn.rho * dphi; tem_g = tn.tem
    thermo_t(an, tn, -pregb * dphi, temgb, real(0.0));
    thermo_t(an2, tn2, real(0.0), real(0.0), pregb);
  }
};

// ----------------------------------------------------------------------------
// total specific energy (EthTotPMP): ethtot bar -> eth, PROG bars
// ----------------------------------------------------------------------------
struct EthTotA {
  static constexpr const char* kName = "ethtot_adj";
  Dims d; Consts c; Geo G;
  const real* prog;
  real* ethtot_b;
  real* eth_b;
  real* prog_b;
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (!pt_of(d, plane, x, p)) return;
    const long NF = d.NF, a = p.a, ks = p.ks;
    const long b = a - (long)p.k * ks;
    const real eb = ethtot_b[a];
    ethtot_b[a] = real(0.0);
    aadd(&eth_b[a], eb);
    auto get = [&](int kk, int q) { return prog[q * NF + b + (long)kk * ks]; };
    auto add = [&](int kk, int q, real v) { aadd(&prog_b[q * NF + b + (long)kk * ks], v); };
    const real kin0 = rhogkin_k(G, d.kmin, d.kmax, p.k, get);
    const real rg = prog[a];
    aadd(&prog_b[a], -eb * kin0 / (rg * rg));
    rhogkin_k_t(G, d.kmin, d.kmax, p.k, get, add, eb / rg);
  }
};


// ----------------------------------------------------------------------------
// small step path1 (Path1P), gather form: the thread of point x adds the
// bars x receives.  The gradient term of an output at level k reads the split
// rhoge at level k + shift(k) (kmin -> kmin+1, ghost rows from the nearest
// physical level's gradient level); the pointwise terms of a ghost-row output
// land on the nearest physical level (the thread there handles the ghost row,
// whose momentum before scaling is the forward output at that level).
// ----------------------------------------------------------------------------
struct Path1A {
  static constexpr const char* kName = "vi_path1_adj";
  Dims d; Consts c; Geo G;
  real dt;
  const real* prog;
  const real* dv[3];       // forward output (halo-filled)
  const real* dv_b[3];     // bars of the output
  real* prog_b;
  real* ps_b[6];
  real* gT_b;
  real* dd_b[3];
  NH_HD int shift(int k) const {
    const int kmin = d.kmin, kmax = d.kmax;
    if (k == kmin || k == kmin - 1) return kmin + 1 - k;
    if (k == kmax + 1) return -1;
    return 0;
  }
  // bar of the gradient (after the projection transpose) of the output at address a, level k
  NH_HD void grad_bar(long a, long ks, int k, const real xx[3], real gb[3]) const {
    const int kmin = d.kmin, kmax = d.kmax;
    real vb[3];
    for (int q = 0; q < 3; ++q) vb[q] = dv_b[q][a];
    if (k < kmin || k > kmax) {
      const bool top = (k == kmax + 1);
      const long an = top ? a - ks : a + ks;
      const real sgn = top ? real(1.0) : -real(1.0);
      for (int q = 0; q < 3; ++q) vb[q] = sgn * vb[q] * prog[a] / prog[an];
    }
    for (int q = 0; q < 3; ++q) gb[q] = -vb[q] * dt;
    proj_t(xx, c.RADIUS, gb);
  }
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (!pt_of(d, plane, x, p)) return;
    const long NF = d.NF, a = p.a, ks = p.ks;
    const int kmin = d.kmin, kmax = d.kmax, k = p.k;
    const real RovCV = c.Rdry / c.CVdry;
    // outputs whose gradient level is k: their level offsets from k
    int nsrc = 0;
    int dk[3];
    if (k >= kmin + 2 && k <= kmax) dk[nsrc++] = 0;
    if (k == kmin + 1) { dk[nsrc++] = 0; dk[nsrc++] = -1; dk[nsrc++] = -2; }
    if (k == kmax) dk[nsrc++] = 1;
    real eb = real(0.0);
    for (int s2 = 0; s2 < nsrc; ++s2) {
      const int ko = k + dk[s2];
      if (p.pole) {
        // the pole centre reads the six pole points; x is its neighbour v
        const long ac = a - p.v + (long)dk[s2] * ks;
        real xx[3], gb[3];
        for (int q = 0; q < 3; ++q) xx[q] = G.gx_pl[p.lp * 3 + q];
        grad_bar(ac, ks, ko, xx, gb);
        for (int dd2 = 0; dd2 < 3; ++dd2) eb += G.cgrad_pl[(p.v * 2 + p.lp) * 3 + dd2] * RovCV * gb[dd2];
      } else {
        for (int m = 0; m < 7; ++m) {
          long ao, co;
          if (!src_of(d, p, m, &ao, &co)) continue;
          ao += (long)dk[s2] * ks;
          real xx[3], gb[3];
          for (int q = 0; q < 3; ++q) xx[q] = G.gx[q * d.nreg + co];
          grad_bar(ao, ks, ko, xx, gb);
          for (int dd2 = 0; dd2 < 3; ++dd2) eb += rc(G.cgrad, d.nreg, dd2 * 7 + m, co) * RovCV * gb[dd2];
        }
      }
    }
    if (nsrc > 0) ps_b[I_RHOGE][a] += eb;
    // pointwise terms at the compute points (ghost rows folded into the nearest
    // level); the damping-force bars are stored (zeros elsewhere): nothing
    // writes them earlier in the reversed small step
    if (!pt_compute(p) || k < kmin || k > kmax) {
      for (int q = 0; q < 3; ++q) dd_b[q][a] = real(0.0);
      return;
    }
    real vt[3];
    for (int q = 0; q < 3; ++q) vt[q] = dv_b[q][a];
    real rnb = real(0.0);
    for (int s2 = 0; s2 < 2; ++s2) {
      if (s2 == 0 ? k != kmin : k != kmax) continue;
      const bool top = (s2 == 1);
      const long ag = top ? a + ks : a - ks;     // the ghost row
      const real sgn = top ? real(1.0) : -real(1.0);
      const real rgn = prog[a], rgg = prog[ag];
      real rgb = real(0.0);
      for (int q = 0; q < 3; ++q) {
        const real vb = dv_b[q][ag];
        const real vn = dv[q][a];
        rgb += sgn * vb * vn / rgn;
        rnb -= sgn * vb * vn * rgg / (rgn * rgn);
        vt[q] += sgn * vb * rgg / rgn;
      }
      prog_b[ag] += rgb;
    }
    prog_b[a] += rnb;
    for (int q = 0; q < 3; ++q) {
      ps_b[1 + q][a] += vt[q];
      gT_b[(1 + q) * NF + a] += vt[q] * dt;
      dd_b[q][a] = vt[q] * dt;
    }
  }
};

// ----------------------------------------------------------------------------
// small step vi_main, horizontal part (ViMainH), gather form: the thread of
// point x adds the bars of dv, eth, ethtot, PROG momentum at x from the
// divergences of the compute points whose stencil holds x, and the bar of
// gT's energy component at x
// ----------------------------------------------------------------------------
struct ViMainHA {
  static constexpr const char* kName = "vi_main_h_adj";
  Dims d; Consts c; Geo G;
  const real* prog;
  const real* dv[3];
  const real* eth;
  const real* ethtot;
  const real* hv_b[3];
  real* dv_b[3];
  real* eth_b;
  real* ethtot_b;
  real* prog_b;
  real* gT_b;
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (!pt_of(d, plane, x, p)) return;
    const long NF = d.NF, a = p.a;
    const int k = p.k;
    if (k < d.kmin || k > d.kmax) return;
    const real rc_ = c.Rdry / c.CVdry;
    real w0[3] = {real(0.0), real(0.0), real(0.0)}, w1[3] = {real(0.0), real(0.0), real(0.0)},
         w2[3] = {real(0.0), real(0.0), real(0.0)};
    auto take = [&](long ao, const real cdir[3]) {
      const real h0 = hv_b[0][ao], h1 = -hv_b[1][ao] * rc_, h2 = hv_b[2][ao];
      for (int q = 0; q < 3; ++q) {
        w0[q] += cdir[q] * h0;
        w1[q] += cdir[q] * h1;
        w2[q] += cdir[q] * h2;
      }
    };
    if (p.pole) {
      real cdir[3];
      for (int q = 0; q < 3; ++q) cdir[q] = G.cdiv_pl[(p.v * 2 + p.lp) * 3 + q];
      take(a - p.v, cdir);
    } else {
      for (int m = 0; m < 7; ++m) {
        long ao, co;
        if (!src_of(d, p, m, &ao, &co)) continue;
        real cdir[3];
        for (int q = 0; q < 3; ++q) cdir[q] = rc(G.cdiv, d.nreg, q * 7 + m, co);
        take(ao, cdir);
      }
    }
    const real e = eth[a], et = ethtot[a];
    real eb = real(0.0), etb = real(0.0);
    for (int q = 0; q < 3; ++q) {
      const real dq = dv[q][a];
      dv_b[q][a] += w0[q] + w1[q] * e + w2[q] * et;
      eb += w1[q] * dq;
      prog_b[(1 + q) * NF + a] += w2[q] * et;
      etb += w2[q] * (prog[(1 + q) * NF + a] + dq);
    }
    eth_b[a] += eb;
    ethtot_b[a] += etb;
    if (pt_compute(p)) gT_b[5 * NF + a] += hv_b[1][a] * rc_;
  }
};

// ----------------------------------------------------------------------------
// small step vi_main, vertical part (ViMainV): the transposed tridiagonal
// solve in FP64, x being the forward solution (the new rhogw split as stored;
// the matrix bars are -y x^T on the band).  Two launches.  ViMainVAy: the
// transposed Thomas sweeps, one thread per compute column, storing the
// solution y of M^T y = xbar into Y at levels kmin+1..kmax (NK > 0, the
// number of unknowns, keeps the elimination in registers; NK = 0 uses the
// FP64 column scratch Y, Gm).  ViMainVAb: one thread per compute point and
// level kmin..kmax adds its own level's bars, from y and x there and from
// level k+1's y (the source row at k+1 also reads level k), so every bar
// element is read and written once, by one thread, and the levels run in
// parallel.
// ----------------------------------------------------------------------------
template <int NK>
struct ViMainVAy {
  static constexpr const char* kName = "vi_main_v_adj";
  static constexpr int kUnroll = NK > 0 ? NK : 4;
  Dims d;
  const real* out_b;   // bar of the new rhogw split
  const double* Mc; const double* Mu; const double* Ml;
  double* Y;           // y (and, for NK = 0, the elimination's u first)
  double* Gm;          // NK = 0: the elimination's multipliers
  NH_HD void operator()(long t) const {
    Cell cl = cell_of(d, t);
    if (cl.pole ? cl.v != 0 : !cl.inner) return;
    const long ks = cl.ks, b = cl.base;
    const int kmin = d.kmin, kmax = d.kmax;
    const int N = (NK > 0) ? NK : kmax - kmin;
    double ub[NK > 0 ? NK : 1], gb[NK > 0 ? NK : 1];
    long a = b + (long)kmin * ks;
    double beta = 0.0, u_prev = 0.0, mu_m = 0.0;
#pragma unroll kUnroll
    for (int n = 0; n < N; ++n) {
      a += ks;
      const double xb = (double)out_b[a];
      const double mc = Mc[a], ml = Ml[a];
      double u;
      if (n == 0) {
        beta = mc;
        u = xb / beta;
      } else {
        const double gm = ml / beta;
        beta = mc - mu_m * gm;
        u = (xb - mu_m * u_prev) / beta;
        if (NK > 0) gb[NK > 0 ? n : 0] = gm;
        else Gm[a] = gm;
      }
      if (NK > 0) ub[NK > 0 ? n : 0] = u;
      else Y[a] = u;
      u_prev = u;
      mu_m = Mu[a];
    }
    double y = u_prev, gm_next = 0.0;
#pragma unroll kUnroll
    for (int n = N - 1; n >= 0; --n) {
      if (n < N - 1) y = ((NK > 0) ? ub[NK > 0 ? n : 0] : Y[a]) - gm_next * y;
      if (NK == 0 && n > 0) gm_next = Gm[a];
      if (NK > 0 && n > 0) gm_next = gb[NK > 0 ? n : 0];
      Y[a] = y;
      a -= ks;
    }
  }
};

struct ViMainVAb {
  static constexpr const char* kName = "vi_main_v_adj_bars";
  Dims d; Consts c; Geo G;
  double dt;
  const real* x;       // forward solution (rhogw split after the step)
  const double* Y;
  real* ps_b[6];
  real* gT_b;
  double* Mc_b; double* Mu_b; double* Ml_b;
  real* hv_b[3];
  NH_HD void operator()(int plane, int xi) const {
    Pt p;
    if (!pt_of(d, plane, xi, p) || !pt_compute(p)) return;
    const int kmin = d.kmin, kmax = d.kmax, k = p.k;
    if (k < kmin || k > kmax) return;
    const long NF = d.NF, ks = p.ks, a = p.a;
    const VProf<double>& V = G.v64;
    const ConstsT<double>& C = c.d;
    const double RovCV = C.Rdry / C.CVdry;
    const double CVovRt2 = C.CVdry / C.Rdry / (dt * dt);
    const int N = kmax - kmin;
    // the terms the source row at level k + 1 adds at level k (zero above kmax)
    double cE = 0.0, cH1 = 0.0, cG = 0.0, cT0 = 0.0, cH0 = 0.0;
    if (k + 1 <= kmax) {
      const double s1 = ldc(&Y[a + ks]) * CVovRt2;
      const double Bb1 = -s1 * dt * V.rdgzh[k + 1], Cb1 = -s1 * dt * C.GRAV;
      const double grm1 = Cb1 * V.bfact[k + 1] * dt;
      cE = -Bb1 * RovCV;
      cH1 = -Bb1 * dt;
      cG = Cb1 * V.bfact[k + 1];
      cT0 = grm1;
      cH0 = -grm1;
    }
    if (k == kmin) {
      // level kmin: the bars the lowest source sent down
      radd(&ps_b[I_RHOGE][a], (real)cE);
      radd(&hv_b[1][a], (real)cH1);
      radd(&ps_b[I_RHOG][a], (real)cG);
      radd(&gT_b[a], (real)cT0);
      radd(&hv_b[0][a], (real)cH0);
      return;
    }
    const int n = k - kmin - 1;
    const double y = ldc(&Y[a]);
    const double xk = (double)ldc(&x[a]);
    const double x_dn = (n >= 1) ? (double)ldc(&x[a - ks]) : 0.0;
    const double x_up = (n <= N - 2) ? (double)ldc(&x[a + ks]) : 0.0;
    radd(&Mc_b[a], -y * xk);
    if (n >= 1) radd(&Ml_b[a], -y * x_dn);
    if (n <= N - 2) radd(&Mu_b[a], -y * x_up);
    const double s = y * CVovRt2;
    const double Ab = s, Bb = -s * dt * V.rdgzh[k], Cb = -s * dt * C.GRAV;
    const double grk = Cb * V.afact[k] * dt;
    radd(&ps_b[I_RHOGW][a], (real)(Ab * C.ALPHA));
    radd(&gT_b[4 * NF + a], (real)(Ab * dt));
    radd(&ps_b[I_RHOGE][a], (real)(Bb * RovCV + cE));
    radd(&hv_b[1][a], (real)(Bb * dt + cH1));
    radd(&ps_b[I_RHOG][a], (real)(Cb * V.afact[k] + cG));
    radd(&gT_b[a], (real)(grk + cT0));
    radd(&hv_b[0][a], (real)(-grk + cH0));
  }
  // every bar element of level k is added to once, by the thread of that
  // point and level, which reads it once before its one store: ld_own
  template <class T>
  NH_HD static void radd(T* p, T v) { *p = ld_own(p) + v; }
};

// ----------------------------------------------------------------------------
// small step mass and energy update (ViMainPost): bars of the new rhog and
// rhoge splits (we[0], we[2], consumed) -> inputs; we_b[1] receives the bar
// of the new rhogw split it read
// ----------------------------------------------------------------------------
// the point of a compute column at level k
NH_HD inline Pt pt_at(const Dims& d, const Cell& cl, int k) {
  Pt p;
  p.pole = cl.pole; p.inner = cl.inner;
  p.i = cl.i; p.j = cl.j; p.l = cl.l; p.v = cl.v; p.lp = cl.lp; p.k = k;
  p.ks = cl.ks;
  p.a = cl.base + (long)k * cl.ks;
  p.ci = cl.ci;
  (void)d;
  return p;
}

struct ViMainPostA {
  static constexpr const char* kName = "vi_main_post_adj";
  // the bar fields it adds to: PROG 0..4, split 0..5, dv 0..2, the new rhogw
  // split, hv 0 and 2, gT 0, ethtot, getot
  enum { F_P = 0, F_S = 5, F_D = 11, F_W1 = 14, F_H0 = 15, F_H2 = 16, F_G0 = 17, F_ET = 18, F_GE = 19, F_N = 20 };
  Dims d; Consts c; Geo G;
  real dt;
  const real* prog;
  const real* ps[6];
  const real* dv[3];
  const real* ethtot;
  const real* gT;
  const real* hv[3];
  const real* we1;
  real* we_b[3];
  real* prog_b;
  real* ps_b[6];
  real* dv_b[3];
  real* ethtot_b;
  real* gT_b;
  real* getot_b;
  real* hv_b[3];
  NH_HD real* fptr(int f) const {
    if (f < F_S) return prog_b + (long)f * d.NF;
    if (f < F_D) return ps_b[f - F_S];
    if (f < F_W1) return dv_b[f - F_D];
    switch (f) {
      case F_W1: return we_b[1];
      case F_H0: return hv_b[0];
      case F_H2: return hv_b[2];
      case F_G0: return gT_b;
      case F_ET: return ethtot_b;
      default: return getot_b;
    }
  }
  // the transpose at one compute point; add(field, dk, value) accumulates the
  // bar of field at level k + dk of the point's column (dk = -1 .. 2).  The
  // forward values are loaded first, at levels clamped into 0..kall-1 (as in
  // ViMainPost), so that no load waits behind a level test.
  template <class ADD>
  NH_HD void point(const Pt& p, ADD add) const {
    const long NF = d.NF, ks = p.ks, a = p.a;
    const int kmin = d.kmin, kmax = d.kmax, k = p.k;
    const long b = a - (long)k * ks;
    long A[4];
    for (int j = 0; j < 4; ++j) {
      int kk = k - 1 + j;
      kk = (kk < 0) ? 0 : ((kk > d.kall - 1) ? d.kall - 1 : kk);
      A[j] = b + (long)kk * ks;
    }
    const real* pg = ps[I_RHOG];
    real P0[3], S0[3], H0[3], T0[3], ET[3], RW[4], P4[2], PQ[3], SQ[3], DQ[3], S4[2];
    for (int j = 0; j < 3; ++j) {
      P0[j] = prog[A[j]];
      S0[j] = pg[A[j]];
      H0[j] = hv[0][A[j]];
      T0[j] = gT[A[j]];
      ET[j] = ethtot[A[j]];
    }
    for (int j = 0; j < 4; ++j) RW[j] = we1[A[j]];
    for (int j = 0; j < 2; ++j) {
      P4[j] = prog[4 * NF + A[1 + j]];
      S4[j] = ps[4][A[1 + j]];
    }
    for (int q = 0; q < 3; ++q) {
      PQ[q] = prog[(1 + q) * NF + a];
      SQ[q] = ps[1 + q][a];
      DQ[q] = dv[q][a];
    }
    const real b0 = we_b[0][a], b2 = we_b[2][a];
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
    // dk: level of the rs1 argument relative to k
    auto rs1_t = [&](int dk, real rb) {
      const int kk = k + dk;
      add(F_S + I_RHOG, dk, rb);
      add(F_G0, dk, rb * dt);
      if (kk >= kmin && kk <= kmax) {
        const real d2b = rb * dt;
        add(F_H0, dk, -d2b);
        if (kk + 1 <= kmax) add(F_W1, dk + 1, -d2b * G.rdgz[kk]);
        if (kk >= kmin + 1) add(F_W1, dk, d2b * G.rdgz[kk]);
      }
    };
    // the energy split of the nearest physical level (three adds of selected
    // values: a single add at a selected level would index the caller's
    // register sums dynamically)
    add(F_S + I_RHOGE, 1, (k < kmin) ? b2 : real(0.0));
    add(F_S + I_RHOGE, -1, (k > kmax) ? b2 : real(0.0));
    add(F_S + I_RHOGE, 0, (k >= kmin && k <= kmax) ? b2 : real(0.0));
    add(F_GE, 0, b2 * dt);
    const real dtotb = b2 * dt;
    add(F_S + I_RHOG, 0, b2 * G.phi[k]);
    rs1_t(0, b0 - b2 * G.phi[k]);
    if (k >= kmin && k <= kmax) {
      add(F_H2, 0, -dtotb);
      const real hib = -dtotb * G.rdgz[k], lob = dtotb * G.rdgz[k];
      if (k + 1 <= kmax) {
        const real W1 = P4[1] + RW[2];
        const real E1 = G.afact[k + 1] * ET[2] + G.bfact[k + 1] * ET[1];
        add(F_P + 4, 1, hib * E1);
        add(F_W1, 1, hib * E1);
        add(F_ET, 1, hib * W1 * G.afact[k + 1]);
        add(F_ET, 0, hib * W1 * G.bfact[k + 1]);
      }
      if (k >= kmin + 1) {
        const real W0 = P4[0] + RW[1];
        const real E0 = G.afact[k] * ET[1] + G.bfact[k] * ET[0];
        add(F_P + 4, 0, lob * E0);
        add(F_W1, 0, lob * E0);
        add(F_ET, 0, lob * W0 * G.afact[k]);
        add(F_ET, -1, lob * W0 * G.bfact[k]);
      }
    }
    // kinetic energies before (PROG + PROG_split) and after (new split) the step
    auto g10 = [&](int kk, int q) {
      const int j = kk - k + 1;
      if (q == 0) return P0[j] + S0[j];
      if (q == 4) return P4[j - 1] + S4[j - 1];
      return PQ[q - 1] + SQ[q - 1];
    };
    auto add10 = [&](int kk, int q, real v) {
      add(F_P + q, kk - k, v);
      add(F_S + q, kk - k, v);
    };
    rhogkin_k_t(G, kmin, kmax, k, g10, add10, b2);
    auto g11 = [&](int kk, int q) {
      const int j = kk - k + 1;
      if (q == 0) return P0[j] + rs1(kk, j);
      if (q == 4) return P4[j - 1] + RW[j];
      return PQ[q - 1] + DQ[q - 1];
    };
    auto add11 = [&](int kk, int q, real v) {
      add(F_P + q, kk - k, v);
      if (q == 0) rs1_t(kk - k, v);
      else if (q == 4) add(F_W1, kk - k, v);
      else add(F_D + q - 1, kk - k, v);
    };
    rhogkin_k_t(G, kmin, kmax, k, g11, add11, -b2);
  }
};

// ViMainPostA's transpose as one gather launch: the thread of a compute point
// at level j re-runs point() for the forward points of its column at levels
// j - 2 .. j + 1 (those whose windows hold level j) and keeps their adds at
// level j, so each bar element is read and written once, by one thread.
// With 4-byte fields the bars of the small step's input split state and of
// the divergences 0, 2 are stored, not added (this is the first kernel of the
// reversed small step to write them), and the divergence bar 1 is cleared: at
// the compute points and, as zeros, at every other point.  With 8-byte fields
// the caller clears them first and all bars are added (measured faster there:
// adjoint step 16.07 -> 15.79 ms at gl06; 9.50 against 9.54 ms in FP32).
struct ViMainPostG {
  static constexpr const char* kName = "vi_main_post_adj";
  static constexpr bool kStore = sizeof(real) == 4;
  ViMainPostA pa;
  NH_HD static bool stored(int f) {
    return (f >= ViMainPostA::F_S && f < ViMainPostA::F_D) || f == ViMainPostA::F_H0 || f == ViMainPostA::F_H2;
  }
  NH_HD void operator()(int plane, int x) const {
    const Dims& d = pa.d;
    Pt p;
    if (!pt_of(d, plane, x, p)) return;
    if (!pt_compute(p)) {
      if (kStore) {
#pragma unroll
        for (int f = 0; f < ViMainPostA::F_N; ++f)
          if (stored(f)) pa.fptr(f)[p.a] = real(0.0);
        pa.hv_b[1][p.a] = real(0.0);
      }
      return;
    }
    const long a = p.a, ks = p.ks;
    const int j = p.k;
    real acc[ViMainPostA::F_N];
    for (int f = 0; f < ViMainPostA::F_N; ++f) acc[f] = real(0.0);
#pragma unroll
    for (int dk = -1; dk <= 2; ++dk) {
      const int k = j - dk;
      if (k < 0 || k > d.kall - 1) continue;
      Pt q = p;
      q.k = k;
      q.a = a - (long)dk * ks;
      pa.point(q, [&](int f, int dd, real v) {
        if (dd == dk) acc[f] += v;
      });
    }
    real cur[ViMainPostA::F_N];
#pragma unroll
    for (int f = 0; f < ViMainPostA::F_N; ++f) cur[f] = (kStore && stored(f)) ? real(0.0) : pa.fptr(f)[a];
#pragma unroll
    for (int f = 0; f < ViMainPostA::F_N; ++f) pa.fptr(f)[a] = cur[f] + acc[f];
    if (kStore) pa.hv_b[1][a] = real(0.0);
  }
};

// ----------------------------------------------------------------------------
// large-step sources and tridiagonal coefficients (ViPath0P)
// ----------------------------------------------------------------------------
// g~(kk) = GRAV - (dpgw(kk) - dbuo(kk)) / rhog_h(kk) of a column, in the
// precision T of the profiles V; rg(kk), pre(kk) read the column's levels
template <class T>
struct GztCol {
  const VProf<T>* V;
  T GRAV;
  int kmin, kmax;
  template <class RG>
  NH_HD T rhog_h(int kk, RG rg) const {
    const int k2 = (kk < kmin) ? kmin : kk;
    return V->c2w0[k2] * rg(k2) + V->c2w1[k2] * rg(k2 - 1);
  }
  NH_HD bool in(int kk) const { return kk >= kmin + 1 && kk <= kmax; }
  template <class RG, class PR>
  NH_HD T value(int kk, RG rg, PR pre) const {
    const T dpgw = in(kk) ? T(1.0) * (pre(kk) * T(1.0) - pre(kk - 1) * T(1.0)) * V->rdgzh[kk] : T(0.0);
    const T dbuo = in(kk) ? -GRAV * (V->c2w0[kk] * rg(kk) + V->c2w1[kk] * rg(kk - 1)) : T(0.0);
    return GRAV - (dpgw - dbuo) / rhog_h(kk, rg);
  }
  // bar v of g~(kk) -> addrg(kk, bar), addpre(kk, bar)
  template <class RG, class PR, class ARG, class APR>
  NH_HD void adj(int kk, RG rg, PR pre, ARG addrg, APR addpre, T v) const {
    const T dpgw = in(kk) ? T(1.0) * (pre(kk) * T(1.0) - pre(kk - 1) * T(1.0)) * V->rdgzh[kk] : T(0.0);
    const T dbuo = in(kk) ? -GRAV * (V->c2w0[kk] * rg(kk) + V->c2w1[kk] * rg(kk - 1)) : T(0.0);
    const T rh = rhog_h(kk, rg);
    const T dpgwb = -v / rh, dbuob = v / rh, rhb = v * (dpgw - dbuo) / (rh * rh);
    if (in(kk)) {
      addpre(kk, dpgwb * V->rdgzh[kk]);
      addpre(kk - 1, -dpgwb * V->rdgzh[kk]);
      addrg(kk, -GRAV * V->c2w0[kk] * dbuob);
      addrg(kk - 1, -GRAV * V->c2w1[kk] * dbuob);
    }
    const int k2 = (kk < kmin) ? kmin : kk;
    addrg(k2, rhb * V->c2w0[k2]);
    addrg(k2 - 1, rhb * V->c2w1[k2]);
  }
};

struct ViPath0A {
  static constexpr const char* kName = "vi_path0_adj";
  Dims d; Consts c; Geo G;
  double dt;
  const real* prog;
  const real* diag;
  const real* eth;
  const real* vt[6];
  real* gT_b;        // consumed
  double* Mc_b; double* Mu_b; double* Ml_b;   // consumed
  real* getot_b;     // consumed
  real* prog_b;
  real* diag_b;
  real* eth_b;
  real* gadv_b;      // 4 fields
  real* vt_b[6];
  real* dd_b[3];
  // the bars the horizontal operators send to the stencil, per compute point
  // and level (zero off the physical levels), gathered by ViPath0G: the two
  // divergences' (-gT0, -gT5 bars) and the pressure gradient's (3)
  real* S[5];
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (!pt_of(d, plane, x, p) || !pt_compute(p)) return;
    const long NF = d.NF, ks = p.ks, a = p.a;
    const int kmin = d.kmin, kmax = d.kmax, k = p.k;
    const real R = c.RADIUS;
    const real* rg = prog;
    const real* rw = prog + 4 * NF;
    const real* pre = diag + I_PRE * NF;
    const long b = a - (long)k * ks;
    auto lev = [&](int kk) { return b + (long)kk * ks; };
    real gb[6];
    for (int q = 0; q < 6; ++q) {
      gb[q] = gT_b[q * NF + a];
      gT_b[q * NF + a] = real(0.0);
    }
    const real geb = getot_b[a];
    getot_b[a] = real(0.0);
    if (k < kmin || k > kmax)
      for (int f = 0; f < 5; ++f) S[f][a] = real(0.0);
    double mcb = 0.0, mub = 0.0, mlb = 0.0;
    if (k >= kmin + 1 && k <= kmax) {
      mcb = Mc_b[a]; mub = Mu_b[a]; mlb = Ml_b[a];
      Mc_b[a] = 0.0; Mu_b[a] = 0.0; Ml_b[a] = 0.0;
    }
    // bars of rhog, pressure, eth and rhogw at the column levels k-2 .. k+1,
    // summed here and added once per level at the end (select-based adds keep
    // the indices static)
    real Arg[4] = {real(0.0), real(0.0), real(0.0), real(0.0)}, Apr[4] = {real(0.0), real(0.0), real(0.0), real(0.0)};
    real Aet[4] = {real(0.0), real(0.0), real(0.0), real(0.0)}, Arw[4] = {real(0.0), real(0.0), real(0.0), real(0.0)};
    auto accw = [&](real* A, int kk, real v) {
      const int jw = kk - k + 2;
      A[0] += (jw <= 0) ? v : real(0.0);
      A[1] += (jw == 1) ? v : real(0.0);
      A[2] += (jw == 2) ? v : real(0.0);
      A[3] += (jw >= 3) ? v : real(0.0);
    };
    auto add_rg = [&](int kk, real v) { accw(Arg, kk, v); };
    auto add_pre = [&](int kk, real v) { accw(Apr, kk, v); };
    auto rgl = [&](int kk) { return ldc(rg + (lev(kk))); };
    auto prl = [&](int kk) { return ldc(pre + (lev(kk))); };
    GztCol<real> gz{&G, c.GRAV, kmin, kmax};
    real xx[3];
    point_pos(G, d, p, xx);
    const real Kh = c.GAMMA_H;
    const real rgk = ldc(rg + (a));
    real rgkb = real(0.0);
    // gT0 = -vt5 + drhog
    aadd(&vt_b[5][a], -gb[0]);
    // gT1..3 = gadv + P(-(vt Kh) rhog) - dpg + dd
    real fvb[3] = {gb[1], gb[2], gb[3]};
    for (int q = 0; q < 3; ++q) {
      aadd(&gadv_b[q * NF + a], gb[1 + q]);
      aadd(&dd_b[q][a], gb[1 + q]);
    }
    proj_t(xx, R, fvb);
    for (int q = 0; q < 3; ++q) {
      aadd(&vt_b[q][a], -fvb[q] * Kh * rgk);
      rgkb += -fvb[q] * ldc(vt[q] + (a)) * Kh;
    }
    // gT4 = gadv_w - (vt3 Kh) rhog_h - dpgw(k) + dbuo(k)
    if (k > kmin && k <= kmax) {
      aadd(&gadv_b[3 * NF + a], gb[4] * G.c2w0[k]);
      aadd(&gadv_b[3 * NF + a - ks], gb[4] * G.c2w1[k]);
    }
    if (k >= kmin) {
      const real rgh = G.c2w0[k] * rgk + G.c2w1[k] * ldc(rg + (a - ks));
      aadd(&vt_b[3][a], -gb[4] * Kh * rgh);
      const real rghb = -gb[4] * ldc(vt[3] + (a)) * Kh;
      rgkb += rghb * G.c2w0[k];
      Arg[1] += rghb * G.c2w1[k];
    }
    if (gz.in(k)) {
      add_pre(k, -gb[4] * G.rdgzh[k]);
      add_pre(k - 1, gb[4] * G.rdgzh[k]);
      add_rg(k, -c.GRAV * G.c2w0[k] * gb[4]);
      add_rg(k - 1, -c.GRAV * G.c2w1[k] * gb[4]);
    }
    // gT5 = -vt4 + drhoge + pw; getot = -vt4
    aadd(&vt_b[4][a], -(gb[5] + geb));
    if (k >= kmin && k <= kmax) {
      // drhog = -div(rhogv) - (rw(k+1) - rw(k)) rdgz
      const real hdb = -gb[0];
      if (k + 1 <= kmax) Arw[3] += -gb[0] * G.rdgz[k];
      if (k >= kmin + 1) Arw[2] += gb[0] * G.rdgz[k];
      // drhoge = -div(rhogv eth) - (ws_hi - ws_lo) rdgz
      const real heb = -gb[5];
      if (k + 1 <= kmax) {
        const real wsb = -gb[5] * G.rdgz[k];
        const real w = ldc(rw + (a + ks));
        Arw[3] += wsb * (G.afact[k + 1] * ldc(eth + (a + ks)) + G.bfact[k + 1] * ldc(eth + (a)));
        Aet[3] += wsb * w * G.afact[k + 1];
        Aet[2] += wsb * w * G.bfact[k + 1];
      }
      if (k >= kmin + 1) {
        const real wsb = gb[5] * G.rdgz[k];
        const real w = ldc(rw + (a));
        Arw[2] += wsb * (G.afact[k] * ldc(eth + (a)) + G.bfact[k] * ldc(eth + (a - ks)));
        Aet[2] += wsb * w * G.afact[k];
        Aet[1] += wsb * w * G.bfact[k];
      }
      // the two divergences: gathered by ViPath0G
      S[0][a] = hdb;
      S[1][a] = heb;
      // pressure gradient (level kmin -> kmin+1) and pw
      const long sh = (k == kmin) ? ks : 0;
      real gr[3];
      for (int dd2 = 0; dd2 < 3; ++dd2) {
        real acc = real(0.0);
        if (p.pole) {
          for (int v = 0; v < 6; ++v) {
            const real pv = G.cgrad_pl[(v * 2 + p.lp) * 3 + dd2] * (ldc(pre + (a + v + sh)) * real(1.0) * real(1.0));
            acc = (v == 0) ? pv : acc + pv;
          }
        } else {
          acc = rc(G.cgrad, d.nreg, dd2 * 7, p.ci) * (ldc(pre + (a + sh)) * real(1.0));
          for (int m = 1; m < 7; ++m)
            acc = acc + rc(G.cgrad, d.nreg, dd2 * 7 + m, p.ci) * (ldc(pre + (a + nb_off(m, d.g) + sh)) * real(1.0));
        }
        gr[dd2] = acc + real(0.0);
      }
      real dpg[3];
      {
        const real prd = (gr[0] * xx[0] + gr[1] * xx[1] + gr[2] * xx[2]) / R;
        for (int q = 0; q < 3; ++q) dpg[q] = gr[q] - prd * xx[q] / R;
      }
      const real pwb = gb[5];
      real dpgb[3];
      for (int q = 0; q < 3; ++q) {
        dpgb[q] = -gb[1 + q] + pwb * ldc(diag + ((I_VX + q) * NF + a));
        aadd(&diag_b[(I_VX + q) * NF + a], pwb * dpg[q]);
      }
      proj_t(xx, R, dpgb);
      // the gradient's stencil: gathered by ViPath0G
      for (int dd2 = 0; dd2 < 3; ++dd2) S[2 + dd2][a] = dpgb[dd2];
      // pw: w2c0(k) pwh(k+1) + w2c1(k) pwh(k), pwh(kk) = -g~(kk) rw(kk)
      for (int s2 = 0; s2 < 2; ++s2) {
        const int kk = (s2 == 0) ? k + 1 : k;
        const real wgt = (s2 == 0) ? G.w2c0[k] : G.w2c1[k];
        const real hb = pwb * wgt;
        const real gzv = gz.value(kk, rgl, prl);
        accw(Arw, kk, -hb * gzv);
        gz.adj(kk, rgl, prl, add_rg, add_pre, -hb * ldc(rw + (lev(kk))));
      }
    }
    Arg[2] += rgkb;
    if (k >= kmin + 1 && k <= kmax) {
      // tridiagonal coefficients (FP64)
      const VProf<double>& V = G.v64;
      const ConstsT<double>& C = c.d;
      GztCol<double> gz64{&V, C.GRAV, kmin, kmax};
      auto rg64 = [&](int kk) { return (double)ldc(rg + (lev(kk))); };
      auto pr64 = [&](int kk) { return (double)ldc(pre + (lev(kk))); };
      auto add_rg64 = [&](int kk, double v) { accw(Arg, kk, (real)v); };
      auto add_pre64 = [&](int kk, double v) { accw(Apr, kk, (real)v); };
      auto eth_h_t = [&](int kk, double v) {
        const int k2 = (kk < kmin) ? kmin : kk;
        accw(Aet, k2, (real)(v * V.afact[k2]));
        accw(Aet, k2 - 1, (real)(v * V.bfact[k2]));
      };
      const double Ca = (1.0 * V.rdgz[k] + 1.0 * V.rdgz[k - 1]) * 1.0;
      const double Cb = V.dfact[k] - V.cfact[k - 1];
      const double rz = V.rdgzh[k];
      eth_h_t(k, mcb * rz * Ca);
      gz64.adj(k, rg64, pr64, add_rg64, add_pre64, -mcb * rz * Cb);
      eth_h_t(k + 1, -mub * rz * V.rdgz[k]);
      gz64.adj(k + 1, rg64, pr64, add_rg64, add_pre64, -mub * rz * V.cfact[k]);
      eth_h_t(k - 1, -mlb * rz * V.rdgz[k]);
      gz64.adj(k - 1, rg64, pr64, add_rg64, add_pre64, mlb * rz * V.dfact[k - 1]);
    }
    for (int jw = 0; jw < 4; ++jw) {
      const int kw = k - 2 + jw;
      if (kw < 0 || kw > d.kall - 1) continue;
      const long aw = lev(kw);
      if (Arg[jw] != real(0.0)) aadd(&prog_b[aw], Arg[jw]);
      if (Apr[jw] != real(0.0)) aadd(&diag_b[I_PRE * NF + aw], Apr[jw]);
      if (Aet[jw] != real(0.0)) aadd(&eth_b[aw], Aet[jw]);
      if (Arw[jw] != real(0.0)) aadd(&prog_b[4 * NF + aw], Arw[jw]);
    }
  }
};

// the horizontal part of ViPath0A's transpose as a gather: each point and
// level collects, from the compute points whose stencil holds it, the bars of
// rhogv and eth (the two divergences, at the same level) and of the pressure
// (the gradient: a point at kmin + 1 also takes the kmin points', which read
// the pressure there)
struct ViPath0G {
  static constexpr const char* kName = "vi_path0_adj2";
  Dims d; Geo G;
  const real* prog;
  const real* eth;
  const real* S[5];
  real* prog_b;
  real* diag_b;
  real* eth_b;
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (!pt_of(d, plane, x, p)) return;
    const long NF = d.NF, a = p.a, ks = p.ks;
    const int kmin = d.kmin, kmax = d.kmax, k = p.k;
    if (k < kmin || k > kmax) return;
    const real e = ldc(eth + a);
    real pv[3];
    for (int q = 0; q < 3; ++q) pv[q] = ldc(prog + (1 + q) * NF + a);
    real rvb[3] = {real(0.0), real(0.0), real(0.0)};
    real eb = real(0.0), preb = real(0.0);
    // Q: the compute point, m: this point's place in its stencil
    auto from = [&](long aQ, const real cdiv3[3], const real cgrad3[3], bool div, bool grad, long gsh) {
      if (div) {
        const real hdb = ldc(S[0] + aQ), heb = ldc(S[1] + aQ);
        for (int q = 0; q < 3; ++q) {
          const real t0 = cdiv3[q] * hdb, t1 = cdiv3[q] * heb;
          rvb[q] += t0 + t1 * e;
          eb += t1 * pv[q];
        }
      }
      if (grad) {
        real sb = real(0.0);
        for (int dd2 = 0; dd2 < 3; ++dd2) sb += cgrad3[dd2] * ldc(S[2 + dd2] + aQ - gsh);
        preb += sb;
      }
    };
    // the gradient's points at this level (none at kmin) and, at kmin + 1, kmin's
    const bool gk = k >= kmin + 1, gk0 = (k == kmin + 1);
    if (p.pole) {
      real cd[3], cg[3];
      for (int q = 0; q < 3; ++q) {
        cd[q] = G.cdiv_pl[(p.v * 2 + p.lp) * 3 + q];
        cg[q] = G.cgrad_pl[(p.v * 2 + p.lp) * 3 + q];
      }
      from(a - p.v, cd, cg, true, gk, 0);
      if (gk0) from(a - p.v, cd, cg, false, true, ks);
    } else {
      for (int m = 0; m < 7; ++m) {
        const int iq = p.i - (((4201 >> (2 * m)) & 3) - 1), jq = p.j - (((421 >> (2 * m)) & 3) - 1);
        if (iq < 1 || iq > d.g - 2 || jq < 1 || jq > d.g - 2) continue;
        const long off = nb_off(m, d.g);
        real cd[3], cg[3];
        for (int q = 0; q < 3; ++q) {
          cd[q] = rc(G.cdiv, d.nreg, q * 7 + m, p.ci - off);
          cg[q] = rc(G.cgrad, d.nreg, q * 7 + m, p.ci - off);
        }
        from(a - off, cd, cg, true, gk, 0);
        if (gk0) from(a - off, cd, cg, false, true, ks);
      }
    }
    // this point's thread is the only one adding to these elements here
    for (int q = 0; q < 3; ++q) prog_b[(1 + q) * NF + a] += rvb[q];
    eth_b[a] += eb;
    diag_b[I_PRE * NF + a] += preb;
  }
};

// ----------------------------------------------------------------------------
// surface pressure (mod_bndcnd.py:521-555, BNDCND_pre_sfc with a flat surface
// z_srf = 0): density extrapolated by the quadratic Lagrange polynomial through
// the lowest three full levels, pressure by hydrostatic balance from kmin.
// One thread per column, FP64 from the stored fields.
// ----------------------------------------------------------------------------
struct SfcPresP {
  static constexpr const char* kName = "sfc_pressure";
  Dims d; Consts c; Geo G;
  const real* prog;
  const real* rhogq;
  double* ps;          // [nreg + 12]
  const double* ps_b;  // adjoint: bar of ps (null for the forward)
  real* prog_b;
  struct Th { double rho, ein, q, cv, tem, Rm, pre; };
  NH_HD Th thermo(long a) const {
    const ConstsT<double>& C = c.d;
    Th t;
    const double rhog = (double)prog[a];
    t.rho = rhog;
    t.ein = (double)prog[5 * d.NF + a] / rhog;
    t.q = (double)rhogq[a] / rhog;
    t.cv = t.q * C.CVW_qv + (1.0 - t.q) * C.CVdry;
    t.tem = t.ein / t.cv;
    t.Rm = (1.0 - t.q) * C.Rdry + t.q * C.Rvap;
    t.pre = t.rho * t.tem * t.Rm;
    return t;
  }
  NH_HD void operator()(long t) const {
    Cell cl = cell_of(d, t);
    const int k1 = d.kmin;
    const VProf<double>& V = G.v64;
    const double z1 = V.gz[k1], z2 = V.gz[k1 + 1], z3 = V.gz[k1 + 2], zs = 0.0;
    const double L1 = ((zs - z2) * (zs - z3)) / ((z1 - z2) * (z1 - z3));
    const double L2 = ((zs - z1) * (zs - z3)) / ((z2 - z1) * (z2 - z3));
    const double L3 = ((zs - z1) * (zs - z2)) / ((z3 - z1) * (z3 - z2));
    const long a1 = cl.base + (long)k1 * cl.ks, a2 = a1 + cl.ks, a3 = a2 + cl.ks;
    const double GRAV = c.d.GRAV;
    const Th t1 = thermo(a1);
    if (!ps_b) {
      const double rho_srf = L1 * (double)prog[a1] + L2 * (double)prog[a2] + L3 * (double)prog[a3];
      ps[t] = t1.pre + 0.5 * (rho_srf + (double)prog[a1]) * GRAV * (z1 - zs);
      return;
    }
    const double pb = ps_b[t];
    const double rsb = 0.5 * GRAV * (z1 - zs) * pb;
    double rg1b = rsb + rsb * L1;
    prog_b[a2] += (real)(rsb * L2);
    prog_b[a3] += (real)(rsb * L3);
    // pre(kmin) = rho tem Rm, tem = ein / cv
    const ConstsT<double>& C = c.d;
    double rhob = pb * t1.tem * t1.Rm;
    const double temb = pb * t1.rho * t1.Rm;
    const double Rmb = pb * t1.rho * t1.tem;
    const double einb = temb / t1.cv;
    const double cvb = -temb * t1.tem / t1.cv;
    const double qb = cvb * (C.CVW_qv - C.CVdry) + Rmb * (C.Rvap - C.Rdry);
    const double rhog = (double)prog[a1];
    prog_b[5 * d.NF + a1] += (real)(einb / rhog);
    rg1b += rhob - einb * t1.ein / rhog - qb * t1.q / rhog;
    prog_b[a1] += (real)rg1b;
  }
};



// ----------------------------------------------------------------------------
// divergence damping, gather form.  Both operators are out(cell) =
// hex_sum(cell, T) over the six triangles around the cell of triangle values
// T = tri_div (+ the vertical rhogw term for the first operator).
// Kernel S: per corner b and level the bars of its two triangles, S(TI(b)) and
// S(TJ(b)), each the sum over the triangle's three cells of the weight
// w_n = sum_dd H_n(dd) o_dd that the cell's output bar o sends to it (H_n: the
// hex_sum coefficient of triangle n of the cell; in the first operator the
// pentagon corner (1,1) of an sgp region sends triangle 4's share to triangle
// 5, as the forward substitutes T5 for T4 there).  The pole centre stores its
// five ring-triangle bars in the pole slots 1..5 of the first field.
// Kernel G: per point x the tri_div transposes at x's corner of each of the
// six triangles around it (and, first operator, the vertical term from the
// triangle sums at x's level and the one below).
// ----------------------------------------------------------------------------
template <int OP>
struct DivDampSA {
  static constexpr const char* kName = OP == 1 ? "divdamp_3d_adj" : "divdamp_post_adj";
  Dims d; Consts c; Geo G;
  const real* ob[3];
  real* S[2];
  // the output bar of cell (ii, jj) at the corner's level, projected and
  // scaled in the second operator; false off the compute cells
  NH_HD bool cell_o(const Pt& p, int ii, int jj, long a, long ci, real o[3]) const {
    if (ii < 1 || ii > d.g - 2 || jj < 1 || jj > d.g - 2) return false;
    for (int q = 0; q < 3; ++q) o[q] = ob[q][a];
    if (OP == 2) {
      real xx[3];
      for (int q = 0; q < 3; ++q) xx[q] = G.gx[q * d.nreg + ci];
      proj_t(xx, c.RADIUS, o);
      for (int q = 0; q < 3; ++q) o[q] = c.ALPHA_D * o[q];
    }
    return true;
  }
  // weight of triangle n of the cell (hex_sum_t's Tb[n])
  NH_HD real wn(long ci, const real o[3], int n) const {
    const long nreg = d.nreg;
    const int m0 = (n == 0) ? 0 : n - 1, m1 = (n == 0) ? 5 : n;
    real t = real(0.0);
    for (int dd = 0; dd < 3; ++dd) {
      const real* cd = G.cdiff + (long)(dd * 6) * nreg + ci;
      t += (cd[m0 * nreg] + cd[m1 * nreg]) * o[dd];
    }
    return t;
  }
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (!pt_of(d, plane, x, p)) return;
    const bool act = p.k >= d.kmin && p.k <= d.kmax;
    if (p.pole) {
      if (p.v != 0) return;
      real sb[5] = {real(0.0), real(0.0), real(0.0), real(0.0), real(0.0)};
      if (act) {
        real o[3];
        for (int q = 0; q < 3; ++q) o[q] = ob[q][p.a];
        if (OP == 2) {
          real xx[3];
          point_pos(G, d, p, xx);
          proj_t(xx, c.RADIUS, o);
          for (int q = 0; q < 3; ++q) o[q] = c.ALPHA_D * o[q];
        }
        pole_ring_t(G, p.lp, o, sb);
      }
      for (int v = 0; v < 5; ++v) S[0][p.a + 1 + v] = sb[v];
      return;
    }
    const int g = d.g, i = p.i, j = p.j;
    if (i > g - 2 || j > g - 2) return;
    const long a = p.a, ci = p.ci;
    // cells b, b+(1,0), b+(1,1), b+(0,1) and the weights they send here
    real w0 = real(0.0), w1 = real(0.0), w2 = real(0.0), w3 = real(0.0), w4 = real(0.0), w5 = real(0.0);
    if (act) {
      const bool sgp = (OP == 1) && G.sgp[p.l];
      real o[3];
      if (cell_o(p, i, j, a, ci, o)) {
        w0 = wn(ci, o, 0);
        w1 = wn(ci, o, 1);
      }
      if (cell_o(p, i + 1, j, a + 1, ci + 1, o)) w2 = wn(ci + 1, o, 2);
      if (cell_o(p, i + 1, j + 1, a + 1 + g, ci + 1 + g, o)) {
        w3 = wn(ci + 1 + g, o, 3);
        w4 = (sgp && i == 0 && j == 0) ? real(0.0) : wn(ci + 1 + g, o, 4);
      }
      if (cell_o(p, i, j + 1, a + g, ci + g, o)) {
        w5 = wn(ci + g, o, 5);
        if (sgp && i == 1 && j == 0) w5 += wn(ci + g, o, 4);
      }
    }
    S[0][a] = w0 + w2 + w4;
    S[1][a] = w1 + w3 + w5;
  }
};

template <int OP>
struct DivDampGA {
  static constexpr const char* kName = OP == 1 ? "divdamp_3d_adj" : "divdamp_post_adj";
  Dims d; Consts c; Geo G;
  const real* S[2];   // the triangle bars of DivDampSA
  real* vb[4];        // OP 1: bars of rhogvx..vz, rhogw (added); OP 2: bars of the first result (set)
  // the six triangles' bars at level offset (0 or -1) and their tri_div transposes at x
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (!pt_of(d, plane, x, p)) return;
    const long a = p.a, ks = p.ks;
    const int kmin = d.kmin, kmax = d.kmax, k = p.k;
    const real sgn = (OP == 1) ? real(1.0) : -real(1.0);
    real acc[3] = {real(0.0), real(0.0), real(0.0)};
    real ssum = real(0.0), ssum_dn = real(0.0);
    const bool vert = (OP == 1) && k >= kmin + 1 && k <= kmax;
    if (p.pole) {
      const long ac = a - p.v;   // pole centre
      const int lp = p.lp;
      auto tri = [&](int v, int corner) {
        const real Sv = S[0][ac + v];
        for (int dd = 0; dd < 3; ++dd) acc[dd] += G.cintp_pl[((v * 2 + lp) * 3 + dd) * 3 + corner] * sgn * Sv;
        ssum += Sv;
        if (vert) ssum_dn += S[0][ac + v - ks];
      };
      if (p.v == 0) {
        for (int v = 1; v <= 5; ++v) tri(v, 0);
      } else {
        tri(p.v, 1);
        tri(p.v == 1 ? 5 : p.v - 1, 2);
      }
    } else {
      const int g = d.g, i = p.i, j = p.j;
      const long nreg = d.nreg;
      // the six triangles with x as a corner: (type, base (bi, bj), corner of x)
      const int tT[6] = {0, 0, 0, 1, 1, 1};
      const int tbi[6] = {0, -1, -1, 0, -1, 0}, tbj[6] = {0, 0, -1, 0, -1, -1};
      const int tcor[6] = {0, 1, 2, 0, 1, 2};
      for (int t = 0; t < 6; ++t) {
        const int bi = i + tbi[t], bj = j + tbj[t];
        // bases on the last row or column hold no compute cell
        if (bi < 0 || bj < 0 || bi > g - 2 || bj > g - 2) continue;
        const long ab = a + (long)tbi[t] + (long)tbj[t] * g;
        const real Sv = S[tT[t]][ab];
        const long cj = p.ci + (long)tbi[t] + (long)tbj[t] * g;
        for (int dd = 0; dd < 3; ++dd)
          acc[dd] += rc(G.cintp, nreg, (dd * 3 + tcor[t]) * 2 + tT[t], cj) * sgn * Sv;
        ssum += Sv;
        if (vert) ssum_dn += S[tT[t]][ab - ks];
      }
    }
    if (OP == 1) {
      for (int dd = 0; dd < 3; ++dd) vb[dd][a] += acc[dd];
      if (vert) vb[3][a] += ssum_dn * G.rdgz[k - 1] / real(3.0) - ssum * G.rdgz[k] / real(3.0);
    } else {
      for (int dd = 0; dd < 3; ++dd) vb[dd][a] = acc[dd];
    }
  }
};

NH_END
