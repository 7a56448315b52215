// LDS-tiled kernels for the triangle-based operators (3-D divergence damping,
// the Kh-weighted diffusion of the second hyperdiffusion pass).
//
// A block owns a tile of one region level: all interior i (W = g-2 points,
// one or more wavefronts) times TH rows.  Phase 1 computes the two triangle
// values of every tile corner (i', j'), i' in 0..W, j' in j0-1..j0+TH-1, once
// into LDS; phase 2 forms each cell's hexagon sum from the six triangles
// around it.  Each triangle value is shared by three cells, so this does a
// third of the work and loads of the per-point form.  Blocks past the
// regional tiles handle the pole levels (the pole centre's five triangles).
// On the GPU the phases are separated by a workgroup barrier; the host test
// build runs them one after the other.
#pragma once

#include "nh_point.h"

NH_BEGIN

struct TileGeom {
  int W, TH;          // tile width (interior points of a row) and height
  int ntp;            // tiles per region level
  int ntiles;         // regional tiles (lall * kall * ntp)
  int nblocks;        // + 2 * kall pole blocks
  int lds;            // LDS elements (real) per block
};

inline TileGeom tile_geom(const Dims& d, int per_corner) {
  TileGeom t;
  t.W = d.g - 2;
  t.TH = 256 / t.W;
  if (t.TH < 1) t.TH = 1;
  t.ntp = (t.W + t.TH - 1) / t.TH;
  t.ntiles = d.lall * d.kall * t.ntp;
  t.nblocks = t.ntiles + 2 * d.kall;
  t.lds = per_corner * (t.W + 1) * (t.TH + 1);
  return t;
}

// tile decode: region l, level k, first row j0
struct TileAt {
  int l, k, j0, rows;
};

NH_HD inline TileAt tile_at(const Dims& d, const TileGeom& tg, int blk) {
  TileAt t;
  const int plane = blk / tg.ntp;
  const int tile = blk - plane * tg.ntp;
  t.l = plane / d.kall;
  t.k = plane - t.l * d.kall;
  t.j0 = 1 + tile * tg.TH;
  t.rows = (d.g - 1 - t.j0 < tg.TH) ? d.g - 1 - t.j0 : tg.TH;   // rows j0..j0+rows-1 <= g-2
  return t;
}

// ----------------------------------------------------------------------------
// stencil point kernels on staged tiles: the fields a point kernel P reads on
// its 7-point stencil (P::kNStage of them; P::stage_ptr(f, k) positions field
// f so that level k's addresses read the level P needs there) are staged for
// the tile's rows j0-1 .. j0+rows (on the levels where P::stage_level holds);
// every cell of the tile then runs P::at with a loader that reads them from
// LDS.  Rows of a region level are contiguous, so
// the staged copy is the address range [base0, base0 + (rows+2)*g) and a
// stencil address b maps to b - base0.  Pole levels run P on the pole centre.
// ----------------------------------------------------------------------------
template <class P>
struct StagedPointT {
  using LdsT = real;
  static constexpr bool kStage = true;
  static constexpr const char* kName = P::kName;
  P pk;
  TileGeom tg;
  NH_HD int sf() const { return (tg.TH + 2) * pk.d.g; }
  NH_HD long base0(const TileAt& t) const {
    const long g = pk.d.g;
    return (long)(t.l * pk.d.kall + t.k) * g * g + (long)(t.j0 - 1) * g;
  }
  NH_HD void phase0(int blk, int tid, real* lds) const {
    if (blk >= tg.ntiles) return;
    const TileAt t = tile_at(pk.d, tg, blk);
    if (!pk.stage_level(t.k)) return;     // P::at does not read the stencil on this level
    const long b0 = base0(t);
    const int npl = (t.rows + 2) * pk.d.g;
    for (int idx = tid; idx < P::kNStage * npl; idx += 256) {
      const int f = idx / npl, e = idx - f * npl;
      lds[f * sf() + e] = pk.stage_ptr(f, t.k)[b0 + e];
    }
  }
  NH_HD void phase1(int, int, real*) const {}
  NH_HD void phase2(int blk, int tid, const real* lds) const {
    const Dims& d = pk.d;
    if (blk >= tg.ntiles) {
      if (tid == 0) pk(d.lall * d.kall + (blk - tg.ntiles), 0);
      return;
    }
    const TileAt t = tile_at(d, tg, blk);
    const int ti = tid % tg.W, tj = tid / tg.W;
    if (tj >= t.rows) return;
    const int g = d.g, i = ti + 1, j = t.j0 + tj;
    Pt p;
    p.pole = false; p.inner = true;
    p.i = i; p.j = j; p.l = t.l; p.k = t.k; p.v = 0; p.lp = 0;
    p.ks = (long)g * g;
    p.a = ((long)(t.l * d.kall + t.k) * g + j) * g + i;
    p.ci = ((long)t.l * g + j) * g + i;
    const long b0 = base0(t);
    const int s = sf();
    pk.at(p, [&](int f, long b) { return lds[f * s + (b - b0)]; });
  }
};

// Staging pays where the stencil loads are 8-byte (the FP64 form: 10-20 %
// faster on MI355X at gl07); with 4-byte fields the direct point kernel is as
// fast or faster, except for kernels that set P::kStageF32.
template <class X, class P>
void run_staged(X& x, const P& pk) {
  if (sizeof(real) == sizeof(float) && !P::kStageF32) {
    x.run_pts(pk, pk.d);
    return;
  }
  StagedPointT<P> f;
  f.pk = pk;
  f.tg = tile_geom(pk.d, 0);
  x.run_tiles(f, f.tg.nblocks, P::kNStage * (f.tg.TH + 2) * pk.d.g);
}

// ----------------------------------------------------------------------------
// first operator: triangle-mean 3-D divergence, gradient back to the cells
// (kernels/oprt3ddivdamp.py:50-173)
// ----------------------------------------------------------------------------
struct DivDamp1T {
  using LdsT = real;
  static constexpr bool kStage = false;
  static constexpr const char* kName = "divdamp_3d";
  static int lds_elems(const TileGeom& tg, int) { return tg.lds; }
  Dims d; Consts c; Geo G; TileGeom tg;
  const real *rvx, *rvy, *rvz, *rw;
  real* out[3];
  NH_HD real rwv(long aa, int kk) const {
    return (kk >= d.kmin + 1 && kk <= d.kmax) ? (real(0.0) * real(1.0) + rw[aa] * real(1.0)) : real(0.0);
  }
  NH_HD bool active(int k) const { return k >= d.kmin && k <= d.kmax; }
  NH_HD void phase1(int blk, int tid, real* lds) const {
    if (blk >= tg.ntiles) return;
    const TileAt t = tile_at(d, tg, blk);
    if (!active(t.k)) return;
    const int g = d.g, W1 = tg.W + 1;
    const long gg = (long)g * g, ks = gg;
    const int nc = W1 * (t.rows + 1);
    for (int cidx = tid; cidx < nc; cidx += 256) {
      const int ci_ = cidx % W1, cj_ = cidx / W1;
      const int i = ci_, j = t.j0 - 1 + cj_;
      const long cj = ((long)t.l * g + j) * g + i;
      const long a = ((long)(t.l * d.kall + t.k) * g + j) * g + i;
      for (int T = 0; T < 2; ++T) {
        const long o1 = (T == 0) ? 1 : 1 + g;
        const long o2 = (T == 0) ? 1 + g : g;
        const real s = tri_div(G, d.nreg, cj, T, rvx, rvy, rvz, a, g, real(1.0));
        const real up = (rwv(a + ks, t.k + 1) + rwv(a + o1 + ks, t.k + 1) + rwv(a + o2 + ks, t.k + 1));
        const real dn = (rwv(a, t.k) + rwv(a + o1, t.k) + rwv(a + o2, t.k));
        lds[(T * (t.rows + 1) + cj_) * W1 + ci_] = s + (up - dn) / real(3.0) * G.rdgz[t.k];
      }
    }
  }
  NH_HD void phase2(int blk, int tid, const real* lds) const {
    const int g = d.g;
    if (blk >= tg.ntiles) {
      pole(blk - tg.ntiles, tid);
      return;
    }
    const TileAt t = tile_at(d, tg, blk);
    const int W1 = tg.W + 1;
    const int ti = tid % tg.W, tj = tid / tg.W;
    if (tj >= t.rows) return;
    const int i = ti + 1, j = t.j0 + tj;
    const long ci = ((long)t.l * g + j) * g + i;
    const long a = ((long)(t.l * d.kall + t.k) * g + j) * g + i;
    if (!active(t.k)) {
      for (int q = 0; q < 3; ++q) out[q][a] = real(0.0);
      return;
    }
    const int R = t.rows + 1;
    auto L = [&](int T, int ii, int jj) { return lds[(T * R + (jj - t.j0 + 1)) * W1 + ii]; };
    real T6[6] = {L(0, i, j), L(1, i, j), L(0, i - 1, j), L(1, i - 1, j - 1), L(0, i - 1, j - 1), L(1, i, j - 1)};
    if (i == 1 && j == 1) {
      const real m0 = G.sgp[t.l] ? real(0.0) : real(1.0), m1 = G.sgp[t.l] ? real(1.0) : real(0.0);
      T6[4] = T6[4] * m0 + T6[5] * m1;
    }
    for (int dd = 0; dd < 3; ++dd) out[dd][a] = hex_sum(G, d.nreg, ci, dd, T6);
  }
  NH_HD void pole(int pb, int tid) const {
    if (tid != 0) return;
    DivDamp1P p{d, c, G, rvx, rvy, rvz, rw, {out[0], out[1], out[2]}};
    p(d.lall * d.kall + pb, 0);
  }
};

// second operator on -(first result), coefficient, projection
// (kernels/oprtdivdamp.py:54-173, kernels/divdamppostcomm.py)
struct DivDamp2T {
  using LdsT = real;
  static constexpr bool kStage = false;
  static constexpr const char* kName = "divdamp_post";
  static int lds_elems(const TileGeom& tg, int) { return tg.lds; }
  Dims d; Consts c; Geo G; TileGeom tg;
  const real* in[3];
  real* out[3];
  NH_HD bool active(int k) const { return k >= d.kmin && k <= d.kmax; }
  NH_HD void phase1(int blk, int tid, real* lds) const {
    if (blk >= tg.ntiles) return;
    const TileAt t = tile_at(d, tg, blk);
    if (!active(t.k)) return;
    const int g = d.g, W1 = tg.W + 1;
    const int nc = W1 * (t.rows + 1);
    for (int cidx = tid; cidx < nc; cidx += 256) {
      const int ci_ = cidx % W1, cj_ = cidx / W1;
      const int i = ci_, j = t.j0 - 1 + cj_;
      const long cj = ((long)t.l * g + j) * g + i;
      const long a = ((long)(t.l * d.kall + t.k) * g + j) * g + i;
      for (int T = 0; T < 2; ++T)
        lds[(T * (t.rows + 1) + cj_) * W1 + ci_] = tri_div(G, d.nreg, cj, T, in[0], in[1], in[2], a, g, -real(1.0));
    }
  }
  NH_HD void phase2(int blk, int tid, const real* lds) const {
    const int g = d.g;
    if (blk >= tg.ntiles) {
      if (tid == 0) {
        DivDamp2P p{d, c, G, {in[0], in[1], in[2]}, {out[0], out[1], out[2]}};
        p(d.lall * d.kall + (blk - tg.ntiles), 0);
      }
      return;
    }
    const TileAt t = tile_at(d, tg, blk);
    const int W1 = tg.W + 1;
    const int ti = tid % tg.W, tj = tid / tg.W;
    if (tj >= t.rows) return;
    const int i = ti + 1, j = t.j0 + tj;
    const long ci = ((long)t.l * g + j) * g + i;
    const long a = ((long)(t.l * d.kall + t.k) * g + j) * g + i;
    real o[3] = {real(0.0), real(0.0), real(0.0)};
    if (active(t.k)) {
      const int R = t.rows + 1;
      auto L = [&](int T, int ii, int jj) { return lds[(T * R + (jj - t.j0 + 1)) * W1 + ii]; };
      const real T6[6] = {L(0, i, j), L(1, i, j), L(0, i - 1, j), L(1, i - 1, j - 1), L(0, i - 1, j - 1),
                            L(1, i, j - 1)};
      for (int dd = 0; dd < 3; ++dd) o[dd] = hex_sum(G, d.nreg, ci, dd, T6);
    }
    const real Rr = c.RADIUS;
    const real gd[3] = {c.ALPHA_D * o[0], c.ALPHA_D * o[1], c.ALPHA_D * o[2]};
    const real gx = G.gx[ci], gy = G.gx[d.nreg + ci], gz = G.gx[2 * d.nreg + ci];
    const real prd = (gd[0] * gx + gd[1] * gy + gd[2] * gz) / Rr;
    out[0][a] = gd[0] - prd * gx / Rr;
    out[1][a] = gd[1] - prd * gy / Rr;
    out[2][a] = gd[2] - prd * gz / Rr;
  }
};


// ----------------------------------------------------------------------------
// second hyperdiffusion pass (mod_numfilter.py:1125-1264): Laplacians of the
// velocity components and w, Kh-weighted diffusion (kernels/oprtdiffusion.py:59-177)
// of temperature and density; the triangle gradients of the two diffused
// fields (3 components x TI/TJ) go through LDS
// ----------------------------------------------------------------------------
struct HDiffPass2T {
  using LdsT = real;
  static constexpr bool kStage = false;
  static constexpr const char* kName = "hdiff_pass2";
  Dims d; Consts c; Geo G; TileGeom tg;
  const real* in[6];
  real* out[6];
  const real* rhog;
  // LDS layout: [field f (0,1)][triangle T][component dd][row][col]
  NH_HD int lidx(const TileAt& t, int f, int T, int dd, int cj_, int ci_) const {
    const int W1 = tg.W + 1, R = t.rows + 1;
    return (((f * 2 + T) * 3 + dd) * R + cj_) * W1 + ci_;
  }
  NH_HD void phase1(int blk, int tid, real* lds) const {
    if (blk >= tg.ntiles) return;
    const TileAt t = tile_at(d, tg, blk);
    const int g = d.g, W1 = tg.W + 1;
    const int nc = W1 * (t.rows + 1);
    for (int cidx = tid; cidx < nc; cidx += 256) {
      const int ci_ = cidx % W1, cj_ = cidx / W1;
      const int i = ci_, j = t.j0 - 1 + cj_;
      const long cj = ((long)t.l * g + j) * g + i;
      const long a = ((long)(t.l * d.kall + t.k) * g + j) * g + i;
      for (int f = 0; f < 2; ++f)
        for (int T = 0; T < 2; ++T) {
          real o3[3];
          tri_grad(G, d.nreg, cj, T, in[4 + f], a, g, o3);
          for (int dd = 0; dd < 3; ++dd) lds[lidx(t, f, T, dd, cj_, ci_)] = o3[dd];
        }
    }
  }
  NH_HD void phase2(int blk, int tid, const real* lds) const {
    const int g = d.g;
    if (blk >= tg.ntiles) {
      if (tid == 0) {
        HDiffPassP<1> p;
        p.d = d; p.c = c; p.G = G; p.rhog = rhog;
        for (int q = 0; q < 6; ++q) { p.in[q] = in[q]; p.out[q] = out[q]; }
        p(d.lall * d.kall + (blk - tg.ntiles), 0);
      }
      return;
    }
    const TileAt t = tile_at(d, tg, blk);
    const int ti = tid % tg.W, tj = tid / tg.W;
    if (tj >= t.rows) return;
    const int i = ti + 1, j = t.j0 + tj;
    const long nreg = d.nreg;
    const long ci = ((long)t.l * g + j) * g + i;
    const long a = ((long)(t.l * d.kall + t.k) * g + j) * g + i;
    for (int q = 0; q < 4; ++q) out[q][a] = -lap7(G, nreg, ci, in[q], a, g);
    const bool pent = (i == 1 && j == 1);
    const real m0 = G.sgp[t.l] ? real(0.0) : real(1.0), m1 = G.sgp[t.l] ? real(1.0) : real(0.0);
    // corner offsets of the six triangles: TI(i,j) TJ(i,j) TI(i-1,j) TJ(i-1,j-1) TI(i-1,j-1) TJ(i,j-1)
    const int cdi[6] = {0, 0, -1, -1, -1, 0}, cdj[6] = {0, 0, 0, -1, -1, -1}, cT[6] = {0, 1, 0, 1, 0, 1};
    for (int f = 0; f < 2; ++f) {
      const real fac = (f == 0) ? c.CVdry : c.HDIFF_FACT_RHO;
      const real gam = c.GAMMA_H;
      auto kh = [&](long x) { return rhog[x] * fac * gam; };
      real tv[6][3];
      for (int n = 0; n < 6; ++n)
        for (int dd = 0; dd < 3; ++dd)
          tv[n][dd] = lds[lidx(t, f, cT[n], dd, j + cdj[n] - t.j0 + 1, i + cdi[n])];
      if (pent)
        for (int dd = 0; dd < 3; ++dd) tv[4][dd] = tv[4][dd] * m0 + tv[5][dd] * m1;
      const real k0 = kh(a);
      const real kf1 = real(0.5) * (k0 + kh(a + 1 + g));
      const real kf2 = real(0.5) * (k0 + kh(a + g));
      const real kf3 = real(0.5) * (k0 + kh(a - 1));
      const real kf4 = real(0.5) * (k0 + kh(a - 1 - g));
      const real kf5 = real(0.5) * (k0 + kh(a - g));
      const real kf6 = real(0.5) * (k0 + kh(a + 1));
      real inner = real(0.0);
      for (int dd = 0; dd < 3; ++dd) {
        const real* cd = G.cdiff + (long)(dd * 6) * nreg + ci;
        const real blk2 = kf1 * cd[0] * (tv[0][dd] + tv[1][dd]) + kf2 * cd[nreg] * (tv[1][dd] + tv[2][dd])
                          + kf3 * cd[2 * nreg] * (tv[2][dd] + tv[3][dd]) + kf4 * cd[3 * nreg] * (tv[3][dd] + tv[4][dd])
                          + kf5 * cd[4 * nreg] * (tv[4][dd] + tv[5][dd]) + kf6 * cd[5 * nreg] * (tv[5][dd] + tv[0][dd]);
        inner = (dd == 0) ? blk2 : inner + blk2;
      }
      out[4 + f][a] = -inner;
    }
  }
};

NH_END
