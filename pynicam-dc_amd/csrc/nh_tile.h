// LDS-tiled kernels.
//
// Level tiles (StagedPointT): a block owns all interior i (W = g-2 points) of
// TH rows of one region level; phase 0 stages the fields a point kernel reads
// on its stencil (the tile's rows and one halo row on each side) into LDS, and
// the point kernel then reads its stencil from there.  Blocks past the
// regional tiles do the pole levels with the point kernel itself.
//
// Column tiles (DivDampK, DiffKhK; the stencil engine of nh_colstencil.h): a
// block owns a 16 x 16 tile of one region over a chunk of levels, each
// thread one corner column with its coefficients in registers; per level the
// corner threads form their two triangle values into double-buffered LDS, one
// barrier, then the tile points form their hexagon sums.
//
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
// divergence damping as column tiles.  A block owns a TX x TY tile of one
// region's points over a chunk of levels; each thread owns one corner column
// of the tile (the (TX+1)(TY+1) corners whose two triangles the tile's hexagons
// use), keeps that column's 2 x 9 triangle interpolation coefficients in
// registers and, for the TX TY corners that are also tile points, the 18
// hexagon coefficients.  Per level: the two triangle values -> LDS (double
// buffered by level parity, so one barrier per level), then the hexagon sums
// -> out.  Per point and level this reads the three velocity components at
// three corners and the new level of w, instead of re-reading 36 coefficients
// per point and level as the tile kernels above do.  (Issuing the next
// level's loads before the barrier measured slower: the registers they hold
// cost more occupancy than the latency they hide.)
// OP 1: first operator, triangle-mean 3-D divergence and its gradient back to
// the cells (kernels/oprt3ddivdamp.py:50-173); OP 2: second operator on the
// negated first result with ALPHA_D and the projection onto the sphere
// (kernels/oprtdivdamp.py:54-173, kernels/divdamppostcomm.py).  The point
// forms DivDamp1P / DivDamp2P do the pole cells.
// ----------------------------------------------------------------------------
struct ColGeom {
  int TX, TY;     // tile width and height in points (divide g - 2, a power of two)
  int nx, ny;     // tiles per region row and column
  int nchunk;     // level chunks
  int KC;         // levels per chunk
  int ntiles;     // lall * nchunk * nx * ny
  int nblocks;    // + 2 * kall pole blocks
};
constexpr int kColThreads = 384;   // (64 + 1) x (4 + 1) corners

inline ColGeom col_geom(const Dims& d, int nchunk) {
  ColGeom c;
  const int W = d.g - 2;
  c.TX = (W < 64) ? W : 64;
  c.TY = (256 / c.TX < W) ? 256 / c.TX : W;
  c.nx = W / c.TX;
  c.ny = W / c.TY;
  c.nchunk = nchunk;
  c.KC = (d.kall + nchunk - 1) / nchunk;
  c.ntiles = d.lall * nchunk * c.nx * c.ny;
  c.nblocks = c.ntiles + 2 * d.kall;
  return c;
}

template <int OP>
struct DivDampK {
  using LdsT = real;
  static constexpr int kThreads = kColThreads;
  static constexpr const char* kName = (OP == 1) ? "divdamp_3d" : "divdamp_post";
  Dims d; Consts c; Geo G; ColGeom cg;
  const real* v[3];     // OP 1: the momentum; OP 2: the first result (negated inside)
  const real* rw;       // OP 1: rhogw
  real* out[3];
  static int lds_elems(const ColGeom& cg) { return 4 * (cg.TX + 1) * (cg.TY + 1); }
  struct State {
    real ci[2][9];      // cintp of the corner column, [T][dd * 3 + vertex]
    real cd[18];        // cdiff of the tile point, [dd * 6 + m]
    real gx[3];         // OP 2: the point's position
    real dn[2];         // OP 1: the lower w sums of the two triangles at this level
    real pv[12];        // v at the corners a, a+1, a+1+g, a+g of this level, [dd * 4 + n]
    real pw[4];         // OP 1: rw at the same corners one level up
    long a;             // the corner's address at this level
    int k, lidx;
    bool corner, outp, pent, sgp, have_dn;
  };
  NH_HD bool active(int k) const { return k >= d.kmin && k <= d.kmax; }
  NH_HD void pole(int pb) const {
    if (OP == 1) {
      DivDamp1P p{d, c, G, v[0], v[1], v[2], rw, {out[0], out[1], out[2]}};
      p(d.lall * d.kall + pb, 0);
    } else {
      DivDamp2P p{d, c, G, {v[0], v[1], v[2]}, {out[0], out[1], out[2]}};
      p(d.lall * d.kall + pb, 0);
    }
  }
  // loads of an active level k at the corner address a (k + 1 <= kall - 1)
  NH_HD void fetch(State& s, long a) const {
    const long g = d.g;
    const long off[4] = {0, 1, 1 + g, g};
    for (int dd = 0; dd < 3; ++dd)
      for (int n = 0; n < 4; ++n) s.pv[dd * 4 + n] = v[dd][a + off[n]];
    if (OP == 1)
      for (int n = 0; n < 4; ++n) s.pw[n] = rw[a + g * g + off[n]];
  }
  // the block's number of levels (uniform over the block)
  NH_HD int init(int blk, int tid, State& s) const {
    s.corner = s.outp = s.pent = s.sgp = s.have_dn = false;
    if (blk >= cg.ntiles) {
      if (tid == 0) pole(blk - cg.ntiles);
      return 0;
    }
    const int g = d.g, W1 = cg.TX + 1, tpr = cg.nx * cg.ny;
    const int tile = blk % tpr, rest = blk / tpr;
    const int chunk = rest % cg.nchunk, l = rest / cg.nchunk;
    const int k0 = chunk * cg.KC;
    const int k1 = (k0 + cg.KC < d.kall) ? k0 + cg.KC : d.kall;
    if (k0 >= k1) return 0;
    const int i0 = 1 + (tile % cg.nx) * cg.TX, j0 = 1 + (tile / cg.nx) * cg.TY;
    const int tx = tid % W1, ty = tid / W1;
    s.k = k0;
    if (tid >= W1 * (cg.TY + 1)) return k1 - k0;
    s.corner = true;
    s.lidx = ty * W1 + tx;
    const int i = i0 - 1 + tx, j = j0 - 1 + ty;
    const long nreg = d.nreg, col = ((long)l * g + j) * g + i;
    s.a = ((long)(l * d.kall + k0) * g + j) * g + i;
    for (int T = 0; T < 2; ++T)
      for (int n = 0; n < 9; ++n) s.ci[T][n] = rc(G.cintp, nreg, n * 2 + T, col);
    if (tx >= 1 && ty >= 1) {
      s.outp = true;
      for (int n = 0; n < 18; ++n) s.cd[n] = G.cdiff[(long)n * nreg + col];
      if (OP == 1) {
        s.pent = (i == 1 && j == 1);
        s.sgp = G.sgp[l] != 0;
      }
      if (OP == 2)
        for (int dd = 0; dd < 3; ++dd) s.gx[dd] = G.gx[(long)dd * nreg + col];
    }
    return k1 - k0;
  }
  // triangle T of the corner from the level's loads (tri_div's arithmetic)
  NH_HD real tri(const State& s, int T) const {
    const int n1 = (T == 0) ? 1 : 2, n2 = (T == 0) ? 2 : 3;
    const real sgn = (OP == 1) ? real(1.0) : -real(1.0);
    real r = real(0.0);
    for (int dd = 0; dd < 3; ++dd) {
      const real t0 = s.ci[T][dd * 3 + 0] * (sgn * s.pv[dd * 4 + 0]);
      const real t1 = s.ci[T][dd * 3 + 1] * (sgn * s.pv[dd * 4 + n1]);
      const real t2 = s.ci[T][dd * 3 + 2] * (sgn * s.pv[dd * 4 + n2]);
      if (dd == 0) r = t0 + t1 + t2;
      else r = r + t0 + t1 + t2;
    }
    return r;
  }
  // the w sums of both triangles (rwv summed over each triangle's corners) from
  // the four corner values r at level kk.  Outside kmin+1..kmax, r * 0 + 0 * 1
  // is +0 as rwv's 0 is, so the level test is a factor.
  NH_HD void wsums(const real r[4], int kk, real w[2]) const {
    const real m = (kk >= d.kmin + 1 && kk <= d.kmax) ? real(1.0) : real(0.0);
    const real r00 = real(0.0) * real(1.0) + r[0] * m;
    const real r10 = real(0.0) * real(1.0) + r[1] * m;
    const real r11 = real(0.0) * real(1.0) + r[2] * m;
    const real r01 = real(0.0) * real(1.0) + r[3] * m;
    w[0] = (r00 + r10 + r11);
    w[1] = (r00 + r11 + r01);
  }
  NH_HD void level_a(int step, State& s, real* lds) const {
    if (!s.corner) return;
    const long g = d.g;
    if (active(s.k)) {
      const int P = (cg.TX + 1) * (cg.TY + 1);
      real* buf = lds + (step & 1) * 2 * P;
      fetch(s, s.a);
      real up[2], dn[2];
      if (OP == 1) {
        wsums(s.pw, s.k + 1, up);
        if (s.have_dn) {
          dn[0] = s.dn[0];
          dn[1] = s.dn[1];
        } else {
          const real r[4] = {rw[s.a], rw[s.a + 1], rw[s.a + 1 + g], rw[s.a + g]};
          wsums(r, s.k, dn);
        }
      }
      for (int T = 0; T < 2; ++T) {
        real val = tri(s, T);
        if (OP == 1) {
          val = val + (up[T] - dn[T]) / real(3.0) * G.rdgz[s.k];
          s.dn[T] = up[T];
        }
        buf[T * P + s.lidx] = val;
      }
      s.have_dn = true;
    }
  }
  NH_HD void level_b(int step, State& s, const real* lds) const {
    if (s.outp) {
      const int W1 = cg.TX + 1, P = W1 * (cg.TY + 1);
      real o[3] = {real(0.0), real(0.0), real(0.0)};
      const bool act = active(s.k);
      if (act) {
        const real* b0 = lds + (step & 1) * 2 * P;
        const real* b1 = b0 + P;
        const int x = s.lidx;
        // {TI(i,j), TJ(i,j), TI(i-1,j), TJ(i-1,j-1), TI(i-1,j-1), TJ(i,j-1)}
        real T6[6] = {b0[x], b1[x], b0[x - 1], b1[x - 1 - W1], b0[x - 1 - W1], b1[x - W1]};
        if (OP == 1 && s.pent) {
          const real m0 = s.sgp ? real(0.0) : real(1.0), m1 = s.sgp ? real(1.0) : real(0.0);
          T6[4] = T6[4] * m0 + T6[5] * m1;
        }
        for (int dd = 0; dd < 3; ++dd) {
          const real* cd = s.cd + dd * 6;
          o[dd] = cd[0] * (T6[0] + T6[1]) + cd[1] * (T6[1] + T6[2]) + cd[2] * (T6[2] + T6[3])
                + cd[3] * (T6[3] + T6[4]) + cd[4] * (T6[4] + T6[5]) + cd[5] * (T6[5] + T6[0]);
        }
      }
      if (OP == 1) {
        for (int dd = 0; dd < 3; ++dd) out[dd][s.a] = act ? o[dd] : real(0.0);
      } else {
        const real Rr = c.RADIUS;
        const real gd[3] = {c.ALPHA_D * o[0], c.ALPHA_D * o[1], c.ALPHA_D * o[2]};
        const real prd = (gd[0] * s.gx[0] + gd[1] * s.gx[1] + gd[2] * s.gx[2]) / Rr;
        for (int dd = 0; dd < 3; ++dd) out[dd][s.a] = gd[dd] - prd * s.gx[dd] / Rr;
      }
    }
    s.a += (long)d.g * d.g;
    s.k += 1;
  }
};

// ----------------------------------------------------------------------------
// the Kh-weighted diffusion of the second hyperdiffusion pass (fields 4 and 5
// of the pass: temperature with CVdry, density with HDIFF_FACT_RHO) as
// column tiles, the same layout as DivDampK: corner threads form the triangle
// gradients of both fields from their 18 held cintp coefficients, tile points
// the Kh-weighted hexagon sums from their 18 held cdiff coefficients.
// Results equal the point form's (diffusion7, kernels/oprtdiffusion.py:59-177).
// ----------------------------------------------------------------------------
struct DiffKhK {
  using LdsT = real;
  static constexpr int kThreads = kColThreads;
  static constexpr const char* kName = "hdiff_pass2_kh";
  Dims d; Consts c; Geo G; ColGeom cg;
  const real* in[6];        // the pass's six inputs (fields 4, 5 diffused here)
  real* out[6];
  const real* rhog;
  // LDS per buffer: [field f][triangle T][component dd][corner]
  static int lds_elems(const ColGeom& cg) { return 2 * 12 * (cg.TX + 1) * (cg.TY + 1); }
  struct State {
    real ci[2][9];
    real cd[18];
    real rg[7];         // rhog at the tile point and its neighbours, loaded before the barrier
    long a;
    int k, lidx;
    bool corner, outp, pent, sgp;
  };
  NH_HD int init(int blk, int tid, State& s) const {
    s.corner = s.outp = s.pent = s.sgp = false;
    if (blk >= cg.ntiles) {
      if (tid == 0) {
        HDiffPassP<1> p;
        p.d = d; p.c = c; p.G = G; p.rhog = rhog;
        for (int q = 0; q < 6; ++q) { p.in[q] = in[q]; p.out[q] = out[q]; }
        p(d.lall * d.kall + (blk - cg.ntiles), 0);
      }
      return 0;
    }
    const int g = d.g, W1 = cg.TX + 1, tpr = cg.nx * cg.ny;
    const int tile = blk % tpr, rest = blk / tpr;
    const int chunk = rest % cg.nchunk, l = rest / cg.nchunk;
    const int k0 = chunk * cg.KC;
    const int k1 = (k0 + cg.KC < d.kall) ? k0 + cg.KC : d.kall;
    if (k0 >= k1) return 0;
    const int i0 = 1 + (tile % cg.nx) * cg.TX, j0 = 1 + (tile / cg.nx) * cg.TY;
    const int tx = tid % W1, ty = tid / W1;
    s.k = k0;
    if (tid >= W1 * (cg.TY + 1)) return k1 - k0;
    s.corner = true;
    s.lidx = ty * W1 + tx;
    const int i = i0 - 1 + tx, j = j0 - 1 + ty;
    const long nreg = d.nreg, col = ((long)l * g + j) * g + i;
    s.a = ((long)(l * d.kall + k0) * g + j) * g + i;
    for (int T = 0; T < 2; ++T)
      for (int n = 0; n < 9; ++n) s.ci[T][n] = rc(G.cintp, nreg, n * 2 + T, col);
    if (tx >= 1 && ty >= 1) {
      s.outp = true;
      for (int n = 0; n < 18; ++n) s.cd[n] = G.cdiff[(long)n * nreg + col];
      s.pent = (i == 1 && j == 1);
      s.sgp = G.sgp[l] != 0;
    }
    return k1 - k0;
  }
  NH_HD void level_a(int step, State& s, real* lds) const {
    if (!s.corner) return;
    const long g = d.g;
    const int P = (cg.TX + 1) * (cg.TY + 1);
    real* buf = lds + (step & 1) * 12 * P;
    if (s.outp) {
      // the point's density loads go out with the triangle loads, so that
      // level_b does not wait for memory after the barrier
      const long a = s.a;
      s.rg[0] = rhog[a]; s.rg[1] = rhog[a + 1 + g]; s.rg[2] = rhog[a + g]; s.rg[3] = rhog[a - 1];
      s.rg[4] = rhog[a - 1 - g]; s.rg[5] = rhog[a - g]; s.rg[6] = rhog[a + 1];
    }
    for (int f = 0; f < 2; ++f) {
      const real* sf = in[4 + f];
      const real v00 = sf[s.a], v10 = sf[s.a + 1], v11 = sf[s.a + 1 + g], v01 = sf[s.a + g];
      for (int T = 0; T < 2; ++T) {
        // tri_grad with the held coefficients
        const real s0 = v00, s1 = (T == 0) ? v10 : v11, s2 = (T == 0) ? v11 : v01;
        for (int dd = 0; dd < 3; ++dd) {
          const real c1 = s.ci[T][dd * 3 + 0], c2 = s.ci[T][dd * 3 + 1], c3 = s.ci[T][dd * 3 + 2];
          buf[((f * 2 + T) * 3 + dd) * P + s.lidx] =
              ((real(2.0) * c1 - c2 - c3) * s0 + (-c1 + real(2.0) * c2 - c3) * s1 + (-c1 - c2 + real(2.0) * c3) * s2)
              / real(3.0);
        }
      }
    }
  }
  NH_HD void level_b(int step, State& s, const real* lds) const {
    if (s.outp) {
      const long g = d.g;
      const int W1 = cg.TX + 1, P = W1 * (cg.TY + 1);
      const real* buf = lds + (step & 1) * 12 * P;
      const long a = s.a;
      const int x = s.lidx;
      // corners of TI(i,j) TJ(i,j) TI(i-1,j) TJ(i-1,j-1) TI(i-1,j-1) TJ(i,j-1)
      const int cx[6] = {x, x, x - 1, x - 1 - W1, x - 1 - W1, x - W1}, cT[6] = {0, 1, 0, 1, 0, 1};
      const real r0 = s.rg[0], r1 = s.rg[1], r2 = s.rg[2], r3 = s.rg[3], r4 = s.rg[4], r5 = s.rg[5], r6 = s.rg[6];
      const real m0 = s.sgp ? real(0.0) : real(1.0), m1 = s.sgp ? real(1.0) : real(0.0);
      for (int f = 0; f < 2; ++f) {
        const real fac = (f == 0) ? c.CVdry : c.HDIFF_FACT_RHO;
        const real gam = c.GAMMA_H;
        real tv[6][3];
        for (int n = 0; n < 6; ++n)
          for (int dd = 0; dd < 3; ++dd) tv[n][dd] = buf[((f * 2 + cT[n]) * 3 + dd) * P + cx[n]];
        if (s.pent)
          for (int dd = 0; dd < 3; ++dd) tv[4][dd] = tv[4][dd] * m0 + tv[5][dd] * m1;
        const real k0 = r0 * fac * gam;
        const real kf1 = real(0.5) * (k0 + r1 * fac * gam);
        const real kf2 = real(0.5) * (k0 + r2 * fac * gam);
        const real kf3 = real(0.5) * (k0 + r3 * fac * gam);
        const real kf4 = real(0.5) * (k0 + r4 * fac * gam);
        const real kf5 = real(0.5) * (k0 + r5 * fac * gam);
        const real kf6 = real(0.5) * (k0 + r6 * fac * gam);
        real inner = real(0.0);
        for (int dd = 0; dd < 3; ++dd) {
          const real* cd = s.cd + dd * 6;
          const real blk2 = kf1 * cd[0] * (tv[0][dd] + tv[1][dd]) + kf2 * cd[1] * (tv[1][dd] + tv[2][dd])
                            + kf3 * cd[2] * (tv[2][dd] + tv[3][dd]) + kf4 * cd[3] * (tv[3][dd] + tv[4][dd])
                            + kf5 * cd[4] * (tv[4][dd] + tv[5][dd]) + kf6 * cd[5] * (tv[5][dd] + tv[0][dd]);
          inner = (dd == 0) ? blk2 : inner + blk2;
        }
        out[4 + f][a] = -inner;
      }
    }
    s.a += (long)d.g * d.g;
    s.k += 1;
  }
};

NH_END
