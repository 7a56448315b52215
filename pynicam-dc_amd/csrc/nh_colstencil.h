// Column tiles for 7-point stencil kernels.
//
// A block owns a TX x TY tile of one region's points (st_geom) over a chunk of
// levels.  Each of the first TX TY threads owns one tile column: the
// per-column coefficients of its operator are loaded once into registers, and
// then, level by level, the thread stages its own values of the operator's
// stencil fields into LDS (the first 2 TX + 2 TY + 2 threads also stage one
// point of the ring around the tile), a barrier, and the thread
// forms its point's result from its neighbours' staged values.  The staging
// planes are double-buffered by level parity, so one barrier per level
// suffices.  The point kernels re-read the 21 stencil coefficients (and the
// neighbours' values) per point and level; here each value is loaded from
// memory once per level and each coefficient once per chunk.
//
// An operator Op supplies
//   kName, kNS (stencil fields), Col (per-column registers), d (Dims),
//   Stg (registers of the column a thread stages: its own and its ring column),
//   init_stage(col index, Stg&), stage_level(k), stage(k, a, Stg, v[kNS]),
//   init_col(col index, Col&),
//   point(k, a, Col, nb) with nb(m, f) = field f at stencil neighbour m
//   (0 (0,0) 1 (+1,0) 2 (+1,+1) 3 (0,+1) 4 (-1,0) 5 (-1,-1) 6 (0,-1)),
//   pole(pb): the pole centre of pole level pb (its point kernel).
#pragma once

#include "nh_tile.h"

NH_BEGIN

// tile shape of the stencil engine: TX x TY = 256 points, 64 x 4 where the
// region is wide enough (a wavefront then reads one 64-point row segment;
// measured faster than 32 x 8 and 16 x 16 at gl07)
struct StGeom {
  int TX, TY;       // tile width and height (points)
  int nx, ny;       // tiles per region row and column
  int nchunk, KC;   // level chunks, levels per chunk
  int ntiles;       // lall * nchunk * nx * ny
  int nblocks;      // + 2 * kall pole blocks
};

inline StGeom st_geom(const Dims& d, int nchunk) {
  StGeom c;
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

template <class Op>
struct ColStencil {
  using LdsT = real;
  static constexpr const char* kName = Op::kName;
  static constexpr int kThreads = 256;
  static constexpr int NS = Op::kNS;
  Op op;
  StGeom cg;
  static int lds_elems(const StGeom& cg) { return 2 * NS * (cg.TX + 2) * (cg.TY + 2); }
  struct State {
    typename Op::Col col;
    typename Op::Stg sg, rsg;   // staging registers of the own and the ring column
    long a, ra;        // own and ring addresses at the current level
    int k, sidx, ridx;
    bool outp, ring;
  };
  NH_HD int init(int blk, int tid, State& s) const {
    const Dims& d = op.d;
    s.outp = s.ring = false;
    if (blk >= cg.ntiles) {
      if (tid == 0) op.pole(blk - cg.ntiles);
      return 0;
    }
    const int g = d.g, TX = cg.TX, TY = cg.TY, SW = TX + 2, tpr = cg.nx * cg.ny;
    const int tile = blk % tpr, rest = blk / tpr;
    const int chunk = rest % cg.nchunk, l = rest / cg.nchunk;
    const int k0 = chunk * cg.KC;
    const int k1 = (k0 + cg.KC < d.kall) ? k0 + cg.KC : d.kall;
    if (k0 >= k1) return 0;
    const int i0 = 1 + (tile % cg.nx) * TX, j0 = 1 + (tile / cg.nx) * TY;
    s.k = k0;
    if (tid < TX * TY) {
      const int tx = tid % TX, ty = tid / TX;
      const int i = i0 + tx, j = j0 + ty;
      s.outp = true;
      s.sidx = (ty + 1) * SW + tx + 1;
      s.a = ((long)(l * d.kall + k0) * g + j) * g + i;
      op.init_col(((long)l * g + j) * g + i, s.col);
      op.init_stage(((long)l * g + j) * g + i, s.sg);
    }
    // the ring: row -1 (x = -1..TX-1), row TY (x = 0..TX), columns -1 and TX (y = 0..TY-1)
    if (tid < 2 * TX + 2 * TY + 2) {
      int x, y;
      if (tid < TX + 1) { x = tid - 1; y = -1; }
      else if (tid < 2 * TX + 2) { x = tid - (TX + 1); y = TY; }
      else if (tid < 2 * TX + TY + 2) { x = -1; y = tid - (2 * TX + 2); }
      else { x = TX; y = tid - (2 * TX + TY + 2); }
      s.ring = true;
      s.ridx = (y + 1) * SW + x + 1;
      s.ra = ((long)(l * d.kall + k0) * g + (j0 + y)) * g + (i0 + x);
      op.init_stage(((long)l * g + (j0 + y)) * g + (i0 + x), s.rsg);
    }
    return k1 - k0;
  }
  NH_HD void level_a(int step, State& s, real* lds) const {
    if (!op.stage_level(s.k)) return;
    const int SW = cg.TX + 2, P = SW * (cg.TY + 2);
    real* buf = lds + (step & 1) * NS * P;
    real v[NS];
    if (s.outp) {
      op.stage(s.k, s.a, s.sg, v);
      for (int f = 0; f < NS; ++f) buf[f * P + s.sidx] = v[f];
    }
    if (s.ring) {
      op.stage(s.k, s.ra, s.rsg, v);
      for (int f = 0; f < NS; ++f) buf[f * P + s.ridx] = v[f];
    }
  }
  NH_HD void level_b(int step, State& s, const real* lds) const {
    if (s.outp) {
      const int SW = cg.TX + 2, P = SW * (cg.TY + 2);
      const real* buf = lds + (step & 1) * NS * P;
      const int x = s.sidx;
      op.point(s.k, s.a, s.col, [&](int m, int f) {
        const int off = (m == 0) ? 0 : (m == 1) ? 1 : (m == 2) ? 1 + SW : (m == 3) ? SW : (m == 4) ? -1
                      : (m == 5) ? -1 - SW : -SW;
        return buf[f * P + x + off];
      });
    }
    const long gg = (long)op.d.g * op.d.g;
    s.a += gg;
    s.ra += gg;
    s.k += 1;
  }
};

// vi_main, horizontal part (ViMainH) as a column operator: the 21 divergence
// coefficients in registers, the 8 stencil fields staged per level
struct ViMainHC {
  static constexpr const char* kName = "vi_main_h";
  static constexpr int kNS = 8;
  ViMainH pk;
  Dims d;
  struct Col {
    real c[3][7];
  };
  struct Stg {};
  NH_HD void init_stage(long, Stg&) const {}
  NH_HD bool stage_level(int k) const { return k >= d.kmin && k <= d.kmax; }
  NH_HD void stage(int, long a, const Stg&, real* v) const {
    for (int f = 0; f < kNS; ++f) v[f] = pk.stage_ptr(f, 0)[a];
  }
  NH_HD void init_col(long ci, Col& col) const {
    for (int dd = 0; dd < 3; ++dd)
      for (int m = 0; m < 7; ++m) col.c[dd][m] = rc(pk.G.cdiv, d.nreg, dd * 7 + m, ci);
  }
  NH_HD void pole(int pb) const { pk(d.lall * d.kall + pb, 0); }
  // ViMainH::at with hdiv_multi<3>, term for term
  template <class NB>
  NH_HD void point(int k, long a, const Col& col, NB nb) const {
    if (k < d.kmin || k > d.kmax) return;
    const long NF = d.NF;
    real s[3][3];
#pragma unroll
    for (int m = 0; m < 7; ++m) {
      real v[3][3];
      const real e = nb(m, 3), et = nb(m, 4);
      for (int q = 0; q < 3; ++q) {
        const real dq = nb(m, q);
        v[0][q] = dq * real(1.0);
        v[1][q] = dq * e * real(1.0);
        v[2][q] = (nb(m, 5 + q) + dq) * et * real(1.0);
      }
      for (int i = 0; i < 3; ++i) {
        const real tx = col.c[0][m] * v[i][0], ty = col.c[1][m] * v[i][1], tz = col.c[2][m] * v[i][2];
        s[i][0] = (m == 0) ? tx : s[i][0] + tx;
        s[i][1] = (m == 0) ? ty : s[i][1] + ty;
        s[i][2] = (m == 0) ? tz : s[i][2] + tz;
      }
    }
    const real divs1 = s[0][0] + s[0][1] + s[0][2];
    const real he = s[1][0] + s[1][1] + s[1][2];
    const real hd = s[2][0] + s[2][1] + s[2][2];
    const real drhoge = -he - (real(0.0) - real(0.0)) * pk.G.rdgz[k];
    const real grhoge1 = pk.gT[5 * NF + a] + drhoge;
    pk.hv[0][a] = divs1;
    pk.hv[1][a] = grhoge1 * pk.c.Rdry / pk.c.CVdry;
    pk.hv[2][a] = hd;
  }
};

// minus the 7-point Laplacian of NFLD fields (HDiffPassP's Laplacians) as a
// column operator: the 7 Laplacian coefficients in registers.  P is the point
// kernel that does the pole centres (all of its fields) unless pole_off.
template <int NFLD, class P>
struct LapC {
  static constexpr const char* kName = (NFLD == 6) ? "hdiff_pass1" : "hdiff_pass2_lap";
  static constexpr int kNS = NFLD;
  P pk;
  Dims d;
  const real* in[NFLD];
  real* out[NFLD];
  bool pole_off;
  struct Col {
    real c[7];
  };
  struct Stg {};
  NH_HD void init_stage(long, Stg&) const {}
  NH_HD bool stage_level(int) const { return true; }
  NH_HD void stage(int k, long a, const Stg&, real* v) const {
    for (int f = 0; f < NFLD; ++f) v[f] = pk.input(f, a, k);
  }
  NH_HD void init_col(long ci, Col& col) const {
    for (int m = 0; m < 7; ++m) col.c[m] = rc(pk.G.clap, d.nreg, m, ci);
  }
  NH_HD void pole(int pb) const {
    if (!pole_off) pk(d.lall * d.kall + pb, 0);
  }
  template <class NB>
  NH_HD void point(int, long a, const Col& col, NB nb) const {
    for (int f = 0; f < NFLD; ++f) {
      real r = col.c[0] * nb(0, f);
#pragma unroll
      for (int m = 1; m < 7; ++m) r = r + col.c[m] * nb(m, f);
      out[f][a] = -r;
    }
  }
};

// the transposed Laplacian of the hyperdiffusion adjoint at the interior
// points: ib(x) += -sum_m clap(x - off(m), m) ob(x - off(m)), summed over
// the interior sources in NegLapGA's order; per column the seven
// coefficients (zero where the source is not interior) in registers, ob
// staged per level.  NegLapGA's 1-D form does the border and pole points.
template <int NFLD>
struct LapTC {
  static constexpr const char* kName = "hdiff_lap_adj";
  static constexpr int kNS = NFLD;
  Dims d; Geo G;
  const real* ob[NFLD];
  real* ib[NFLD];
  struct Col {
    real c[7];
  };
  struct Stg {};
  NH_HD void init_stage(long, Stg&) const {}
  NH_HD bool stage_level(int) const { return true; }
  NH_HD void stage(int, long a, const Stg&, real* v) const {
    for (int f = 0; f < NFLD; ++f) v[f] = ob[f][a];
  }
  NH_HD void init_col(long ci, Col& col) const {
    const long gg = (long)d.g * d.g;
    const long r = ci % gg;
    const int j = (int)(r / d.g), i = (int)(r - (long)j * d.g);
    for (int m = 0; m < 7; ++m) {
      const int di = ((4201 >> (2 * m)) & 3) - 1, dj = ((421 >> (2 * m)) & 3) - 1;
      const int is = i - di, js = j - dj;
      const bool in = is >= 1 && is <= d.g - 2 && js >= 1 && js <= d.g - 2;
      col.c[m] = in ? rc(G.clap, d.nreg, m, ci - (long)di - (long)dj * d.g) : real(0.0);
    }
  }
  NH_HD void pole(int) const {}
  template <class NB>
  NH_HD void point(int, long a, const Col& col, NB nb) const {
    // the source x - off(m) is staged neighbour opp(m)
    const int opp[7] = {0, 4, 5, 6, 1, 2, 3};
    for (int f = 0; f < NFLD; ++f) {
      real acc = real(0.0);
#pragma unroll
      for (int m = 0; m < 7; ++m) acc += col.c[m] * nb(opp[m], f);
      ib[f][a] += -acc;
    }
  }
};

// momentum advection and Coriolis (AdvMomP) as a column operator: the 21
// divergence coefficients and the column's position in registers; staged per
// level are the merged velocity v + w_full r/|r| (formed once per point, where
// the point kernel forms it once per neighbour) and the three momentum
// components.  Term for term the point kernel's arithmetic; pole centres run
// the point kernel.
struct AdvMomC {
  static constexpr const char* kName = "adv_mom";
  static constexpr int kNS = 6;
  AdvMomP pk;
  Dims d;
  struct Col {
    real c[3][7];
    real x[3];
  };
  struct Stg {
    real x[3];
  };
  NH_HD void init_stage(long ci, Stg& sg) const {
    for (int q = 0; q < 3; ++q) sg.x[q] = pk.G.gx[q * d.nreg + ci];
  }
  NH_HD bool stage_level(int k) const { return k >= d.kmin && k <= d.kmax; }
  NH_HD void stage(int k, long a, const Stg& sg, real* v) const {
    const long NF = d.NF, gg = (long)d.g * d.g;
    const real* wd = pk.diag + I_W * NF;
    const real R = pk.c.RADIUS;
    const real wc = pk.G.cfact[k] * wd[a + gg] + pk.G.dfact[k] * wd[a];
    v[0] = pk.diag[I_VX * NF + a] + wc * sg.x[0] / R;
    v[1] = pk.diag[I_VY * NF + a] + wc * sg.x[1] / R;
    v[2] = pk.diag[I_VZ * NF + a] + wc * sg.x[2] / R;
    v[3] = pk.prog[1 * NF + a];
    v[4] = pk.prog[2 * NF + a];
    v[5] = pk.prog[3 * NF + a];
  }
  NH_HD void init_col(long ci, Col& col) const {
    for (int dd = 0; dd < 3; ++dd)
      for (int m = 0; m < 7; ++m) col.c[dd][m] = rc(pk.G.cdiv, d.nreg, dd * 7 + m, ci);
    for (int q = 0; q < 3; ++q) col.x[q] = pk.G.gx[q * d.nreg + ci];
  }
  NH_HD void pole(int pb) const { pk(d.lall * d.kall + pb, 0); }
  template <class NB>
  NH_HD void point(int k, long a, const Col& col, NB nb) const {
    const long NF = d.NF, ks = (long)d.g * d.g;
    const int kmin = d.kmin, kmax = d.kmax;
    real* out = pk.out;
    if (k < kmin || k > kmax) {
      for (int q = 0; q < 4; ++q) out[q * NF + a] = real(0.0);
      return;
    }
    const Geo& G = pk.G;
    const real R = pk.c.RADIUS;
    const real* rw = pk.prog + 4 * NF;
    const real* rg = pk.prog;
    const real* vxd = pk.diag + I_VX * NF;
    const real* vyd = pk.diag + I_VY * NF;
    const real* vzd = pk.diag + I_VZ * NF;
    const real* wd = pk.diag + I_W * NF;
    const real* x0 = col.x;
    auto vv3 = [&](long b, int kk, real* o) {
      const real wc = G.cfact[kk] * wd[b + ks] + G.dfact[kk] * wd[b];
      o[0] = vxd[b] + wc * x0[0] / R;
      o[1] = vyd[b] + wc * x0[1] / R;
      o[2] = vzd[b] + wc * x0[2] / R;
    };
    real sd[3][3];
#pragma unroll
    for (int m = 0; m < 7; ++m) {
      const real vv[3] = {nb(m, 0), nb(m, 1), nb(m, 2)};
      const real fx0 = nb(m, 3), fy0 = nb(m, 4), fz0 = nb(m, 5);
      const real cx = col.c[0][m], cy = col.c[1][m], cz = col.c[2][m];
      for (int q = 0; q < 3; ++q) {
        const real tx = cx * (fx0 * vv[q] * real(1.0));
        const real ty = cy * (fy0 * vv[q] * real(1.0));
        const real tz = cz * (fz0 * vv[q] * real(1.0));
        sd[0][q] = (m == 0) ? tx : sd[0][q] + tx;
        sd[1][q] = (m == 0) ? ty : sd[1][q] + ty;
        sd[2][q] = (m == 0) ? tz : sd[2][q] + tz;
      }
    }
    real div[3];
    for (int q = 0; q < 3; ++q) div[q] = sd[0][q] + sd[1][q] + sd[2][q];
    const real vvm0[3] = {nb(0, 0), nb(0, 1), nb(0, 2)};
    real hi[3] = {real(0.0), real(0.0), real(0.0)}, lo[3] = {real(0.0), real(0.0), real(0.0)};
    if (k + 1 <= kmax) {
      real up[3];
      vv3(a + ks, k + 1, up);
      const real wup = rw[a + ks];
      for (int q = 0; q < 3; ++q)
        hi[q] = real(0.0) * real(1.0) + wup * (G.afact[k + 1] * up[q] + G.bfact[k + 1] * vvm0[q]) * real(1.0);
    }
    if (k >= kmin + 1) {
      real dn[3];
      vv3(a - ks, k - 1, dn);
      const real w0 = rw[a];
      for (int q = 0; q < 3; ++q)
        lo[q] = real(0.0) * real(1.0) + w0 * (G.afact[k] * vvm0[q] + G.bfact[k] * dn[q]) * real(1.0);
    }
    real dvv[3];
    for (int q = 0; q < 3; ++q) dvv[q] = -div[q] - (hi[q] - lo[q]) * G.rdgz[k];
    const real gxc = x0[0] / R, gyc = x0[1] / R, gzc = x0[2] / R;
    const real rh = rg[a];
    const real dx = dvv[0] - (-real(2.0) * rh * (pk.c.OHM * vvm0[1]));
    const real dy = dvv[1] - (real(2.0) * rh * (pk.c.OHM * vvm0[0]));
    const real dz = dvv[2];
    const real prd = dx * gxc + dy * gyc + dz * gzc;
    out[0 * NF + a] = dx - prd * gxc;
    out[1 * NF + a] = dy - prd * gyc;
    out[2 * NF + a] = dz - prd * gzc;
    out[3 * NF + a] = prd * pk.c.ALPHA;
  }
};

NH_END
