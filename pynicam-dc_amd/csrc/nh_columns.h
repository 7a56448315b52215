// Column kernels of the dynamical-core step.
//
// Every kernel is a functor `f(t)` over a flat thread index t.  Regional
// columns come first (t < lall*g*g, i fastest), pole columns after.  A thread
// marches its own column in k, reading the six neighbour columns of the
// icosahedral stencil at the same level, so vertical recurrences (half-level
// interpolation, the tridiagonal solve) stay in registers and every level of a
// wavefront is a coalesced row of i.
//
// Field layout in memory (one field = one array of NF doubles):
//   regional  ((l*kall + k)*g + j)*g + i        pole  R + (lp*kall + k)*6 + v
// so a column is (base, stride) with stride g*g (regional) or 6 (pole).
//
// The functors are plain structs of pointers and scalars, compiled for the GPU
// (hipcc) and, in the test build, for the host.
//
// Precision: `real` is the type of the fields and of the stencil arithmetic.
// The sources are compiled once per variant, each in its own namespace
// (nh::f64 with real = double, the drop-in form; nh::f32 with real = float,
// the mixed form).  Values the mixed form keeps in FP64 (the tridiagonal
// matrix and solve, the global sums, time steps) are written as `double`.
#pragma once

#include <cmath>
#include <cstdint>

#ifndef NH_HD
#define NH_HD __host__ __device__
#endif

#ifndef NH_REAL
#define NH_REAL double
#define NH_VARIANT f64
#endif

#define NH_KMAX 80   // largest supported kall

#define NH_BEGIN namespace nh { namespace NH_VARIANT {
#define NH_END } }

NH_BEGIN

using real = NH_REAL;

// Load of a value that no thread writes during the kernel (coefficients, and
// input fields a kernel only reads).  On the device it reads through the
// constant address space, which the compiler knows is never stored to, so
// such loads may be issued ahead of the kernel's stores and atomics instead
// of waiting behind each one.
template <class T>
NH_HD inline T ldc(const T* p) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef const __attribute__((address_space(4))) T* cptr;
  return *(cptr)p;
#else
  return *p;
#endif
}

// Load of an element that the calling thread alone reads and writes during
// the kernel, once each and the load first (a read-modify-write of a value
// the thread owns, such as a column's adjoint bar).  No other thread stores
// to it, so like ldc it may read through the constant address space and need
// not wait behind the thread's stores to other elements.  A second load of
// the element after its store would break this: the caller keeps one load
// and one store per element.
template <class T>
NH_HD inline T ld_own(const T* p) {
  return ldc(p);
}

// type-exact absolute value and sign transfer (no promotion of float arguments)
NH_HD inline float nabs(float x) { return fabsf(x); }
NH_HD inline double nabs(double x) { return fabs(x); }
NH_HD inline float nsign(float m, float s) { return copysignf(m, s); }
NH_HD inline double nsign(double m, double s) { return copysign(m, s); }

struct Dims {
  int g;       // gall_1d
  int kall, kmin, kmax;
  int lall;
  long nreg;   // lall*g*g   regional columns
  long R;      // lall*kall*g*g regional points
  long NF;     // R + 2*kall*6 points per field
};

template <class T>
struct ConstsT {
  T RADIUS, OHM, GRAV, Rdry, CVdry, Rvap, CVW_qv, ALPHA;
  T GAMMA_H, HDIFF_FACT_RHO, ALPHA_D;
};

// physical constants in the field precision, and in FP64 (d) for the parts
// the mixed form keeps in double
struct Consts : ConstsT<real> {
  ConstsT<double> d;
  int THUBURN_LIM;
};

// vertical profiles [kall]
template <class T>
struct VProf {
  const T *afact, *bfact, *cfact, *dfact, *rdgz, *rdgzh;
  const T *c2w0, *c2w1, *w2c0, *w2c1, *phi;
  const T *dgz, *gz, *gzh;
};

struct Geo : VProf<real> {
  // regional, structure of arrays [comp][nreg]
  const real* cdiv;    // 21 = d*7 + m
  const real* cgrad;   // 21
  const real* clap;    // 7
  const real* cintp;   // 18 = (d*3 + vertex)*2 + T
  const real* cdiff;   // 18 = d*6 + arc
  const real* gx;      // 3  GRD_x [m]
  // pole, [v][lp][...]
  const real* cdiv_pl;   // (v*2+lp)*3+d
  const real* cgrad_pl;  // (v*2+lp)*3+d
  const real* clap_pl;   // v*2+lp
  const real* cintp_pl;  // ((v*2+lp)*3+d)*3+vertex
  const real* cdiff_pl;  // (v*2+lp)*3+d
  const real* gx_pl;     // (v*2+lp)*3+d
  const int* sgp;        // [lall] 1 on pentagon regions
  // tracer transport: regional [comp][nreg], pole [v][lp][...]
  const real* twt;      // 6 = T*3 + w
  const real* ahn;      // 9 = arc*3 + d
  const real* rarea;    // 1
  const real* axr;      // 9 = arc*3 + d
  const real* twt_pl;   // (v*2+lp)*3+w
  const real* ahn_pl;   // (v*2+lp)*3+d
  const real* rarea_pl; // lp
  const real* axr_pl;   // (v*2+lp)*3+d
  // FP64: budgets (control-volume area per column), prescribed winds
  // (latitude, longitude per column and at the poles)
  const double* area;     // 1
  const double* lat;      // 1
  const double* lon;      // 1
  const double* lat_pl;   // v*2+lp
  const double* lon_pl;   // v*2+lp
  // the vertical profiles in FP64
  VProf<double> v64;
  // hydrostatic basic state (Bsst.pre_bs / tem_bs / rho_bs, mod_bsstate.py:159),
  // FP64 k-profiles in every form; all zero for ref_type NOBASE
  const double* pre_bs;
  const double* tem_bs;
  const double* rho_bs;
  bool bs_on;
};

enum { I_RHOG = 0, I_RHOGVX, I_RHOGVY, I_RHOGVZ, I_RHOGW, I_RHOGE };
enum { I_PRE = 0, I_TEM, I_VX, I_VY, I_VZ, I_W };

// ----------------------------------------------------------------------------
// column addressing
// ----------------------------------------------------------------------------
struct Cell {
  bool pole;    // pole centre (or pole point for all-column kernels)
  bool inner;   // regional interior point (stencil output location)
  int i, j, l, v, lp;
  long base;    // column base address
  long ks;      // level stride
  long ci;      // regional cell index (l*g + j)*g + i
};

NH_HD inline Cell cell_of(const Dims& d, long t) {
  Cell c;
  if (t < d.nreg) {
    long gg = (long)d.g * d.g;
    c.l = (int)(t / gg);
    long r = t - (long)c.l * gg;
    c.j = (int)(r / d.g);
    c.i = (int)(r - (long)c.j * d.g);
    c.pole = false;
    c.inner = (c.i >= 1 && c.i <= d.g - 2 && c.j >= 1 && c.j <= d.g - 2);
    c.base = ((long)c.l * d.kall * d.g + c.j) * d.g + c.i;
    c.ks = gg;
    c.ci = t;
    c.v = 0;
    c.lp = 0;
  } else {
    long p = t - d.nreg;
    c.lp = (int)(p / 6);
    c.v = (int)(p - (long)c.lp * 6);
    c.pole = true;
    c.inner = false;
    c.base = d.R + (long)c.lp * d.kall * 6 + c.v;
    c.ks = 6;
    c.ci = -1;
    c.i = c.j = c.l = 0;
  }
  return c;
}

// ----------------------------------------------------------------------------
// point addressing: point kernels run on a 2-D grid, y = plane (region level
// l*kall + k, then the 2*kall pole levels lp*kall + k), x = point in the plane
// (j*g + i, or the pole point v).  The field address is plane*g*g + x
// (regional) or R + plane_pole*6 + v.
// ----------------------------------------------------------------------------
struct Pt {
  bool pole, inner;
  int i, j, l, v, lp, k;
  long a;       // field address of the point
  long ks;      // level stride
  long ci;      // regional column index (l*g + j)*g + i
};

NH_HD inline bool pt_of(const Dims& d, int plane, int x, Pt& p) {
  // every field is computed into a local on both paths and stored once, so
  // the struct stays in registers
  const int g = d.g;
  const int nrp = d.lall * d.kall;
  const bool pole = plane >= nrp;
  if (x >= (pole ? 6 : g * g)) return false;
  int l = 0, k, i = 0, j = 0, v = 0, lp = 0;
  long a, ks, ci;
  if (!pole) {
    l = plane / d.kall;
    k = plane - l * d.kall;
    j = x / g;
    i = x - j * g;
    ks = (long)g * g;
    a = (long)plane * g * g + x;
    ci = (long)l * g * g + x;
  } else {
    const int pp = plane - nrp;
    lp = pp / d.kall;
    k = pp - lp * d.kall;
    v = x;
    ks = 6;
    a = d.R + (long)pp * 6 + x;
    ci = -1;
  }
  p.pole = pole;
  p.inner = !pole && i >= 1 && i <= g - 2 && j >= 1 && j <= g - 2;
  p.i = i; p.j = j; p.l = l; p.v = v; p.lp = lp; p.k = k;
  p.a = a; p.ks = ks; p.ci = ci;
  return true;
}

// the column of a point
NH_HD inline Cell cell_of_pt(const Pt& p) {
  Cell c;
  c.pole = p.pole;
  c.inner = p.inner;
  c.i = p.i; c.j = p.j; c.l = p.l; c.v = p.v; c.lp = p.lp;
  c.ks = p.ks;
  c.base = p.a - (long)p.k * p.ks;
  c.ci = p.ci;
  return c;
}

// the compute points of the horizontal operators: regional interior and pole centres
NH_HD inline bool pt_compute(const Pt& p) { return p.pole ? p.v == 0 : p.inner; }

// stencil offsets in column-base units: (0,0) (+1,0) (+1,+1) (0,+1) (-1,0) (-1,-1) (0,-1)
NH_HD inline void stencil_offsets(int g, long off[7]) {
  off[0] = 0; off[1] = 1; off[2] = 1 + g; off[3] = g;
  off[4] = -1; off[5] = -1 - g; off[6] = -g;
}

template <class T>
NH_HD inline T rc(const T* a, long nreg, int comp, long ci) { return ldc(a + comp * nreg + ci); }
// the same as a plain load for 8-byte values (the tracer transport's flux
// and remap kernels measured slower with ldc in the FP64 form, faster in FP32)
template <class T>
NH_HD inline T rcp(const T* a, long nreg, int comp, long ci) {
  if constexpr (sizeof(T) == 4) return ldc(a + comp * nreg + ci);
  else return a[comp * nreg + ci];
}

// 7-point divergence of (ax, ay, az)(point) at one level for a regional cell
template <class FX, class FY, class FZ>
NH_HD inline real div7(const Geo& G, long nreg, long ci, FX fx, FY fy, FZ fz) {
  real sx, sy, sz;
  // same association as the reference: sum over m for each direction, then x + y + z
  sx = rc(G.cdiv, nreg, 0, ci) * fx(0);
  for (int m = 1; m < 7; ++m) sx = sx + rc(G.cdiv, nreg, m, ci) * fx(m);
  sy = rc(G.cdiv, nreg, 7, ci) * fy(0);
  for (int m = 1; m < 7; ++m) sy = sy + rc(G.cdiv, nreg, 7 + m, ci) * fy(m);
  sz = rc(G.cdiv, nreg, 14, ci) * fz(0);
  for (int m = 1; m < 7; ++m) sz = sz + rc(G.cdiv, nreg, 14 + m, ci) * fz(m);
  return sx + sy + sz;
}

// pole-centre divergence: sum over v = 0..5 of (cx*ax + cy*ay + cz*az)
template <class FX, class FY, class FZ>
NH_HD inline real div_pl(const Geo& G, int lp, FX fx, FY fy, FZ fz) {
  real s = real(0.0);
  for (int v = 0; v < 6; ++v) {
    const real* c = G.cdiv_pl + (v * 2 + lp) * 3;
    real pv = c[0] * fx(v) + c[1] * fy(v) + c[2] * fz(v);
    s = (v == 0) ? pv : s + pv;
  }
  return s;
}

// ----------------------------------------------------------------------------
// halo gather
// ----------------------------------------------------------------------------
// field address of level 0 of halo-map column c (regional l*g*g + j*g + i,
// then the 12 pole columns lp*6 + v) and its level stride
NH_HD inline long halo_colbase(const Dims& d, long c, long* ks) {
  if (c < d.nreg) {
    long gg = (long)d.g * d.g;
    long l = c / gg;
    long r = c - l * gg;
    *ks = gg;
    return l * d.kall * gg + r;
  }
  long p = c - d.nreg;
  long lp = p / 6;
  *ks = 6;
  return d.R + lp * d.kall * 6 + (p - lp * 6);
}

struct HaloGather {
  static constexpr const char* kName = "halo_gather";
  Dims d;
  const int* dst;   // level-0 field addresses of the halo columns (halo_colbase)
  const int* src;
  long n;           // number of halo columns
  real* f[16];
  int nf;
  // level stride of a column from its level-0 address: regional below R
  NH_HD inline long stride(long b) const { return (b < d.R) ? (long)d.g * d.g : 6L; }
  // t over n * kall: halo column fastest so a wavefront copies consecutive
  // halo slots (32-bit index arithmetic: n * kall < 2^31 is checked at upload)
  NH_HD inline void operator()(long t) const {
    const unsigned tt = (unsigned)t, nn = (unsigned)n;
    const unsigned k = tt / nn, h = tt - k * nn;
    const long bd0 = dst[h], bs0 = src[h];
    const long bd = bd0 + (long)k * stride(bd0);
    const long bs = bs0 + (long)k * stride(bs0);
    for (int q = 0; q < nf; ++q) f[q][bd] = f[q][bs];
  }
};

NH_END
