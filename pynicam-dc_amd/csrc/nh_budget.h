// Global mass and energy sums of the prognostic state
// (Embudget._store_var, share/mod_embudget.py:82-139).
//
// One thread per interior point of the physical levels (i, j in 1..g-2,
// k in kmin..kmax, all regions; pole cells are left out as in _gsum,
// mod_embudget.py:74-80).  Phase 1 writes the six volume-weighted terms of
// the point into LDS; phase 2 lets six threads sum one term each over the
// block with Neumaier's compensated summation and store (sum, correction) as
// the block's partial.  The host adds the partials, again compensated.  All
// arithmetic is FP64 in every form (field values are widened on load).
#pragma once

#include "nh_step.h"

NH_BEGIN

// Neumaier's improved Kahan summation: s + c carries the running sum
NH_HD inline void neu_add(double& s, double& c, double v) {
  const double t = s + v;
  if (nabs(s) >= nabs(v)) c += (s - t) + v;
  else c += (v - t) + s;
  s = t;
}

enum BudgetTerm { B_POT = 0, B_INT, B_KIN, B_DRY, B_VAP, B_NTERMS };

struct BudgetT {
  using LdsT = double;
  static constexpr bool kStage = false;
  static constexpr const char* kName = "budget";
  Dims d; Consts c; Geo G;
  const real* prog;
  const real* rhogq;     // qv
  double lhv;
  long npts;             // interior points of the physical levels
  double* part;          // [nblocks][B_NTERMS][2]

  NH_HD void phase1(int blk, int t, double* lds) const {
    const long n = (long)blk * 256 + t;
    double v[B_NTERMS] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (n < npts) {
      const int W = d.g - 2;
      const int nz = d.kmax - d.kmin + 1;
      const int i = 1 + (int)(n % W);
      const int j = 1 + (int)((n / W) % W);
      const long r = n / ((long)W * W);
      const int k = d.kmin + (int)(r % nz);
      const int l = (int)(r / nz);
      const long ks = (long)d.g * d.g;
      const long a = (((long)l * d.kall + k) * d.g + j) * d.g + i;
      const long col = ((long)l * d.g + j) * d.g + i;
      const long NF = d.NF;
      const ConstsT<double>& C = c.d;
      const double rhog = prog[a];
      const double rho = rhog / 1.0;                  // GSGAM2 = 1
      const double ein = (double)prog[5 * NF + a] / rhog;
      const double q = (double)rhogq[a] / rhog;
      const double qd = 1.0 - q;
      const double cv = q * C.CVW_qv + qd * C.CVdry;
      const double tem = ein / cv;
      const double vol = G.area[col] * 1.0 * G.v64.dgz[k];  // area * GSGAM2 * dgz
      double e = qd * C.CVdry * tem;
      e = e + q * C.CVW_qv * tem;
      e = e + q * lhv;
      const double kin = rhogkin_k(G.v64, d.kmin, d.kmax, k, [&](int kk, int qq) {
        return (double)prog[qq * NF + a + (long)(kk - k) * ks];
      });
      v[B_POT] = rho * G.v64.phi[k] * vol;
      v[B_INT] = rho * e * vol;
      v[B_KIN] = kin * 1.0 * vol;                     // RGSGAM2 = 1
      v[B_DRY] = rho * qd * vol;
      v[B_VAP] = rho * q * vol;
    }
    for (int m = 0; m < B_NTERMS; ++m) lds[m * 256 + t] = v[m];
  }
  NH_HD void phase2(int blk, int t, double* lds) const {
    if (t >= B_NTERMS) return;
    double s = 0.0, cc = 0.0;
    for (int u = 0; u < 256; ++u) neu_add(s, cc, lds[t * 256 + u]);
    part[((long)blk * B_NTERMS + t) * 2] = s;
    part[((long)blk * B_NTERMS + t) * 2 + 1] = cc;
  }
};

// Second stage on the device (the allocation-free budget): one block sums the
// nblocks partials of every term, each thread a strided share compensated, then
// six threads the 256 thread sums; fin[term] = (sum, correction).
struct BudgetSumT {
  using LdsT = double;
  static constexpr bool kStage = false;
  static constexpr const char* kName = "budget_sum";
  long nblocks;
  const double* part;    // [nblocks][B_NTERMS][2]
  double* fin;           // [B_NTERMS][2]
  NH_HD void phase1(int, int t, double* lds) const {
    for (int m = 0; m < B_NTERMS; ++m) {
      double s = 0.0, c = 0.0;
      for (long b = t; b < nblocks; b += 256) {
        neu_add(s, c, part[(b * B_NTERMS + m) * 2]);
        neu_add(s, c, part[(b * B_NTERMS + m) * 2 + 1]);
      }
      lds[(m * 256 + t) * 2] = s;
      lds[(m * 256 + t) * 2 + 1] = c;
    }
  }
  NH_HD void phase2(int, int t, double* lds) const {
    if (t >= B_NTERMS) return;
    double s = 0.0, c = 0.0;
    for (int u = 0; u < 256; ++u) {
      neu_add(s, c, lds[(t * 256 + u) * 2]);
      neu_add(s, c, lds[(t * 256 + u) * 2 + 1]);
    }
    fin[t * 2] = s;
    fin[t * 2 + 1] = c;
  }
};

NH_END
