// Held & Suarez (1994) forcing of the prognostic state (AF_TYPE =
// 'HELD-SUAREZ'): Rayleigh friction of the low-level winds and Newtonian
// relaxation of the temperature toward a radiative-equilibrium profile,
// applied after the dynamics step (Dyn.forcing_step,
// nhm/dynamics/mod_dynamics.py:597-720 -> Frc.forcing_step_hs / hs_apply_core,
// nhm/forcing/mod_forcing.py:12-29, 329-350 -> AfHeldsuarez.AF_heldsuarez,
// nhm/forcing/mod_af_heldsuarez.py:43-95).
//
// Two launches: HsSfcRef stores each regional column's reference pressure
// 0.5 (p(kmin) + p(kmin-1)) from the unforced state, then HeldSuarezP forces
// every regional point of the physical levels in place (the pole points are
// not forced, as in the reference); the caller refills the halos.  The
// diagnostics are those of compute_diagnostics (kernels/diag.py:53) with
// tracer 0 the only water tracer, flat topography (GSGAM2 = 1).
#pragma once

#include "nh_engine.h"

NH_BEGIN

// the forcing's parameters (mod_af_heldsuarez.py:17-27, AF_heldsuarez_init)
struct HsParams {
  double sigma_b, Kf, Dth_z, Ka, Ks, T_eq0, DT_y, PRE00;
};

// dry (Held and Suarez 1994) or moist (Thatcher and Jablonowski 2016) case
inline HsParams hs_params(int moist_case) {
  HsParams h;
  h.sigma_b = 0.7;
  h.Kf = 1.0 / (1.0 * 86400.0);
  h.Dth_z = 10.0;
  h.Ka = 1.0 / (40.0 * 86400.0);
  h.Ks = 1.0 / (4.0 * 86400.0);
  h.T_eq0 = moist_case ? 294.0 : 315.0;
  h.DT_y = moist_case ? 65.0 : 60.0;
  h.PRE00 = 1.0e5;
  return h;
}

template <class T>
NH_HD inline T hs_pre(const ConstsT<T>& c, T rhog, T rhoge, T rhogq) {
  const T rho = rhog / T(1.0);
  const T ein = rhoge / rhog;
  const T q = rhogq / rhog;
  const T qd = T(1.0) - q;
  const T cv = q * c.CVW_qv + qd * c.CVdry;
  const T tem = ein / cv;
  return rho * tem * (qd * c.Rdry + q * c.Rvap);
}

struct HsSfcRef {
  static constexpr const char* kName = "held_suarez_sfc";
  Dims d; Consts c;
  const real* prog;
  const real* rhogq;
  real* psref;          // [nreg]
  NH_HD void operator()(long col) const {
    const long gg = (long)d.g * d.g;
    const long l = col / gg;
    const long a1 = (l * d.kall + d.kmin) * gg + (col - l * gg);
    const long a0 = a1 - gg;
    const long NF = d.NF;
    const real p1 = hs_pre<real>(c, prog[a1], prog[5 * NF + a1], rhogq[a1]);
    const real p0 = hs_pre<real>(c, prog[a0], prog[5 * NF + a0], rhogq[a0]);
    psref[col] = real(0.5) * (p1 + p0);
  }
};

struct HeldSuarezP {
  static constexpr const char* kName = "held_suarez";
  Dims d; Consts c; Geo G;
  HsParams hp;
  real dt;
  real* prog;
  const real* rhogq;
  const real* psref;
  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (!pt_of(d, plane, x, p)) return;
    if (p.pole || p.k < d.kmin || p.k > d.kmax) return;
    const long NF = d.NF, a = p.a;
    const real rhog = prog[a];
    const real rho = rhog / real(1.0);
    const real vx = prog[1 * NF + a] / rhog, vy = prog[2 * NF + a] / rhog, vz = prog[3 * NF + a] / rhog;
    const real ein = prog[5 * NF + a] / rhog;
    const real q = rhogq[a] / rhog;
    const real qd = real(1.0) - q;
    const real cv = q * c.CVW_qv + qd * c.CVdry;
    const real tem = ein / cv;
    const real pre = rho * tem * (qd * c.Rdry + q * c.Rvap);
    const real sigma_b = (real)hp.sigma_b;
    const real sigma = pre / psref[p.ci];
    const real fs = (sigma - sigma_b) / (real(1.0) - sigma_b);
    const real factor = (fs > real(0.0)) ? fs : real(0.0);
    const real Kf = (real)hp.Kf;
    const real fvx = -Kf * factor * vx;
    const real fvy = -Kf * factor * vy;
    const real fvz = -Kf * factor * vz;
    // the equilibrium temperature in FP64 in every form (latitude, log, power)
    const double lat = G.lat[p.ci];
    const double sinlat = fabs(sin(lat)), coslat = fabs(cos(lat));
    const double ap0 = fabs((double)pre / hp.PRE00);
    const double RovCp = c.d.Rdry / (c.d.Rdry + c.d.CVdry);
    const double teq = (hp.T_eq0 - hp.DT_y * (sinlat * sinlat)
                        - hp.Dth_z * log(ap0 > 1e-30 ? ap0 : 1e-30) * (coslat * coslat)) * pow(ap0, RovCp);
    const real T_eq = (real)(teq > 200.0 ? teq : 200.0);
    const real Ka = (real)hp.Ka, Ks = (real)hp.Ks;
    const real Kt = Ka + (Ks - Ka) * factor * (real)pow(coslat, 4.0);
    const real fe = -Kt * (tem - T_eq) * c.CVdry;
    prog[1 * NF + a] = prog[1 * NF + a] + dt * fvx * rho * real(1.0);
    prog[2 * NF + a] = prog[2 * NF + a] + dt * fvy * rho * real(1.0);
    prog[3 * NF + a] = prog[3 * NF + a] + dt * fvz * rho * real(1.0);
    prog[5 * NF + a] = prog[5 * NF + a] + dt * fe * rho * real(1.0);
  }
};

// one forcing step of dt on prog (tracer 0 of rhogq is qv), then the halo
// fill of the six prognostics; the surface reference pressures go to a
// workspace field (the forcing runs between steps, when it is free)
template <class X>
int forcing_hs(X& x, const Context& ctx, real* prog, const real* rhogq, double dt, int moist_case) {
  const Dims& d = ctx.d;
  real* psref = ctx.field(W_DIAG);
  {
    HsSfcRef f{d, ctx.c, prog, rhogq, psref};
    x.run(f, d.nreg);
  }
  {
    HeldSuarezP f;
    f.d = d; f.c = ctx.c; f.G = ctx.G;
    f.hp = hs_params(moist_case);
    f.dt = (real)dt;
    f.prog = prog; f.rhogq = rhogq; f.psref = psref;
    x.run_pts(f, d);
  }
  halo_block(x, ctx, prog, 6);
  return x.check("forcing_hs");
}

NH_END
