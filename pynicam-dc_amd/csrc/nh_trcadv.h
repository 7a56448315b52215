// Prescribed winds of the offline tracer runs (Forcing.forcing_update,
// nhm/forcing/mod_forcing.py:113-206): the velocity of an analytic flow at
// time t is written into prog as rhogv = v rhog, rhogw = w rhog at every
// point (halo and ghost levels included); the density is left unchanged.
// The flow is evaluated in FP64 and rounded to the field precision on store.
#pragma once

#include "nh_columns.h"

NH_BEGIN

struct TrcadvWindsP {
  static constexpr const char* kName = "trcadv_winds";
  Dims d; Geo G;
  real* prog;
  double time;
  int kind;              // 0: DCMIP 1-1, 1: deformational flow
  double par[4];         // kind 1: period, umax, w0, ztop

  // horizontal (u, v) at full levels and w at half levels of level k
  NH_HD void dcmip11(double lat, double lon, int k, double& u, double& v, double& w) const {
    // Trcadv.test11_velocity, nhm/forcing/mod_af_trcadv.py:23-98
    const double pi = 3.141592653589793238;
    const double a = 6371220.0, Rd = 287.0, grav = 9.80616;
    const double tau = 12.0 * 86400.0;
    const double u0 = 2.0 * pi * a / tau;
    const double k0 = 10.0 * a / tau;
    const double omega0 = 23000.0 * pi / tau;
    const double T0 = 300.0;
    const double H = Rd * T0 / grav;
    const double p0 = 1000.e2;
    const double dlon = 2.0 * pi * time / tau;
    const double lonp = lon - dlon;
    const double bs = 0.2;
    const double ptop = p0 * exp(-12000.0 / H);
    const double bp = bs * ptop;
    {
      const double p = p0 * exp(-G.v64.gz[k] / H);
      const double cl = cos(lat);
      const double ud = (omega0 * a) / bp * cos(lonp) * (cl * cl) * cos(dlon) *
                        (-exp((p - p0) / bp) + exp((ptop - p) / bp));
      const double sl = sin(lonp);
      u = k0 * sin(2.0 * lat) * (sl * sl) * cos(0.5 * dlon) + u0 * cl + ud;
      v = k0 * sin(2.0 * lonp) * cl * cos(0.5 * dlon);
    }
    {
      const double p = p0 * exp(-G.v64.gzh[k] / H);
      const double s = 1.0 + exp((ptop - p0) / bp) - exp((p - p0) / bp) - exp((ptop - p) / bp);
      w = -(Rd * T0) / (grav * p) * omega0 * sin(lonp) * cos(lat) * cos(dlon) * s;
    }
  }

  // non-divergent deformation (Nair and Lauritzen 2010, case 1), reversing at
  // half period, peak speed umax; w = w0 sin(pi z / ztop) cos(2 pi t / T) on
  // the inner half levels
  NH_HD void deform(double lat, double lon, int k, double& u, double& v, double& w) const {
    const double pi = 3.141592653589793238;
    const double T = par[0], umax = par[1], w0 = par[2], ztop = par[3];
    const double ct = cos(pi * time / T);
    const double sh = sin(0.5 * lon);
    u = umax * (sh * sh) * sin(2.0 * lat) * ct;
    v = 0.5 * umax * sin(lon) * cos(lat) * ct;
    w = (k > d.kmin && k <= d.kmax) ? w0 * sin(pi * G.v64.gzh[k] / ztop) * cos(2.0 * pi * time / T) : 0.0;
  }

  NH_HD void operator()(int plane, int x) const {
    Pt p;
    if (!pt_of(d, plane, x, p)) return;
    double lat, lon;
    if (p.pole) {
      lat = G.lat_pl[p.v * 2 + p.lp];
      lon = G.lon_pl[p.v * 2 + p.lp];
    } else {
      lat = G.lat[p.ci];
      lon = G.lon[p.ci];
    }
    double u, v, w;
    if (kind == 0) dcmip11(lat, lon, p.k, u, v, w);
    else deform(lat, lon, p.k, u, v, w);
    // local east and north unit vectors (Sp_Unit_East/North, mod_af_trcadv.py:151-170)
    const double ex = -sin(lon), ey = cos(lon), ez = 0.0;
    const double nx = -sin(lat) * cos(lon), ny = -sin(lat) * sin(lon), nz = cos(lat);
    const long NF = d.NF, a = p.a;
    const double rhog = prog[a];
    prog[1 * NF + a] = (real)((ex * u + nx * v) * rhog);
    prog[2 * NF + a] = (real)((ey * u + ny * v) * rhog);
    prog[3 * NF + a] = (real)((ez * u + nz * v) * rhog);
    prog[4 * NF + a] = (real)(w * rhog);
  }
};

// zero a field (density tendency of the offline tracer step)
struct FillZero {
  static constexpr const char* kName = "fill_zero";
  real* p;
  NH_HD void operator()(long t) const { p[t] = real(0.0); }
};

NH_END
