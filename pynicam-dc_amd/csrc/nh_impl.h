// One precision variant of the library behind the interface of nh_variant.h,
// generic over the executor type NHX defined by the including translation unit:
//   NHX(int device, void* stream, void** prof_slot)  sets the device, binds the
//   stream and the context's kernel-timer state;
//   alloc/release/zero/h2d/d2h/d2d, run/run_pts/run_tiles, check, message,
//   prof_enable/prof_report/prof_release.
#pragma once

#include <new>
#include <string>

#include "nh_variant.h"
#include "nh_engine.h"
#include "nh_adjstep.h"
#include "nh_forcing.h"
#include "nh_ops.h"

NH_BEGIN

namespace {
inline int failv(std::string* err, int code, const std::string& msg) {
  if (err) *err = msg;
  return code;
}
}  // namespace

int v_create(const nh_dims* dims, const nh_consts* consts, int device, Context** out, std::string* err) {
  Context* c = new (std::nothrow) Context();
  if (!c) return failv(err, NH_ERR_ALLOC, "host allocation");
  set_dims(c->d, dims);
  set_consts(c->c, consts);
  c->device = device;
  NHX x(device, nullptr, nullptr);
  if (x.check("nh_create")) { delete c; return failv(err, NH_ERR_DEVICE, x.message()); }
  const size_t bytes = (size_t)W_NFIELDS * c->d.NF * sizeof(real);
  const size_t dbytes = (size_t)D_NFIELDS * c->d.NF * sizeof(double);
  c->ws = (real*)x.alloc(bytes);
  c->dws = (double*)x.alloc(dbytes);
  if (!c->ws || !c->dws) {
    if (c->ws) x.release(c->ws);
    if (c->dws) x.release(c->dws);
    delete c;
    return failv(err, NH_ERR_ALLOC, "workspace allocation of " + std::to_string(bytes + dbytes) + " bytes");
  }
  x.zero(c->ws, bytes);
  x.zero(c->dws, dbytes);
  if (x.check("nh_create")) {
    x.release(c->ws);
    x.release(c->dws);
    delete c;
    return failv(err, NH_ERR_DEVICE, x.message());
  }
  *out = c;
  return 0;
}

void v_destroy(Context* c) {
  if (!c) return;
  NHX x(c->device, nullptr, &c->prof);
  x.sync();
  void* bufs[9] = {c->ws, c->dws, c->geo_buf, c->geo64_buf, c->int_buf, c->halo_dst, c->halo_src, c->aws, c->adws};
  for (void* b : bufs)
    if (b) x.release(b);
  x.prof_release();
  delete c;
}

bool v_geo_ready(const Context* c) { return c->geo_ready; }
long v_halo_n(const Context* c) { return c->halo_n; }

int v_upload_geometry(Context* c, const nh_geometry* g, std::string* err) {
  NHX x(c->device, nullptr, &c->prof);
  int rc = upload_geometry(x, *c, g);
  if (rc == NH_ERR_ARG) return failv(err, rc, "null coefficient pointer");
  if (rc == NH_ERR_ALLOC) return failv(err, rc, "device allocation");
  if (rc) return failv(err, rc, x.message());
  return 0;
}

int v_upload_halo(Context* c, const int32_t* dst, const int32_t* src, long n, std::string* err) {
  NHX x(c->device, nullptr, &c->prof);
  int rc = upload_halo(x, *c, dst, src, n);
  if (rc == NH_ERR_ARG) return failv(err, rc, "column id out of range");
  if (rc == NH_ERR_ALLOC) return failv(err, rc, "device allocation");
  if (rc) return failv(err, rc, x.message());
  return 0;
}

int v_halo_fill(Context* c, void* const* fields, int nf, void* stream, std::string* err) {
  NHX x(c->device, stream, &c->prof);
  real* const* f = reinterpret_cast<real* const*>(fields);
  for (int q0 = 0; q0 < nf; q0 += 16) halo(x, *c, f + q0, (nf - q0 < 16) ? nf - q0 : 16);
  if (x.check("nh_halo_fill")) return failv(err, NH_ERR_DEVICE, x.message());
  return 0;
}

int v_laplacian(Context* c, const void* scl, void* out, void* stream, std::string* err) {
  NHX x(c->device, stream, &c->prof);
  OpLapP f{c->d, c->G, (const real*)scl, (real*)out};
  x.run_pts(f, c->d);
  if (x.check("nh_laplacian")) return failv(err, NH_ERR_DEVICE, x.message());
  return 0;
}

int v_divergence(Context* c, const void* vx, const void* vy, const void* vz, void* out, void* stream,
                 std::string* err) {
  NHX x(c->device, stream, &c->prof);
  OpDivP f{c->d, c->G, (const real*)vx, (const real*)vy, (const real*)vz, (real*)out};
  x.run_pts(f, c->d);
  if (x.check("nh_divergence")) return failv(err, NH_ERR_DEVICE, x.message());
  return 0;
}

int v_gradient(Context* c, const void* scl, void* gx, void* gy, void* gz, void* stream, std::string* err) {
  NHX x(c->device, stream, &c->prof);
  OpGradP f{c->d, c->G, (const real*)scl, (real*)gx, (real*)gy, (real*)gz};
  x.run_pts(f, c->d);
  if (x.check("nh_gradient")) return failv(err, NH_ERR_DEVICE, x.message());
  return 0;
}

// standalone operators (nh_ops.h); in[] and out[] as listed per op in nh_variant.h
int v_op(Context* c, int op, const void* const* in, void* const* out, double dt, void* stream, std::string* err) {
  NHX x(c->device, stream, &c->prof);
  const Dims& d = c->d;
  const long NF = d.NF;
  auto R = [&](int q) { return (const real*)in[q]; };
  switch (op) {
    case OP_VI_MATRIX: {
      ViMatrixOp f{d, c->c, c->G, dt, R(0), R(1), (double*)out[0], (double*)out[1], (double*)out[2]};
      x.run(f, ncols(d));
      break;
    }
    case OP_VI_SOLVE: {
      auto solve = [&](auto f) {
        f.d = d; f.c = c->c; f.G = c->G; f.dt = dt;
        f.rhogw = R(0); f.rhogw0 = R(1); f.preg0 = R(2); f.rhog0 = R(3);
        f.Srho = R(4); f.Sw = R(5); f.Spre = R(6);
        f.Mc = (const double*)in[7]; f.Mu = (const double*)in[8]; f.Ml = (const double*)in[9];
        f.out = (real*)out[0];
        x.run(f, ncols(d));
      };
      if (d.kmax - d.kmin == 39) solve(ViSolveOp<39>{});
      else solve(ViSolveOp<0>{});
      break;
    }
    case OP_PREPOST: {
      PrePostP f{d, c->c, c->G, (real*)out[0], R(1), (real*)out[1], (real*)out[2]};
      x.run_pts(f, d);
      break;
    }
    case OP_HDIFF: {
      // the two passes into the step workspace (cleared, so that columns no
      // halo fill reaches hold zeros), a halo fill after each, then the tendency
      const real* prog = R(0);
      const real* diag = R(1);
      real* hda = c->field(W_HDA);
      real* hdb = c->field(W_HDB);
      x.run(FillZero{hda}, 12 * NF);
      const StageFields sf{hda, hdb, c->field(W_GT), c->field(W_ETHTOT), c->dfield(D_MC)};
      hdiff_passes(x, *c, prog, diag, sf);
      HDiffTendOp f;
      f.d = d; f.c = c->c; f.G = c->G;
      f.rhog = prog;
      for (int q = 0; q < 6; ++q) { f.vt[q] = hdb + q * NF; f.out[q] = (real*)out[q]; }
      x.run_pts(f, d);
      break;
    }
    case OP_DIVDAMP: {
      divdamp(x, *c, R(0), R(1), R(2), R(3), (real*)out[0], (real*)out[1], (real*)out[2]);
      break;
    }
    case OP_LIMITER_V: {
      LimVOp f{d, (real*)out[0], R(0), R(1), R(2), R(3)};
      x.run(f, ncols(d));
      break;
    }
    case OP_LIMITER_H: {
      // q_a is read by the neighbours of each cell: limit from a copy of it
      real* qa_in = c->field(W_HDA);
      real* qo = c->field(W_QOUT);
      x.d2d(qa_in, out[0], 6 * NF * sizeof(real));
      const LimHGeom H{d, c->G.sgp, R(0), R(3)};
      LimHQoutOp a{H, R(1), R(2), {qo, qo + NF}};
      x.run_pts(a, d);
      halo(x, *c, a.qout, 2);
      LimHApplyOp b{H, qa_in, {qo, qo + NF}, (real*)out[0]};
      x.run_pts(b, d);
      break;
    }
    default:
      return failv(err, NH_ERR_ARG, "unknown operator");
  }
  if (x.check("operator")) return failv(err, NH_ERR_DEVICE, x.message());
  return 0;
}

int v_dyn_step(Context* c, void* prog, void* progq, int nq, double dtl, int sstep_max, void* record, void* tape,
               void* stream, std::string* err) {
  NHX x(c->device, stream, &c->prof);
  int rc = dyn_step(x, *c, (real*)prog, (real*)progq, nq, dtl, sstep_max, (real*)record, tape);
  if (rc == NH_ERR_ARG) return failv(err, rc, "invalid argument");
  if (rc) return failv(err, NH_ERR_DEVICE, x.message());
  return 0;
}

int v_tracer_step(Context* c, const void* prog, void* progq, int nq, double dtl, void* stream, std::string* err) {
  NHX x(c->device, stream, &c->prof);
  int rc = tracer_step(x, *c, (const real*)prog, (real*)progq, nq, dtl);
  if (rc) return failv(err, NH_ERR_DEVICE, x.message());
  return 0;
}

int v_trcadv_winds(Context* c, void* prog, double time, int kind, const double* params, void* stream,
                   std::string* err) {
  NHX x(c->device, stream, &c->prof);
  int rc = trcadv_winds(x, *c, (real*)prog, time, kind, params);
  if (rc) return failv(err, NH_ERR_DEVICE, x.message());
  return 0;
}

int v_forcing_hs(Context* c, void* prog, const void* progq, double dt, int moist_case, void* stream,
                 std::string* err) {
  NHX x(c->device, stream, &c->prof);
  int rc = forcing_hs(x, *c, (real*)prog, (const real*)progq, dt, moist_case);
  if (rc) return failv(err, NH_ERR_DEVICE, x.message());
  return 0;
}

int v_budget(Context* c, const void* prog, const void* progq, double lhv, double* out, void* stream,
             std::string* err) {
  NHX x(c->device, stream, &c->prof);
  int rc = budget(x, *c, (const real*)prog, (const real*)progq, lhv, out);
  if (rc == NH_ERR_ARG) return failv(err, rc, "grid too small for the partial-sum buffer");
  if (rc) return failv(err, NH_ERR_DEVICE, x.message());
  return 0;
}

int v_budget_ws(Context* c, const void* prog, const void* progq, double lhv, double* out, double* ws, long ws_n,
                void* stream, std::string* err) {
  NHX x(c->device, stream, &c->prof);
  int rc = budget_ws(x, *c, (const real*)prog, (const real*)progq, lhv, out, ws, ws_n);
  if (rc == NH_ERR_ARG)
    return failv(err, rc, "grid needs more than the " + std::to_string(ws_n) + " FP64 values of the budget scratch");
  if (rc) return failv(err, NH_ERR_DEVICE, x.message());
  return 0;
}

namespace {
// the adjoint workspace with nck small-step checkpoint slots (grown on demand)
template <class X>
int adj_workspace(X& x, Context* c, int nck, std::string* err) {
  if (c->aws && c->adj_ck >= nck) return 0;
  x.sync();
  if (c->aws) x.release(c->aws);
  if (c->adws) x.release(c->adws);
  c->aws = nullptr;
  c->adws = nullptr;
  c->adj_ck = 0;
  const size_t bytes = (size_t)adj_fields(nck) * c->d.NF * sizeof(real);
  const size_t dbytes = (size_t)AD_NFIELDS * c->d.NF * sizeof(double);
  c->aws = (real*)x.alloc(bytes);
  c->adws = (double*)x.alloc(dbytes);
  if (!c->aws || !c->adws)
    return failv(err, NH_ERR_ALLOC, "adjoint workspace allocation of " + std::to_string(bytes + dbytes) + " bytes");
  x.zero(c->aws, bytes);
  x.zero(c->adws, dbytes);
  c->adj_ck = nck;
  return 0;
}
}  // namespace

long v_tape_bytes(const Context* c, int sstep_max) { return (long)TapeLayout::bytes(c->d.NF, sstep_max); }

int v_dyn_step_adj_tape(Context* c, const void* tape, const void* progq, int nq, double dtl, int sstep_max,
                        void* prog_b, void* stream, std::string* err) {
  if (sstep_max < 3 || sstep_max > 64) return failv(err, NH_ERR_ARG, "sstep_max out of range for the adjoint");
  if (!c->halo_adj_ok)
    return failv(err, NH_ERR_STATE, "halo map has a destination listed twice or used as a source");
  if (c->G.bs_on) return failv(err, NH_ERR_STATE, "the adjoint supports the NOBASE basic state only");
  NHX x(c->device, stream, &c->prof);
  if (int rc = adj_workspace(x, c, 0, err)) return rc;
  int rc = dyn_step_adj_tape(x, *c, tape, (const real*)progq, nq, dtl, sstep_max, (real*)prog_b);
  if (rc == NH_ERR_ARG) return failv(err, rc, "invalid argument");
  if (rc) return failv(err, NH_ERR_DEVICE, x.message());
  return 0;
}

int v_dyn_step_adj(Context* c, const void* prog, const void* progq, int nq, double dtl, int sstep_max, void* prog_b,
                   const void* stage_in, void* stream, std::string* err) {
  if (sstep_max < 3 || sstep_max > 64) return failv(err, NH_ERR_ARG, "sstep_max out of range for the adjoint");
  if (!c->halo_adj_ok)
    return failv(err, NH_ERR_STATE, "halo map has a destination listed twice or used as a source");
  if (c->G.bs_on) return failv(err, NH_ERR_STATE, "the adjoint supports the NOBASE basic state only");
  NHX x(c->device, stream, &c->prof);
  if (int rc = adj_workspace(x, c, sstep_max + 1, err)) return rc;
  int rc = dyn_step_adj(x, *c, (const real*)prog, (const real*)progq, nq, dtl, sstep_max, (real*)prog_b,
                        (const real*)stage_in);
  if (rc == NH_ERR_ARG) return failv(err, rc, "invalid argument");
  if (rc == NH_ERR_STATE) return failv(err, rc, "adjoint workspace missing");
  if (rc) return failv(err, NH_ERR_DEVICE, x.message());
  return 0;
}

int v_tracer_step_adj(Context* c, const void* prog, const void* progq, int nq, double dtl, void* progq_b,
                      void* stream, std::string* err) {
  if (!c->halo_adj_ok)
    return failv(err, NH_ERR_STATE, "halo map has a destination listed twice or used as a source");
  NHX x(c->device, stream, &c->prof);
  if (int rc = adj_workspace(x, c, 0, err)) return rc;
  int rc = tracer_step_adj(x, *c, (const real*)prog, (const real*)progq, nq, dtl, (real*)progq_b);
  if (rc == NH_ERR_ARG) return failv(err, rc, "invalid argument");
  if (rc == NH_ERR_STATE) return failv(err, rc, "adjoint workspace missing");
  if (rc) return failv(err, NH_ERR_DEVICE, x.message());
  return 0;
}

int v_sfc_pressure(Context* c, const void* prog, const void* progq, double* ps, const double* ps_b, void* prog_b,
                   void* stream, std::string* err) {
  NHX x(c->device, stream, &c->prof);
  int rc = sfc_pressure(x, *c, (const real*)prog, (const real*)progq, ps, ps_b, (real*)prog_b);
  if (rc == NH_ERR_ARG) return failv(err, rc, "invalid argument (null pointer or fewer than three levels)");
  if (rc) return failv(err, NH_ERR_DEVICE, x.message());
  return 0;
}

int v_prof_enable(Context* c, bool on, std::string* err) {
  NHX x(c->device, nullptr, &c->prof);
  if (x.prof_enable(on)) return failv(err, NH_ERR_DEVICE, x.message());
  return 0;
}

int v_prof_only(Context* c, const char* name, std::string* err) {
  NHX x(c->device, nullptr, &c->prof);
  if (x.prof_only(name)) return failv(err, NH_ERR_DEVICE, x.message());
  return 0;
}

int v_prof_report(Context* c, int max_kernels, char* names, double* total_ms, int64_t* launches,
                  std::string* err) {
  NHX x(c->device, nullptr, &c->prof);
  int n = x.prof_report(max_kernels, names, total_ms, launches);
  if (n < 0) return failv(err, NH_ERR_DEVICE, x.message());
  return n;
}

NH_END
