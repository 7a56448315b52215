// extern "C" entry points of include/nicamhip.h, generic over the executor
// type NHX defined by the including translation unit:
//   NHX(int device, void* stream, Context* ctx)  sets the device, binds the
//   stream and the context's kernel timers;
//   prof_enable(bool), prof_report(...), prof_release() for the timers.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <new>
#include <string>

#include "nh_engine.h"

struct nh_context {
  nh::Context c;
};

namespace nh {

static thread_local std::string g_err;

inline int fail(int code, const char* where, const std::string& msg) {
  g_err = std::string(where) + ": " + msg;
  return code;
}

inline int check_ready(const nh_context* ctx, const char* where, bool need_halo) {
  if (!ctx) return fail(NH_ERR_ARG, where, "null context");
  if (!ctx->c.geo_ready) return fail(NH_ERR_STATE, where, "geometry not uploaded");
  if (need_halo && ctx->c.halo_n == 0) return fail(NH_ERR_STATE, where, "halo map not uploaded");
  return 0;
}

}  // namespace nh

extern "C" {

const char* nh_last_error(void) { return nh::g_err.c_str(); }

int64_t nh_field_size(const nh_dims* dims) {
  if (!dims) return nh::fail(NH_ERR_ARG, "nh_field_size", "null dims");
  nh::Dims d;
  nh::set_dims(d, dims);
  return d.NF;
}

int nh_create(const nh_dims* dims, const nh_consts* consts, int device, nh_context** out) {
  const char* where = "nh_create";
  if (!dims || !consts || !out) return nh::fail(NH_ERR_ARG, where, "null argument");
  *out = nullptr;
  if (dims->gall_1d < 4 || dims->lall < 1 || dims->kmin != 1 || dims->kmax != dims->kall - 2 ||
      dims->kall > NH_KMAX || dims->kmax - dims->kmin < 2)
    return nh::fail(NH_ERR_ARG, where, "unsupported dimensions (need gall_1d>=4, kmin=1, kmax=kall-2, kall<=80)");
  nh_context* ctx = new (std::nothrow) nh_context();
  if (!ctx) return nh::fail(NH_ERR_ALLOC, where, "host allocation");
  nh::set_dims(ctx->c.d, dims);
  nh::set_consts(ctx->c.c, consts);
  ctx->c.device = device;
  NHX x(device, nullptr, nullptr);
  if (x.check(where)) { delete ctx; return nh::fail(NH_ERR_DEVICE, where, x.message()); }
  size_t bytes = (size_t)nh::W_NFIELDS * ctx->c.d.NF * sizeof(double);
  ctx->c.ws = (double*)x.alloc(bytes);
  if (!ctx->c.ws) {
    delete ctx;
    return nh::fail(NH_ERR_ALLOC, where, "workspace allocation of " + std::to_string(bytes) + " bytes");
  }
  x.zero(ctx->c.ws, bytes);
  if (x.check(where)) {
    x.release(ctx->c.ws);
    delete ctx;
    return nh::fail(NH_ERR_DEVICE, where, x.message());
  }
  *out = ctx;
  return 0;
}

int nh_destroy(nh_context* ctx) {
  if (!ctx) return 0;
  NHX x(ctx->c.device, nullptr, &ctx->c);
  x.sync();
  if (ctx->c.ws) x.release(ctx->c.ws);
  if (ctx->c.geo_buf) x.release(ctx->c.geo_buf);
  if (ctx->c.int_buf) x.release(ctx->c.int_buf);
  if (ctx->c.halo_dst) x.release(ctx->c.halo_dst);
  if (ctx->c.halo_src) x.release(ctx->c.halo_src);
  x.prof_release();
  delete ctx;
  return 0;
}

int nh_upload_geometry(nh_context* ctx, const nh_geometry* geo) {
  const char* where = "nh_upload_geometry";
  if (!ctx || !geo) return nh::fail(NH_ERR_ARG, where, "null argument");
  NHX x(ctx->c.device, nullptr, &ctx->c);
  int rc = nh::upload_geometry(x, ctx->c, geo);
  if (rc == NH_ERR_ARG) return nh::fail(rc, where, "null coefficient pointer");
  if (rc == NH_ERR_ALLOC) return nh::fail(rc, where, "device allocation");
  if (rc) return nh::fail(rc, where, x.message());
  return 0;
}

int nh_upload_halo_map(nh_context* ctx, const int32_t* dst, const int32_t* src, int64_t n) {
  const char* where = "nh_upload_halo_map";
  if (!ctx || n < 0 || (n > 0 && (!dst || !src))) return nh::fail(NH_ERR_ARG, where, "invalid argument");
  NHX x(ctx->c.device, nullptr, &ctx->c);
  int rc = nh::upload_halo(x, ctx->c, dst, src, (long)n);
  if (rc == NH_ERR_ARG) return nh::fail(rc, where, "column id out of range");
  if (rc == NH_ERR_ALLOC) return nh::fail(rc, where, "device allocation");
  if (rc) return nh::fail(rc, where, x.message());
  return 0;
}

int nh_halo_fill(nh_context* ctx, double* const* fields, int32_t nfields, void* stream) {
  const char* where = "nh_halo_fill";
  if (!ctx || nfields < 0 || (nfields > 0 && !fields)) return nh::fail(NH_ERR_ARG, where, "invalid argument");
  if (ctx->c.halo_n == 0) return nh::fail(NH_ERR_STATE, where, "halo map not uploaded");
  for (int q = 0; q < nfields; ++q)
    if (!fields[q]) return nh::fail(NH_ERR_ARG, where, "null field pointer");
  NHX x(ctx->c.device, stream, &ctx->c);
  for (int q0 = 0; q0 < nfields; q0 += 16) {
    int m = (nfields - q0 < 16) ? nfields - q0 : 16;
    nh::halo(x, ctx->c, fields + q0, m);
  }
  if (x.check(where)) return nh::fail(NH_ERR_DEVICE, where, x.message());
  return 0;
}

int nh_laplacian(nh_context* ctx, const double* scl, double* out, void* stream) {
  const char* where = "nh_laplacian";
  if (int rc = nh::check_ready(ctx, where, false)) return rc;
  if (!scl || !out) return nh::fail(NH_ERR_ARG, where, "null field pointer");
  NHX x(ctx->c.device, stream, &ctx->c);
  nh::OpLapP f{ctx->c.d, ctx->c.G, scl, out};
  x.run_pts(f, ctx->c.d);
  if (x.check(where)) return nh::fail(NH_ERR_DEVICE, where, x.message());
  return 0;
}

int nh_divergence(nh_context* ctx, const double* vx, const double* vy, const double* vz, double* out,
                  void* stream) {
  const char* where = "nh_divergence";
  if (int rc = nh::check_ready(ctx, where, false)) return rc;
  if (!vx || !vy || !vz || !out) return nh::fail(NH_ERR_ARG, where, "null field pointer");
  NHX x(ctx->c.device, stream, &ctx->c);
  nh::OpDivP f{ctx->c.d, ctx->c.G, vx, vy, vz, out};
  x.run_pts(f, ctx->c.d);
  if (x.check(where)) return nh::fail(NH_ERR_DEVICE, where, x.message());
  return 0;
}

int nh_gradient(nh_context* ctx, const double* scl, double* gx, double* gy, double* gz, void* stream) {
  const char* where = "nh_gradient";
  if (int rc = nh::check_ready(ctx, where, false)) return rc;
  if (!scl || !gx || !gy || !gz) return nh::fail(NH_ERR_ARG, where, "null field pointer");
  NHX x(ctx->c.device, stream, &ctx->c);
  nh::OpGradP f{ctx->c.d, ctx->c.G, scl, gx, gy, gz};
  x.run_pts(f, ctx->c.d);
  if (x.check(where)) return nh::fail(NH_ERR_DEVICE, where, x.message());
  return 0;
}

int nh_dyn_step(nh_context* ctx, double* prog, double* progq, int32_t nq, double dtl, int32_t sstep_max,
                void* stream) {
  const char* where = "nh_dyn_step";
  if (int rc = nh::check_ready(ctx, where, true)) return rc;
  if (!prog || !progq || nq < 1) return nh::fail(NH_ERR_ARG, where, "need prog and at least one tracer field");
  if (sstep_max < 3 || sstep_max % 6 != 0 || !(dtl > 0.0))
    return nh::fail(NH_ERR_ARG, where, "sstep_max must be a positive multiple of 6 and dtl > 0");
  NHX x(ctx->c.device, stream, &ctx->c);
  int rc = nh::dyn_step(x, ctx->c, prog, progq, nq, dtl, sstep_max);
  if (rc == NH_ERR_ARG) return nh::fail(rc, where, "invalid argument");
  if (rc) return nh::fail(NH_ERR_DEVICE, where, x.message());
  return 0;
}

int nh_tracer_step(nh_context* ctx, const double* prog, double* progq, int32_t nq, double dtl, void* stream) {
  const char* where = "nh_tracer_step";
  if (int rc = nh::check_ready(ctx, where, true)) return rc;
  if (!prog || !progq || nq < 1) return nh::fail(NH_ERR_ARG, where, "need prog and at least one tracer field");
  if (!(dtl > 0.0)) return nh::fail(NH_ERR_ARG, where, "dtl must be > 0");
  NHX x(ctx->c.device, stream, &ctx->c);
  int rc = nh::tracer_step(x, ctx->c, prog, progq, nq, dtl);
  if (rc) return nh::fail(NH_ERR_DEVICE, where, x.message());
  return 0;
}

int nh_trcadv_winds(nh_context* ctx, double* prog, double time, int32_t kind, const double* params,
                    void* stream) {
  const char* where = "nh_trcadv_winds";
  if (int rc = nh::check_ready(ctx, where, false)) return rc;
  if (!prog) return nh::fail(NH_ERR_ARG, where, "null prog");
  if (kind != NH_WIND_DCMIP11 && kind != NH_WIND_DEFORM) return nh::fail(NH_ERR_ARG, where, "unknown wind kind");
  if (kind == NH_WIND_DEFORM && (!params || !(params[0] > 0.0) || !(params[3] > 0.0)))
    return nh::fail(NH_ERR_ARG, where, "deformational flow needs period > 0 and ztop > 0");
  NHX x(ctx->c.device, stream, &ctx->c);
  int rc = nh::trcadv_winds(x, ctx->c, prog, time, kind, params);
  if (rc) return nh::fail(NH_ERR_DEVICE, where, x.message());
  return 0;
}

int nh_budget(nh_context* ctx, const double* prog, const double* progq, int32_t nq, double lhv, double* out,
              void* stream) {
  const char* where = "nh_budget";
  if (int rc = nh::check_ready(ctx, where, false)) return rc;
  if (!prog || !progq || nq < 1 || !out) return nh::fail(NH_ERR_ARG, where, "need prog, qv and an output array");
  NHX x(ctx->c.device, stream, &ctx->c);
  int rc = nh::budget(x, ctx->c, prog, progq, lhv, out);
  if (rc == NH_ERR_ARG) return nh::fail(rc, where, "grid too small for the partial-sum buffer");
  if (rc) return nh::fail(NH_ERR_DEVICE, where, x.message());
  return 0;
}

int nh_prof_enable(nh_context* ctx, int32_t on) {
  if (!ctx) return nh::fail(NH_ERR_ARG, "nh_prof_enable", "null context");
  NHX x(ctx->c.device, nullptr, &ctx->c);
  if (x.prof_enable(on != 0)) return nh::fail(NH_ERR_DEVICE, "nh_prof_enable", x.message());
  return 0;
}

int nh_prof_report(nh_context* ctx, int32_t max_kernels, char* names, double* total_ms, int64_t* launches,
                   int32_t* n_kernels) {
  const char* where = "nh_prof_report";
  if (!ctx || max_kernels < 0 || !n_kernels || (max_kernels > 0 && (!names || !total_ms || !launches)))
    return nh::fail(NH_ERR_ARG, where, "invalid argument");
  NHX x(ctx->c.device, nullptr, &ctx->c);
  int n = x.prof_report(max_kernels, names, total_ms, launches);
  if (n < 0) return nh::fail(NH_ERR_DEVICE, where, x.message());
  *n_kernels = n;
  return 0;
}

}  // extern "C"
