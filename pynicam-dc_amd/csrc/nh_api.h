// extern "C" entry points of include/nicamhip.h: argument checks, then
// dispatch to the context's precision variant (nh_variant.h).  Included by
// exactly one translation unit of a library build.
#pragma once

#include <new>
#include <string>

#include "nh_variant.h"

struct nh_context {
  int precision;            // NH_PREC_FP64 or NH_PREC_MIXED
  nh::f64::Context* c64;
  nh::f32::Context* c32;
  long NF;                  // values per field
  int device;               // the device the context lives on
};

// values per field of a context whose pointer has been checked
inline long nh_field_size_of(const nh_context* ctx) { return ctx->NF; }

namespace nh {

static thread_local std::string g_err;

inline int fail(int code, const char* where, const std::string& msg) {
  g_err = std::string(where) + ": " + msg;
  return code;
}

// the field precision an entry point was called for: 0 double, 1 float
inline int check_ctx(const nh_context* ctx, const char* where, int want_f32) {
  if (!ctx) return fail(NH_ERR_ARG, where, "null context");
  const int is_f32 = (ctx->precision == NH_PREC_MIXED);
  if (want_f32 >= 0 && want_f32 != is_f32)
    return fail(NH_ERR_STATE, where,
                is_f32 ? "the context holds float fields: call the _f32 entry point"
                       : "the context holds double fields: call the entry point without _f32");
  return 0;
}

inline bool geo_ready(const nh_context* ctx) {
  return ctx->c64 ? f64::v_geo_ready(ctx->c64) : f32::v_geo_ready(ctx->c32);
}

inline long halo_n(const nh_context* ctx) { return ctx->c64 ? f64::v_halo_n(ctx->c64) : f32::v_halo_n(ctx->c32); }

inline int check_ready(const nh_context* ctx, const char* where, int want_f32, bool need_halo) {
  if (int rc = check_ctx(ctx, where, want_f32)) return rc;
  if (!geo_ready(ctx)) return fail(NH_ERR_STATE, where, "geometry not uploaded");
  if (need_halo && halo_n(ctx) == 0) return fail(NH_ERR_STATE, where, "halo map not uploaded");
  return 0;
}

inline int done(int rc, const char* where, const std::string& err) { return rc < 0 ? fail(rc, where, err) : 0; }

// call the variant of ctx with the same arguments
#define NH_DISPATCH(ctx, fn, ...) \
  ((ctx)->c64 ? nh::f64::fn((ctx)->c64, __VA_ARGS__) : nh::f32::fn((ctx)->c32, __VA_ARGS__))

inline int halo_fill(nh_context* ctx, void* const* fields, int32_t nfields, void* stream, int f32) {
  const char* where = f32 ? "nh_halo_fill_f32" : "nh_halo_fill";
  if (int rc = check_ctx(ctx, where, f32)) return rc;
  if (nfields < 0 || (nfields > 0 && !fields)) return fail(NH_ERR_ARG, where, "invalid argument");
  if (halo_n(ctx) == 0) return fail(NH_ERR_STATE, where, "halo map not uploaded");
  for (int q = 0; q < nfields; ++q)
    if (!fields[q]) return fail(NH_ERR_ARG, where, "null field pointer");
  std::string err;
  return done(NH_DISPATCH(ctx, v_halo_fill, fields, nfields, stream, &err), where, err);
}

inline int laplacian(nh_context* ctx, const void* scl, void* out, void* stream, int f32) {
  const char* where = f32 ? "nh_laplacian_f32" : "nh_laplacian";
  if (int rc = check_ready(ctx, where, f32, false)) return rc;
  if (!scl || !out) return fail(NH_ERR_ARG, where, "null field pointer");
  std::string err;
  return done(NH_DISPATCH(ctx, v_laplacian, scl, out, stream, &err), where, err);
}

inline int divergence(nh_context* ctx, const void* vx, const void* vy, const void* vz, void* out, void* stream,
                      int f32) {
  const char* where = f32 ? "nh_divergence_f32" : "nh_divergence";
  if (int rc = check_ready(ctx, where, f32, false)) return rc;
  if (!vx || !vy || !vz || !out) return fail(NH_ERR_ARG, where, "null field pointer");
  std::string err;
  return done(NH_DISPATCH(ctx, v_divergence, vx, vy, vz, out, stream, &err), where, err);
}

inline int gradient(nh_context* ctx, const void* scl, void* gx, void* gy, void* gz, void* stream, int f32) {
  const char* where = f32 ? "nh_gradient_f32" : "nh_gradient";
  if (int rc = check_ready(ctx, where, f32, false)) return rc;
  if (!scl || !gx || !gy || !gz) return fail(NH_ERR_ARG, where, "null field pointer");
  std::string err;
  return done(NH_DISPATCH(ctx, v_gradient, scl, gx, gy, gz, stream, &err), where, err);
}

inline int dyn_step(nh_context* ctx, void* prog, void* progq, int32_t nq, double dtl, int32_t sstep_max,
                    void* stream, int f32, void* record = nullptr, const char* name = nullptr,
                    void* tape = nullptr) {
  const char* where = name ? name : (f32 ? "nh_dyn_step_f32" : "nh_dyn_step");
  if (int rc = check_ready(ctx, where, f32, true)) return rc;
  if (!prog || !progq || nq < 1) return fail(NH_ERR_ARG, where, "need prog and at least one tracer field");
  if (sstep_max < 3 || sstep_max % 6 != 0 || !(dtl > 0.0))
    return fail(NH_ERR_ARG, where, "sstep_max must be a positive multiple of 6 and dtl > 0");
  std::string err;
  return done(NH_DISPATCH(ctx, v_dyn_step, prog, progq, nq, dtl, sstep_max, record, tape, stream, &err), where, err);
}

inline int tracer_step(nh_context* ctx, const void* prog, void* progq, int32_t nq, double dtl, void* stream,
                       int f32) {
  const char* where = f32 ? "nh_tracer_step_f32" : "nh_tracer_step";
  if (int rc = check_ready(ctx, where, f32, true)) return rc;
  if (!prog || !progq || nq < 1) return fail(NH_ERR_ARG, where, "need prog and at least one tracer field");
  if (!(dtl > 0.0)) return fail(NH_ERR_ARG, where, "dtl must be > 0");
  std::string err;
  return done(NH_DISPATCH(ctx, v_tracer_step, prog, progq, nq, dtl, stream, &err), where, err);
}

inline int tracer_step_adj(nh_context* ctx, const void* prog, const void* progq, int32_t nq, double dtl,
                           void* progq_b, void* stream, int f32) {
  const char* where = f32 ? "nh_tracer_step_adj_f32" : "nh_tracer_step_adj";
  if (int rc = check_ready(ctx, where, f32, true)) return rc;
  if (!prog || !progq || nq < 1 || !progq_b)
    return fail(NH_ERR_ARG, where, "need prog, progq_b and at least one tracer field");
  if (!(dtl > 0.0)) return fail(NH_ERR_ARG, where, "dtl must be > 0");
  std::string err;
  return done(NH_DISPATCH(ctx, v_tracer_step_adj, prog, progq, nq, dtl, progq_b, stream, &err), where, err);
}

inline int trcadv_winds(nh_context* ctx, void* prog, double time, int32_t kind, const double* params, void* stream,
                        int f32) {
  const char* where = f32 ? "nh_trcadv_winds_f32" : "nh_trcadv_winds";
  if (int rc = check_ready(ctx, where, f32, false)) return rc;
  if (!prog) return fail(NH_ERR_ARG, where, "null prog");
  if (kind != NH_WIND_DCMIP11 && kind != NH_WIND_DEFORM) return fail(NH_ERR_ARG, where, "unknown wind kind");
  if (kind == NH_WIND_DEFORM && (!params || !(params[0] > 0.0) || !(params[3] > 0.0)))
    return fail(NH_ERR_ARG, where, "deformational flow needs period > 0 and ztop > 0");
  std::string err;
  return done(NH_DISPATCH(ctx, v_trcadv_winds, prog, time, kind, params, stream, &err), where, err);
}

inline int budget(nh_context* ctx, const void* prog, const void* progq, int32_t nq, double lhv, double* out,
                  void* stream, int f32) {
  const char* where = f32 ? "nh_budget_f32" : "nh_budget";
  if (int rc = check_ready(ctx, where, f32, false)) return rc;
  if (!prog || !progq || nq < 1 || !out) return fail(NH_ERR_ARG, where, "need prog, qv and an output array");
  std::string err;
  return done(NH_DISPATCH(ctx, v_budget, prog, progq, lhv, out, stream, &err), where, err);
}

inline int forcing_hs(nh_context* ctx, void* prog, const void* progq, double dt, int32_t moist_case, void* stream,
                      int f32) {
  const char* where = f32 ? "nh_forcing_hs_f32" : "nh_forcing_hs";
  if (int rc = check_ready(ctx, where, f32, true)) return rc;
  if (!prog || !progq) return fail(NH_ERR_ARG, where, "need prog and the qv field");
  if (!(dt > 0.0) || (moist_case != 0 && moist_case != 1))
    return fail(NH_ERR_ARG, where, "need dt > 0 and moist_case 0 or 1");
  std::string err;
  return done(NH_DISPATCH(ctx, v_forcing_hs, prog, progq, dt, moist_case, stream, &err), where, err);
}

inline int dyn_step_adj(nh_context* ctx, const void* prog, const void* progq, int32_t nq, double dtl,
                        int32_t sstep_max, void* prog_b, void* stream, int f32, const void* stage_in = nullptr,
                        const char* name = nullptr) {
  const char* where = name ? name : (f32 ? "nh_dyn_step_adj_f32" : "nh_dyn_step_adj");
  if (int rc = check_ready(ctx, where, f32, true)) return rc;
  if (!prog || !progq || nq < 1 || !prog_b)
    return fail(NH_ERR_ARG, where, "need prog, prog_b and at least one tracer field");
  if (sstep_max < 3 || sstep_max % 6 != 0 || !(dtl > 0.0))
    return fail(NH_ERR_ARG, where, "sstep_max must be a positive multiple of 6 and dtl > 0");
  std::string err;
  return done(NH_DISPATCH(ctx, v_dyn_step_adj, prog, progq, nq, dtl, sstep_max, prog_b, stage_in, stream, &err),
              where, err);
}

inline int dyn_step_adj_tape(nh_context* ctx, const void* tape, const void* progq, int32_t nq, double dtl,
                             int32_t sstep_max, void* prog_b, void* stream, int f32) {
  const char* where = f32 ? "nh_dyn_step_adj_tape_f32" : "nh_dyn_step_adj_tape";
  if (int rc = check_ready(ctx, where, f32, true)) return rc;
  if (!tape || !progq || nq < 1 || !prog_b)
    return fail(NH_ERR_ARG, where, "need the tape, prog_b and at least one tracer field");
  if (sstep_max < 3 || sstep_max % 6 != 0 || !(dtl > 0.0))
    return fail(NH_ERR_ARG, where, "sstep_max must be a positive multiple of 6 and dtl > 0");
  std::string err;
  return done(NH_DISPATCH(ctx, v_dyn_step_adj_tape, tape, progq, nq, dtl, sstep_max, prog_b, stream, &err), where,
              err);
}

inline int sfc_pressure(nh_context* ctx, const void* prog, const void* progq, double* ps, const double* ps_b,
                        void* prog_b, void* stream, int f32, const char* where) {
  if (int rc = check_ready(ctx, where, f32, false)) return rc;
  if (!prog || !progq || (ps_b ? !prog_b : !ps)) return fail(NH_ERR_ARG, where, "null pointer");
  std::string err;
  return done(NH_DISPATCH(ctx, v_sfc_pressure, prog, progq, ps, ps_b, prog_b, stream, &err), where, err);
}

inline int op(nh_context* ctx, int which, const void* const* in, int nin, void* const* out, int nout, double dt,
              void* stream, int f32, const char* name, bool need_halo) {
  std::string where = std::string(name) + (f32 ? "_f32" : "");
  if (int rc = check_ready(ctx, where.c_str(), f32, need_halo)) return rc;
  for (int q = 0; q < nin; ++q)
    if (!in[q]) return fail(NH_ERR_ARG, where.c_str(), "null field pointer");
  for (int q = 0; q < nout; ++q)
    if (!out[q]) return fail(NH_ERR_ARG, where.c_str(), "null field pointer");
  if ((which == OP_VI_MATRIX || which == OP_VI_SOLVE) && !(dt > 0.0))
    return fail(NH_ERR_ARG, where.c_str(), "dt must be > 0");
  std::string err;
  return done(NH_DISPATCH(ctx, v_op, which, in, out, dt, stream, &err), where.c_str(), err);
}

}  // namespace nh

extern "C" {

const char* nh_last_error(void) { return nh::g_err.c_str(); }

int64_t nh_field_size(const nh_dims* dims) {
  if (!dims) return nh::fail(NH_ERR_ARG, "nh_field_size", "null dims");
  const long nreg = (long)dims->lall * dims->gall_1d * dims->gall_1d;
  return nreg * dims->kall + 2L * dims->kall * 6;
}

int nh_create_ex(const nh_dims* dims, const nh_consts* consts, int device, int32_t precision, nh_context** out) {
  const char* where = "nh_create";
  if (!dims || !consts || !out) return nh::fail(NH_ERR_ARG, where, "null argument");
  *out = nullptr;
  if (precision != NH_PREC_FP64 && precision != NH_PREC_MIXED)
    return nh::fail(NH_ERR_ARG, where, "precision must be NH_PREC_FP64 or NH_PREC_MIXED");
  if (dims->gall_1d < 4 || dims->lall < 1 || dims->kmin != 1 || dims->kmax != dims->kall - 2 ||
      dims->kall > NH_KMAX || dims->kmax - dims->kmin < 2)
    return nh::fail(NH_ERR_ARG, where, "unsupported dimensions (need gall_1d>=4, kmin=1, kmax=kall-2, kall<=80)");
  // a region is 2^(glevel - rlevel) points wide; the column tiles (st_geom, col_geom:
  // sides min(W, 64) and min(256 / TX, W)) then cover it exactly
  const int W = dims->gall_1d - 2;
  if ((W & (W - 1)) != 0)
    return nh::fail(NH_ERR_ARG, where, "unsupported dimensions (gall_1d - 2 must be a power of two)");
  const long nreg = (long)dims->lall * dims->gall_1d * dims->gall_1d;
  nh_context* ctx = new (std::nothrow) nh_context{precision, nullptr, nullptr, nreg * dims->kall + 2L * dims->kall * 6, device};
  if (!ctx) return nh::fail(NH_ERR_ALLOC, where, "host allocation");
  std::string err;
  const int rc = (precision == NH_PREC_FP64) ? nh::f64::v_create(dims, consts, device, &ctx->c64, &err)
                                             : nh::f32::v_create(dims, consts, device, &ctx->c32, &err);
  if (rc) {
    delete ctx;
    return nh::fail(rc, where, err);
  }
  ctx->device = device;
  // set-up of the allocation-free entry point: the scratch array's address on
  // this device (looked up once per device; a failure is reported by that entry)
  double* sp = nullptr;
  long sn = 0;
  (void)nh::scratch_lookup(device, &sp, &sn, nullptr);
  *out = ctx;
  return 0;
}

int nh_create(const nh_dims* dims, const nh_consts* consts, int device, nh_context** out) {
  return nh_create_ex(dims, consts, device, NH_PREC_FP64, out);
}

int32_t nh_precision(const nh_context* ctx) {
  if (!ctx) return nh::fail(NH_ERR_ARG, "nh_precision", "null context");
  return ctx->precision;
}

int nh_destroy(nh_context* ctx) {
  if (!ctx) return 0;
  if (ctx->c64) nh::f64::v_destroy(ctx->c64);
  if (ctx->c32) nh::f32::v_destroy(ctx->c32);
  delete ctx;
  return 0;
}

int nh_upload_geometry(nh_context* ctx, const nh_geometry* geo) {
  const char* where = "nh_upload_geometry";
  if (!ctx || !geo) return nh::fail(NH_ERR_ARG, where, "null argument");
  std::string err;
  return nh::done(NH_DISPATCH(ctx, v_upload_geometry, geo, &err), where, err);
}

int nh_upload_halo_map(nh_context* ctx, const int32_t* dst, const int32_t* src, int64_t n) {
  const char* where = "nh_upload_halo_map";
  if (!ctx || n < 0 || (n > 0 && (!dst || !src))) return nh::fail(NH_ERR_ARG, where, "invalid argument");
  std::string err;
  return nh::done(NH_DISPATCH(ctx, v_upload_halo, dst, src, (long)n, &err), where, err);
}

int nh_halo_fill(nh_context* ctx, double* const* fields, int32_t nfields, void* stream) {
  return nh::halo_fill(ctx, (void* const*)fields, nfields, stream, 0);
}
int nh_halo_fill_f32(nh_context* ctx, float* const* fields, int32_t nfields, void* stream) {
  return nh::halo_fill(ctx, (void* const*)fields, nfields, stream, 1);
}

int nh_laplacian(nh_context* ctx, const double* scl, double* out, void* stream) {
  return nh::laplacian(ctx, scl, out, stream, 0);
}
int nh_laplacian_f32(nh_context* ctx, const float* scl, float* out, void* stream) {
  return nh::laplacian(ctx, scl, out, stream, 1);
}

int nh_divergence(nh_context* ctx, const double* vx, const double* vy, const double* vz, double* out,
                  void* stream) {
  return nh::divergence(ctx, vx, vy, vz, out, stream, 0);
}
int nh_divergence_f32(nh_context* ctx, const float* vx, const float* vy, const float* vz, float* out,
                      void* stream) {
  return nh::divergence(ctx, vx, vy, vz, out, stream, 1);
}

int nh_gradient(nh_context* ctx, const double* scl, double* gx, double* gy, double* gz, void* stream) {
  return nh::gradient(ctx, scl, gx, gy, gz, stream, 0);
}
int nh_gradient_f32(nh_context* ctx, const float* scl, float* gx, float* gy, float* gz, void* stream) {
  return nh::gradient(ctx, scl, gx, gy, gz, stream, 1);
}

int nh_dyn_step(nh_context* ctx, double* prog, double* progq, int32_t nq, double dtl, int32_t sstep_max,
                void* stream) {
  return nh::dyn_step(ctx, prog, progq, nq, dtl, sstep_max, stream, 0);
}
int nh_dyn_step_f32(nh_context* ctx, float* prog, float* progq, int32_t nq, double dtl, int32_t sstep_max,
                    void* stream) {
  return nh::dyn_step(ctx, prog, progq, nq, dtl, sstep_max, stream, 1);
}

int nh_tracer_step(nh_context* ctx, const double* prog, double* progq, int32_t nq, double dtl, void* stream) {
  return nh::tracer_step(ctx, prog, progq, nq, dtl, stream, 0);
}
int nh_tracer_step_f32(nh_context* ctx, const float* prog, float* progq, int32_t nq, double dtl, void* stream) {
  return nh::tracer_step(ctx, prog, progq, nq, dtl, stream, 1);
}

int nh_trcadv_winds(nh_context* ctx, double* prog, double time, int32_t kind, const double* params,
                    void* stream) {
  return nh::trcadv_winds(ctx, prog, time, kind, params, stream, 0);
}
int nh_trcadv_winds_f32(nh_context* ctx, float* prog, double time, int32_t kind, const double* params,
                        void* stream) {
  return nh::trcadv_winds(ctx, prog, time, kind, params, stream, 1);
}

int nh_budget(nh_context* ctx, const double* prog, const double* progq, int32_t nq, double lhv, double* out,
              void* stream) {
  return nh::budget(ctx, prog, progq, nq, lhv, out, stream, 0);
}
int nh_budget_f32(nh_context* ctx, const float* prog, const float* progq, int32_t nq, double lhv, double* out,
                  void* stream) {
  return nh::budget(ctx, prog, progq, nq, lhv, out, stream, 1);
}

int nh_forcing_hs(nh_context* ctx, double* prog, const double* progq, double dt, int32_t moist_case,
                  void* stream) {
  return nh::forcing_hs(ctx, prog, progq, dt, moist_case, stream, 0);
}
int nh_forcing_hs_f32(nh_context* ctx, float* prog, const float* progq, double dt, int32_t moist_case,
                      void* stream) {
  return nh::forcing_hs(ctx, prog, progq, dt, moist_case, stream, 1);
}

int nh_prof_only(nh_context* ctx, const char* kernel) {
  if (int rc = nh::check_ctx(ctx, "nh_prof_only", -1)) return rc;
  std::string err;
  return nh::done(NH_DISPATCH(ctx, v_prof_only, kernel, &err), "nh_prof_only", err);
}

int nh_prof_enable(nh_context* ctx, int32_t on) {
  if (!ctx) return nh::fail(NH_ERR_ARG, "nh_prof_enable", "null context");
  std::string err;
  return nh::done(NH_DISPATCH(ctx, v_prof_enable, on != 0, &err), "nh_prof_enable", err);
}

int nh_prof_report(nh_context* ctx, int32_t max_kernels, char* names, double* total_ms, int64_t* launches,
                   int32_t* n_kernels) {
  const char* where = "nh_prof_report";
  if (!ctx || max_kernels < 0 || !n_kernels || (max_kernels > 0 && (!names || !total_ms || !launches)))
    return nh::fail(NH_ERR_ARG, where, "invalid argument");
  std::string err;
  const int n = NH_DISPATCH(ctx, v_prof_report, max_kernels, names, total_ms, launches, &err);
  if (n < 0) return nh::fail(n, where, err);
  *n_kernels = n;
  return 0;
}

int nh_tracer_step_adj(nh_context* ctx, const double* prog, const double* progq, int32_t nq, double dtl,
                       double* progq_b, void* stream) {
  return nh::tracer_step_adj(ctx, prog, progq, nq, dtl, progq_b, stream, 0);
}
int nh_tracer_step_adj_f32(nh_context* ctx, const float* prog, const float* progq, int32_t nq, double dtl,
                           float* progq_b, void* stream) {
  return nh::tracer_step_adj(ctx, prog, progq, nq, dtl, progq_b, stream, 1);
}

int nh_dyn_step_adj(nh_context* ctx, const double* prog, const double* progq, int32_t nq, double dtl,
                    int32_t sstep_max, double* prog_b, void* stream) {
  return nh::dyn_step_adj(ctx, prog, progq, nq, dtl, sstep_max, prog_b, stream, 0);
}
int nh_dyn_step_adj_f32(nh_context* ctx, const float* prog, const float* progq, int32_t nq, double dtl,
                        int32_t sstep_max, float* prog_b, void* stream) {
  return nh::dyn_step_adj(ctx, prog, progq, nq, dtl, sstep_max, prog_b, stream, 1);
}

int nh_dyn_step_record(nh_context* ctx, double* prog, double* progq, int32_t nq, double dtl, int32_t sstep_max,
                       double* stage_in, void* stream) {
  if (!stage_in) return nh::fail(NH_ERR_ARG, "nh_dyn_step_record", "null stage_in");
  return nh::dyn_step(ctx, prog, progq, nq, dtl, sstep_max, stream, 0, stage_in, "nh_dyn_step_record");
}
int nh_dyn_step_record_f32(nh_context* ctx, float* prog, float* progq, int32_t nq, double dtl, int32_t sstep_max,
                           float* stage_in, void* stream) {
  if (!stage_in) return nh::fail(NH_ERR_ARG, "nh_dyn_step_record_f32", "null stage_in");
  return nh::dyn_step(ctx, prog, progq, nq, dtl, sstep_max, stream, 1, stage_in, "nh_dyn_step_record_f32");
}
int nh_dyn_step_adj_rec(nh_context* ctx, const double* stage_in, const double* progq, int32_t nq, double dtl,
                        int32_t sstep_max, double* prog_b, void* stream) {
  return nh::dyn_step_adj(ctx, stage_in, progq, nq, dtl, sstep_max, prog_b, stream, 0, stage_in,
                          "nh_dyn_step_adj_rec");
}
int nh_dyn_step_adj_rec_f32(nh_context* ctx, const float* stage_in, const float* progq, int32_t nq, double dtl,
                            int32_t sstep_max, float* prog_b, void* stream) {
  return nh::dyn_step_adj(ctx, stage_in, progq, nq, dtl, sstep_max, prog_b, stream, 1, stage_in,
                          "nh_dyn_step_adj_rec_f32");
}

int64_t nh_tape_bytes(const nh_context* ctx, int32_t sstep_max) {
  if (int rc = nh::check_ctx(ctx, "nh_tape_bytes", -1)) return rc;
  if (sstep_max < 3 || sstep_max % 6 != 0) return nh::fail(NH_ERR_ARG, "nh_tape_bytes", "sstep_max must be a positive multiple of 6");
  return ctx->c64 ? nh::f64::v_tape_bytes(ctx->c64, sstep_max) : nh::f32::v_tape_bytes(ctx->c32, sstep_max);
}
int nh_dyn_step_tape(nh_context* ctx, double* prog, double* progq, int32_t nq, double dtl, int32_t sstep_max,
                     void* tape, void* stream) {
  if (!tape) return nh::fail(NH_ERR_ARG, "nh_dyn_step_tape", "null tape");
  return nh::dyn_step(ctx, prog, progq, nq, dtl, sstep_max, stream, 0, nullptr, "nh_dyn_step_tape", tape);
}
int nh_dyn_step_tape_f32(nh_context* ctx, float* prog, float* progq, int32_t nq, double dtl, int32_t sstep_max,
                         void* tape, void* stream) {
  if (!tape) return nh::fail(NH_ERR_ARG, "nh_dyn_step_tape_f32", "null tape");
  return nh::dyn_step(ctx, prog, progq, nq, dtl, sstep_max, stream, 1, nullptr, "nh_dyn_step_tape_f32", tape);
}
int nh_dyn_step_adj_tape(nh_context* ctx, const void* tape, const double* progq, int32_t nq, double dtl,
                         int32_t sstep_max, double* prog_b, void* stream) {
  return nh::dyn_step_adj_tape(ctx, tape, progq, nq, dtl, sstep_max, prog_b, stream, 0);
}
int nh_dyn_step_adj_tape_f32(nh_context* ctx, const void* tape, const float* progq, int32_t nq, double dtl,
                             int32_t sstep_max, float* prog_b, void* stream) {
  return nh::dyn_step_adj_tape(ctx, tape, progq, nq, dtl, sstep_max, prog_b, stream, 1);
}

int nh_sfc_pressure(nh_context* ctx, const double* prog, const double* progq, double* ps, void* stream) {
  return nh::sfc_pressure(ctx, prog, progq, ps, nullptr, nullptr, stream, 0, "nh_sfc_pressure");
}
int nh_sfc_pressure_f32(nh_context* ctx, const float* prog, const float* progq, double* ps, void* stream) {
  return nh::sfc_pressure(ctx, prog, progq, ps, nullptr, nullptr, stream, 1, "nh_sfc_pressure_f32");
}
int nh_sfc_pressure_adj(nh_context* ctx, const double* prog, const double* progq, const double* ps_b,
                        double* prog_b, void* stream) {
  return nh::sfc_pressure(ctx, prog, progq, nullptr, ps_b, prog_b, stream, 0, "nh_sfc_pressure_adj");
}
int nh_sfc_pressure_adj_f32(nh_context* ctx, const float* prog, const float* progq, const double* ps_b,
                            float* prog_b, void* stream) {
  return nh::sfc_pressure(ctx, prog, progq, nullptr, ps_b, prog_b, stream, 1, "nh_sfc_pressure_adj_f32");
}

int nh_budget_noalloc(nh_context* ctx, const void* prog, const void* progq, int32_t nq, double lhv, double* out,
                      void* stream) {
  const char* where = "nh_budget_noalloc";
  if (int rc = nh::check_ready(ctx, where, -1, false)) return rc;
  if (!prog || !progq || nq < 1 || !out) return nh::fail(NH_ERR_ARG, where, "need prog, qv and an output array");
  std::string err;
  double* ws = nullptr;
  long ws_n = 0;
  if (int rc = nh::scratch_lookup(ctx->device, &ws, &ws_n, &err)) return nh::fail(rc, where, err);
  return nh::done(NH_DISPATCH(ctx, v_budget_ws, prog, progq, lhv, out, ws, ws_n, stream, &err), where, err);
}

/* standalone operators of the reference's kernel API */
int nh_vi_matrix(nh_context* ctx, const double* eth_h, const double* g_tilde, double dt, double* Mc, double* Mu,
                 double* Ml, void* stream) {
  const void* in[2] = {eth_h, g_tilde};
  void* out[3] = {Mc, Mu, Ml};
  return nh::op(ctx, OP_VI_MATRIX, in, 2, out, 3, dt, stream, 0, "nh_vi_matrix", false);
}
int nh_vi_matrix_f32(nh_context* ctx, const float* eth_h, const float* g_tilde, double dt, double* Mc, double* Mu,
                     double* Ml, void* stream) {
  const void* in[2] = {eth_h, g_tilde};
  void* out[3] = {Mc, Mu, Ml};
  return nh::op(ctx, OP_VI_MATRIX, in, 2, out, 3, dt, stream, 1, "nh_vi_matrix", false);
}

int nh_vi_solve(nh_context* ctx, const double* rhogw, const double* rhogw0, const double* preg0, const double* rhog0,
                const double* Srho, const double* Sw, const double* Spre, const double* Mc, const double* Mu,
                const double* Ml, double dt, double* out, void* stream) {
  const void* in[10] = {rhogw, rhogw0, preg0, rhog0, Srho, Sw, Spre, Mc, Mu, Ml};
  void* o[1] = {out};
  return nh::op(ctx, OP_VI_SOLVE, in, 10, o, 1, dt, stream, 0, "nh_vi_solve", false);
}
int nh_vi_solve_f32(nh_context* ctx, const float* rhogw, const float* rhogw0, const float* preg0, const float* rhog0,
                    const float* Srho, const float* Sw, const float* Spre, const double* Mc, const double* Mu,
                    const double* Ml, double dt, float* out, void* stream) {
  const void* in[10] = {rhogw, rhogw0, preg0, rhog0, Srho, Sw, Spre, Mc, Mu, Ml};
  void* o[1] = {out};
  return nh::op(ctx, OP_VI_SOLVE, in, 10, o, 1, dt, stream, 1, "nh_vi_solve", false);
}

int nh_prepost(nh_context* ctx, double* prog, const double* progq, double* diag, double* eth, void* stream) {
  const void* in[2] = {prog, progq};
  void* out[3] = {prog, diag, eth};
  return nh::op(ctx, OP_PREPOST, in, 2, out, 3, 0.0, stream, 0, "nh_prepost", false);
}
int nh_prepost_f32(nh_context* ctx, float* prog, const float* progq, float* diag, float* eth, void* stream) {
  const void* in[2] = {prog, progq};
  void* out[3] = {prog, diag, eth};
  return nh::op(ctx, OP_PREPOST, in, 2, out, 3, 0.0, stream, 1, "nh_prepost", false);
}

int nh_hdiffusion(nh_context* ctx, const double* prog, const double* diag, double* tend, void* stream) {
  if (int rc = nh::check_ctx(ctx, "nh_hdiffusion", 0)) return rc;
  const long NF = nh_field_size_of(ctx);
  const void* in[2] = {prog, diag};
  void* out[6];
  for (int q = 0; q < 6; ++q) out[q] = tend ? tend + q * NF : nullptr;
  return nh::op(ctx, OP_HDIFF, in, 2, out, 6, 0.0, stream, 0, "nh_hdiffusion", true);
}
int nh_hdiffusion_f32(nh_context* ctx, const float* prog, const float* diag, float* tend, void* stream) {
  if (int rc = nh::check_ctx(ctx, "nh_hdiffusion_f32", 1)) return rc;
  const long NF = nh_field_size_of(ctx);
  const void* in[2] = {prog, diag};
  void* out[6];
  for (int q = 0; q < 6; ++q) out[q] = tend ? tend + q * NF : nullptr;
  return nh::op(ctx, OP_HDIFF, in, 2, out, 6, 0.0, stream, 1, "nh_hdiffusion", true);
}

int nh_divdamp3d(nh_context* ctx, const double* rhogvx, const double* rhogvy, const double* rhogvz,
                 const double* rhogw, double* gdx, double* gdy, double* gdz, void* stream) {
  const void* in[4] = {rhogvx, rhogvy, rhogvz, rhogw};
  void* out[3] = {gdx, gdy, gdz};
  return nh::op(ctx, OP_DIVDAMP, in, 4, out, 3, 0.0, stream, 0, "nh_divdamp3d", true);
}
int nh_divdamp3d_f32(nh_context* ctx, const float* rhogvx, const float* rhogvy, const float* rhogvz,
                     const float* rhogw, float* gdx, float* gdy, float* gdz, void* stream) {
  const void* in[4] = {rhogvx, rhogvy, rhogvz, rhogw};
  void* out[3] = {gdx, gdy, gdz};
  return nh::op(ctx, OP_DIVDAMP, in, 4, out, 3, 0.0, stream, 1, "nh_divdamp3d", true);
}

int nh_limiter_v(nh_context* ctx, double* q_h, const double* q, const double* d, const double* ck, void* stream) {
  if (int rc = nh::check_ctx(ctx, "nh_limiter_v", 0)) return rc;
  const long NF = nh_field_size_of(ctx);
  const void* in[4] = {q, d, ck, ck ? ck + NF : nullptr};
  void* out[1] = {q_h};
  return nh::op(ctx, OP_LIMITER_V, in, 4, out, 1, 0.0, stream, 0, "nh_limiter_v", false);
}
int nh_limiter_v_f32(nh_context* ctx, float* q_h, const float* q, const float* d, const float* ck, void* stream) {
  if (int rc = nh::check_ctx(ctx, "nh_limiter_v_f32", 1)) return rc;
  const long NF = nh_field_size_of(ctx);
  const void* in[4] = {q, d, ck, ck ? ck + NF : nullptr};
  void* out[1] = {q_h};
  return nh::op(ctx, OP_LIMITER_V, in, 4, out, 1, 0.0, stream, 1, "nh_limiter_v", false);
}

int nh_limiter_h(nh_context* ctx, double* q_a, const double* q, const double* d, const double* ch,
                 const double* cmask, void* stream) {
  const void* in[4] = {q, d, ch, cmask};
  void* out[1] = {q_a};
  return nh::op(ctx, OP_LIMITER_H, in, 4, out, 1, 0.0, stream, 0, "nh_limiter_h", true);
}
int nh_limiter_h_f32(nh_context* ctx, float* q_a, const float* q, const float* d, const float* ch,
                     const float* cmask, void* stream) {
  const void* in[4] = {q, d, ch, cmask};
  void* out[1] = {q_a};
  return nh::op(ctx, OP_LIMITER_H, in, 4, out, 1, 0.0, stream, 1, "nh_limiter_h", true);
}

}  // extern "C"
