// The interface each precision variant of the library implements (nh_impl.h,
// compiled once per variant) and the C-ABI layer (nh_api.h) dispatches to.
// Field pointers are untyped here: the variant knows its field precision.
// Every function returns 0 or a negative NH_ERR_* code and sets *err.
#pragma once

#include <cstdint>
#include <string>

#include "../../include/nicamhip.h"

// standalone operators of v_op: inputs in[], outputs out[] (field pointers)
enum NhOp {
  OP_VI_MATRIX = 0,   // in: eth_h, g_tilde; out: Mc, Mu, Ml (double)
  OP_VI_SOLVE = 1,    // in: rhogw, rhogw0, preg0, rhog0, Srho, Sw, Spre, Mc, Mu, Ml (double); out: rhogw
  OP_PREPOST = 2,     // in: -, progq; out: prog (ghost rows, in place), diag (6), eth
  OP_HDIFF = 3,       // in: prog, diag; out: 6 tendencies
  OP_DIVDAMP = 4,     // in: rhogvx, rhogvy, rhogvz, rhogw; out: gdx, gdy, gdz
  OP_LIMITER_V = 5,   // in: q, d, ck0, ck1; out: q_h (in place)
  OP_LIMITER_H = 6    // in: q, d, ch (6), cmask (6); out: q_a (6, in place)
};

#define NH_VARIANT_API(NS)                                                                                  \
  namespace nh {                                                                                            \
  namespace NS {                                                                                            \
  struct Context;                                                                                           \
  int v_create(const nh_dims* dims, const nh_consts* consts, int device, Context** out, std::string* err); \
  void v_destroy(Context* c);                                                                               \
  bool v_geo_ready(const Context* c);                                                                       \
  long v_halo_n(const Context* c);                                                                          \
  int v_upload_geometry(Context* c, const nh_geometry* g, std::string* err);                               \
  int v_upload_halo(Context* c, const int32_t* dst, const int32_t* src, long n, std::string* err);         \
  int v_halo_fill(Context* c, void* const* fields, int nf, void* stream, std::string* err);                \
  int v_laplacian(Context* c, const void* scl, void* out, void* stream, std::string* err);                 \
  int v_divergence(Context* c, const void* vx, const void* vy, const void* vz, void* out, void* stream,     \
                   std::string* err);                                                                       \
  int v_gradient(Context* c, const void* scl, void* gx, void* gy, void* gz, void* stream, std::string* err); \
  int v_op(Context* c, int op, const void* const* in, void* const* out, double dt, void* stream,              \
           std::string* err);                                                                               \
  int v_dyn_step(Context* c, void* prog, void* progq, int nq, double dtl, int sstep_max, void* record,      \
                 void* tape, void* stream, std::string* err);                                               \
  long v_tape_bytes(const Context* c, int sstep_max);                                                      \
  int v_dyn_step_adj_tape(Context* c, const void* tape, const void* progq, int nq, double dtl,              \
                          int sstep_max, void* prog_b, void* stream, std::string* err);                      \
  int v_tracer_step(Context* c, const void* prog, void* progq, int nq, double dtl, void* stream,            \
                    std::string* err);                                                                      \
  int v_trcadv_winds(Context* c, void* prog, double time, int kind, const double* params, void* stream,     \
                     std::string* err);                                                                     \
  int v_forcing_hs(Context* c, void* prog, const void* progq, double dt, int moist_case, void* stream,      \
                   std::string* err);                                                                       \
  int v_budget(Context* c, const void* prog, const void* progq, double lhv, double* out, void* stream,      \
               std::string* err);                                                                           \
  int v_budget_ws(Context* c, const void* prog, const void* progq, double lhv, double* out, double* ws,       \
                  long ws_n, void* stream, std::string* err);                                               \
  int v_dyn_step_adj(Context* c, const void* prog, const void* progq, int nq, double dtl, int sstep_max,      \
                     void* prog_b, const void* stage_in, void* stream, std::string* err);                   \
  int v_tracer_step_adj(Context* c, const void* prog, const void* progq, int nq, double dtl, void* progq_b,    \
                        void* stream, std::string* err);                                                      \
  int v_sfc_pressure(Context* c, const void* prog, const void* progq, double* ps, const double* ps_b,       \
                     void* prog_b, void* stream, std::string* err);                                         \
  int v_prof_enable(Context* c, bool on, std::string* err);                                                 \
  int v_prof_only(Context* c, const char* name, std::string* err);                                          \
  int v_prof_report(Context* c, int max_kernels, char* names, double* total_ms, int64_t* launches,          \
                    std::string* err);                                                                      \
  }                                                                                                         \
  }

NH_VARIANT_API(f64)
NH_VARIANT_API(f32)

namespace nh {
// the library's fixed FP64 scratch on a device (the allocation-free entry
// points): *p its device address, *n its size in values.  Defined beside the
// scratch array by each build (gfx950: a file-scope device array whose address
// is looked up once per device; host test build: a static array).
int scratch_lookup(int device, double** p, long* n, std::string* err);
}
