/*
 * nicamhip: C-ABI of the MI355X (gfx950) backend of the pyNICAM-DC
 * dynamical-core time step.
 *
 * Plain C types only: dimensions and constants as POD structs, fields as
 * device pointers to arrays in the backend layout, streams as opaque
 * hipStream_t handles (NULL = default stream).
 *
 * Two forms, chosen when a context is created:
 *   NH_PREC_FP64   fields are double; the drop-in for the reference's float64
 *                  path (backend precision "float64");
 *   NH_PREC_MIXED  fields are float: stencils, advection, hyperdiffusion and
 *                  tracer transport run in FP32 with FP32 accumulation, while
 *                  the tridiagonal matrix and rhogw solve, the global sums and
 *                  the time steps stay FP64.  Entry points that take fields
 *                  come in a double form and an _f32 form; calling the form
 *                  that does not match the context returns NH_ERR_STATE.  Every entry point returns 0 on
 * success or a negative NH_ERR_* code; nh_last_error() then describes the
 * failure (thread-local, valid until the next call on that thread).  The
 * library never frees caller buffers.
 *
 * Field layout (one field = NF values):
 *   regional point (i, j, k, l):  ((l*kall + k)*gall_1d + j)*gall_1d + i
 *   pole point (v, k, lpl):       R + (lpl*kall + k)*6 + v,  R = lall*kall*gall_1d^2
 *   NF = R + 2*kall*6
 * i.e. the reference's (i, j, k, l) / (v, k, lpl) arrays with the axes
 * reversed, so that a column is strided by gall_1d^2 and a level of a region
 * is contiguous in i.  Multi-field arguments are NF-strided blocks.
 *
 * Regional column id c = (l*gall_1d + j)*gall_1d + i; pole column id
 * lall*gall_1d^2 + lpl*6 + v.
 */
#ifndef NICAMHIP_H
#define NICAMHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  NH_OK = 0,
  NH_ERR_ARG = -1,     /* invalid argument (null pointer, size out of range) */
  NH_ERR_STATE = -2,   /* context not ready (geometry or halo map missing)   */
  NH_ERR_ALLOC = -3,   /* device allocation failed                           */
  NH_ERR_DEVICE = -4   /* HIP runtime or kernel launch error                 */
};

/* Grid dimensions (mod_adm.py:121-182: ADM_gall_1d, ADM_kall, ADM_kmin,
 * ADM_kmax, ADM_lall). kmin must be 1 and kmax = kall-2 <= 78. */
typedef struct {
  int32_t gall_1d;
  int32_t kall;
  int32_t kmin;
  int32_t kmax;
  int32_t lall;
} nh_dims;

/* Physical constants and numerical-filter coefficients
 * (mod_const.py:9-205, mod_runconf.py:13-19, mod_numfilter.py:304-366). */
typedef struct {
  double radius;          /* CONST_RADIUS                                     */
  double ohm;             /* CONST_OHM                                        */
  double grav;            /* CONST_GRAV                                       */
  double rdry;            /* CONST_Rdry                                       */
  double cvdry;           /* CONST_CVdry                                      */
  double rvap;            /* CONST_Rvap                                       */
  double cvw_qv;          /* CVW[I_QV]                                        */
  double alpha;           /* NON_HYDRO_ALPHA                                  */
  double gamma_h;         /* hyperdiffusion coefficient, hdiff_type DIRECT    */
  double hdiff_fact_rho;  /* density factor of the hyperdiffusion             */
  double alpha_d;         /* 3-D divergence damping coefficient, DIRECT       */
  int32_t thuburn_lim;    /* THUBURN_LIM: tracer limiters on (1) or off (0)   */
} nh_consts;

/* Host pointers to the operator coefficients and vertical profiles
 * (mod_oprt.py:102-110, mod_grd.py:190-222, mod_vmtr.py:30-258).
 * Regional arrays are structure-of-arrays [component][column]:
 *   coef_div, coef_grad  21 = dir*7 + stencil point
 *   coef_lap              7
 *   coef_intp            18 = (dir*3 + vertex)*2 + triangle (TI, TJ)
 *   coef_diff            18 = dir*6 + arc
 *   grd_x                 3   position (m)
 * Pole arrays: coef_div_pl/coef_grad_pl/coef_diff_pl/grd_x_pl [v][lpl][dir],
 *   coef_lap_pl [v][lpl], coef_intp_pl [v][lpl][dir][vertex].
 * Vertical profiles have kall entries; c2w0/c2w1 and w2c0/w2c1 are the two
 * columns of C2Wfact/W2Cfact, phi the geopotential.  sgp[l] is 1 on the
 * regions whose (gmin-1, gmin-1) corner is a pentagon. */
typedef struct {
  const double *coef_div, *coef_grad, *coef_lap, *coef_intp, *coef_diff, *grd_x;
  const double *coef_div_pl, *coef_grad_pl, *coef_lap_pl, *coef_intp_pl, *coef_diff_pl, *grd_x_pl;
  const int32_t* sgp;
  const double *afact, *bfact, *cfact, *dfact, *rdgz, *rdgzh;
  const double *c2w0, *c2w1, *w2c0, *w2c1, *phi;
  /* tracer transport geometry (mod_gmtr.py GMTR_t / GMTR_a / GMTR_p, mod_grd.py GRD_xr):
   *   tri_w   6 = triangle*3 + w     interpolation weights W1..W3 of TI, TJ
   *   arc_hn  9 = arc*3 + dir        hexagon-arc normals (AI, AIJ, AJ) times length
   *   rarea   1                      1 / control-volume area
   *   arc_xr  9 = arc*3 + dir        hexagon-arc chord midpoints
   * pole: tri_w_pl [v][lpl][w], arc_hn_pl / arc_xr_pl [v][lpl][dir], rarea_pl [lpl] */
  const double *tri_w, *arc_hn, *rarea, *arc_xr;
  const double *tri_w_pl, *arc_hn_pl, *rarea_pl, *arc_xr_pl;
  /* control-volume area (GMTR_p P_AREA, one per column) and layer thickness
   * dgz [kall]: VOLUME = area * GSGAM2 * dgz (mod_vmtr.py) for the budgets */
  const double *area, *dgz;
  /* prescribed tracer winds: latitude/longitude [rad] per regional column
   * (GRD_LAT/GRD_LON) and [v][lpl] at the poles, full- and half-level
   * heights gz, gzh [kall] (GRD_vz Z/ZH under flat topography) */
  const double *lat, *lon, *lat_pl, *lon_pl, *gz, *gzh;
  /* hydrostatic basic state [kall] (Bsst.bsstate_setup with ref_type 'INPUT',
   * mod_bsstate.py:20-90, 159-200, flat topography: k-profiles): pressure,
   * temperature and density, kept in FP64 in both forms; the step forms
   * pre - pre_bs and rho - rho_bs for the pressure gradient and buoyancy and
   * tem - tem_bs, rho - rho_bs for the hyperdiffusion.  NULL = ref_type 'NOBASE'
   * (zeros). */
  const double *pre_bs, *tem_bs, *rho_bs;
} nh_geometry;

typedef struct nh_context nh_context;

enum { NH_PREC_FP64 = 0, NH_PREC_MIXED = 1 };

/* Description of the last failure on the calling thread. */
const char* nh_last_error(void);

/* Number of values per field (NF) for the given dimensions. */
int64_t nh_field_size(const nh_dims* dims);

/* Create a context on a device and allocate its step workspace.
 * Replaces the per-process setup chain Adm/Comm/Grd/Gmtr/Oprt/Vmtr
 * (mod_adm.py:78, mod_comm.py:119, mod_vmtr.py:30) for the step's buffers. */
int nh_create(const nh_dims* dims, const nh_consts* consts, int device, nh_context** out);

/* nh_create with the form of the fields: NH_PREC_FP64 (as nh_create) or
 * NH_PREC_MIXED.  The precision switch of Backend.configure(name, precision)
 * (share/mod_backend.py:175-283); the geometry is uploaded from the same
 * double arrays and rounded to the field precision on the device, with the
 * vertical profiles kept in FP64 as well. */
int nh_create_ex(const nh_dims* dims, const nh_consts* consts, int device, int32_t precision, nh_context** out);

/* The context's form (NH_PREC_FP64 or NH_PREC_MIXED), or NH_ERR_ARG. */
int32_t nh_precision(const nh_context* ctx);

/* Release a context and everything it allocated. */
int nh_destroy(nh_context* ctx);

/* Copy operator coefficients and vertical profiles to the device
 * (OPRT_setup mod_oprt.py:125-1386, VMTR_setup mod_vmtr.py:30-258). */
int nh_upload_geometry(nh_context* ctx, const nh_geometry* geo);

/* Upload the halo gather map: for h < n, column dst[h] <- column src[h] on
 * every level. Replaces the plan of Comm._build_comm_plan (mod_comm.py:1400-1561)
 * built from copy_r2r / copy_p2r / copy_r2p / singular lists. */
int nh_upload_halo_map(nh_context* ctx, const int32_t* dst, const int32_t* src, int64_t n);

/* In-place halo and pole fill of nfields fields (each NF doubles).
 * Replaces Comm.COMM_data_transfer (mod_comm.py:1066, 1563-1644). */
int nh_halo_fill(nh_context* ctx, double* const* fields, int32_t nfields, void* stream);
int nh_halo_fill_f32(nh_context* ctx, float* const* fields, int32_t nfields, void* stream);

/* 7-point Laplacian on interior points and pole centres; other points of out
 * are left unchanged.  Replaces OPRT_laplacian (mod_oprt.py:1429,
 * kernels/oprtlaplacian.py:41-85). */
int nh_laplacian(nh_context* ctx, const double* scl, double* out, void* stream);
int nh_laplacian_f32(nh_context* ctx, const float* scl, float* out, void* stream);

/* 7-point divergence of (vx, vy, vz) on interior points and pole centres.
 * Replaces the horizontal part of src_flux_convergence (mod_src.py:631,
 * kernels/fluxconv.py:238 _divergence). */
int nh_divergence(nh_context* ctx, const double* vx, const double* vy, const double* vz, double* out,
                  void* stream);
int nh_divergence_f32(nh_context* ctx, const float* vx, const float* vy, const float* vz, float* out,
                      void* stream);

/* 7-point gradient of scl on interior points and pole centres (3 output
 * fields).  Replaces OPRT_gradient (mod_oprt.py:1387, kernels/oprtgradient.py:45-88). */
int nh_gradient(nh_context* ctx, const double* scl, double* gx, double* gy, double* gz, void* stream);
int nh_gradient_f32(nh_context* ctx, const float* scl, float* gx, float* gy, float* gz, void* stream);

/* ---- Standalone operators of the reference's kernel API.  Each runs one
 * operator of the step on caller fields (device layout, NF values per field;
 * multi-field arguments are NF-strided blocks) on the given stream.  Vertical
 * metrics are those of flat topography (the context's k-profiles), where
 * every metric factor of the reference expression is exactly 1. */

/* Tridiagonal matrix of the vertical implicit rhogw solve, rows kmin+1..kmax of
 * Mc, Mu, Ml (FP64 in both forms; other rows untouched), from eth and g_tilde at
 * half levels.  Replaces compute_rhow_matrix_reg/_pl with precombine_matrix_reg/_pl
 * (nhm/dynamics/kernels/vimatrix.py:67-165; called from mod_vi.py:1776-1922
 * vi_rhow_update_matrix). */
int nh_vi_matrix(nh_context* ctx, const double* eth_h, const double* g_tilde, double dt, double* Mc, double* Mu,
                 double* Ml, void* stream);
int nh_vi_matrix_f32(nh_context* ctx, const float* eth_h, const float* g_tilde, double dt, double* Mc, double* Mu,
                     double* Ml, void* stream);

/* rhogw column solve (FP64 arithmetic in both forms): source assembly, the two
 * boundary rows and the Thomas sweeps on rows kmin+1..kmax; out rows below kmin
 * and above kmax+1 are copies of rhogw; out may be rhogw.  Replaces
 * compute_rhow_solver_reg/_pl (nhm/dynamics/kernels/virhowsolver.py:67-172;
 * called from mod_vi.py:2433-2504 vi_rhow_solver). */
int nh_vi_solve(nh_context* ctx, const double* rhogw, const double* rhogw0, const double* preg0, const double* rhog0,
                const double* Srho, const double* Sw, const double* Spre, const double* Mc, const double* Mu,
                const double* Ml, double dt, double* out, void* stream);
int nh_vi_solve_f32(nh_context* ctx, const float* rhogw, const float* rhogw0, const float* preg0, const float* rhog0,
                    const float* Srho, const float* Sw, const float* Spre, const double* Mc, const double* Mu,
                    const double* Ml, double dt, float* out, void* stream);

/* Diagnostics, boundary conditions and eth of the large step: diag (6 fields:
 * pre, tem, vx, vy, vz, w) and eth from prog and qv (progq's first field); the
 * ghost levels of prog (rhog, rhogvx..vz, rhogw, rhoge) are set in place.
 * Replaces compute_diagnostics (kernels/diag.py:53), Bndc.BNDCND_all
 * (mod_bndcnd.py:182-266, kernels/bndcnd.py:54-296) and THRMDYN_eth
 * (kernels/thrmdyn.py:56) as called at mod_dynamics.py:965-1283. */
int nh_prepost(nh_context* ctx, double* prog, const double* progq, double* diag, double* eth, void* stream);
int nh_prepost_f32(nh_context* ctx, float* prog, const float* progq, float* diag, float* eth, void* stream);

/* Fourth-order horizontal hyperdiffusion (lap order 2, DIRECT coefficient):
 * the two iterated-Laplacian passes with a halo fill after each, then the
 * tendency, six fields in PROG order (rhog, rhogvx, rhogvy, rhogvz, rhogw,
 * rhoge), momentum horizontalised; the tracer tendency of the reference is zero
 * under MIURA2004 transport and is not produced.  prog and diag as left by
 * nh_prepost.  Uses the step workspace.  Replaces Numf.numfilter_hdiffusion
 * (nhm/dynamics/mod_numfilter.py:845-1469, tendency :1519-1578). */
int nh_hdiffusion(nh_context* ctx, const double* prog, const double* diag, double* tend, void* stream);
int nh_hdiffusion_f32(nh_context* ctx, const float* prog, const float* diag, float* tend, void* stream);

/* 3-D divergence damping (lap order 2): gdx, gdy, gdz on the interior points
 * and pole centres; the vertical component gdvz of the reference is zero when
 * its vertical damping is off (every reference configuration) and is not
 * produced.  Uses the step workspace.  Replaces Numf.numfilter_divdamp
 * (nhm/dynamics/mod_numfilter.py:2032-2268: OPRT3D_divdamp, kernels/oprt3ddivdamp.py:240,
 * halo, kernels/divdamppostcomm.py:29). */
int nh_divdamp3d(nh_context* ctx, const double* rhogvx, const double* rhogvy, const double* rhogvz,
                 const double* rhogw, double* gdx, double* gdy, double* gdz, void* stream);
int nh_divdamp3d_f32(nh_context* ctx, const float* rhogvx, const float* rhogvy, const float* rhogvz,
                     const float* rhogw, float* gdx, float* gdy, float* gdz, void* stream);

/* Vertical Thuburn limiter of the face values q_h (rows kmin+1..kmax limited in
 * place) from the cell values q, d and the two Courant numbers ck (2 fields).
 * Replaces compute_vertical_limiter (nhm/dynamics/kernels/verticallimiter.py:41-92;
 * mod_src_tracer.py:1849-2113 vertical_limiter_thuburn). */
int nh_limiter_v(nh_context* ctx, double* q_h, const double* q, const double* d, const double* ck, void* stream);
int nh_limiter_v_f32(nh_context* ctx, float* q_h, const float* q, const float* d, const float* ck, void* stream);

/* Horizontal Thuburn limiter of the face values q_a (6 fields, in place): the
 * outflow bounds Qout from q, d, the Courant numbers ch (6 fields) and the
 * upwind masks cmask (6 fields; faces 0..2 are read), their halo fill, then the
 * limited faces.  Pole points use the first field of ch, cmask and q_a.  Uses the
 * step workspace.  Replaces compute_horizontal_limiter_qout/_apply and their _pl
 * forms (nhm/dynamics/kernels/horizontallimiter.py:51-255) with the Qout
 * exchange between them (mod_src_tracer.py:2115-2662
 * horizontal_limiter_thuburn). */
int nh_limiter_h(nh_context* ctx, double* q_a, const double* q, const double* d, const double* ch,
                 const double* cmask, void* stream);
int nh_limiter_h_f32(nh_context* ctx, float* q_a, const float* q, const float* d, const float* ch,
                     const float* cmask, void* stream);

/* One dynamical-core large step, in place: prog holds the 6 prognostic
 * fields (rhog, rhogvx, rhogvy, rhogvz, rhogw, rhoge), progq the nq >= 1
 * tracer mass fields (tracer 0 = qv).  RK3 with sstep_max/3, sstep_max/2 and
 * sstep_max HEVI small steps of dtl/sstep_max; the tracers are transported
 * in the last stage with the time-mean mass fluxes (MIURA2004, Thuburn
 * limiters when consts.thuburn_lim).
 * Replaces Dyn.dynamics_step (mod_dynamics.py:1725-2741) with
 * Srctr.src_tracer_advection (mod_src_tracer.py:83-1228). */
int nh_dyn_step(nh_context* ctx, double* prog, double* progq, int32_t nq, double dtl, int32_t sstep_max,
                void* stream);
int nh_dyn_step_f32(nh_context* ctx, float* prog, float* progq, int32_t nq, double dtl, int32_t sstep_max,
                    void* stream);

/* Offline tracer step (TIME_integ_type 'TRCADV', Dyn.dynamics_step,
 * mod_dynamics.py:2030-2076): MIURA2004 transport of progq
 * (src_tracer_advection, mod_src_tracer.py:83-1228) by the mass fluxes held in
 * prog (rhog, rhogvx..vz, rhogw; rhog is both the initial and the mean
 * density), with zero density tendency, then the halo fill of progq. */
int nh_tracer_step(nh_context* ctx, const double* prog, double* progq, int32_t nq, double dtl, void* stream);
int nh_tracer_step_f32(nh_context* ctx, const float* prog, float* progq, int32_t nq, double dtl, void* stream);

/* Adjoint of one nh_tracer_step with respect to the tracer masses (no
 * counterpart in the reference, which has no adjoint).  prog (the mass
 * fluxes) and progq (the tracers the forward step starts from) are unchanged;
 * progq_b holds on entry the gradient of a scalar with respect to the tracers
 * after the step, on exit its gradient with respect to progq, every point
 * including the halo points.  The fluxes are held fixed and the limiters'
 * min/max/clip selections and upwind switches are those of the forward pass
 * from progq, which makes the step linear in progq: the transpose is the
 * exact derivative wherever no selection is tied.  Uses the adjoint
 * workspace of nh_dyn_step_adj_tape (allocated by the first call); needs the
 * halo map with unique destinations that are never sources (else
 * NH_ERR_STATE). */
int nh_tracer_step_adj(nh_context* ctx, const double* prog, const double* progq, int32_t nq, double dtl,
                       double* progq_b, void* stream);
int nh_tracer_step_adj_f32(nh_context* ctx, const float* prog, const float* progq, int32_t nq, double dtl,
                           float* progq_b, void* stream);

/* Prescribed winds of the offline tracer runs (Forcing.forcing_update,
 * mod_forcing.py:113-206): rhogv = v(t) rhog, rhogw = w(t) rhog at every
 * point, prog's density unchanged.
 * kind NH_WIND_DCMIP11: Trcadv.test11_velocity (mod_af_trcadv.py:23-98),
 *   params unused;
 * kind NH_WIND_DEFORM: non-divergent, time-reversing deformational flow
 *   with vertical motion, params = {period [s], umax [m/s], w0 [m/s],
 *   ztop [m]}. */
enum { NH_WIND_DCMIP11 = 0, NH_WIND_DEFORM = 1 };
int nh_trcadv_winds(nh_context* ctx, double* prog, double time, int32_t kind, const double* params,
                    void* stream);
int nh_trcadv_winds_f32(nh_context* ctx, float* prog, double time, int32_t kind, const double* params,
                        void* stream);

/* Held-Suarez forcing step (AF_TYPE 'HELD-SUAREZ'; Dyn.forcing_step,
 * nhm/dynamics/mod_dynamics.py:597-720 -> Frc.forcing_step_hs / hs_apply_core,
 * nhm/forcing/mod_forcing.py:12-29 -> AfHeldsuarez.AF_heldsuarez,
 * nhm/forcing/mod_af_heldsuarez.py:43-95): diagnostics of prog, Rayleigh
 * friction of rhogvx/vy/vz and Newtonian relaxation of rhoge on the regional
 * physical levels (the pole points are not forced, as in the reference),
 * PROG += dt * f * rho, then the halo fill of the six prognostics.  Tracer 0 of
 * progq is qv.  moist_case 0: Held and Suarez (1994), 1: Thatcher and
 * Jablonowski (2016) equilibrium temperature. */
int nh_forcing_hs(nh_context* ctx, double* prog, const double* progq, double dt, int32_t moist_case,
                  void* stream);
int nh_forcing_hs_f32(nh_context* ctx, float* prog, const float* progq, double dt, int32_t moist_case,
                      void* stream);

/* Global mass and energy sums (Embudget._store_var, mod_embudget.py:82-139):
 * sum of var * VOLUME over the regional interior points of the physical
 * levels (pole cells omitted, as in _gsum, mod_embudget.py:74-80), with
 * compensated FP64 summation.  Tracer 0 of progq is qv, lhv its latent heat.
 * out[0..5] = potential, internal, kinetic and total energy [J], dry-air
 * mass and water-vapour mass [kg].  Synchronises the stream. */
int nh_budget(nh_context* ctx, const double* prog, const double* progq, int32_t nq, double lhv, double* out,
              void* stream);
int nh_budget_f32(nh_context* ctx, const float* prog, const float* progq, int32_t nq, double lhv, double* out,
                  void* stream);

/* nh_budget without any device allocation (no counterpart in the reference;
 * for callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops).  The same sums, in either form (prog and progq
 * are double or float as the context's fields), with both reduction stages on
 * the device in the library's fixed 8 MiB FP64 scratch array (2^20 values,
 * held by each GPU the library runs on until the process ends): at most
 * (2^20 - 10) / 10 blocks of 256 interior points, i.e. 26,843,136 interior
 * points of the physical levels (gl08 z40 has 26,214,400); a larger grid is
 * refused with NH_ERR_ARG.  Every partial is written before it is read, so no
 * state carries over between calls; calls that overlap on one GPU (two streams
 * or threads) share the array and race.  Synchronises the stream. */
int nh_budget_noalloc(nh_context* ctx, const void* prog, const void* progq, int32_t nq, double lhv, double* out,
                      void* stream);

/* Adjoint of one nh_dyn_step (no counterpart in the reference, which has no
 * adjoint; the forward call it transposes is Dyn.dynamics_step,
 * nhm/dynamics/mod_dynamics.py:1725).  prog, progq: the state the forward step
 * starts from (unchanged).  prog_b: on entry the gradient of a scalar with
 * respect to the state after the step, on exit its gradient with respect to
 * prog.  The step is recomputed inside from prog (checkpoints of every RK
 * stage and small step in a workspace the first call allocates).  Tracer
 * masses are held at progq: the result is the full derivative when they are
 * zero (dry runs).  Needs the halo map with unique destinations that are
 * never sources (else NH_ERR_STATE). */
int nh_dyn_step_adj(nh_context* ctx, const double* prog, const double* progq, int32_t nq, double dtl,
                    int32_t sstep_max, double* prog_b, void* stream);
int nh_dyn_step_adj_f32(nh_context* ctx, const float* prog, const float* progq, int32_t nq, double dtl,
                        int32_t sstep_max, float* prog_b, void* stream);

/* The same pair with the stage inputs recorded (no counterpart in the
 * reference): nh_dyn_step_record is nh_dyn_step that also writes the input
 * state of each of the three RK stages to stage_in (3 x 6 fields of NF
 * values); nh_dyn_step_adj_rec is nh_dyn_step_adj from those records (the
 * first six fields being the step's input state), which saves it the
 * recomputation of stages 0 and 1. */
int nh_dyn_step_record(nh_context* ctx, double* prog, double* progq, int32_t nq, double dtl, int32_t sstep_max,
                       double* stage_in, void* stream);
int nh_dyn_step_record_f32(nh_context* ctx, float* prog, float* progq, int32_t nq, double dtl, int32_t sstep_max,
                           float* stage_in, void* stream);
int nh_dyn_step_adj_rec(nh_context* ctx, const double* stage_in, const double* progq, int32_t nq, double dtl,
                        int32_t sstep_max, double* prog_b, void* stream);
int nh_dyn_step_adj_rec_f32(nh_context* ctx, const float* stage_in, const float* progq, int32_t nq, double dtl,
                            int32_t sstep_max, float* prog_b, void* stream);

/* Taped step and its adjoint (no counterpart in the reference): the tape
 * holds, per RK stage, its input state and the fields its small steps read
 * (hyperdiffusion passes, large-step tendencies, total energy, the FP64
 * tridiagonal matrix) and every small step's split state and divergences;
 * nh_tape_bytes gives its size for one step (device memory the caller owns).
 * nh_dyn_step_tape is nh_dyn_step writing its intermediate fields to the tape
 * instead of the workspace; nh_dyn_step_adj_tape is nh_dyn_step_adj from that
 * tape, recomputing only the diagnostics. */
int64_t nh_tape_bytes(const nh_context* ctx, int32_t sstep_max);
int nh_dyn_step_tape(nh_context* ctx, double* prog, double* progq, int32_t nq, double dtl, int32_t sstep_max,
                     void* tape, void* stream);
int nh_dyn_step_tape_f32(nh_context* ctx, float* prog, float* progq, int32_t nq, double dtl, int32_t sstep_max,
                         void* tape, void* stream);
int nh_dyn_step_adj_tape(nh_context* ctx, const void* tape, const double* progq, int32_t nq, double dtl,
                         int32_t sstep_max, double* prog_b, void* stream);
int nh_dyn_step_adj_tape_f32(nh_context* ctx, const void* tape, const float* progq, int32_t nq, double dtl,
                             int32_t sstep_max, float* prog_b, void* stream);

/* Surface pressure per column (Bndc.BNDCND_pre_sfc, nhm/share/mod_bndcnd.py:521-555,
 * with a flat surface z_srf = 0): density extrapolated quadratically from the
 * lowest three levels, hydrostatic pressure from kmin.  ps: nreg + 12 doubles
 * (regional columns (l*g + j)*g + i, then the pole columns lp*6 + v).  The _adj
 * form adds the transpose applied to ps_b (per column) into prog_b. */
int nh_sfc_pressure(nh_context* ctx, const double* prog, const double* progq, double* ps, void* stream);
int nh_sfc_pressure_f32(nh_context* ctx, const float* prog, const float* progq, double* ps, void* stream);
int nh_sfc_pressure_adj(nh_context* ctx, const double* prog, const double* progq, const double* ps_b,
                        double* prog_b, void* stream);
int nh_sfc_pressure_adj_f32(nh_context* ctx, const float* prog, const float* progq, const double* ps_b,
                            float* prog_b, void* stream);

/* Kernel timers (the per-kernel counterpart of the section timers
 * PROF_rapstart / PROF_rapend / PROF_rapreport, share/mod_prof.py:81-140).
 * While enabled, every launch of a context is bracketed by two events on its
 * stream.  nh_prof_report waits for the recorded launches, writes up to
 * max_kernels entries (names: 32 bytes each, NUL-terminated; summed duration
 * in ms; launch count), sets *n_kernels and clears the record. */
int nh_prof_enable(nh_context* ctx, int32_t on);
/* Restrict the timers to the launches of one kernel (its report name; NULL or
 * "" = every kernel), so that a timed region carries one event pair per launch
 * of that kernel only. */
int nh_prof_only(nh_context* ctx, const char* kernel);
int nh_prof_report(nh_context* ctx, int32_t max_kernels, char* names, double* total_ms, int64_t* launches,
                   int32_t* n_kernels);

#ifdef __cplusplus
}
#endif

#endif /* NICAMHIP_H */
