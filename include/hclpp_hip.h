/*
 * hclpp_hip.h -- C-ABI of libhclpp_hip.so, the MI355X (gfx950) HIP
 * implementation of the HERACLES++ hydro step.
 *
 * Plain C: opaque context, plain device pointers and sizes, int status codes
 * (HCLPP_OK = 0) with a thread-local message from hclpp_last_error().  No
 * exception crosses this boundary.
 *
 * Data layout (identical to the reference's Kokkos::LayoutLeft views, so a
 * caller hands over View::data() unchanged): every array spans the
 * ghost-inclusive extents nwg[0] x nwg[1] x nwg[2] (nwg[d] = n[d] + 2 ng on
 * active dims, n[d] = 1 on inactive dims), flat index
 * i + nwg0 * (j + nwg1 * k), and component indices are slowest:
 *   u, rhou          [wg, ndim]                 (component plane stride = nwg0*nwg1*nwg2)
 *   fx               [wg, nfx]
 *   var_rec, P_rec   [wg, side(2), idim(ndim)]
 *   rho_rec, E_rec   [wg, side(2), idim(ndim)]
 *   u_rec, rhou_rec  [wg, side(2), idim(ndim), component(ndim)]
 *   fx_rec           [wg, side(2), idim(ndim), nfx]
 * All array pointers are device pointers.  Work is enqueued on the
 * context's stream (hclpp_set_stream); functions that return a host scalar
 * synchronise that stream.
 *
 * Citations are to the reference tree (src/...).
 */
#ifndef HCLPP_HIP_H
#define HCLPP_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define HCLPP_OK 0
#define HCLPP_ERR_ARG 1      /* invalid argument or configuration */
#define HCLPP_ERR_HIP 2      /* HIP runtime error */
#define HCLPP_ERR_STATE 3    /* invalid physical state detected */
#define HCLPP_ERR_UNSUPPORTED 4

enum { HCLPP_GEOM_CARTESIAN = 0, HCLPP_GEOM_SPHERICAL = 1 };
enum { HCLPP_LIM_CONSTANT = 0, HCLPP_LIM_VANLEER = 1, HCLPP_LIM_MINMOD = 2, HCLPP_LIM_VANALBADA = 3 };
enum { HCLPP_RS_HLL = 0, HCLPP_RS_HLLC = 1, HCLPP_RS_LOWMACH = 2 };
enum { HCLPP_GRAV_UNIFORM = 0, HCLPP_GRAV_POINT_MASS = 1, HCLPP_GRAV_INTERNAL_MASS = 2 };
enum { HCLPP_BC_PERIODIC = 0, HCLPP_BC_NULLGRADIENT = 1, HCLPP_BC_REFLEXIVE = 2 };
enum { HCLPP_FP64 = 0, HCLPP_FP32_STORAGE = 1 };

typedef struct hclpp_ctx hclpp_ctx;

/* Replaces the compile-time ndim/geom/EoS/gravity choices and the run-time
 * Grid, PerfectGas, gravity and boundary objects built in main.cpp:140-303
 * (Grid: mesh/grid.cpp:48-76; PerfectGas(gamma, mmw): perfect_gas.cpp:13-22;
 * make_uniform_gravity / make_point_mass_gravity / make_internal_mass_gravity: gravity/gravity.cpp:29-128;
 * bc choices and priority: main.cpp:175-186, boundary_distribute.cpp:33-71). */
typedef struct {
    int ndim;          /* 1, 2 or 3 */
    int geom;          /* HCLPP_GEOM_* */
    int limiter;       /* HCLPP_LIM_*: used by the fused step */
    int riemann;       /* HCLPP_RS_*: used by the fused step */
    int gravity;       /* HCLPP_GRAV_* */
    int nfx;           /* number of passive scalars */
    int ng;            /* ghost layers on active dims (the reference default is 2) */
    int n[3];          /* interior cells per dim (1 on inactive dims) */
    int bc[6];         /* per face 2*dim+side: HCLPP_BC_* */
    int bc_order[6];   /* order of border BCs (0..2*ndim-1) */
    double gamma;      /* adiabatic index (> 1) */
    double mmw;        /* mean molecular weight (> 0) */
    double g[3];       /* uniform gravity */
    double central_mass; /* point-mass and internal-mass gravity */
    const double* x[3];  /* HOST node coordinates, nwg[d] + 1 entries (ghosts included) */
    int precision;     /* HCLPP_FP64 or HCLPP_FP32_STORAGE (fused step only) */
} hclpp_config;

/* Cell range [begin, end) per dim in ghost-inclusive indices (mesh/range.hpp:19-55). */
typedef struct {
    int begin[3];
    int end[3];
} hclpp_range;

const char* hclpp_last_error(void);
int hclpp_ctx_create(const hclpp_config* cfg, hclpp_ctx** out);
int hclpp_ctx_destroy(hclpp_ctx* ctx);
/* stream is a hipStream_t (NULL: the null stream). */
int hclpp_set_stream(hclpp_ctx* ctx, void* stream);
/* Range::with_ghosts(ngeff) of the whole local domain (mesh/range.cpp:89-107, :120-135). */
int hclpp_range_with_ghosts(const hclpp_ctx* ctx, int ngeff, hclpp_range* out);
/* name -> enum as the reference factories map them; -1 if unknown
 * (face_reconstruction.cpp:31-50, godunov_scheme.hpp:305-321, boundary_factory.hpp:31-47). */
int hclpp_limiter_from_name(const char* name);
int hclpp_riemann_from_name(const char* name);
int hclpp_bc_from_name(const char* name);

/* ---- reference operator interfaces (FP64), one kernel per reference kernel ---- */

/* conv_prim_to_cons: hydro/array_conversion.hpp:32-97 */
int hclpp_conv_prim_to_cons(hclpp_ctx* ctx, const hclpp_range* range, const double* rho, const double* u, const double* P,
                            double* rhou, double* E);
/* conv_cons_to_prim: hydro/array_conversion.hpp:107-172 */
int hclpp_conv_cons_to_prim(hclpp_ctx* ctx, const hclpp_range* range, const double* rho, const double* rhou, const double* E,
                            double* u, double* P);
/* IFaceReconstruction::execute with the LimitedLinearReconstruction<limiter> made by
 * factory_face_reconstruction: hydro/face_reconstruction.hpp:21-43,
 * limited_linear_reconstruction.hpp:34-65 */
int hclpp_face_reconstruction(hclpp_ctx* ctx, int limiter, const hclpp_range* range, const double* var, double* var_rec);
/* IExtrapolationReconstruction::execute of HancockExtrapolationReconstruction:
 * hydro/extrapolation_reconstruction.hpp:27-54, :65-338.  dt is the half step. */
int hclpp_hancock_extrapolation(hclpp_ctx* ctx, const hclpp_range* range, double dt, const double* u_rec, const double* P_rec,
                                double* rho_rec, double* rhou_rec, double* E_rec, double* fx_rec);
/* The face-state conversion of MUSCLHancockHydroReconstruction::execute,
 * hydro/hydro_reconstruction.hpp:111-126: conv_prim_to_cons of every face state (iside, idim)
 * over the range, from rho_rec, u_rec, P_rec (layouts [cell][side][dim] and
 * [cell][side][dim][component]) to rhou_rec and E_rec.  For callers that compose the
 * reconstruction from their own IFaceReconstruction / IExtrapolationReconstruction. */
int hclpp_face_prim_to_cons(hclpp_ctx* ctx, const hclpp_range* range, const double* rho_rec, const double* u_rec, const double* P_rec,
                            double* rhou_rec, double* E_rec);
/* IHydroReconstruction::execute of MUSCLHancockHydroReconstruction (P_rec, u_rec are
 * its scratch views): hydro/hydro_reconstruction.hpp:30-59, :85-129 */
int hclpp_hydro_reconstruction(hclpp_ctx* ctx, int limiter, const hclpp_range* range, double dt, const double* rho, const double* u,
                               const double* P, const double* fx, double* rho_rec, double* rhou_rec, double* E_rec, double* fx_rec,
                               double* P_rec, double* u_rec);
/* IGodunovScheme::execute of RiemannBasedGodunovScheme<solver> made by
 * factory_godunov_scheme: hydro/godunov_scheme.hpp:31-64, :77-302, :305-321 */
int hclpp_godunov(hclpp_ctx* ctx, int riemann, const hclpp_range* range, double dt, const double* rho, const double* rhou,
                  const double* E, const double* fx, const double* rho_rec, const double* rhou_rec, const double* E_rec,
                  const double* fx_rec, double* rho_new, double* rhou_new, double* E_new, double* fx_new);
/* time_step<EoS>: hydro/time_step.hpp:35-82; *dt_out = 1 / max(sum_d (|u_d|+c)/dx_geom_d)
 * (the caller multiplies by cfl, as main.cpp:355 does).  Synchronises. */
int hclpp_time_step(hclpp_ctx* ctx, const hclpp_range* range, const double* rho, const double* u, const double* P, double* dt_out);
/* hclpp_time_step for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  Same kernels and the same result, bit for bit, written to the device double
 * *dt_dev on the context's stream; no allocation and no synchronisation.  The workgroup
 * partials live in a static device array of 2^20 doubles (8 MiB, held by each GPU the library
 * runs on until the process ends): a range needing more than 2^20 workgroups of 64 x 4 cells
 * (about 268 M cells) fails with HCLPP_ERR_ARG.  Every call on one GPU shares that array, so
 * calls that overlap in time on one GPU (other streams or threads) race. */
int hclpp_time_step_noalloc(hclpp_ctx* ctx, const hclpp_range* range, const double* rho, const double* u, const double* P,
                            double* dt_dev);
/* temperature: analysis/temperature.hpp:18-26, T = compute_temp_from_pres(rho, P) over the range
 * (perfect_gas.hpp:69-72, IEEE division). */
int hclpp_temperature(hclpp_ctx* ctx, const hclpp_range* range, const double* rho, const double* P, double* T);
/* minimum_internal_energy: analysis/internal_energy.cpp:20-45.  Synchronises. */
int hclpp_min_internal_energy(hclpp_ctx* ctx, const hclpp_range* range, const double* rho, const double* rhou, const double* E,
                              double* out);
/* integrate: analysis/integration.cpp:22-37 (sum of var * dV).  Synchronises. */
int hclpp_integrate(hclpp_ctx* ctx, const hclpp_range* range, const double* var, double* out);
/* pressure_fix: hydro/pressure_fix.hpp:19-172 (main.cpp:459-475, range no_ghosts); rec arrays are the
 * MUSCL-Hancock face states of the step, E_new is updated in place. */
int hclpp_pressure_fix(hclpp_ctx* ctx, const hclpp_range* range, double dt, double eps, const double* rho, const double* rhou,
                       const double* E, const double* rho_rec, const double* rhou_rec, const double* E_rec, const double* rho_new,
                       const double* rhou_new, double* E_new);
/* make_internal_mass_gravity: gravity/gravity.cpp:59-128 (one rank along dim 0).  Rebuilds the
 * context's radial gravity table from rho (ghost-filled), as main.cpp:309 and :528 do after
 * every step; the fused step does it by itself.  Needs gravity = HCLPP_GRAV_INTERNAL_MASS. */
int hclpp_internal_mass_gravity(hclpp_ctx* ctx, const double* rho);
/* Copy the context's radial gravity table g(x0 centre) (nwg[0] doubles, point-mass or internal-mass
 * gravity) to host memory.  Synchronises. */
int hclpp_gravity_table(hclpp_ctx* ctx, double* host_out);
/* DistributedBoundaryCondition::operator() on one rank: boundary_distribute.hpp:42-76,
 * boundary_distribute.cpp:75-120, boundary.cpp:33-108 (conservative arrays). */
int hclpp_fill_ghosts(hclpp_ctx* ctx, double* rho, double* rhou, double* E, double* fx);

/* ---- fused step with a device-resident clock (replaces the loop body of main.cpp:354-531) ----
 * The clock lives in device memory so consecutive steps run without host
 * synchronisation: each step reads dt from the clock, and its epilogue
 * computes the next dt = cfl * (1 / max CFL rate over all cells including
 * ghosts), clips it to t_end, counts the iteration and raises `done` when
 * t_end or max_iter is reached (main.cpp:355-385).  A step launched after
 * `done` is a no-op.  min_eint holds the minimum internal energy of the state
 * each step consumed (main.cpp:397-425); status becomes HCLPP_ERR_STATE when it
 * was NaN or negative, and every later step is a no-op.
 * With an output schedule (hclpp_clock_set_outputs), dt is also clipped to the next
 * output time and output iterations are counted as main.cpp:362-375 does; the clock
 * then pauses after every step that ends at an output (done = 1, output_due = 1, t and
 * iter advanced): the caller writes the output and starts the clock again. */
typedef struct {
    double t;        /* time of the current state */
    double dt;       /* step the next call will take */
    double t_end;
    double cfl;
    double min_eint; /* minimum internal energy seen by the last step (interior) */
    double mass;     /* integrals over the interior of the last state produced */
    double momentum[3];
    double energy;
    int iter;
    int max_iter;
    int done;
    int status;          /* 0, HCLPP_ERR_STATE (invalid internal energy) or HCLPP_ERR_NEGATIVE_DT */
    int iter_output_id;  /* output counters of main.cpp:362-375 (-1 before the first output) */
    int time_output_id;
    int output_due;      /* the last step ended at an output (main.cpp:533-557: make_output) */
    int exiting;         /* the last step was the final one (t_end or max_iter reached) */
} hclpp_clock;

/* status of a clock whose next output time is not after the current time
 * (main.cpp:369-371: "Error: the time step is negative") */
#define HCLPP_ERR_NEGATIVE_DT 5

/* The pressure fix of main.cpp:459-475 (hydro/pressure_fix.hpp:19-172, "Pressure fix" = On,
 * eps_pf) in every following hclpp_step / hclpp_step_fx: after the update, the energy of each
 * cell whose state at time n has E - e_k < eps e_k is rewritten from the step's face states, as
 * hclpp_pressure_fix does, and the CFL rate and energy total take the rewritten energies.  FP64
 * storage only; on = 0 turns it off.  The reverse step refuses a context with the fix on. */
int hclpp_set_pressure_fix(hclpp_ctx* ctx, int on, double eps);

/* Opt-in fast FP64 arithmetic for the fused step (on = 1; on = 0 restores the reference
 * arithmetic).  No counterpart in the reference, whose arithmetic hclpp_step reproduces bit for
 * bit; this mode keeps the FP64 state and every formula of the step but not their rounding:
 * P = (gamma-1) e without the temperature round trip through kb and mp (perfect_gas.hpp:46-78),
 * divisions as a product with a shared reciprocal, the slope limiters (slope_limiters.hpp:36-86)
 * in closed form, and fused multiply-adds.  Its results drift from the reference's by rounding,
 * growing with the step count.  Affects hclpp_step and hclpp_step_dt on an FP64 context; the
 * operator entry points keep the reference arithmetic.  HCLPP_ERR_UNSUPPORTED with passive
 * scalars (nfx > 0) or the pressure fix on. */
int hclpp_set_fast_arithmetic(hclpp_ctx* ctx, int on);

/* Output schedule of main.cpp:362-375 (utils/param.cpp:44-47), applied from the next
 * hclpp_clock_start on: iteration outputs every iter_output_frequency steps (0: none), time
 * outputs at time_first_output + k time_output_frequency (frequency 0: none), and the output
 * counters to continue from (-1 for a new run; a restart resumes the file's counters).  The
 * final step always ends at an output, as in the reference. */
int hclpp_clock_set_outputs(hclpp_ctx* ctx, int iter_output_frequency, double time_first_output, double time_output_frequency,
                            int iter_output_id, int time_output_id);

/* Start the clock on a ghost-filled conservative state (FP64 arrays) at time t0 and
 * iteration iter0 (0 for a new run; a restart resumes the iteration count, main.cpp:223-238).
 * The run stops when t reaches t_end or the iteration count reaches max_iter, counted from
 * iteration 0 as in main.cpp:382-385. */
int hclpp_clock_start(hclpp_ctx* ctx, double t0, int iter0, double t_end, double cfl, int max_iter, const double* rho,
                      const double* rhou, const double* E);
/* Continue a clock paused at an output (hclpp_clock.output_due) on the state it stopped at:
 * t, iter, the output counters and the schedule carry over and dt is computed from the state,
 * as the reference's next loop iteration does (main.cpp:355).  HCLPP_ERR_STATE after an error
 * status, HCLPP_ERR_ARG once the run has ended. */
int hclpp_clock_resume(hclpp_ctx* ctx, const double* rho, const double* rhou, const double* E);
int hclpp_clock_resume_f32(hclpp_ctx* ctx, const float* rho, const float* rhou, const float* E);
/* Copy the device clock to the host (synchronises). */
int hclpp_clock_read(hclpp_ctx* ctx, hclpp_clock* out);
/* One fused step: reads (rho, rhou, E) ghost-filled, writes the new state,
 * ghosts included, to the *_new arrays (which must not alias the inputs).
 * FP64 storage; HCLPP_ERR_UNSUPPORTED when the context has nfx > 0 (see hclpp_step_fx). */
int hclpp_step(hclpp_ctx* ctx, const double* rho, const double* rhou, const double* E, double* rho_new, double* rhou_new,
               double* E_new);
/* hclpp_step carrying the context's passive scalars (nfx >= 1; fx and fx_new hold nfx arrays of
 * mass fractions, layout [f][cell], ghosts filled on input as hclpp_fill_ghosts leaves them):
 * the main loop body of main.cpp:426-457 with nfx > 0, i.e. each scalar reconstructed with the
 * state (hydro_reconstruction.hpp:99-109), carried through the Hancock half step as a partial
 * density (extrapolation_reconstruction.hpp:87-88, 105-108, 315-337), upwinded with the mass
 * flux and clamped to [0, 1] in the update (godunov_scheme.hpp:122, 263-301).  The fused kernel
 * carries one scalar per launch: nfx > 1 launches it nfx times (the first also writes the
 * hydro update), so the hydro arithmetic is repeated per scalar. */
int hclpp_step_fx(hclpp_ctx* ctx, const double* rho, const double* rhou, const double* E, const double* fx, double* rho_new,
                  double* rhou_new, double* E_new, double* fx_new);

/* ---- FP32 storage (mixed precision) fused step ----
 * Same step as hclpp_step with the state stored and streamed in FP32 (half
 * the HBM bytes) and FP32 arithmetic, except the pressure recovery
 * (gamma-1)(E - |m|^2 / 2 rho), the CFL reduction and the conserved totals,
 * which are FP64.  The context must be created with precision =
 * HCLPP_FP32_STORAGE.  Not bit-identical to the FP64 path: the accuracy is
 * stated as an L1 distance to it. */
int hclpp_clock_start_f32(hclpp_ctx* ctx, double t0, int iter0, double t_end, double cfl, int max_iter, const float* rho,
                          const float* rhou, const float* E);
int hclpp_step_f32(hclpp_ctx* ctx, const float* rho, const float* rhou, const float* E, float* rho_new, float* rhou_new, float* E_new);
/* hclpp_step_fx with FP32 storage (the scalar is stored and computed in FP32). */
int hclpp_step_fx_f32(hclpp_ctx* ctx, const float* rho, const float* rhou, const float* E, const float* fx, float* rho_new,
                      float* rhou_new, float* E_new, float* fx_new);
/* element-wise conversions between FP64 and FP32 device arrays of n values */
int hclpp_cast_f64_to_f32(hclpp_ctx* ctx, const double* in, float* out, long long n);
int hclpp_cast_f32_to_f64(hclpp_ctx* ctx, const float* in, double* out, long long n);

/* ---- reverse mode (no counterpart in the reference: the derivative of hclpp_step) ----
 * The step is differentiated with dt held fixed; the caller replays the forward dt
 * schedule (recorded with hclpp_clock_record_dt) through hclpp_step_dt.  Cartesian or
 * spherical geometry (the geometric sources of the predictor and of the update are
 * differentiated), HLL or HLLC, any limiter, uniform or point-mass gravity, any boundary
 * kinds, FP64 storage, nfx = 0, 2 <= ng <= 4; otherwise HCLPP_ERR_UNSUPPORTED. */

/* hclpp_step with the step length dt instead of the clock's CFL dt (t and iter advance
 * by it; the t_end / max_iter exit rules apply as in hclpp_step). */
int hclpp_step_dt(hclpp_ctx* ctx, double dt, const double* rho, const double* rhou, const double* E, double* rho_new,
                  double* rhou_new, double* E_new);
/* From the next step on, each step stores its dt at log[iter] (device array of cap doubles,
 * iter < cap); log = NULL stops recording.  hclpp_clock_start resets it. */
int hclpp_clock_record_dt(hclpp_ctx* ctx, double* log, int cap);
/* Vector-Jacobian product of one step U -> U' with step length dt:
 *   (rho, rhou, E)            the step's input, ghosts filled (as hclpp_step takes it);
 *   (lrho_new, lrhou_new, lE_new)  dL/dU' on the interior (ghost entries are not read);
 *   (lrho, lrhou, lE)         receive dL/dU on the interior (ghost entries are not written);
 *   dgamma                    device double; dL/dgamma of this step is added to it.
 * Deterministic: repeated calls give bit-identical results.  Scratch of
 * (2 ndim + 2)(ndim + 2) doubles per cell is allocated on the first call. */
int hclpp_step_adjoint(hclpp_ctx* ctx, double dt, const double* rho, const double* rhou, const double* E, const double* lrho_new,
                       const double* lrhou_new, const double* lE_new, double* lrho, double* lrhou, double* lE, double* dgamma);

/* Batched forms (one call per run of steps, to keep host overhead off the critical path):
 * hclpp_steps_dt runs n steps, step k reading state k and writing state k + 1, with the
 * state pointers given as arrays of n + 1 entries per variable (entries k and k + 1 must
 * differ; ping-pong buffers are fine).  hclpp_steps_adjoint runs the n reverse steps from
 * last to first: reverse step k uses state k (arrays of n entries) and maps adjoint k + 1
 * to adjoint k (arrays of n + 1 entries); every step adds to *dgamma.  On an error return,
 * the steps already run are complete, their adjoints written and their *dgamma terms added. */
int hclpp_steps_dt(hclpp_ctx* ctx, int n, const double* dts, double* const* rho, double* const* rhou, double* const* E);
int hclpp_steps_adjoint(hclpp_ctx* ctx, int n, const double* dts, const double* const* rho, const double* const* rhou,
                        const double* const* E, double* const* lrho, double* const* lrhou, double* const* lE, double* dgamma);

/* ---- instrumentation (no reference counterpart; for benchmarks) ----
 * Between begin and end, each hclpp_step records a HIP event pair around its
 * fused kernel on the context's stream (at most max_steps pairs).  end
 * synchronises and returns the summed kernel time and the number of timed
 * launches. */
int hclpp_profile_begin(hclpp_ctx* ctx, int max_steps);
int hclpp_profile_end(hclpp_ctx* ctx, double* kernel_ms, int* launches);
/* Launch plan of the fused kernel chosen at context creation: planes per workgroup segment,
 * workgroups per launch, and workgroups of that kernel resident per CU. */
int hclpp_fused_plan(hclpp_ctx* ctx, int* seglen, long long* workgroups, int* blocks_per_cu);

#ifdef __cplusplus
}
#endif
#endif
