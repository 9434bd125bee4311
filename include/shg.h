/*
 * shg.h — C ABI of the MI355X focused-Gaussian-beam SHG split-step solver.
 *
 * The reference program has no library interface; its solve sits behind
 *   (1) the CLI  `SHG <P W> <waist um> <T1 C> <T2 C> <save_only_last> <save_temperature>`
 *       (src/SHG.cpp:37-50), and
 *   (2) the in-process C++ call
 *       Solver::run(Power, waist, focalpoint, Temp, T_inf, Tpeltier1, Tpeltier2,
 *                   save_only_last, save_temperature)            (src/headers/Solver.h:152,279)
 *       over raw host arrays Efields::Pump/Signal (src/headers/Efields.h:150).
 * Every entry point below states which part of that path it replaces.
 *
 * Conventions
 *   - Plain C linkage, POD structs, no exceptions cross the boundary.
 *   - Every function returns SHG_OK (0) or a negative SHG_E* code; the message of
 *     the last failure on the calling thread is returned by shg_last_error().
 *   - Complex planes are interleaved complex64 (re, im float pairs, layout of
 *     std::complex<float> / fftwf_complex), indexed IDX(x,y) = y*nx + x, x fastest
 *     (src/headers/Common.h:27-29).  Host buffers are owned by the caller and copied.
 *   - A context owns all of its device memory and one HIP stream.  Calls on one
 *     context are not reentrant; use one context per host thread.  Contexts on
 *     different threads run concurrently: the library copies only through its own
 *     non-blocking streams, so one thread's copies never break another's graph
 *     capture (SHG_sweep runs THERMAL points this way).
 *   - Units are the reference's: um, ps, V, W, degC.  Fields are in V/um.
 */
#ifndef SHG_H
#define SHG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SHG_OK              0
#define SHG_E_INVALID     (-1)   /* bad argument or unsupported parameter */
#define SHG_E_DEVICE      (-2)   /* HIP runtime error */
#define SHG_E_NOMEM       (-3)   /* device or host allocation failed */
#define SHG_E_STATE       (-4)   /* call out of order (e.g. forward before set_fields) */
#define SHG_E_NONFINITE   (-5)   /* NaN/Inf found in a field */
#define SHG_E_NOCONVERGE  (-6)   /* Jacobi hit its sweep cap before the tolerance */
#define SHG_E_INTERNAL    (-7)   /* unexpected host-side failure (message says which) */

#define SHG_CRYSTAL_SPPLT   0    /* MgO:sPPLT, src/headers/Crystal.h:55-71 */
#define SHG_CRYSTAL_PPLN    1    /* MgO:PPLN (Gayer et al. 2008); absent from the reference */

/* Heat-equation boundary (THERMAL).  TOP_OPEN: heatEquationTopOpen with the two ovens
 * under y = 0 (src/headers/Tfield.h:48-68, 113-221; the reference's run).  OVEN_SURROUNDED:
 * kernelHeatEquationOvenSurrounded with the ovens on the four lateral faces
 * (src/headers/Tfield.h:71-110, 224-258; reachable in the reference only through the
 * commented call at src/headers/Solver.h:290). */
#define SHG_BC_TOP_OPEN        0
#define SHG_BC_OVEN_SURROUNDED 1

#define SHG_SCHEME_REFERENCE 0   /* RK4(dz) - L(dz) - RK4(dz), src/headers/Solver.h:236-242 */
#define SHG_SCHEME_STRANG    1   /* RK4(dz/2) - L(dz) - RK4(dz/2) */

/* Every constant the reference hard-codes (src/SHG.cpp:7-20, src/headers/Crystal.h:18-53)
 * plus the argv values (src/SHG.cpp:37-50).  shg_default_params() fills the
 * reference's values; the solver rounds them to float exactly as the reference's
 * `const real_t` declarations do. */
typedef struct shg_params {
    int32_t nx, ny, nz;        /* transverse grid (power of two, 64..512, nx == ny) and z points */
    int32_t crystal;           /* SHG_CRYSTAL_* */
    int32_t thermal;           /* 0: plain SSFM, 1: THERMAL (heat equation + dk(T) feedback) */
    int32_t scheme;            /* SHG_SCHEME_* */
    double  lx, ly, lz;        /* crystal box [um]; dx = lx/(nx-1), dz = lz/(nz-1) */
    double  lambda_p;          /* pump wavelength [um]; SH at lambda_p/2 */
    double  deff;              /* [um/V] */
    double  qpm_period;        /* grating period Lambda [um] */
    double  alpha_p, alpha_s;  /* linear absorption [1/um] */
    double  beta_s;            /* SH two-photon absorption [um/W] */
    double  rho_p, rho_s;      /* walk-off angles [rad] */
    double  k_th;              /* thermal conductivity [W/(um K)] */
    double  h_conv;            /* heat-transfer coefficient [W/(um^2 K)] */
    double  power;             /* pump power [W]            (argv[1]) */
    double  waist;             /* pump waist [um]           (argv[2]) */
    double  focus;             /* focal position [um]       (reference: 0.5*lz) */
    double  temp;              /* crystal temperature [C]   (argv[3]) */
    double  t_inf;             /* ambient temperature [C]   (reference: 25) */
    double  t_peltier1;        /* oven under z < L/2 [C]    (argv[3]) */
    double  t_peltier2;        /* oven under z >= L/2 [C]   (argv[4]) */
    /* THERMAL solver controls (src/headers/Solver.h:295-336) */
    int32_t heat_nz;           /* heat-grid z points; 0 = nz (the reference's coupling) */
    int32_t jacobi_check;      /* residual check interval; reference 5000 */
    int32_t max_outer;         /* outer fixed-point iterations; reference 5 */
    int32_t jacobi_max;        /* cap on sweeps per outer iteration; 0 = no cap */
    double  jacobi_tol;        /* mean squared update tolerance; 0 = reference 5e-4/(nx*ny*heat_nz) */
    int32_t thermal_bc;        /* SHG_BC_* (default TOP_OPEN) */
} shg_params;

typedef struct shg_ctx shg_ctx;

/* Reference constants and the SHG.sh default run (src/SHG.sh:31-37). */
void shg_default_params(shg_params *p);

/* Last error message on this thread ("" if none). */
const char *shg_last_error(void);

/* Replaces `new Efields/Tfield/PhaseMatching` + the constant setup in Solver::run
 * (src/headers/Solver.h:283-287, 343-347): validates the grid, evaluates the crystal
 * constants with the reference's float/double mix, builds the propagators, the
 * per-step phase table and the twiddles, and allocates the device state. */
int shg_create(const shg_params *p, shg_ctx **out);
void shg_destroy(shg_ctx *ctx);

/* Reference initial pump (setPump, src/headers/Efields.h:74-104) evaluated on the
 * host in the reference's arithmetic, written to `fh` (nx*ny complex64).  `t_centre`
 * is the crystal temperature used for n_p (reference: Tfinal at the grid centre). */
int shg_make_pump(const shg_params *p, double t_centre, float *fh);

/* Replaces setInputPump/setNoisyField (src/headers/Efields.h:174-184): copies the
 * caller's z = 0 fields (FH = pump, SH = signal) into the context. */
int shg_set_fields(shg_ctx *ctx, const float *fh, const float *sh);

/* Replaces Solver::CWES (src/headers/Solver.h:245-276): marches the shg_set_fields
 * fields through the crystal on the GPU.  Afterwards the context holds the state the
 * reference writes as Pump_out / Signal_out (after nz-1 steps, z = lz).  A plain
 * context executes nz-1 steps (the reference's last step has no observable effect); a
 * THERMAL context executes all nz steps with the current temperature, updating the heat
 * source after every step as the reference does. */
int shg_forward(shg_ctx *ctx);

/* Restores the device fields to the last shg_set_fields state (z = 0). */
int shg_reset_fields(shg_ctx *ctx);

/* Advances the current device fields by z-steps [first_step, first_step + nsteps) of the
 * reference's z-loop (Solver::SSFM per step, src/headers/Solver.h:236-242, 265), with
 * first_step + nsteps <= nz - 1.  Used for the reference's per-slice output mode
 * (save_only_last = 0, src/headers/Solver.h:257-262), which saves the fields before
 * every step. */
int shg_march(shg_ctx *ctx, int first_step, int nsteps);

/* Copies the current device fields to caller buffers (nx*ny complex64 each). */
int shg_get_fields(shg_ctx *ctx, float *fh, float *sh);

/* The save_only_last = 0 output of CWES (src/headers/Solver.h:257-262): runs the march of
 * shg_forward from the shg_set_fields state (THERMAL: at the current temperature) and
 * writes the fields before every step s = 0 .. nz-1 (the reference's Pump_<s>,
 * Signal_<s>) to fh_steps + 2*s*nx*ny and sh_steps + 2*s*nx*ny (complex64).  Holds
 * nz*nx*ny*16 bytes of device memory during the call.  Leaves the context as
 * shg_forward does. */
int shg_forward_snapshots(shg_ctx *ctx, float *fh_steps, float *sh_steps);

/* FP64 powers: P_j = 0.5*eps0*c*n_j*sum|A_j|^2*dx*dy for the z = 0 pump and the
 * current SH, and eta = P_SH / P_FH(z=0).  New: the reference never reduces powers. */
int shg_powers(shg_ctx *ctx, double *p_fh_in, double *p_sh_out, double *eta);

/* Milliseconds of GPU time of the last shg_forward (HIP events on the context stream). */
int shg_last_forward_ms(shg_ctx *ctx, double *ms);

/* Number of transverse cell-steps executed by one shg_forward (nx*ny*steps). */
int64_t shg_cell_steps(const shg_ctx *ctx);

/* ----- THERMAL build (src/headers/Solver.h:289-340, src/headers/Tfield.h) ------------- */

typedef struct shg_thermal_stats {
    int32_t outer_iterations;  /* global iterations run (reference: at most 5) */
    int32_t reserved;
    int64_t sweeps_last;       /* Jacobi sweeps of the last outer iteration */
    int64_t sweeps_total;
    double  residual_last;     /* mean squared update at the last residual check (FP64) */
    double  jacobi_ms;         /* device time of all Jacobi sweeps */
    double  march_ms;          /* device time of all z-marches */
} shg_thermal_stats;

/* Replaces the THERMAL branch of Solver::run (src/headers/Solver.h:289-340) on a context
 * created with thermal = 1: temperature initialised to `temp` with the bottom ovens
 * (Tfield.h:31-68) and Q = 0, then up to max_outer iterations of [Jacobi to the tolerance
 * with the residual checked every jacobi_check sweeps and the reference's stagnation and
 * early-stop rules; dk(T) and the z-march with the heat source updated after every step].
 * The SH input is the one given to shg_set_fields; the pump is rebuilt each iteration by
 * setPump at the grid-centre temperature when pump_from_temperature != 0 (as the
 * reference does), else the shg_set_fields pump is reused.  Afterwards shg_get_fields
 * returns the Pump_out/Signal_out state of the last iteration. */
int shg_thermal_run(shg_ctx *ctx, int pump_from_temperature, shg_thermal_stats *stats);

/* Progress of shg_thermal_run, reported at the points where the reference's THERMAL run
 * prints to stdout (src/headers/Solver.h:299-333): the "Global iteration #" banner and
 * its tolerance, the residual line at every check, the "N iterations -> steady-state."
 * line and the start of the CWES march.  The callback runs on the calling thread, inside
 * shg_thermal_run; it must not call into the same context. */
#define SHG_THERMAL_EVENT_OUTER   0  /* outer iteration `outer` begins; value = Jacobi tolerance */
#define SHG_THERMAL_EVENT_CHECK   1  /* residual check after sweep number `sweeps` (0, check, 2 check, ...);
                                        value = mean squared update of that sweep (FP64) */
#define SHG_THERMAL_EVENT_STEADY  2  /* the Jacobi of this iteration stopped after `sweeps` sweeps;
                                        value = the last residual */
#define SHG_THERMAL_EVENT_MARCH   3  /* the z-march begins; value = grid-centre temperature of slice 0 */
typedef void (*shg_thermal_log_fn)(void *user, int event, int32_t outer, int64_t sweeps, double value);

/* Installs (fn != NULL) or removes the progress callback of shg_thermal_run. */
int shg_set_thermal_log(shg_ctx *ctx, shg_thermal_log_fn fn, void *user);

/* `nsweeps` Jacobi sweeps of heatEquationTopOpen (src/headers/Tfield.h:113-221) on the
 * current temperature and heat source; *mean_sq_update (if non-null) receives
 * sum((Tf-Ti)^2)/(nx*ny*heat_nz) of the last sweep in FP64 (reducedSum, Tfield.h:13-28). */
int shg_jacobi(shg_ctx *ctx, int64_t nsweeps, double *mean_sq_update);

/* No counterpart in the reference: shg_jacobi for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * Runs `nsweeps` Jacobi sweeps of the heat grid described by `p` (nx == ny in 64..512,
 * heat_nz or nz slices, the box, k_th, h_conv, t_inf, thermal_bc) on caller-owned device
 * buffers, with no context: T (nx*ny*heat_nz floats, IDX layout) in and out, T_work the
 * same size (clobbered), Q the heat source.  `layout` is a Jacobi layout of
 * shg_set_jacobi_layout, or -1 for single sweeps; every layout gives the same bits.  If
 * `mean_sq_update` (a device double) is not NULL it receives the last sweep's mean
 * squared update, as shg_jacobi returns it.  Work is enqueued on `stream` (a
 * hipStream_t; NULL: the null stream) and the call returns without waiting.  It calls no
 * allocating routine: the residual's partial sums live in a fixed 256 KiB array of the
 * library on each device (resolved on the device's first call), which limits the heat
 * grid to 2^15 single-sweep partials, 512 x 512 x 256 voxels (larger grids with a
 * residual are rejected with SHG_E_INVALID).  Every call on a device shares that array:
 * calls that overlap on one GPU (other threads, other streams) race. */
int shg_jacobi_noalloc(const shg_params *p, float *T, float *T_work, const float *Q, int64_t nsweeps, int layout,
                       double *mean_sq_update, void *stream);

/* Temperature and heat source (nx*ny*heat_nz floats, IDX(x,y,z) = z*nx*ny + y*nx + x).
 * shg_set_temperature also rebuilds the propagators at the grid-centre temperature. */
int shg_get_temperature(shg_ctx *ctx, float *T);
int shg_set_temperature(shg_ctx *ctx, const float *T);
int shg_get_heat_source(shg_ctx *ctx, float *Q);
int shg_set_heat_source(shg_ctx *ctx, const float *Q);

/* ----- Gradients (new: the reference has none) --------------------------------------- */

typedef struct shg_adjoint_result {
    double  eta;                    /* P_SH(out) / P_FH(z = 0) of the forward solve */
    double  d_eta_d_focus;          /* through the reference pump setPump (Efields.h:74-104), per um */
    double  d_eta_d_temp;           /* total, per degC: crystal path + pump path */
    double  d_eta_d_temp_crystal;   /* crystal path only (coupling, phase mismatch, TPA, propagator,
                                       the n of both powers) with the z = 0 fields held fixed */
    double  forward_ms;             /* device time of the checkpointed forward */
    double  adjoint_ms;             /* device time of the taped recomputation + adjoint march */
    int32_t checkpoint_interval;    /* K: a checkpoint every K steps */
    int32_t segments;
} shg_adjoint_result;

/* Discrete adjoint of the z-march (Solver::CWES, src/headers/Solver.h:245-276) for
 * J = eta.  Checkpoints every `checkpoint_interval` steps, with per-segment recomputation;
 * <= 0 tapes the whole march once (no recomputation) when that tape, about
 * (nz-1)*nx*ny*32 bytes (36 in THERMAL), takes at most a quarter of the free device memory,
 * and otherwise checkpoints every ceil(sqrt(nz-1)) steps; res->checkpoint_interval reports
 * the choice.  d_eta_d_fh (nx*ny complex64, may be null) receives the
 * gradient with respect to the z = 0 pump field, dJ = Re sum conj(g) dA.  The focus and
 * the pump part of the temperature gradient assume the z = 0 pump is setPump at the
 * context's power, waist, focus and temperature (plus additive noise).
 * On a THERMAL context the march runs at the context's current temperature field, which
 * is held fixed (the heat source is not updated); the temperature derivatives are those
 * of a uniform shift of that field (the propagators, powers and pump at the grid-centre
 * temperature of slice 0, as the forward evaluates them). */
int shg_adjoint(shg_ctx *ctx, int checkpoint_interval, float *d_eta_d_fh, shg_adjoint_result *res);

/* The field part of shg_adjoint only (no reference counterpart): eta and d eta / d FH(z=0)
 * into `d_eta_d_fh` (required) — eta exactly, the gradient to rounding (~1e-6) of
 * shg_adjoint's — without forming the focus and temperature sensitivities (nor, THERMAL, the dJ/dT map, whose
 * shg_adjoint_temperature_map is then unavailable).  For callers that need the field
 * gradient alone, such as gradient ascent of a pump phase mask. */
int shg_adjoint_field(shg_ctx *ctx, int checkpoint_interval, float *d_eta_d_fh, double *eta);

/* After shg_adjoint on a THERMAL context (no reference counterpart): d eta / d T of every
 * heat voxel (nx*ny*heat_nz doubles, IDX layout) with the rest of the field held fixed,
 * crystal path only (coupling, phase mismatch through DKint, TPA; the propagator and
 * power terms sit on the grid-centre voxel of slice 0).  Sums to d_eta_d_temp_crystal. */
int shg_adjoint_temperature_map(shg_ctx *ctx, double *map);

/* ----- Batched sweeps ---------------------------------------------------------------
 * Replace the reference's sweep driver, which runs one `./SHG` process per parameter
 * point (src/SHG.sh:39-81): `count` independent plain instances (each its own grid size
 * 64..512, nz, waist, power, temperature, focus, ...) march in one device schedule
 * captured in a hipGraph.  Instances are addressed by their index in `params`. */
typedef struct shg_batch shg_batch;

int shg_batch_create(const shg_params *params, int count, shg_batch **out);
void shg_batch_destroy(shg_batch *b);

/* z = 0 fields of every instance on the device: the reference pump (setPump,
 * src/headers/Efields.h:74-104) plus complex Gaussian noise of standard deviation
 * noise_rel * |A(0,0)| per component (seeded, reproducible), SH = 0. */
int shg_batch_init_fields(shg_batch *b, double noise_rel, uint64_t seed);

/* Caller-supplied z = 0 fields of one instance (nx*ny complex64 each). */
int shg_batch_set_fields(shg_batch *b, int index, const float *fh, const float *sh);

/* Marches every instance through its crystal; afterwards each holds its Pump_out /
 * Signal_out state (after its own nz-1 steps). */
int shg_batch_run(shg_batch *b);

/* Schedule of shg_batch_run (no counterpart in the reference): every grid size is cut
 * into groups of at most group_mib MiB of fields, each group marches all its steps
 * before the next one of its stream (its fields stay in the last-level cache), and the
 * groups are spread over `streams` (1..4) concurrent streams.  Defaults: 128 MiB, 3. */
int shg_batch_set_schedule(shg_batch *b, int group_mib, int streams);

/* Current fields (after shg_batch_run: the outputs) and the z = 0 fields of one instance. */
int shg_batch_get_fields(shg_batch *b, int index, float *fh, float *sh);
int shg_batch_get_input(shg_batch *b, int index, float *fh, float *sh);

/* FP64 powers and efficiency per instance (arrays of `count`; any may be null). */
int shg_batch_powers(shg_batch *b, double *p_fh_in, double *p_sh_out, double *eta);

/* Sum over instances of nx*ny*(nz-1), and the device time of the last shg_batch_run. */
int64_t shg_batch_cell_steps(const shg_batch *b);
int shg_batch_last_ms(shg_batch *b, double *ms);

/* ----- Batched THERMAL sweeps ----------------------------------------------------------
 * Replace the sweep driver's per-point `./SHG` processes (src/SHG.sh:39-81) for a THERMAL
 * build: `count` instances of the THERMAL branch of Solver::run
 * (src/headers/Solver.h:289-340) that share the crystal, grid and heat-equation constants
 * and differ only in power, waist, focus, temp, t_peltier1 and t_peltier2 (others are
 * rejected with SHG_E_INVALID).  Their outer loops run in lockstep on one device state:
 * one Jacobi launch sweeps every instance still converging, every instance is checked at
 * the sweep counts its own run checks at and stops by its own residual, stagnation,
 * sweep-cap and early-stop rules, and the z-marches of all instances that march in an
 * outer iteration share each launch.  Each instance's result equals shg_thermal_run's on
 * a context with the same parameters (pump from the temperature), bit for bit with the
 * same row and column layouts. */
typedef struct shg_thermal_batch shg_thermal_batch;

int shg_thermal_batch_create(const shg_params *params, int count, shg_thermal_batch **out);
void shg_thermal_batch_destroy(shg_thermal_batch *b);

/* SH input (z = 0, nx*ny complex64) of one instance; zero until set. */
int shg_thermal_batch_set_signal(shg_thermal_batch *b, int index, const float *sh);

/* Row (0..6) and column (0..9) layouts of the batched march; tuned on the first run
 * otherwise (SHG_NO_TUNE=1: 0, 0).  Get returns -1, -1 while untuned. */
int shg_thermal_batch_set_variants(shg_thermal_batch *b, int row_variant, int col_variant);
int shg_thermal_batch_get_variants(const shg_thermal_batch *b, int *row_variant, int *col_variant);

/* Jacobi layout of the batched sweeps (0..SHG_JACOBI_LAYOUTS-1, as shg_set_jacobi_layout:
 * every layout gives the same sweeps bit for bit); tuned on the first run otherwise. */
int shg_thermal_batch_set_jacobi_layout(shg_thermal_batch *b, int layout);
int shg_thermal_batch_get_jacobi_layout(const shg_thermal_batch *b, int *layout);

/* Runs every instance (stats: array of `count`, may be null).  jacobi_ms and march_ms are
 * the batch's device times, the same in every entry. */
int shg_thermal_batch_run(shg_thermal_batch *b, shg_thermal_stats *stats);

/* After shg_thermal_batch_run: one instance's Pump_out/Signal_out fields, its temperature
 * (nx*ny*heat_nz floats), and the FP64 powers and efficiency of every instance (arrays of
 * `count`; any may be null). */
int shg_thermal_batch_get_fields(shg_thermal_batch *b, int index, float *fh, float *sh);
int shg_thermal_batch_get_temperature(shg_thermal_batch *b, int index, float *T);
int shg_thermal_batch_powers(shg_thermal_batch *b, double *p_fh_in, double *p_sh_out, double *eta);

/* ----- Launch-shape autotuning (no reference counterpart) ------------------------------
 * The row and column passes exist in several thread layouts (FFT elements per thread,
 * rows or columns per workgroup).  shg_forward / shg_batch_run tune them on first use
 * (timing each (row, column) pair, replayed as the march runs them, on the context's own
 * data) unless the environment sets SHG_NO_TUNE=1; these calls tune explicitly or pin a
 * choice.  Rows 0..4 (THERMAL contexts: 0..6), columns 0..9. */
int shg_tune(shg_ctx *ctx, int reps, int *row_variant, int *col_variant);
int shg_set_variants(shg_ctx *ctx, int row_variant, int col_variant);
/* The layouts the next march runs (-1 while untuned). */
int shg_get_variants(const shg_ctx *ctx, int *row_variant, int *col_variant);

/* Layout of the THERMAL Jacobi sweeps (no reference counterpart; the reference's sweep is
 * heatEquationTopOpen, src/headers/Tfield.h:113-221).  Sweeps run two per pass, marching
 * in z through LDS tiles; layouts 0..8 are 3 tile shapes x 3 workgroup-wave counts (the
 * z-chunk depth follows from the device's resident slots), 9 is two single sweeps.  Every layout gives the reference's float sweeps bit for bit.  shg_jacobi and
 * shg_thermal_run tune it on first use (SHG_NO_TUNE=1: layout 0); -1 while untuned. */
#define SHG_JACOBI_LAYOUTS 10
int shg_set_jacobi_layout(shg_ctx *ctx, int layout);
int shg_get_jacobi_layout(const shg_ctx *ctx, int *layout);

/* Measurement hook (no reference counterpart): average device time of one row pass and
 * one column pass (the tuned layouts), each replayed `reps` times back to back from a
 * graph (the fastest of three replays).  Clobbers the current fields; the next shg_forward starts again from the
 * shg_set_fields state. */
int shg_profile_kernels(shg_ctx *ctx, int reps, double *row_ms, double *col_ms);

#ifdef __cplusplus
}
#endif

#endif /* SHG_H */
