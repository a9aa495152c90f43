/*
 * clode_hip.h - C ABI of the MI355X ensemble integrator (libclode_hip.so).
 *
 * One opaque handle per simulator, plain pointers and sizes, an int status on
 * every call and a thread-local last-error string.  The entry points replace
 * the C++ classes that the reference's bindings call; each declaration names
 * the reference member it stands in for.  Reference paths are relative to
 * the clODE repository (patrickfletcher/clODE):
 *   clode/cpp/CLODE.hpp, CLODE.cpp                 (base simulator)
 *   clode/cpp/CLODEfeatures.hpp, CLODEfeatures.cpp (feature simulator)
 *   clode/cpp/CLODEtrajectory.hpp, CLODEtrajectory.cpp
 *   clode/cpp/CLODEpython.cpp                      (pybind11 module)
 *   matlab/clODEmex.cpp, clODEfeaturesmex.cpp, clODEtrajectorymex.cpp
 *
 * Arrays cross the boundary as double in the reference's column-major
 * structure-of-arrays layout: x0/xf [nVar][nPts], pars [nPar][nPts],
 * F [nFeatures][nPts], t [maxStore][nPts], x/dx [maxStore][nVar][nPts],
 * aux [maxStore][nAux][nPts] (1 element when nAux == 0), nStored [nPts].
 *
 * Status codes: CLODE_OK, CLODE_ERR_INVALID (bad argument; the reference
 * throws std::invalid_argument -> Python ValueError), CLODE_ERR_RUNTIME
 * (device / compiler failure; the reference rethrows cl::Error -> Python
 * RuntimeError).  clode_last_error() describes the last failure on the
 * calling thread.
 */
#ifndef CLODE_HIP_H
#define CLODE_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CLODE_OK 0
#define CLODE_ERR_INVALID 1
#define CLODE_ERR_RUNTIME 2

/* precision modes: SINGLE and DOUBLE follow clode/cpp/realtype.cl; FP32C is
 * float state with Kahan-compensated updates and double time/error norm */
#define CLODE_PREC_SINGLE 1
#define CLODE_PREC_DOUBLE 2
#define CLODE_PREC_FP32C 3

/* simulator kinds: CLODE, CLODEfeatures, CLODEtrajectory */
#define CLODE_KIND_BASE 0
#define CLODE_KIND_FEATURES 1
#define CLODE_KIND_TRAJECTORY 2

typedef struct clode_sim clode_sim;

/* SolverParams<double> (clode/cpp/clODE_struct_defs.cl:11-20) */
typedef struct clode_solver_params {
    double dt;
    double dtmax;
    double abstol;
    double reltol;
    unsigned int max_steps;
    unsigned int max_store;
    unsigned int nout;
} clode_solver_params;

/* ObserverParams<double> (clode/cpp/observers.cl:25-46) */
typedef struct clode_observer_params {
    unsigned int eVarIx;
    unsigned int fVarIx;
    unsigned int maxEventCount;
    unsigned int maxEventTimestamps;
    double minXamp;
    double minIMI;
    double nHoodRadius;
    double xUpThresh;
    double xDownThresh;
    double dxUpThresh;
    double dxDownThresh;
    double eps_dx;
} clode_observer_params;

/* ProblemInfo (clode/cpp/CLODE.hpp:22-91); names are '\n'-separated */
typedef struct clode_problem_info {
    const char *src_file;
    int num_var;
    int num_par;
    int num_aux;
    int num_noise;
    const char *var_names;
    const char *par_names;
    const char *aux_names;
} clode_problem_info;

/* device description (stands in for OpenCLResource / deviceInfo,
 * clode/cpp/OpenCLResource.hpp:33-71) */
typedef struct clode_device_info {
    char name[256];
    char arch[64];
    int compute_units;
    int max_clock_khz;
    int max_work_group_size;
    unsigned long long device_memory_size;
    unsigned long long max_memory_alloc_size;
    int double_support;
} clode_device_info;

const char *clode_last_error(void);
const char *clode_version(void);

/* ---- devices (OpenCLResource ctor + getDeviceInfo, OpenCLResource.cpp:20-128,367-419) */
int clode_device_count(int *count);
int clode_device_query(int device, clode_device_info *info);

/* ---- lifetime (CLODE::CLODE CLODE.cpp:13-37; CLODEfeatures::CLODEfeatures
 *      CLODEfeatures.cpp:9-29; CLODEtrajectory::CLODEtrajectory CLODEtrajectory.cpp:6-20).
 *      observer/op are used by CLODE_KIND_FEATURES only (may be NULL otherwise).
 *      clode_root: directory for the on-disk code-object cache (may be NULL). */
int clode_create(clode_sim **out, int kind, const clode_problem_info *pi, const char *stepper,
                 const char *observer, const clode_observer_params *op, int precision, int device,
                 const char *clode_root);
int clode_destroy(clode_sim *sim);

/* ---- setters that change the program (CLODE.hpp:148-154) */
int clode_set_problem_info(clode_sim *sim, const clode_problem_info *pi);
int clode_set_stepper(clode_sim *sim, const char *stepper);     /* unknown name: warning, unchanged */
int clode_set_precision(clode_sim *sim, int precision);
int clode_set_device(clode_sim *sim, int device);                /* setOpenCL */
int clode_build(clode_sim *sim);                                 /* buildCL */

/* ---- data (CLODE.hpp:157-164) */
int clode_set_problem_data(clode_sim *sim, const double *x0, size_t nx0, const double *pars, size_t npars);
int clode_set_tspan(clode_sim *sim, const double *tspan, size_t n);
int clode_set_x0(clode_sim *sim, const double *x0, size_t n);
int clode_set_pars(clode_sim *sim, const double *pars, size_t n);
int clode_set_solver_params(clode_sim *sim, const clode_solver_params *sp);
int clode_seed_rng(clode_sim *sim);                 /* seedRNG(): random_device */
int clode_seed_rng_int(clode_sim *sim, int seed);   /* seedRNG(int) */

/* ---- simulation (CLODE.hpp:167-174) */
int clode_transient(clode_sim *sim);
int clode_shift_tspan(clode_sim *sim);
int clode_shift_x0(clode_sim *sim);

/* ---- getters (CLODE.hpp:177-188); *_n return element counts */
int clode_get_n_pts(clode_sim *sim, int *n);
int clode_get_tspan(clode_sim *sim, double *tspan2);
int clode_get_solver_params(clode_sim *sim, clode_solver_params *sp);
int clode_get_pars(clode_sim *sim, double *out, size_t n);
int clode_get_x0(clode_sim *sim, double *out, size_t n);
int clode_get_xf(clode_sim *sim, double *out, size_t n);
int clode_get_dt(clode_sim *sim, double *out, size_t n);
int clode_get_tf(clode_sim *sim, double *out, size_t n);
/* string getters: copy up to cap bytes (NUL-terminated), *needed = full length + 1 */
int clode_get_available_steppers(clode_sim *sim, char *buf, size_t cap, size_t *needed);
int clode_get_program_string(clode_sim *sim, char *buf, size_t cap, size_t *needed);
int clode_get_problem_info(clode_sim *sim, int *nvar, int *npar, int *naux, int *nnoise);
int clode_get_precision(clode_sim *sim, int *precision);
int clode_get_stepper(clode_sim *sim, char *buf, size_t cap, size_t *needed);
int clode_print_status(clode_sim *sim);

/* ---- features (CLODEfeatures.hpp:54-75) */
int clode_set_observer_params(clode_sim *sim, const clode_observer_params *op);
int clode_set_observer(clode_sim *sim, const char *observer);   /* unknown name: warning, unchanged */
int clode_initialize_observer(clode_sim *sim);
int clode_is_observer_initialized(clode_sim *sim, int *initialized);
int clode_features(clode_sim *sim, int reinitialize);          /* reinitialize: -1 = features(), 0/1 = features(bool) */
int clode_get_observer_params(clode_sim *sim, clode_observer_params *op);
int clode_get_observer_name(clode_sim *sim, char *buf, size_t cap, size_t *needed);
int clode_get_n_features(clode_sim *sim, int *n);
int clode_get_f(clode_sim *sim, double *out, size_t n);
/* F member-major, out[member * nFeatures + feature] (n >= nFeatures * nPts):
 * the layout the reference's Python layer builds from getF by a host-side
 * reshape (features.py:467-507), transposed here on the device before the
 * copy to the host */
int clode_get_f_members(clode_sim *sim, double *out, size_t n);
int clode_get_feature_names(clode_sim *sim, char *buf, size_t cap, size_t *needed);
int clode_get_available_observers(clode_sim *sim, char *buf, size_t cap, size_t *needed);

/* transient over transient_tspan2, x0 <- xf, and features over
 * features_tspan2 with the observer re-initialised, in one kernel launch
 * (each member runs its three passes back to back).  No reference
 * counterpart: the reference's callers make the calls transient (CLODE.cpp:
 * 468-493), shiftX0 (:502-515) and features(true) (CLODEfeatures.cpp:215-259),
 * which this reproduces bit for bit.  The simulator's tspan is left at
 * features_tspan2.  Global spans only (CLODE_ERR_INVALID with per-member
 * spans set); always a grid launch. */
int clode_transient_features(clode_sim *sim, const double *transient_tspan2, const double *features_tspan2);

/* ---- trajectory (CLODEtrajectory.hpp:36-50) */
int clode_trajectory(clode_sim *sim);
int clode_get_trajectory_sizes(clode_sim *sim, size_t *nt, size_t *nx, size_t *naux);
int clode_get_t(clode_sim *sim, double *out, size_t n);
int clode_get_x(clode_sim *sim, double *out, size_t n);
int clode_get_dx(clode_sim *sim, double *out, size_t n);
int clode_get_aux(clode_sim *sim, double *out, size_t n);
int clode_get_n_stored(clode_sim *sim, int *out, size_t n);

/* ---- logging (CLODEpython.cpp:43-95 LogLevel / get_logger):
 *      levels 0 trace, 1 debug, 2 info, 3 warn, 4 err, 5 critical, 6 off */
int clode_set_log_level(int level);
int clode_get_log_level(int *level);

/* ---- extensions with no reference counterpart ----------------------- */
/* Work distribution for every launch: mode 0 = one member per lane in a 1-D
 * grid (the reference's NDRange), 1 = persistent work queue: as many
 * workgroups as fit on the device at once, whose lanes claim the next member
 * with an atomic add as soon as their member finishes.  Results are identical
 * in both modes.  block_size 0 = 64. */
int clode_set_launch(clode_sim *sim, int mode, int block_size);
/* Work-queue mode: loop iterations a lane runs between queue checks
 * (default 8). */
int clode_set_queue_slice(clode_sim *sim, int slice);
/* Per-member time spans: tspan_m is [2][nPts] (start row, then end row) or
 * NULL to use the global span again.  Every launch honours them; the work
 * queue then hands out members longest span first. */
int clode_set_member_tspan(clode_sim *sim, const double *tspan_m, size_t n);
/* Refill the device step-size buffer with sp.dt on set_solver_params (the
 * reference refills it only when nPts changes). */
int clode_set_dt_refresh(clode_sim *sim, int enable);
/* Time of the last kernel launch on the simulator's stream, milliseconds
 * (HIP events around the launch). */
int clode_last_kernel_ms(clode_sim *sim, double *ms);
/* Which compilation of the program the last launch used: 0 for the default,
 * w > 0 for the variant compiled for at least w waves per SIMD, which grid
 * launches of ensembles with more waves of members than the default kernel
 * keeps resident per SIMD use (no reference counterpart: a diagnostic). */
int clode_last_launch_variant(clode_sim *sim, int *waves_per_simd);
/* Features replay: HBM capacity spent instead of a second integration.  With
 * a budget of `bytes` > 0, a feature simulator whose observer has a warm-up
 * pass (thresh2, nhood2) and whose system has no noise records each member's
 * accepted warm-up steps (time, state, derivative, aux values and next step
 * size, as doubles: 8 (2 + 2 nVar + nAux) bytes per member-step, 8 nVar more
 * in fp32c) up to the budget, and the features pass that follows the warm-up
 * with nothing changed in between reads those steps back instead of
 * integrating them a second time; a member whose steps did not all fit
 * integrates on from its last record.  Results are bit-identical.  0 (the
 * default) turns it off and frees the buffer.  Grid launches only (the work
 * queue mode integrates both passes).  No reference counterpart: the
 * reference integrates both passes (CLODEfeatures.cpp:180-259). */
int clode_set_replay_budget(clode_sim *sim, unsigned long long bytes);
/* Records per member of the last warm-up recording (0: none) and whether the
 * last features pass replayed it (1) or integrated (0). */
int clode_get_replay_info(clode_sim *sim, unsigned int *records_per_member, int *replayed);
/* Compile-only check of the current configuration (no device needed). */
int clode_compile(clode_sim *sim);
/* Per-member count of accepted steps (member-steps), accumulated over every
 * launch while enabled: transient, observer warm-up, features, trajectory. */
int clode_set_step_counter(clode_sim *sim, int enable);
int clode_reset_step_counts(clode_sim *sim);
int clode_get_step_counts(clode_sim *sim, unsigned long long *out, size_t n);
/* Per-member count of step trials over the same launches: one per stepper
 * call, i.e. the accepted steps plus the rejected trials of the adaptive
 * controller. */
int clode_get_trial_counts(clode_sim *sim, unsigned long long *out, size_t n);
/* The members' noise-generator state, [2][nPts] words (s0 row, then s1): the
 * reference keeps it on the device only (clode/cpp/CLODE.cpp:420-464 writes
 * it, the kernels advance it); exposed so that callers can check or save the
 * stream position. */
int clode_get_rng_state(clode_sim *sim, unsigned long long *out, size_t n);

/* ---- batches: several simulators in one work-queue launch ------------- */
/* No reference counterpart (the reference runs one right-hand side per
 * program and one time span per launch, clode/cpp/CLODE.cpp:148,
 * transient.cl:10).  A batch holds 1 to 8 base or feature simulators of one
 * precision on one device; their systems (right-hand sides, dimensions,
 * steppers, observers) may all differ.  They are compiled into one program and
 * integrated by dispatching work-queue kernels; each simulator keeps its own
 * buffers and host semantics, and its results are bit-identical to running it
 * alone.  The simulators must outlive the batch. */
typedef struct clode_batch clode_batch;
int clode_batch_create(clode_batch **out, clode_sim *const *sims, size_t n);
int clode_batch_destroy(clode_batch *batch);
/* compile-only check of the batched program for these simulators (no device
 * needed) */
int clode_batch_compile(clode_sim *const *sims, size_t n);
/* workgroup size 64..256 (multiple of 64), loop iterations between queue
 * checks >= 1; defaults 64 and 8 */
int clode_batch_set_launch(clode_batch *batch, int block_size, int slice);
int clode_batch_get_launch(clode_batch *batch, int *block_size, int *slice);
/* transient / initialize_observer / features / shift_x0 of every simulator,
 * each as one launch (features: reinitialize -1 continue or initialise where
 * needed, 0 continue, 1 re-initialise) */
int clode_batch_transient(clode_batch *batch);
int clode_batch_initialize_observer(clode_batch *batch);
int clode_batch_features(clode_batch *batch, int reinitialize);
int clode_batch_shift_x0(clode_batch *batch);
/* rounds x (transient, x0 <- xf, features, x0 <- xf); use_graph != 0 captures
 * the whole sequence into one graph and launches it once */
int clode_batch_sweep(clode_batch *batch, int rounds, int use_graph);
/* Time the transient launch over every stride-th queued member, each member
 * capped at max_steps steps (0: no cap; a cap keeps a heavy tail of long
 * members from dominating the trials), for each block size in {64, 128, 256}
 * and slice in {1, 4, 16, 64}, three times in interleaved rounds, restoring
 * the member state after each trial, and keep the shape with the fastest
 * median.  The choice is stored in the first simulator's code-object cache
 * directory, keyed by the batched program, the device, the ensemble sizes,
 * stride and max_steps; a later call with the same key takes it from there
 * without trials unless the environment sets CLODE_AUTOTUNE_FRESH=1. */
int clode_batch_autotune(clode_batch *batch, int stride, unsigned int max_steps, int *block_size, int *slice);
/* The last autotune's trials, one row of 5 doubles per launch shape: block
 * size, slice, median, fastest and slowest device time (ms).  Writes at most
 * capacity rows, their number to *n_rows (0 when the shape came from the
 * cache: *from_cache is then 1). */
int clode_batch_autotune_table(clode_batch *batch, double *rows, int capacity, int *n_rows, int *from_cache);
/* device time of the last batch operation, milliseconds */
int clode_batch_last_ms(clode_batch *batch, double *ms);

#ifdef __cplusplus
}
#endif

#endif /* CLODE_HIP_H */
