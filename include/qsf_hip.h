/* qsf_hip.h — C ABI of libqsf_hip.so, the MI355X (gfx950) split-operator step of QSF.
 *
 * The reference has no runtime ABI: a user program instantiates header templates and the grid
 * type provides the member-template surface that SplitPropagator calls. Each entry point below
 * replaces one member of that surface (cited), so a grid backend (qsf_amd/host/qsf/device_grid.h)
 * can forward SplitPropagator's calls here unchanged.
 *
 * Conventions: every function returns 0 on success and a negative code on error
 * (qsf_last_error(ctx) returns the message); handles are owned by the library; host buffers are
 * owned by the caller; wave functions cross the boundary as interleaved FP64 (re, im) in the
 * reference's row-major layout (X rep) or FFTW's transposed P layout (D>1 main grids,
 * wf/grid.h:238-248). One HIP stream per context; calls on one context are not thread-safe,
 * different contexts may be used from different threads (the library keeps no process-wide
 * mutable state apart from the per-device scratch of qsf_average_noalloc).
 */
#ifndef QSF_HIP_H
#define QSF_HIP_H

#include "qsf_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qsf_ctx qsf_ctx;
typedef struct qsf_set qsf_set;
typedef struct qsf_grid qsf_grid;
typedef struct qsf_prop qsf_prop;

enum {
  QSF_OK = 0,
  QSF_ERR_ARG = -1,        /* invalid argument                         */
  QSF_ERR_UNSUPPORTED = -2,/* configuration the device path lacks      */
  QSF_ERR_HIP = -3,        /* HIP runtime error                        */
  QSF_ERR_STATE = -4       /* call not valid in the current state      */
};

/* ---- context (one device, one stream) ---- */
int qsf_ctx_create(int device, qsf_ctx** out);
int qsf_ctx_destroy(qsf_ctx* ctx);
const char* qsf_last_error(const qsf_ctx* ctx);
int qsf_ctx_synchronize(qsf_ctx* ctx);
/* Non-blocking: *idle = 1 when all work enqueued on the context's stream has finished, else 0
 * (no counterpart in the reference; used by schedulers that keep several contexts busy). */
int qsf_ctx_query(qsf_ctx* ctx, int* idle);
/* Device time (ms) between the two most recent qsf_timer_mark calls, on the context stream. */
int qsf_timer_mark(qsf_ctx* ctx);
int qsf_timer_elapsed_ms(qsf_ctx* ctx, double* ms);

/* Per-kernel timing with HIP events recorded on the context stream around every axis-pass
 * launch between qsf_prof_begin and qsf_prof_end. Entries are keyed by kernel
 * ("pass_<prec>_<row|col>_n<N>"); bytes = algorithmic HBM bytes per launch (state read + write). */
int qsf_prof_begin(qsf_ctx* ctx);
int qsf_prof_end(qsf_ctx* ctx, int* nentries);
int qsf_prof_entry(qsf_ctx* ctx, int i, char* name, int name_len, int64_t* launches, double* total_ms,
                   double* bytes_per_launch);

/* Autotuner: times every compiled variant (tile width x elements per thread) of the axis
 * passes this grid's step uses (reps launches each, HIP events) and keeps the fastest for
 * (precision, pass class, axis length, grid size) in the context. The class (`row` below) is
 * 1 for a row pass (the contiguous last axis), 0 for a column pass of an inner axis and 2 for a
 * column pass of the outermost axis. Entries can be read back and set again (e.g. from a cache
 * file kept by the caller). */
int qsf_autotune(qsf_grid* g, int reps);
int qsf_tune_count(qsf_ctx* ctx, int* count);
int qsf_tune_entry(qsf_ctx* ctx, int i, int* precision, int* row, int64_t* n, int64_t* grid_size,
                   int* variant, double* best_ms);
int qsf_tune_set(qsf_ctx* ctx, int precision, int row, int64_t n, int64_t grid_size, int variant);

/* ---- grid sets: one grid, or the 2^D regions of MultiRegionsReduced (mpi_logic.h:246-263) ---- */
/* Replaces LocalGrid(Section&) construction (grid.h:115-123, multigrid.h:45-46, 124-202).
 * Multigrid sets: 2D (4 regions) or 3D (8 regions; a cubic grid whose side is a multiple of nCAP
 * and at least 2*nCAP, as the reference's transfer boxes assume, and a reduced-dimension
 * potential). CoulombInteraction over two or more bounded coordinates is applied as a pass of
 * its own, so propagators over such sets, and over 3D multigrid sets, run unfused. */
int qsf_set_create(qsf_ctx* ctx, const qsf_grid_desc* desc, qsf_set** out);
/* Sets and the propagators created on them may be destroyed in either order; a propagator whose
 * set is gone answers its scalar queries and qsf_prop_destroy, and QSF_ERR_STATE otherwise. */
int qsf_set_destroy(qsf_set* set);
int qsf_set_count(const qsf_set* set, int* count);
int qsf_set_grid(qsf_set* set, int region, qsf_grid** out);
/* LocalGrid<...,true>::postCompute: transfer() then maskRegion() (multigrid.h:617-657). */
int qsf_set_post_compute(qsf_set* set);

/* ---- one grid: the HamWF member surface SplitPropagator calls ---- */
int qsf_grid_size(const qsf_grid* g, int64_t* m);                       /* localSize grid.h:359 */
int qsf_grid_rep(const qsf_grid* g, int* rep);
/* operator[] access in bulk (grid.h:296). rep = representation the data is in. */
int qsf_grid_upload(qsf_grid* g, const double* interleaved, int64_t count, int rep);
int qsf_grid_download(qsf_grid* g, double* interleaved, int64_t count);
/* fourier<REP>() grid.h:369-378 (main) / multigrid.h:666-680 (partial FFT of regions). */
int qsf_fourier(qsf_grid* g, int to_rep);
/* evolve<MODE,REP>(delta) grid.h:394-423 with Spin0 hamiltonian.h:44-97. coupling3 holds
 * _coupling[Axis<d>] (coupling.h:63-67); couples_in = couplesInRep (QSF_REP_X | QSF_REP_P). */
int qsf_evolve(qsf_grid* g, int mode, int rep, double delta, const double* coupling3, int couples_in);
/* average<REP,Op>() grid.h:425-441; op 0 = Identity, 1 = PotentialEnergy, 2 = KineticEnergy. */
int qsf_average(qsf_grid* g, int rep, int op, double* out);
/* qsf_average that allocates nothing. No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). The observation pass's block sums and their reduction go to a static
 * per-device scratch array of 8 MiB (held by each device until the process ends), so grids whose
 * last-axis pass has more than (8 MiB / 8 - 776) / 3 = 349,266 blocks are refused (QSF_ERR_ARG;
 * every grid up to 512^3 fits). The grid must be at a step boundary: QSF_ERR_STATE while a
 * propagator step is pending on its set (any other grid entry point or qsf_prop_finish settles
 * it). All calls on one device share the scratch: calls that overlap on one device (two
 * contexts, or two threads) race. Same kernels and reduction order as qsf_average. */
int qsf_average_noalloc(qsf_grid* g, int rep, int op, double* out);
/* operation<M,R,Normalize>() grid.h:599-606. */
int qsf_normalize(qsf_grid* g);
/* current_map() + velocityGaugeCorrection + flux<R,F>() (grid.h:467-573, propagator.h:249-265).
 * vg_coupling3 = A per axis for the velocity gauge, or NULL. */
int qsf_flux(qsf_grid* g, const double* vg_coupling3, const qsf_border_desc* borders, int nb,
             double* out);
/* maskRegion() grid.h:582-597 (bit-exact: host-built FP64 mask table). */
int qsf_mask_region(qsf_grid* g);
/* operation<M, X, AntiSymmetrize<particles>> (wf/grid.h:611-723), in place on the X state:
 * particles 2 swaps the two coordinates of a 2D grid (the last two of a 3D grid), particles 3
 * sums the six permutations of a cubic 3D grid; nothing for dim < particles. */
int qsf_antisymmetrize(qsf_grid* g, int particles);
/* croossOut(au_distance) (wf/grid.h:890-907): in place on the X state, keeps nodes with exactly
 * two coordinates beyond the multiIonizationFilter<2_D, 2_D> radius (:862-888) and ramps the rest
 * down with 0.5 (1 + cos(pi k / 30)). */
int qsf_cross_out(qsf_grid* g, double au_distance);
/* orthogonalizeWith (wf/grid.h:909-946) with the stored state already read: s = <phi|psi> dV
 * (FP64), psi -= s * phi; `phi` is count = grid size complex FP64 values (interleaved), s goes
 * to overlap[2] when given. */
int qsf_orthogonalize(qsf_grid* g, const double* phi, int64_t count, double* overlap);

/* ---- propagator: SplitPropagator<MODE, MultiProductSplit<Split3Base<X,P,X>,order>, ...>
 *      (propagator.h:51-499) over a grid set, with the step executed as fused axis passes ---- */
int qsf_prop_create(qsf_set* set, const qsf_prop_desc* desc, const qsf_field_desc* fields, int nfields,
                    qsf_prop** out);
int qsf_prop_destroy(qsf_prop* p);
int qsf_prop_max_steps(const qsf_prop* p, int64_t* out);      /* propagator.h:153 */
int qsf_prop_start(qsf_prop* p);                              /* computeEach<AT_START> :439-441 */
/* Resume in place of qsf_prop_start: makeFakeStep until step == at_step (propagator.h:347-366,
 * 425-436): each chain's precalc side effects and timer increments, no evolution, no chain
 * timer restore (so an order-2 MultiProductSplit clock advances by both chains' sums per fake
 * step, as in the reference). No AT_START row. The caller then uploads the saved state. */
int qsf_prop_fast_forward(qsf_prop* p, int64_t at_step);
int qsf_prop_should_continue(qsf_prop* p, int* out);          /* shouldContinue :304-309 */
int qsf_prop_step(qsf_prop* p);                               /* one loop body :445-449 */
int qsf_prop_run_steps(qsf_prop* p, int64_t k, int64_t* done); /* up to k loop bodies, no host sync */
int qsf_prop_finish(qsf_prop* p);                             /* final step + AT_END :451-454 */
int qsf_prop_step_count(const qsf_prop* p, int64_t* out);
int qsf_prop_time(const qsf_prop* p, double* out);
int qsf_prop_coupling(const qsf_prop* p, double* out3);
/* Observable rows (one per computeEach): step, time, norm, epot, ekin, etot, ediff, flux[nb]. */
int qsf_prop_record_width(const qsf_prop* p, int* out);
int qsf_prop_nrecords(qsf_prop* p, int* out);
int qsf_prop_records(qsf_prop* p, double* out, int max_rows);
/* FP64 time integral of each flux family, sum over steps of flux*dt (device-resident). */
int qsf_prop_flux_integral(qsf_prop* p, double* out);
/* 0 = fused axis passes (default), 1 = one standalone call per reference member. */
int qsf_prop_set_exec(qsf_prop* p, int exec_mode);

/* ---- batches: many single-region 1D and 2D solves, mixed in one batch (real-time Strang,
 *      velocity gauge, one field each, one border list for all) advanced together. One step =
 *      one batched launch per (size group, pass: one pass for 1D, two for 2D) driven by a
 *      device step counter; with use_graph it is captured once and replayed as a hipGraph.
 *      Each solve follows its own SplitPropagator loop (max_steps+2 steps, propagator.h:153,
 *      443-451); finished solves drop out of the launches. Within a step the size groups and
 *      their observations run on the context stream and streams of the batch's own, forked from
 *      the context stream and joined back into it before the step ends: work queued on it after
 *      qsf_batch_run sees every step complete. ---- */
typedef struct qsf_batch qsf_batch;
int qsf_batch_create(qsf_ctx* ctx, const qsf_grid_desc* grids, const qsf_field_desc* fields, int nsolves,
                     const qsf_prop_desc* desc, qsf_batch** out);
int qsf_batch_destroy(qsf_batch* b);
int qsf_batch_upload(qsf_batch* b, int solve, const double* interleaved);
int qsf_batch_download(qsf_batch* b, int solve, double* interleaved);
int qsf_batch_start(qsf_batch* b);                                  /* computeEach<AT_START> */
int qsf_batch_run(qsf_batch* b, int64_t steps, int use_graph, int* all_done);
int qsf_batch_finish(qsf_batch* b);                                 /* end states + AT_END rows */
int qsf_batch_info(qsf_batch* b, int solve, int64_t* max_steps, int64_t* grid_size);
/* rows: step, time, norm, flux[nb] per observed step boundary */
int qsf_batch_records(qsf_batch* b, int solve, double* out, int max_rows, int* rows);
int qsf_batch_flux_integral(qsf_batch* b, int solve, double* out);

#ifdef __cplusplus
}
#endif
#endif
