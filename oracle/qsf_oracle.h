/* qsf_oracle.h — C entry points of the CPU FP64 oracle.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the correctness oracle for the HIP path: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only as
 * the checker or as the timed CPU baseline. The product (qsf_amd/) never links or calls it.
 *
 * It restates, on one CPU process, what the reference computes with MPI + FFTW. Each function
 * in qsf_oracle.cpp cites the reference file:line it follows.
 */
#ifndef QSF_ORACLE_H
#define QSF_ORACLE_H
#include "../include/qsf_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_grid orc_grid;
typedef struct orc_set orc_set;
typedef struct orc_prop orc_prop;

/* FFT pins: own mixed-radix FFT, and a long-double brute-force DFT (sign -1 = forward). */
void orc_fft1d(int64_t n, int sign, const double* in, double* out);
int orc_set_threads(int k);
void orc_dft_bruteforce(int64_t n, int sign, const double* in, double* out);

/* Grid set: one grid (multigrid=0) or the 2^D regions of MultiRegionsReduced (multigrid=1). */
orc_set* orc_set_create(const qsf_grid_desc* desc);
void orc_set_destroy(orc_set* s);
int orc_set_count(const orc_set* s);
orc_grid* orc_set_grid(orc_set* s, int region);

int64_t orc_grid_size(const orc_grid* g);
int orc_grid_rep(const orc_grid* g);
void orc_grid_set_psi(orc_grid* g, const double* interleaved);
void orc_grid_get_psi(const orc_grid* g, double* interleaved);
void orc_fourier(orc_grid* g, int to_rep);
void orc_evolve(orc_grid* g, int mode, int rep, double delta, const double* coupling3);
double orc_average(orc_grid* g, int rep, int op); /* op: 0 Identity, 1 Potential, 2 Kinetic */
void orc_normalize(orc_grid* g);
/* current_map + (optional) velocity-gauge correction + flux for each border. */
void orc_flux(orc_grid* g, const double* vg_coupling3, const qsf_border_desc* b, int nb,
              double* out);
void orc_current_map(orc_grid* g, const double* vg_coupling3, double* curr_out /* m*dim */);
void orc_mask_region(orc_grid* g);
void orc_set_transfer(orc_set* s);
void orc_set_post_compute(orc_set* s);
/* Debug: masked edge slice that `transfer` would FFT for receiver `region` and source index. */
int orc_set_transfer_slice(orc_set* s, int region, int source_index, double* slice_out);

/* Tables that must be bit-exact on the device. */
double orc_absorber_mask(const orc_grid* g, int nd, const int64_t* nodes);
double orc_absorber_absorb(const orc_grid* g, double delta, int nd, const int64_t* nodes);
void orc_border_vec(int kind, int size, int dim, const int64_t* idx, double* out3);
double orc_potential(const orc_grid* g, const double* coords);
double orc_potential_partial(const orc_grid* g, int free_axes, const double* coords);

/* Fields: value of every coupling after a precalc(t) sequence (coupling.h, field.h). */
void orc_coupling_sequence(const qsf_field_desc* f, int nf, int gauge, const double* times,
                           int nt, double* out /* nt*3 */);
double orc_max_pulse_duration(const qsf_field_desc* f, int nf);

/* Propagator (propagator.h:51-499) over a grid set. */
orc_prop* orc_prop_create(orc_set* s, const qsf_prop_desc* p, const qsf_field_desc* f, int nf);
void orc_prop_destroy(orc_prop* p);
int64_t orc_prop_max_steps(const orc_prop* p);
void orc_prop_start(orc_prop* p);             /* computeEach<AT_START> (propagator.h:439-441) */
int orc_prop_should_continue(orc_prop* p);    /* propagator.h:304-309 */
void orc_prop_step(orc_prop* p);              /* one loop body (propagator.h:445-449) */
void orc_prop_finish(orc_prop* p);            /* final step (propagator.h:451-454) */
void orc_prop_make_step(orc_prop* p);         /* makeStep only (propagator.h:334-346) */
int64_t orc_prop_step_count(const orc_prop* p);
double orc_prop_time(const orc_prop* p);
int orc_prop_nrecords(const orc_prop* p);
int orc_prop_record_width(const orc_prop* p);
void orc_prop_records(const orc_prop* p, double* out);
void orc_prop_coupling(const orc_prop* p, double* out3);

#ifdef __cplusplus
}
#endif
#endif
