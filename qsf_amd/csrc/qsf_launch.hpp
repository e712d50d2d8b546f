// qsf_launch.hpp — host-callable launchers of the HIP kernels (internal to libqsf_hip.so).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pass_params.hpp"

namespace qsf {

bool axis_size_supported(int64_t n);
// Launches one fused axis pass with kernel variant v (0 = default; see axis_variants);
// returns the number of blocks (or -1 if n is unsupported).
int launch_axis_pass(int precision, const PassParams& p, hipStream_t s, int v = 0);

// The flux stencil values of an X-representation state ahead of the pass that would gather
// them: out[k] = psi[row[k], pos[k]] times the pending X factor xa at that point (when has_xa),
// the same product the pass forms. Rows run along the last axis.
struct GatherXArgs {
  const void* psi;
  void* out;
  const int32_t* row;
  const int32_t* pos;
  int64_t nvals;
  int64_t n[3];
  int32_t dim;
  int32_t has_xa;
  XFactor xa;
};
void launch_gather_x(int precision, const GatherXArgs& a, hipStream_t s);
int64_t axis_pass_tiles(int precision, const PassParams& p, int v = 0);
// Number of compiled variants for an axis length (tile width x elements per thread).
int axis_variants(int64_t n);
// Fused 2D multigrid transfer into one receiving region (-1 if the length is unsupported).
int launch_transfer(int precision, const TransferParams& p, hipStream_t s);

// Per-axis phase tables of one substep, built on the device from FP64 scalars.
//   X table: expOp(δ·(Ne(x_i)·use_ne + x_i·C)) · absorber(δ, i)
//   P table: expOp(δ·(½p_i² + p_i·C)) · norm
struct AxisTableArgs {
  void* out;             // complex (precision) table of n entries
  int64_t n;
  int32_t kind;          // 0 = X, 1 = P
  int32_t mode;          // QSF_MODE_IM | QSF_MODE_RE
  double delta;
  double C;              // coupling value for this axis (0 if the gauge acts elsewhere)
  // X coordinates / potential / absorber
  double xmin, dx;
  int32_t use_ne;        // add the per-axis electron-nucleus term
  double NE, Nsoft;
  int32_t absorber;      // QSF_ABS_*; applied only when apply_abs
  int32_t apply_abs;
  int64_t nCAP;
  double eta, tau;
  // P coordinates (storage-axis parameters: dp, n, n/2)
  double dp;
  int64_t pn, pn2;
  double norm;
};
void launch_axis_tables(int precision, const AxisTableArgs* a, int count, hipStream_t s);

// Static pair tables exp(-iδ ee): kind PAIR_DIFF (n entries, |i-j|) or PAIR_FULL (n*n).
struct PairTableArgs {
  void* out;
  int64_t n;
  int32_t kind;          // PAIR_DIFF | PAIR_FULL
  int32_t model;         // QSF_POT_EBERLY | QSF_POT_ECKHARDT_SACHA
  int32_t mode;
  double delta, EE, Esoft, xmin, dx;
};
void launch_pair_table(int precision, const PairTableArgs& a, hipStream_t s);

// psi *= exp(-i delta V) (RE) or exp(-delta V) (IM) with V = NE / sqrt(sum x_d^2 + Nsoft) over the
// bounded coordinates (CoulombInteraction, potential.h:40-44), in place on an X-rep state.
struct RadialArgs {
  void* psi;
  int precision;
  int64_t n[3];
  double xmin[3], dx[3];
  int bounded_mask;
  double NE, Nsoft, delta;
  int mode;
};
void launch_radial_x(const RadialArgs& a, hipStream_t s);

// roots[j] = exp(-2πi j/n), computed in FP64 and rounded to the precision.
void launch_roots(int precision, void* out, int64_t n, hipStream_t s);

// Sum of `count` blocks of `width` doubles into out[width] (fixed order). Large counts use a
// second level through `scratch` (256*width doubles). With ncols >= 0 only the first ncols
// doubles of each block are summed, into out[ncols] (a column reduced on its own: pass the
// partials from that column on).
void launch_reduce_partials(const double* partials, int64_t count, int width, double* out, double* scratch,
                            hipStream_t s, int ncols = -1);

// Flux through border families from gathered boundary values.
struct FluxArgs {
  const void* gathered;  // complex (precision)
  int32_t precision;
  const int32_t* pts;    // per point: [slot_c, slot_plus[d], slot_minus[d] for d < D] (1 + 2D ints)
  const double* bvec;    // per point: nb * D border components
  int64_t npts;
  int32_t D, nb;
  double inv_2dx[3];
  double vgC[3];         // velocity-gauge correction (0 if not velocity gauge)
  double scale;          // dV / dx0
  double* partials;      // per block: nb sums
  double* out;           // nb sums (after reduction), may be accumulated into
  double* accum;         // if non-null: accum[b] += out[b] * accum_weight
  double accum_weight;
};
void launch_flux(const FluxArgs& a, hipStream_t s);

// Element-wise multiply of a complex (precision) array by a real FP64 table value per point.
void launch_scale(int precision, void* psi, int64_t m, double s, hipStream_t st);
void launch_copy_convert(int from_prec, const void* src, int to_prec, void* dst, int64_t m, hipStream_t s);

}  // namespace qsf

namespace qsf {
// 2D multigrid transfer pieces (multigrid.h:454-616): gather + mask the edge slice of the
// source into `slice` (zero elsewhere), and dst += slice.
struct SliceArgs {
  const void* src;
  void* slice;
  int64_t n0, n1, n2;    // n2 = 1 for 2D
  int32_t dirFree;       // 0, 1 (2D, 3D) or 2 (3D)
  int64_t nCAP;
  const double* inv_tab; // inv_mask(node) for node in [0, nCAP]
  const double* corr_tab;// 2D: mask_correction by the other axis' node, or nullptr (not group 1)
  const double* m1_tab;  // 3D: mask(node) for node in [0, nCAP] when the receiver is group 1
};
void launch_slice_2d(int precision, const SliceArgs& a, hipStream_t s);
// 3D slices of every box at once: the boxes of multigrid.h:454-569 tile the sliding axis, and
// the transform along the freed axis never crosses them.
void launch_slice_3d(int precision, const SliceArgs& a, hipStream_t s);
void launch_add(int precision, void* dst, const void* src, int64_t m, hipStream_t s);
}  // namespace qsf

namespace qsf {
// ---- batched solves (one launch covers every solve of a size group; blockIdx.y = solve) ----
struct BatchTableArgs {       // P tables of one solve for the current device step
  void* out[2];
  int64_t n[2], pn[2], pn2[2];
  double dp[2];
  double delta;               // dt of the P substep
  const double* coupling;     // [total_steps][3]: A per axis of each step's P substep
  int64_t total_steps;
  int32_t storage_axis[2];    // coupling / P coordinates of the reference's storage axis
};
void launch_batch_ptables(int precision, const BatchTableArgs* dev_args, int nsolves, int64_t nmax,
                          const int64_t* step_ctr, hipStream_t s);

struct BatchObsArgs {         // observables of one solve at the boundary the pass observed
  const double* partials;
  int64_t tiles;
  const void* gathered;
  const double* coupling;     // [total_steps][3]
  int64_t total_steps;
  double* rec;                // [total_steps + 1][1 + nb]: norm, flux[nb]
  double* flux_int;           // [nb]
  double dV;
};
struct BatchFluxGeo {
  const int32_t* pts;
  const double* bvec;
  int64_t npts;
  int32_t D, nb, precision;
  double inv_2dx[3], scale, dt;
};
// Observes the boundary after step *step_ctr (called before the counter advances).
void launch_batch_obs(const BatchObsArgs* dev_args, int nsolves, const BatchFluxGeo& geo,
                      const int64_t* step_ctr, hipStream_t s);
void launch_step_increment(int64_t* step_ctr, hipStream_t s);
}  // namespace qsf
