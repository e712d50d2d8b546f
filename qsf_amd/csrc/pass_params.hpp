// pass_params.hpp — argument blocks shared by the host runtime and the HIP kernels.
#pragma once
#include <stdint.h>

namespace qsf {

// What one axis pass does, in this order (each step optional):
//   load (x scale) -> inverse FFT -> x A-factor -> observe -> x B-factor -> forward FFT
//   -> x P-factor -> inverse FFT -> store
enum PassFlags : uint32_t {
  F_INV1 = 1u << 0,     // inverse transform first (input is in P along the axis)
  F_XA = 1u << 1,       // multiply by the X factor A (end of the previous X substep)
  F_OBS = 1u << 2,      // observables of the step-boundary state (see ObsParams)
  F_XB = 1u << 3,       // multiply by the X factor B (start of the next X substep)
  F_FWD = 1u << 4,      // forward transform
  F_P = 1u << 5,        // multiply by the P factor (product of per-axis tables)
  F_INV2 = 1u << 6,     // inverse transform after the P factor
  F_SCALE = 1u << 7,    // multiply by `scale` at load
  F_MASK = 1u << 8,     // multiply by the region mask (after observing, before the B factor)
  F_NOSTORE = 1u << 9   // read-only pass (observables only)
};

// Pair interaction part of an X factor: none, a table indexed by |i-j| (Eberly), or a full n*n
// table (EckhardSacha; measured faster than one FP64 sincos per element).
enum PairKind : int32_t { PAIR_NONE = 0, PAIR_DIFF = 1, PAIR_FULL = 2 };

// X factor = prod_a axis[a][i_a] * prod_{active pairs (p,q)} pair(i_p, i_q)
struct XFactor {
  const void* axis[3];   // complex tables of n_a entries (nullptr = 1)
  const void* pair;      // PAIR_DIFF: n entries indexed by |i-j|; PAIR_FULL: n*n, [i*n + j]
  int32_t pair_kind;
  int32_t pair_mask;     // bit0: (0,1), bit1: (1,2), bit2: (2,0)
  int64_t pair_n;        // row length of a PAIR_FULL table
  // what the pair table holds: exp(-iδ·ee) (RE) or exp(-δ·ee) (IM), ee of potential.h:130-136
  // (identifies the table when two factors are compared)
  double delta, EE, Esoft, xmin, dx;
  int32_t mode, model;   // QSF_MODE_*, QSF_POT_EBERLY | QSF_POT_ECKHARDT_SACHA
};

// Observation of the step-boundary state (after the A factor).
struct ObsParams {
  double* partials;      // per block: [norm, epot, ekin]
  const double* kin_w[3];  // per-axis additive weights for <½p²> (nullptr = 0)
  // potential for <V> (closed forms of potential.h evaluated on the fly)
  int32_t pot_kind, free_mask;
  double NE, EE, Nsoft, Esoft;
  double xmin[3], dx[3];
  int32_t want_epot;
  // sparse gather of the points flux needs (row-major rows of the last axis)
  const int32_t* gather_rowptr;   // rows+1
  const int32_t* gather_pos;      // position within the row of each gathered point
  void* gather_out;               // complex, one per gathered point
  // full copy of the boundary state (kinetic energy on record steps)
  void* copy_out;
};

// One solve of a batched pass (blockIdx.y selects the entry).
struct BatchEntry {
  void* psi;               // the solve's state (same shape for every entry of a launch)
  const void* pf[3];       // its P factor tables of the current step
  double* partials;        // its per-block observable sums
  void* gather_out;        // its gathered flux points
  int64_t total_steps;     // steps the solve takes; later steps of the batch skip it
};

struct PassParams {
  const BatchEntry* batch; // non-null: batched launch, per-solve pointers from batch[blockIdx.y]
  const int64_t* step_ctr; // device step counter of a batched launch
  int32_t nbatch;          // number of solves (gridDim.y) of a batched launch
  const void* src;
  void* dst;
  int64_t n[3];          // padded 3D shape (n0, n1, n2)
  int32_t axis;          // transform axis
  int32_t nbatch_tiles;  // number of tiles (blocks) along the batch dimension
  int64_t inner;         // prod_{d>axis} n_d
  int64_t outer;         // prod_{d<axis} n_d
  const void* roots;     // N roots exp(-2πi j/N)
  uint32_t flags;
  double scale;
  XFactor xa, xb;
  const void* pf[3];     // P factor tables (n_a entries; nullptr = 1)
  ObsParams obs;
  // region mask (F_MASK): FP64 table over per-axis CAP nodes clamped to [0, nCAP]
  const double* mask_tab;
  int64_t mask_stride[3];  // 0 for free axes
  int64_t nCAP;
  // 2D multigrid sources (written with the observation, before the mask): edge rows
  // [2nCAP][n1] and edge columns [n0][2nCAP] of the step-boundary state
  void* er_out;
  void* ec_out;
};

// One transfer into a 2D receiving region along its freed axis (see transfer_pass.hip).
struct TransferParams {
  void* dst;                 // receiver state (n0, n1)
  int64_t n0, n1;
  int32_t axis;              // freed axis f: 0 = column pass, 1 = row pass
  const void* roots;         // roots of length n_f
  int64_t nCAP;
  const void* edge_in;       // source edge nodes along f
  int32_t edge_in_layout;    // 0: [2nCAP][n_other] (edge rows), 1: [n_other][2nCAP] (edge columns)
  int32_t edge_out_layout;   // layout of edge_out (0 rows, 1 columns)
  void* edge_out;            // receiver edge nodes for the next group, or nullptr
  const double* inv_tab;     // inv_mask(node) of the receiver
  const double* corr_tab;    // group-1 mask_correction(node of the other axis), or nullptr
  int32_t has_xa;            // apply the receiver's pending X factor before the add
  XFactor xa;
  const double* mask_tab;    // receiver region mask, or nullptr
  int64_t mask_stride_axis, mask_stride_other;
};

}  // namespace qsf
