// Device-side argument blocks shared by kernels.hip and api.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "layout.h"

namespace otfem {

// Per-body CG state, written by the last tile of a body to finish a phase and
// read by every tile of that body after the next grid barrier.
struct BodyState {
  double bb;      // b.b over free dofs
  double rr0;     // r.r after the warm start
  double d;       // r.r before the current iteration's update
  double d2;      // p.Ap of the current iteration (clamped in reference mode)
  double alpha;   // d / d2
  double beta;    // d1 / d (clamped)
  double rr;      // latest r.r
  int32_t iterations;
  int32_t active;
  int32_t converged;
  int32_t flags;  // bit 0: non-finite scalar
};

// Grid-wide control words for the persistent kernel.  All start at zero and
// return to zero at the end of a well-formed step.
struct Control {
  unsigned int bar_count;
  unsigned int bar_gen;
  unsigned int abort;          // set on a barrier time-out; every wave then exits
  int active_bodies;
  unsigned int pad[4];
};

struct StepArgs {
  // nodes, internal order, xyz interleaved
  double *x;
  const double *x0;
  double *v;
  double *fext;
  const double *mass;
  const uint8_t *fixed;
  double *rhs;
  double *r;
  double *p0;
  double *p1;
  float *y;
  // tiles / slots
  const TileInfo *tiles;
  const int32_t *halo;
  const ushort4 *conn;
  const int32_t *slot_tet;
  const double *rest64;  // [9][S]  grad N_k (k = 1..3) at rest
  const double *vd64;    // [3][S]  |V| * (D0, D1, D2)
  const float *vd32;     // [3][S]
  float *bt;             // [9][S]  R * grad N_k, refreshed each step
  // per tet (caller order)
  double *Re;            // [9][T] or null
  double *plastic;       // [6][T]
  const double *plastic_param;  // [3][T] yield, creep, max
  // bodies / reductions
  const int32_t *body_tile_begin;
  BodyState *body;
  double *tile_partial;  // [2 * ntiles]
  unsigned int *body_cnt;  // [2 * B]
  Control *ctrl;
  // sizes
  int64_t S;
  int64_t T;
  int32_t ntiles;
  int32_t B;
  // step scalars
  double dt;
  double cM;   // 1 + alpha*dt  (mass coefficient of A)
  double cK;   // dt*dt + beta*dt (stiffness coefficient of A)
  double tol;
  double tiny;
  int32_t tol_mode;
  int32_t cg_min;
  int32_t cg_max;
  int32_t warp;
  int32_t rotation_mode;
  int32_t store_rotations;
  int32_t plastic_on;
  int32_t nblocks;       // grid size of the persistent kernel
  long long spin_limit;  // barrier time-out in wall-clock ticks
};

// Host launchers (kernels.hip).
hipError_t launch_element(const StepArgs &a, int block, size_t lds, hipStream_t s);
hipError_t launch_cg(const StepArgs &a, int grid, int block, size_t lds, bool cooperative,
                     hipStream_t s);
hipError_t cg_occupancy(int block, size_t lds, int *blocks_per_cu);
size_t element_lds_bytes(int max_nodes, int max_owned, int plastic_on);
size_t cg_lds_bytes(int max_nodes, int max_owned);

}  // namespace otfem
