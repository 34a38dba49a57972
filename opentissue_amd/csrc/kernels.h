// Device-side argument blocks shared by kernels.hip and api.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "layout.h"

namespace otfem {

// Per-body CG record.  Written only by the body's first tile (body_rank 0);
// read by other workgroups in later kernels of the same step.
struct BodyState {
  double bb;      // b.b over free dofs
  double rr0;     // r.r after the warm start
  double rr;      // r.r at the last recorded point
  double d;       // r.r entering the current iteration (for alpha)
  int32_t iterations;
  int32_t stopped;  // 0 while running; else 1 + index of the CG kernel that recorded the stop
  int32_t converged;
  int32_t flags;  // bit 0: non-finite scalar
  int32_t q32_from;  // first CG kernel of the step that stored q in FP32 (INT32_MAX until one does)
  double alpha_hist[2];  // alpha of iteration j-1, written by (even) CG kernel j at [j & 1]
};

// SpMV arithmetic inside CG (element data is FP32 in memory in every mode).
enum SpmvMode { SPMV_F32 = 0, SPMV_F32_ACC64 = 1, SPMV_F64 = 2, SPMV_F64_STORE64 = 3 };

struct StepArgs {
  // nodes, internal order, xyz interleaved
  double *x;
  const double *x0;
  double *v;
  const double *vw;      // CG start (== v for the plain start, else a separate buffer)
  double *fext;
  const double *mass;
  const uint8_t *fixed;
  double *rhs;
  double *r;             // [2][3N] residual, alternating by CG iteration
  double *p;             // [2][3N] search direction (FP64), or [2][3N] FP32 in its first half
                         //         with `p32`
  double *y;             // [2][3N] q = A p = cM m p + cK K' p (0 at fixed nodes)
  float *y32;            // [2][3N] q in FP32, for bodies whose residual is small (with p32)
  float *r32;            // [2][3N] r in FP32, likewise (a body switches both at once)
  // tiles / slots
  const TileInfo *tiles;
  const int32_t *halo;
  // slot corners: four tile-local node indices of 12 bits each (tile tables
  // hold at most kMaxTileNodes nodes), c0 | c1 << 12 | (c2 & 0xff) << 24 in
  // conn_lo and c2 >> 8 | c3 << 4 in conn_hi: 6 bytes a slot
  const uint32_t *conn_lo;
  const uint16_t *conn_hi;
  const int32_t *slot_tet;
  const double *rest64;  // [9][S]  grad N_k (k = 1..3) at rest
  const double *vd64;    // [3][S]  |V| * (D0, D1, D2)
  float4 *bt4;           // [3][S]  (R grad N_k, |V| D_k), refreshed each step; with
                         // `folded`, the same storage holds [2][S] float4 + [S] float:
                         // the nine values sqrt(E |V|) R grad N_k (k = 1..3)
  double4 *bt8;          // [3][S]  same in FP64 (only with SPMV_F64_STORE64)
  // per tet (caller order)
  double *Re;            // [T][9]
  double *plastic;       // [T][6]  plastic strain at the start of the step (read only in k_element)
  double *plastic_next;  // [T][6]  the step's updated strain, written by each tet's primary
                         //         copy and committed to `plastic` after k_element
  const double *plastic_param;  // [T][3] yield, creep, max
  const double4 *tile_d; // [ntiles] (D0, D1, D2, 0) per unit volume and unit E of a tile's one nu
  const double *tile_kb; // [ntiles] max over owned nodes of sqrt(12) sum of trace(Ke) of its tets
  // bodies / reductions (no atomics except done_count)
  BodyState *body;
  double *part_bb;       // [ntiles]  b.b (element kernel)
  double *part_rr;       // [ntiles]  r_0.r_0 (element kernel)
  double *part4;         // [2][ntiles][4]  p.q, r.q, q.q, r.r per CG kernel
  int *done_count;       // [2]: bodies whose CG stopped in this step, active tiles
  int32_t *act;          // [ntiles] tiles of running bodies (k_compact)
  int32_t compact;       // CG workgroups map onto act[] (batched meshes)
  unsigned long long *iter_acc;  // [B] CG iterations summed over steps
  unsigned long long *unconv_acc;  // [B] steps whose CG stopped unconverged
  // sizes
  int64_t N;
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
  int32_t store64;       // element kernel also writes bt8
  int32_t folded;        // slot records in folded form (every tile has one Poisson ratio)
  int32_t p32;           // the CG direction p is stored in FP32 (CG to a relative residual in
                         // the default SpMV mode): the SpMV, the dot products and the updates
                         // of v and r all use the stored value, so r stays the residual of v
  double q32_rr;         // with p32: a body's q = A p and r are stored in FP32 (y32, r32) from
                         // the first CG kernel of the step at which r.r <= q32_rr b.b (0: never).
                         // r then drifts from the true residual by the rounding of alpha q and
                         // of r, which stays far below the stopping tolerance once |r| is small
  // Ordered sums (deterministic mode): slots store their per-corner terms in
  // `ordered_terms` (slot-major: one contiguous record per slot) and every
  // owned node adds its terms in slot order, so results do not depend on
  // wave scheduling.
  int32_t ordered;
  const int32_t *ord_begin;  // [N+1] internal node -> first entry
  const int32_t *ord_ent;    // [4T] entries 4 * slot + corner, ascending per node
  double *ordered_terms;     // [S][36] element terms (3 x [4][3]); CG uses [S][12]
  // Subspace (Galerkin) CG start: v_start = v^n + sum_j c_j d_j with the
  // directions d_0 = v^n, d_j = v^(n-j+1) - v^(n-j) (j = 1..nh) and, with a
  // network attached, d = v_net - v^n; c minimises the A-norm error of the
  // start over their span, per body.  Off when sub_hist == 0.
  float4 *vring;             // [sub_ring][N] velocity differences v^j - v^(j-1) (xyz, FP32)
  double *vbuf;              // [2][3N] v^(n-1) and v^n (halves alternate by step)
  int32_t *sub_state;        // [3] on the device: ring slot of this step's difference,
                             //     earlier steps since the reset, half of vbuf with v^(n-1)
  int32_t sub_ring;          // ring slots (>= sub_hist)
  int32_t sub_hist;          // past velocities used (nh <= sub_hist)
  int32_t sub_net;           // the network's start is one more direction (vnet)
  const double *vnet;        // [3N] the network's CG start (with sub_net)
  double *sub_part;          // [ntiles][kSubPart] per-tile Gram partials
  const int32_t *body_tiles; // [B+1] body b owns tiles [body_tiles[b], body_tiles[b+1])
  double *sub_coef;          // [B][kSubMax] the step's coefficients c per body
  const int32_t *prim_begin; // [ntiles+1] tile t's primary slots are prim_slot[prim_begin[t] ..]
  const int32_t *prim_slot;  // [T] slot of each tet's primary copy, grouped by tile
};

// Subspace start: at most kSubMax directions; a tile's partial record holds
// the upper triangle of D^T A D (kSubPairs, entry (i, j), i <= j, at
// j (j + 1) / 2 + i) followed by D^T r_plain (kSubMax).
constexpr int kSubMax = 12;
// Largest tile node table (owned + halo) the 12-bit slot corners address.
constexpr int kMaxTileNodes = 4096;
constexpr int kSubPairs = kSubMax * (kSubMax + 1) / 2;
constexpr int kSubPart = kSubPairs + kSubMax;

struct LaunchConfig {
  int cg_block = 256;   // 256, 512 or 1024 threads per tile in the CG kernels
  int spmv_mode = SPMV_F64;
};

// One step = enqueue_element, then enqueue_cg for iterations [0, cg_max)
// (one kernel per iteration; kernels of bodies that already stopped exit at
// once, so the iteration range may be issued in chunks and cut short once
// every body has stopped), then enqueue_finish.
hipError_t enqueue_element(const StepArgs &a, size_t lds_elem, hipStream_t s);
// With a.compact, the range starts with k_compact and the CG kernels run
// `grid` workgroups over the active-tile list (grid >= active tiles).
hipError_t enqueue_cg(const StepArgs &a, const LaunchConfig &lc, size_t lds_cg, hipStream_t s,
                      int it0, int it1, int grid);
hipError_t enqueue_finish(const StepArgs &a, const LaunchConfig &lc, hipStream_t s);
// Subspace start (after enqueue_element, before the CG): per-tile Gram
// partials of the directions, then per body the A-optimal coefficients, the
// start v^n + D c, its residual r_plain - A D c (one SpMV) and the tile
// partials of r.r; then the ring advances.  lds_gram / lds_start from
// sub_gram_lds_bytes / cg_lds_bytes.
hipError_t enqueue_subspace(const StepArgs &a, size_t lds_gram, size_t lds_start, hipStream_t s);
size_t sub_gram_lds_bytes(int max_nodes, int max_owned, int directions);
// Static LDS of k_sub_gram (a launch needs dynamic + static within the CU's LDS).
size_t sub_gram_static_lds_bytes();
// Surface coupling: out[3i..] = sum_k w[i].k x[3 ids[i].k ..] for bound
// vertices (ids[i].x >= 0, internal node ids), rest[3i..] for the others.
hipError_t enqueue_surface_update(int64_t n, const int4 *ids, const double4 *w, const double *rest,
                                  const double *x, double *out, hipStream_t s);
// Surface normals: out[3i..] = Re[tet[i]] n0[3i..] (Re row-major [9T]) for
// bound vertices when use_re, n0 otherwise.
hipError_t enqueue_surface_normals(int64_t n, const int32_t *tet, const double *Re, int use_re,
                                   const double *n0, double *out, hipStream_t s);
size_t element_lds_bytes(int max_nodes, int max_owned, int plastic_on);
// LDS of the CG kernels (and of k_sub_start with p_bytes = 8); p_bytes is the
// size of the direction's storage type (8, or 4 with p32).
size_t cg_lds_bytes(int max_nodes, int max_owned, int p_bytes = 8);

}  // namespace otfem
