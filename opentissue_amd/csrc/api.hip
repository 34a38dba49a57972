// otfem C ABI (include/otfem/otfem.h): handle, device memory, launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <tuple>
#include <string>
#include <thread>
#include <vector>

#include "kernels.h"
#include "otfem/mesh_coupling.hpp"
#include "layout.h"
#include "mlp.h"
#include "otfem/otfem.h"

using namespace otfem;

namespace {
const char *kVersion = "otfem 0.2.0 (gfx950)";
}  // namespace

struct otfem_solver {
  Layout L;
  std::string err;
  // host copies, caller order
  std::vector<double> x_init, x0_host, v_init;
  std::vector<int32_t> tets;
  std::vector<uint8_t> fixed_host;
  std::vector<double> young, poisson, density;  // per tet, from the last init
  double plastic_c[3] = {0.0, 0.0, 0.0};        // yield, creep, max of the last init
  std::vector<double> mass_host, vol_host;
  // staging of node arrays in internal order for the set/get calls, kept
  // between calls so that per-step transfers reuse resident pages
  std::vector<double> stage;
  // two pinned halves of pin_half doubles for device-to-host reads of node
  // and tet arrays (pinned_read)
  double *pin_stage = nullptr;
  size_t pin_half = 0;
  hipEvent_t pin_ev[2] = {nullptr, nullptr};
  bool device_ready = false;
  bool initialized = false;
  bool plastic_on = false;
  bool rotations_valid = false;
  int device = -1;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
  // device buffers
  double *x = nullptr, *x0 = nullptr, *v = nullptr, *fext = nullptr, *mass = nullptr;
  double *rhs = nullptr, *r = nullptr, *p = nullptr, *y = nullptr;
  float *y32 = nullptr, *r32 = nullptr;  // [2][3N] q = A p and r in FP32 (small residual, FP32 p)
  uint8_t *fixed = nullptr;
  TileInfo *tiles = nullptr;
  int32_t *halo = nullptr;
  uint32_t *conn_lo = nullptr;  // slot corners, 12 bits each (see StepArgs)
  uint16_t *conn_hi = nullptr;
  int32_t *slot_tet = nullptr;
  double *rest64 = nullptr, *vd64 = nullptr;
  float4 *bt4 = nullptr;
  double4 *bt8 = nullptr;
  // deterministic mode (otfem_set_deterministic)
  bool ordered = false;
  int32_t *ord_begin = nullptr, *ord_ent = nullptr;
  double *ordered_terms = nullptr;
  double4 *tile_d = nullptr;  // [ntiles] D per unit volume and unit E of a tile's one nu
  double *tile_kb = nullptr;  // [ntiles] bound of |K'p| / max|p| over the tile's owned nodes
  bool one_nu_per_tile = false;  // slot records may use the folded form
  double *Re = nullptr, *plastic = nullptr, *plastic_next = nullptr, *plastic_param = nullptr;
  BodyState *body = nullptr;
  double *part_bb = nullptr, *part_rr = nullptr, *part4 = nullptr;
  int *done_count = nullptr;  // [2]: stopped bodies, active tiles
  int32_t *act = nullptr;     // [ntiles] active-tile list of batched meshes
  unsigned long long *iter_acc = nullptr, *unconv_acc = nullptr;
  size_t lds_elem = 0, lds_cg = 0, lds_cg32 = 0;  // lds_cg32: CG kernels with FP32 directions
  LaunchConfig lc;
  // captured step graphs (element part, CG part) and the inputs they bake in
  bool use_graphs = true;
  hipGraphExec_t g_elem = nullptr, g_fin = nullptr;
  // (first iteration, iterations, grid) -> captured CG iterations
  std::map<std::tuple<int, int, int>, hipGraphExec_t> g_chunks;
  int cg_stop_seen = 0;  // end of the chunk in which the last step's CG was seen to stop
                                                          // [j C, (j+1) C), built on first use
  hipStream_t cap_stream = nullptr;      // stream used only for graph capture
  int *done_host = nullptr;              // [2][2] pinned copies of done_count
  hipEvent_t chunk_ev[2] = {nullptr, nullptr};
  std::vector<unsigned char> graph_key;
  std::vector<BodyState> last_body;
  bool pending = false;  // asynchronous steps enqueued since the last sync
  double last_elem_ms = 0.0, last_cg_ms = 0.0;
  double acc_elem_ms = 0.0, acc_cg_ms = 0.0, acc_ws_ms = 0.0, acc_sub_ms = 0.0;
  long long timed_steps = 0;
  bool timing = false;
  // per-step event triples of asynchronous timed steps, resolved at sync
  std::vector<hipEvent_t> pool;
  size_t pool_used = 0;
  // learned warm start
  otfem_model *model = nullptr;
  bool recording = false;
  bool ws_suspended = false;     // autotune trials run the plain start
  std::vector<float> kd_host;    // [4N] rest stiffness diagonal, internal order
  int64_t n_free = 0;
  double *vw = nullptr;          // [3N] CG start of learned steps
  float4 *hist1 = nullptr, *hist2 = nullptr, *kd = nullptr;
  int32_t *free_rank = nullptr;
  int *ws_bad = nullptr;
  // subspace start: ring of past velocities, its state, Gram partials
  float4 *vring = nullptr;
  double *vbuf = nullptr;
  int ring_slots = 0;
  int32_t *sub_state = nullptr;
  double *sub_part = nullptr;
  int32_t *body_tiles = nullptr;  // [B+1]
  double *sub_coef = nullptr;     // [B][kSubMax]
  int32_t *prim_begin = nullptr;  // [ntiles+1] primary slots per tile
  int32_t *prim_slot = nullptr;   // [T]
  size_t lds_sub = 0;
  hipGraphExec_t g_sub = nullptr;
};

namespace {

otfem_status fail(otfem_solver *h, otfem_status st, const std::string &msg) {
  if (h) h->err = msg;
  return st;
}

otfem_status hip_fail(otfem_solver *h, hipError_t e, const char *what) {
  return fail(h, OTFEM_HIP_ERROR, std::string(what) + ": " + hipGetErrorString(e));
}

// Runs body(): a failed host allocation inside it becomes
// OTFEM_OUT_OF_MEMORY and any other C++ exception OTFEM_INVALID_ARG (with
// the message on h when there is a handle), so no exception crosses the C
// ABI.
template <class F>
otfem_status no_throw(otfem_solver *h, F &&body) {
  auto note = [h](const char *msg) {
    try {
      if (h) h->err = msg;
    } catch (...) {
    }
  };
  try {
    return body();
  } catch (const std::bad_alloc &) {
    note("out of host memory");
    return OTFEM_OUT_OF_MEMORY;
  } catch (const std::exception &e) {
    note(e.what());
    return OTFEM_INVALID_ARG;
  }
}

#define HIPCHK(h, call)                                  \
  do {                                                   \
    hipError_t e_ = (call);                              \
    if (e_ != hipSuccess) return hip_fail((h), e_, #call); \
  } while (0)

template <typename T>
hipError_t dalloc(T **p, size_t n) {
  if (n == 0) n = 1;
  return hipMalloc(reinterpret_cast<void **>(p), n * sizeof(T));
}

template <typename T>
void dfree(T *&p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

constexpr int kCgChunk = 32;  // CG iterations per captured graph
constexpr int kCgTail = 8;    // iterations per graph near the expected stop
// Batched meshes shrink the grid to the tiles of running bodies once per
// chunk, and a body that stops inside a chunk keeps its workgroups in the
// grid until the chunk ends (each issues its prologue loads, then exits):
// shorter chunks there (the combined 2,048-body set: 154.7 -> 150.2 ms of CG
// per learned step with 8 instead of 32; a single 64^3 body: 1% slower with
// 16, from the extra graph launches).
constexpr int kCgChunkBatched = 8;
inline int cg_chunk(const StepArgs &a) { return a.compact ? kCgChunkBatched : kCgChunk; }

void drop_graphs(otfem_solver *h) {
  if (h->g_elem) (void)hipGraphExecDestroy(h->g_elem);
  if (h->g_fin) (void)hipGraphExecDestroy(h->g_fin);
  if (h->g_sub) (void)hipGraphExecDestroy(h->g_sub);
  h->g_sub = nullptr;
  for (auto &kv : h->g_chunks)
    if (kv.second) (void)hipGraphExecDestroy(kv.second);
  h->g_elem = h->g_fin = nullptr;
  h->g_chunks.clear();
  h->graph_key.clear();
}

void free_device(otfem_solver *h) {
  dfree(h->x); dfree(h->x0); dfree(h->v); dfree(h->fext); dfree(h->mass);
  dfree(h->rhs); dfree(h->r); dfree(h->p); dfree(h->y); dfree(h->y32); dfree(h->r32); dfree(h->fixed);
  dfree(h->tiles); dfree(h->halo); dfree(h->conn_lo); dfree(h->conn_hi); dfree(h->slot_tet);
  dfree(h->rest64); dfree(h->vd64); dfree(h->bt4); dfree(h->bt8); dfree(h->tile_d); dfree(h->tile_kb);
  dfree(h->Re); dfree(h->plastic); dfree(h->plastic_next); dfree(h->plastic_param);
  dfree(h->body); dfree(h->part_bb); dfree(h->part_rr);
  dfree(h->part4); dfree(h->done_count); dfree(h->iter_acc); dfree(h->unconv_acc);
  dfree(h->vw); dfree(h->hist1); dfree(h->hist2); dfree(h->kd); dfree(h->free_rank);
  dfree(h->ws_bad); dfree(h->act);
  dfree(h->ord_begin); dfree(h->ord_ent); dfree(h->ordered_terms);
  dfree(h->vring); dfree(h->vbuf); dfree(h->sub_state); dfree(h->sub_part);
  dfree(h->body_tiles); dfree(h->sub_coef); dfree(h->prim_begin); dfree(h->prim_slot);
  h->ring_slots = 0;
  drop_graphs(h);
  if (h->done_host) (void)hipHostFree(h->done_host);
  h->done_host = nullptr;
  if (h->pin_stage) (void)hipHostFree(h->pin_stage);
  h->pin_stage = nullptr;
  h->pin_half = 0;
  for (hipEvent_t &e : h->pin_ev) {
    if (e) (void)hipEventDestroy(e);
    e = nullptr;
  }
  for (hipEvent_t &e : h->chunk_ev) {
    if (e) (void)hipEventDestroy(e);
    e = nullptr;
  }
  if (h->cap_stream) (void)hipStreamDestroy(h->cap_stream);
  h->cap_stream = nullptr;
  for (hipEvent_t e : h->pool) (void)hipEventDestroy(e);
  h->pool.clear();
  h->pool_used = 0;
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->ev2) (void)hipEventDestroy(h->ev2);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  h->ev0 = h->ev1 = h->ev2 = nullptr;
  h->stream = nullptr;
  h->device_ready = false;
}

// caller order [3N] -> internal order
void to_internal3(const Layout &L, const double *src, std::vector<double> &dst) {
  dst.resize(3 * L.N);
  for (int64_t g = 0; g < L.N; ++g) {
    int64_t u = L.node_of_internal[g];
    dst[3 * g] = src[3 * u];
    dst[3 * g + 1] = src[3 * u + 1];
    dst[3 * g + 2] = src[3 * u + 2];
  }
}

void to_caller3(const Layout &L, const std::vector<double> &src, double *dst) {
  for (int64_t g = 0; g < L.N; ++g) {
    int64_t u = L.node_of_internal[g];
    dst[3 * u] = src[3 * g];
    dst[3 * u + 1] = src[3 * g + 1];
    dst[3 * u + 2] = src[3 * g + 2];
  }
}

// Reads n doubles from device memory through the pinned halves: chunk k+1
// copies into one half while `sink(offset, chunk, count)` consumes chunk k
// from the other.  Halves hold at most kPinHalf doubles, a multiple of 9, so
// chunks hold whole nodes (3 doubles) and whole rotations (9).
constexpr size_t kPinHalf = size_t(9) << 18;  // 18 MiB

template <typename Sink>
otfem_status pinned_read(otfem_solver *h, const double *src, size_t n, Sink &&sink) {
  if (n == 0) return OTFEM_OK;
  const size_t want = std::min(kPinHalf, (n + 8) / 9 * 9);
  if (h->pin_half < want) {
    if (h->pin_stage) HIPCHK(h, hipHostFree(h->pin_stage));
    h->pin_stage = nullptr;
    h->pin_half = 0;
    HIPCHK(h, hipHostMalloc(reinterpret_cast<void **>(&h->pin_stage), 2 * want * sizeof(double),
                            hipHostMallocDefault));
    h->pin_half = want;
  }
  for (hipEvent_t &e : h->pin_ev)
    if (!e) HIPCHK(h, hipEventCreateWithFlags(&e, hipEventDisableTiming));
  const size_t half = h->pin_half, chunks = (n + half - 1) / half;
  auto issue = [&](size_t k) -> hipError_t {
    const size_t off = k * half, len = std::min(half, n - off);
    hipError_t e = hipMemcpyAsync(h->pin_stage + (k & 1) * half, src + off, len * sizeof(double),
                                  hipMemcpyDeviceToHost, h->stream);
    return e != hipSuccess ? e : hipEventRecord(h->pin_ev[k & 1], h->stream);
  };
  HIPCHK(h, issue(0));
  for (size_t k = 0; k < chunks; ++k) {
    if (k + 1 < chunks) HIPCHK(h, issue(k + 1));  // its half was consumed at k - 1
    HIPCHK(h, hipEventSynchronize(h->pin_ev[k & 1]));
    const size_t off = k * half;
    sink(off, h->pin_stage + (k & 1) * half, std::min(half, n - off));
  }
  return OTFEM_OK;
}

// Reads a [3N] node array (internal order) into caller order.
otfem_status read_nodes3(otfem_solver *h, const double *src, double *dst) {
  const int32_t *caller = h->L.node_of_internal.data();
  return pinned_read(h, src, 3 * (size_t)h->L.N, [&](size_t off, const double *p, size_t len) {
    for (size_t i = 0; i < len; i += 3) {
      const int64_t u = caller[(off + i) / 3];
      dst[3 * u] = p[i];
      dst[3 * u + 1] = p[i + 1];
      dst[3 * u + 2] = p[i + 2];
    }
  });
}

// Free-node ranks (internal order) for the sample store.
otfem_status ws_upload_free(otfem_solver *h) {
  const Layout &L = h->L;
  // a step's samples follow the caller's order of its free nodes
  std::vector<int32_t> by_caller(L.N), rank(L.N);
  int64_t n = 0;
  for (int64_t u = 0; u < L.N; ++u) by_caller[u] = h->fixed_host[u] ? -1 : (int32_t)n++;
  for (int64_t g = 0; g < L.N; ++g) rank[g] = by_caller[L.node_of_internal[g]];
  h->n_free = n;
  HIPCHK(h, hipMemcpyAsync(h->free_rank, rank.data(), L.N * sizeof(int32_t), hipMemcpyHostToDevice,
                           h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return OTFEM_OK;
}

// Velocity history of the subspace start: none before this state.
otfem_status sub_reset_history(otfem_solver *h) {
  if (!h->sub_state) return OTFEM_OK;
  HIPCHK(h, hipMemsetAsync(h->sub_state, 0, 3 * sizeof(int32_t), h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return OTFEM_OK;
}

// Velocity histories of the learned and subspace starts: zero / none = at
// rest before this state.
otfem_status ws_reset_history(otfem_solver *h) {
  otfem_status st = sub_reset_history(h);
  if (st != OTFEM_OK || !h->hist1) return st;
  HIPCHK(h, hipMemsetAsync(h->hist1, 0, h->L.N * sizeof(float4), h->stream));
  HIPCHK(h, hipMemsetAsync(h->hist2, 0, h->L.N * sizeof(float4), h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return OTFEM_OK;
}

otfem_status ws_upload_kd(otfem_solver *h) {
  if (!h->kd || h->kd_host.empty()) return OTFEM_OK;
  HIPCHK(h, hipMemcpyAsync(h->kd, h->kd_host.data(), h->L.N * sizeof(float4), hipMemcpyHostToDevice,
                           h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return OTFEM_OK;
}

otfem_status ensure_device(otfem_solver *h) {
  if (h->device_ready) return OTFEM_OK;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    return fail(h, OTFEM_NO_DEVICE, "no HIP device");
  int dev = 0;
  HIPCHK(h, hipGetDevice(&dev));
  hipDeviceProp_t prop;
  HIPCHK(h, hipGetDeviceProperties(&prop, dev));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(h, OTFEM_NO_DEVICE, std::string("otfem is built for gfx950, device is ") +
                                        prop.gcnArchName);
  h->device = dev;
  const Layout &L = h->L;
  const int64_t N = L.N, S = (int64_t)L.slot_tet.size(), T = L.T, B = L.B;
  const int64_t nt = (int64_t)L.tiles.size();
  HIPCHK(h, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  HIPCHK(h, hipStreamCreateWithFlags(&h->cap_stream, hipStreamNonBlocking));
  HIPCHK(h, hipHostMalloc(reinterpret_cast<void **>(&h->done_host), 4 * sizeof(int),
                          hipHostMallocDefault));
  for (hipEvent_t &e : h->chunk_ev) HIPCHK(h, hipEventCreateWithFlags(&e, hipEventDisableTiming));
  HIPCHK(h, hipEventCreate(&h->ev0));
  HIPCHK(h, hipEventCreate(&h->ev1));
  HIPCHK(h, hipEventCreate(&h->ev2));
  HIPCHK(h, dalloc(&h->x, 3 * N));
  HIPCHK(h, dalloc(&h->x0, 3 * N));
  HIPCHK(h, dalloc(&h->v, 3 * N));
  HIPCHK(h, dalloc(&h->fext, 3 * N));
  HIPCHK(h, dalloc(&h->mass, N));
  HIPCHK(h, dalloc(&h->rhs, 3 * N));
  HIPCHK(h, dalloc(&h->r, 6 * N));
  HIPCHK(h, dalloc(&h->p, 6 * N));
  HIPCHK(h, dalloc(&h->y, 6 * N));
  HIPCHK(h, dalloc(&h->y32, 6 * N));
  HIPCHK(h, dalloc(&h->r32, 6 * N));
  HIPCHK(h, dalloc(&h->fixed, N));
  HIPCHK(h, dalloc(&h->tiles, nt));
  HIPCHK(h, dalloc(&h->halo, L.halo.size()));
  HIPCHK(h, dalloc(&h->conn_lo, S));
  HIPCHK(h, dalloc(&h->conn_hi, S));
  HIPCHK(h, dalloc(&h->slot_tet, S));
  HIPCHK(h, dalloc(&h->rest64, 9 * S));
  HIPCHK(h, dalloc(&h->vd64, 3 * S));
  HIPCHK(h, dalloc(&h->bt4, 3 * S));
  HIPCHK(h, dalloc(&h->tile_d, nt));
  HIPCHK(h, dalloc(&h->tile_kb, nt));
  HIPCHK(h, dalloc(&h->body_tiles, B + 1));
  HIPCHK(h, dalloc(&h->Re, 9 * T));
  HIPCHK(h, dalloc(&h->plastic, 6 * T));
  HIPCHK(h, dalloc(&h->plastic_next, 6 * T));
  HIPCHK(h, dalloc(&h->plastic_param, 3 * T));
  HIPCHK(h, dalloc(&h->body, B));
  HIPCHK(h, dalloc(&h->part_bb, nt));
  HIPCHK(h, dalloc(&h->part_rr, nt));
  HIPCHK(h, dalloc(&h->part4, 8 * nt));
  HIPCHK(h, dalloc(&h->done_count, 2));
  HIPCHK(h, dalloc(&h->act, nt));
  HIPCHK(h, dalloc(&h->iter_acc, B));
  HIPCHK(h, dalloc(&h->unconv_acc, B));
  hipStream_t s = h->stream;
  HIPCHK(h, hipMemcpyAsync(h->tiles, L.tiles.data(), nt * sizeof(TileInfo), hipMemcpyHostToDevice, s));
  HIPCHK(h, hipMemcpyAsync(h->body_tiles, L.body_tile_begin.data(), (B + 1) * sizeof(int32_t),
                           hipMemcpyHostToDevice, s));
  if (!L.halo.empty())
    HIPCHK(h, hipMemcpyAsync(h->halo, L.halo.data(), L.halo.size() * sizeof(int32_t),
                             hipMemcpyHostToDevice, s));
  std::vector<uint32_t> conn_lo;  // packed corners, alive until the sync below
  std::vector<uint16_t> conn_hi;
  if (S > 0) {  // a mesh may have no tets (every node fixed)
    conn_lo.resize(S);
    conn_hi.resize(S);
    for (int64_t k = 0; k < S; ++k) {
      const uint32_t c0 = L.slot_conn[4 * k], c1 = L.slot_conn[4 * k + 1];
      const uint32_t c2 = L.slot_conn[4 * k + 2], c3 = L.slot_conn[4 * k + 3];
      conn_lo[k] = c0 | (c1 << 12) | ((c2 & 0xffu) << 24);
      conn_hi[k] = (uint16_t)((c2 >> 8) | (c3 << 4));
    }
    HIPCHK(h, hipMemcpyAsync(h->conn_lo, conn_lo.data(), S * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    HIPCHK(h, hipMemcpyAsync(h->conn_hi, conn_hi.data(), S * sizeof(uint16_t), hipMemcpyHostToDevice, s));
    HIPCHK(h, hipMemcpyAsync(h->slot_tet, L.slot_tet.data(), S * sizeof(int32_t), hipMemcpyHostToDevice, s));
  }
  std::vector<double> tmp;
  to_internal3(L, h->x_init.data(), tmp);
  HIPCHK(h, hipMemcpyAsync(h->x, tmp.data(), 3 * N * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(h, hipStreamSynchronize(s));
  to_internal3(L, h->x0_host.data(), tmp);
  HIPCHK(h, hipMemcpyAsync(h->x0, tmp.data(), 3 * N * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(h, hipStreamSynchronize(s));
  to_internal3(L, h->v_init.data(), tmp);
  HIPCHK(h, hipMemcpyAsync(h->v, tmp.data(), 3 * N * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(h, hipStreamSynchronize(s));
  std::vector<uint8_t> fx(N);
  for (int64_t g = 0; g < N; ++g) fx[g] = h->fixed_host[L.node_of_internal[g]];
  HIPCHK(h, hipMemcpyAsync(h->fixed, fx.data(), N, hipMemcpyHostToDevice, s));
  HIPCHK(h, hipMemsetAsync(h->fext, 0, 3 * N * sizeof(double), s));
  HIPCHK(h, hipMemsetAsync(h->r, 0, 6 * N * sizeof(double), s));
  HIPCHK(h, hipMemsetAsync(h->p, 0, 6 * N * sizeof(double), s));
  HIPCHK(h, hipMemsetAsync(h->y, 0, 6 * N * sizeof(double), s));
  HIPCHK(h, hipMemsetAsync(h->plastic, 0, 6 * T * sizeof(double), s));
  HIPCHK(h, hipMemsetAsync(h->body, 0, B * sizeof(BodyState), s));
  HIPCHK(h, hipMemsetAsync(h->part_bb, 0, nt * sizeof(double), s));
  HIPCHK(h, hipMemsetAsync(h->part_rr, 0, nt * sizeof(double), s));
  HIPCHK(h, hipMemsetAsync(h->part4, 0, 8 * nt * sizeof(double), s));
  HIPCHK(h, hipMemsetAsync(h->done_count, 0, 2 * sizeof(int), s));
  HIPCHK(h, hipMemsetAsync(h->iter_acc, 0, B * sizeof(unsigned long long), s));
  HIPCHK(h, hipMemsetAsync(h->unconv_acc, 0, B * sizeof(unsigned long long), s));
  HIPCHK(h, hipStreamSynchronize(s));
  h->device_ready = true;
  return OTFEM_OK;
}

// Per-tet element data, following fem_compute_ke.h:41-77 (d1..d18) and
// fem_compute_volume.h:27-39.  Gradients are stored with the true sign,
// gradN_k = -B_k of the reference; Ke is quadratic in B, so it is unchanged.
struct ElemData {
  double grad[9];
  double vd[3];
  double V;  // signed volume (compute_volume edge form)
};

// Isotropic material constants of fem_compute_ke.h:74-77 (d16, d17, d18):
// D per unit volume.
void material_d(double E, double nu, double d[3]) {
  const double d15 = E / (1.0 + nu) / (1.0 - 2 * nu);
  d[0] = (1.0 - nu) * d15;
  d[1] = nu * d15;
  d[2] = E / 2 / (1.0 + nu);
}

ElemData element_data(const double *p0, const double *p1, const double *p2, const double *p3,
                      double E, double nu) {
  ElemData e;
  double d = p0[0], d4 = p0[1], d8 = p0[2];
  double d1 = p1[0] - d, d5 = p1[1] - d4, d9 = p1[2] - d8;
  double d2 = p2[0] - d, d6 = p2[1] - d4, d10 = p2[2] - d8;
  double d3 = p3[0] - d, d7 = p3[1] - d4, d11 = p3[2] - d8;
  double d12 = (d1 * d6 * d11 + d2 * d7 * d9 + d3 * d5 * d10) - d1 * d7 * d10 - d2 * d5 * d11 -
               d3 * d6 * d9;
  double d13 = 1.0 / d12;
  double d14 = std::fabs(d12 / 6.0);
  double B1[3] = {(d10 * d7 - d6 * d11) * d13, (d2 * d11 - d10 * d3) * d13, (d6 * d3 - d2 * d7) * d13};
  double B2[3] = {(d5 * d11 - d9 * d7) * d13, (d9 * d3 - d1 * d11) * d13, (d1 * d7 - d5 * d3) * d13};
  double B3[3] = {(d9 * d6 - d5 * d10) * d13, (d1 * d10 - d9 * d2) * d13, (d5 * d2 - d1 * d6) * d13};
  for (int c = 0; c < 3; ++c) {
    e.grad[c] = -B1[c];
    e.grad[3 + c] = -B2[c];
    e.grad[6 + c] = -B3[c];
  }
  double dm[3];
  material_d(E, nu, dm);
  e.vd[0] = dm[0] * d14;
  e.vd[1] = dm[1] * d14;
  e.vd[2] = dm[2] * d14;
  // e10 . (e20 % e30) / 6 with operator% of math_vector3.h:193
  double cx = d6 * d11 - d7 * d10, cy = d3 * d10 - d2 * d11, cz = d2 * d7 - d3 * d6;
  e.V = (d1 * cx + d5 * cy + d9 * cz) / 6.0;
  return e;
}

otfem_status init_common(otfem_solver *h, const double *E, const double *nu, const double *rho,
                         bool uniform, double c_yield, double c_creep, double c_max) {
  const Layout &L = h->L;
  const int64_t T = L.T, S = (int64_t)L.slot_tet.size(), N = L.N;
  if (!(c_yield >= 0) || !(c_creep >= 0) || !(c_max >= 0))
    return fail(h, OTFEM_INVALID_ARG, "plastic parameters must be non-negative");
  // everything is checked before the solver changes: a rejected init leaves
  // the previous materials, element data and masses in place
  // The loops below run on host threads, in chunks of tets, slots or nodes;
  // per-node sums run per body (a body's tets touch only its own nodes), each
  // in tet order, so every value is the one a single thread computes.
  const int nth = host_threads();
  constexpr int64_t kChunk = int64_t(1) << 16;
  auto par = [&](int64_t n, auto &&f) {  // f(i0, i1) over [0, n)
    const int64_t nc = (n + kChunk - 1) / kChunk;
    parallel_for(nc, nth, [&](int64_t c) { f(c * kChunk, std::min(n, (c + 1) * kChunk)); });
  };
  const int64_t nb = std::max<int64_t>(L.B, 1);
  bool by_body = L.B > 1 && (int64_t)L.body_tet_begin.size() == L.B + 1 &&
                 (int64_t)L.body_node_begin.size() == L.B + 1;
  if (by_body) {
    std::atomic<bool> inside{true};
    parallel_for(nb, nth, [&](int64_t b) {
      const int64_t n0 = L.body_node_begin[b], n1 = L.body_node_begin[b + 1];
      for (int64_t t = L.body_tet_begin[b]; t < L.body_tet_begin[b + 1]; ++t)
        for (int k = 0; k < 4; ++k) {
          const int64_t g = h->tets[4 * t + k];
          if (g < n0 || g >= n1) inside.store(false, std::memory_order_relaxed);
        }
    });
    by_body = inside.load();
  }
  // per_node(f): f(t) for every tet, in tet order within each body (or all
  // tets in order on one thread), for sums into the tet's nodes
  auto per_node = [&](auto &&f) {
    if (!by_body) {
      for (int64_t t = 0; t < T; ++t) f(t);
      return;
    }
    parallel_for(nb, nth, [&](int64_t b) {
      for (int64_t t = L.body_tet_begin[b]; t < L.body_tet_begin[b + 1]; ++t) f(t);
    });
  };
  std::vector<double> young(T), poisson(T), density(T);
  std::atomic<bool> bad_material{false};
  par(T, [&](int64_t t0, int64_t t1) {
    for (int64_t t = t0; t < t1; ++t) {
      double e = uniform ? E[0] : E[t], n = uniform ? nu[0] : nu[t], r = uniform ? rho[0] : rho[t];
      // fem_uniform_young.h:32, fem_uniform_poisson.h:32-33, density > 0; nu = 0.5
      // is rejected as well since it makes D infinite.
      if (!(e >= 0) || !(n >= 0) || !(n < 0.5) || !(r > 0) || !std::isfinite(e) || !std::isfinite(r))
        bad_material.store(true, std::memory_order_relaxed);
      young[t] = e;
      poisson[t] = n;
      density[t] = r;
    }
  });
  if (bad_material.load())
    return fail(h, OTFEM_INVALID_ARG, "material out of range (E>=0, 0<=nu<0.5, rho>0)");
  // 104 B per tet (6.1 GB on 59 M tets): left uninitialised, every entry is
  // written by the thread that computes it
  std::unique_ptr<ElemData[]> ed(new ElemData[std::max<int64_t>(T, 1)]);
  std::vector<double> vol(T);
  std::atomic<int64_t> inverted{0};
  const double *x0 = h->x0_host.data();
  par(T, [&](int64_t a0, int64_t a1) {
    int64_t inv = 0;
    for (int64_t t = a0; t < a1; ++t) {
      const int32_t *tt = &h->tets[4 * t];
      ed[t] = element_data(x0 + 3 * tt[0], x0 + 3 * tt[1], x0 + 3 * tt[2], x0 + 3 * tt[3],
                           young[t], poisson[t]);
      vol[t] = ed[t].V;
      if (!(ed[t].V > 0)) ++inv;
    }
    inverted.fetch_add(inv, std::memory_order_relaxed);
  });
  if (inverted.load())
    return fail(h, OTFEM_INVERTED_TET,
                std::to_string(inverted.load()) + " tetrahedra have non-positive rest volume");
  // fem_compute_mass.h:154-171 in caller order
  std::vector<double> mass(N, 0.0);
  for (int64_t i = 0; i < N; ++i) mass[i] = h->fixed_host[i] ? std::numeric_limits<double>::max() : 0.0;
  per_node([&](int64_t t) {
    double amount = density[t] * ed[t].V * .25;
    for (int k = 0; k < 4; ++k) mass[h->tets[4 * t + k]] += amount;
  });
  // A free node outside every tetrahedron has zero mass and an empty row of A;
  // the reference would divide by zero in CG.  Reject it (fix or drop it).
  for (int64_t i = 0; i < N; ++i)
    if (!h->fixed_host[i] && !(mass[i] > 0.0))
      return fail(h, OTFEM_INVALID_ARG,
                  "free node " + std::to_string(i) + " belongs to no tetrahedron (zero mass)");
  otfem_status st = ensure_device(h);
  if (st != OTFEM_OK) return st;
  const size_t lds_elem = element_lds_bytes(L.max_tile_nodes, L.max_tile_owned,
                                            (c_creep > 0.0 && c_max > 0.0) ? 1 : 0);
  const size_t lds_cg = cg_lds_bytes(L.max_tile_nodes, L.max_tile_owned);
  {
    hipDeviceProp_t prop;
    HIPCHK(h, hipGetDeviceProperties(&prop, h->device));
    // the kernels' static LDS (block reductions) is under 1 KiB
    constexpr size_t kStaticLds = 1024;
    if (lds_elem + kStaticLds > (size_t)prop.sharedMemPerBlock ||
        lds_cg + kStaticLds > (size_t)prop.sharedMemPerBlock)
      return fail(h, OTFEM_INVALID_ARG, "tile too large for LDS; lower target_tile_nodes");
  }
  h->young.swap(young);
  h->poisson.swap(poisson);
  h->density.swap(density);
  h->vol_host.swap(vol);
  h->mass_host.swap(mass);
  // Diagonal of the rest stiffness K_ii (R = I): per tet and corner k with
  // gradient b, (Ke_kk)_dd = D0 b_d^2 + D2 (|b|^2 - b_d^2) (fem_compute_ke.h:78-99,
  // with D = V (d16, d17, d18)); feature of the learned warm start.
  {
    std::vector<double> kdc(3 * N, 0.0);
    per_node([&](int64_t t) {
      const double *gr = ed[t].grad;
      double b[4][3];
      for (int d = 0; d < 3; ++d) {
        b[1][d] = gr[d];
        b[2][d] = gr[3 + d];
        b[3][d] = gr[6 + d];
        b[0][d] = -(b[1][d] + b[2][d] + b[3][d]);
      }
      for (int k = 0; k < 4; ++k) {
        const double n2 = b[k][0] * b[k][0] + b[k][1] * b[k][1] + b[k][2] * b[k][2];
        for (int d = 0; d < 3; ++d)
          kdc[3 * (int64_t)h->tets[4 * t + k] + d] +=
              ed[t].vd[0] * b[k][d] * b[k][d] + ed[t].vd[2] * (n2 - b[k][d] * b[k][d]);
      }
    });
    h->kd_host.assign(4 * N, 0.0f);
    par(N, [&](int64_t g0, int64_t g1) {
      for (int64_t g = g0; g < g1; ++g)
        for (int d = 0; d < 3; ++d) h->kd_host[4 * g + d] = (float)kdc[3 * L.node_of_internal[g] + d];
    });
  }
  // One Poisson ratio per tile: D of a tet is then E |V| times constants of
  // the tile, so the CG's slot records fold sqrt(E |V|) into the gradients
  // and take the constants from the tile (44 instead of 56 bytes per slot).
  // (td outlives the stream synchronisation below: the copy is asynchronous)
  std::vector<double4> td(L.tiles.size(), make_double4(0.0, 0.0, 0.0, 0.0));
  {
    const int64_t nt = (int64_t)L.tiles.size();
    std::atomic<bool> one{true};
    parallel_for(nt, nth, [&](int64_t i) {
      const TileInfo &ti = L.tiles[i];
      if (ti.n_slots == 0) return;
      const int64_t t0 = L.slot_tet[ti.slot_begin] & 0x7fffffff;
      for (int64_t s = ti.slot_begin; s < ti.slot_begin + ti.n_slots; ++s) {
        const int64_t t = L.slot_tet[s] & 0x7fffffff;
        if (h->poisson[t] != h->poisson[t0]) {
          one.store(false, std::memory_order_relaxed);
          break;
        }
      }
      double dm[3];
      material_d(1.0, h->poisson[t0], dm);
      td[i] = make_double4(dm[0], dm[1], dm[2], 0.0);
    });
    h->one_nu_per_tile = one.load();
    if (nt > 0)
      HIPCHK(h, hipMemcpyAsync(h->tile_d, td.data(), nt * sizeof(double4), hipMemcpyHostToDevice,
                               h->stream));
  }
  // Bound of the deterministic CG's fixed-point node sums: per tile, the
  // largest over owned nodes of sqrt(12) sum over the node's tets of
  // trace(Ke) = sum_k (D0 + 2 D2) |b_k|^2 (D with |V|), since
  // |(K'p)_i| <= sum_e ||Ke||_2 ||p_e||_2 and the rotation keeps the norm.
  std::vector<double> kb(L.tiles.size(), 0.0);
  {
    std::vector<double> kn(N, 0.0);
    per_node([&](int64_t t) {
      const double *gr = ed[t].grad;
      double tr = 0.0;
      for (int k = 0; k < 4; ++k) {
        double n2 = 0.0;
        for (int d = 0; d < 3; ++d) {
          const double b = k == 0 ? -(gr[d] + gr[3 + d] + gr[6 + d]) : gr[3 * (k - 1) + d];
          n2 += b * b;
        }
        tr += (ed[t].vd[0] + 2.0 * ed[t].vd[2]) * n2;
      }
      for (int k = 0; k < 4; ++k) kn[h->tets[4 * t + k]] += std::sqrt(12.0) * tr;
    });
    parallel_for((int64_t)L.tiles.size(), nth, [&](int64_t i) {
      const TileInfo &ti = L.tiles[i];
      for (int l = 0; l < ti.n_owned; ++l)
        kb[i] = std::max(kb[i], kn[L.node_of_internal[ti.node_begin + l]]);
      kb[i] *= 1.0 + 1e-6;  // the slot records' FP32 rounding
    });
    if (!kb.empty())
      HIPCHK(h, hipMemcpyAsync(h->tile_kb, kb.data(), kb.size() * sizeof(double),
                               hipMemcpyHostToDevice, h->stream));
  }
  // per-slot element data (SoA), filled by the threads that first touch it
  std::unique_ptr<double[]> rest(new double[std::max<int64_t>(9 * S, 1)]);
  std::unique_ptr<double[]> vd(new double[std::max<int64_t>(3 * S, 1)]);
  par(S, [&](int64_t s0, int64_t s1) {
    for (int64_t s = s0; s < s1; ++s) {
      int64_t t = L.slot_tet[s] & 0x7fffffff;
      for (int k = 0; k < 9; ++k) rest[k * S + s] = ed[t].grad[k];
      for (int k = 0; k < 3; ++k) vd[k * S + s] = ed[t].vd[k];
    }
  });
  std::vector<double> mint(N), pp(3 * T);
  par(N, [&](int64_t g0, int64_t g1) {
    for (int64_t g = g0; g < g1; ++g) mint[g] = h->mass_host[L.node_of_internal[g]];
  });
  par(T, [&](int64_t t0, int64_t t1) {
    for (int64_t t = t0; t < t1; ++t) {
      pp[3 * t] = c_yield;
      pp[3 * t + 1] = c_creep;
      pp[3 * t + 2] = c_max;
    }
  });
  h->plastic_on = (c_creep > 0.0 && c_max > 0.0);
  h->plastic_c[0] = c_yield;
  h->plastic_c[1] = c_creep;
  h->plastic_c[2] = c_max;
  hipStream_t s = h->stream;
  if (S > 0) {
    HIPCHK(h, hipMemcpyAsync(h->rest64, rest.get(), 9 * S * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(h, hipMemcpyAsync(h->vd64, vd.get(), 3 * S * sizeof(double), hipMemcpyHostToDevice, s));
  }
  HIPCHK(h, hipMemcpyAsync(h->mass, mint.data(), N * sizeof(double), hipMemcpyHostToDevice, s));
  if (T > 0)
    HIPCHK(h, hipMemcpyAsync(h->plastic_param, pp.data(), 3 * T * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(h, hipMemsetAsync(h->plastic, 0, 6 * T * sizeof(double), s));
  HIPCHK(h, hipStreamSynchronize(s));

  h->lds_elem = lds_elem;
  h->lds_cg = lds_cg;
  h->lds_cg32 = cg_lds_bytes(L.max_tile_nodes, L.max_tile_owned, 4);
  // material data changed: captured graphs are rebuilt on the next step
  drop_graphs(h);
  h->initialized = true;
  h->rotations_valid = false;
  st = ws_upload_kd(h);
  if (st != OTFEM_OK) return st;
  return ws_reset_history(h);
}

StepArgs make_args(otfem_solver *h, double dt, int warp, const otfem_params &P) {
  StepArgs a{};
  a.x = h->x; a.x0 = h->x0; a.v = h->v; a.vw = h->v; a.fext = h->fext; a.mass = h->mass; a.fixed = h->fixed;
  a.rhs = h->rhs; a.r = h->r; a.p = h->p; a.y = h->y; a.y32 = h->y32; a.r32 = h->r32;
  a.tiles = h->tiles; a.halo = h->halo; a.conn_lo = h->conn_lo; a.conn_hi = h->conn_hi; a.slot_tet = h->slot_tet;
  a.rest64 = h->rest64; a.vd64 = h->vd64; a.bt4 = h->bt4; a.bt8 = h->bt8;
  a.Re = h->Re; a.plastic = h->plastic; a.plastic_next = h->plastic_next;
  a.plastic_param = h->plastic_param;
  a.body = h->body;
  a.part_bb = h->part_bb; a.part_rr = h->part_rr; a.part4 = h->part4;
  a.done_count = h->done_count; a.iter_acc = h->iter_acc; a.unconv_acc = h->unconv_acc;
  a.act = h->act;
  a.compact = h->L.B > 1 ? 1 : 0;
  a.N = h->L.N;
  a.S = (int64_t)h->L.slot_tet.size();
  a.T = h->L.T;
  a.ntiles = (int32_t)h->L.tiles.size();
  a.B = (int32_t)h->L.B;
  a.dt = dt;
  a.cM = 1.0 + P.alpha * dt;
  a.cK = dt * dt + P.beta * dt;
  a.tol = P.tol;
  a.tiny = P.tiny;
  a.tol_mode = P.tol_mode;
  a.cg_min = P.cg_min_iterations;
  a.cg_max = P.cg_max_iterations;
  a.warp = warp ? 1 : 0;
  a.rotation_mode = P.rotation_mode;
  a.store_rotations = P.store_rotations ? 1 : 0;
  a.plastic_on = h->plastic_on ? 1 : 0;
  a.store64 = (h->lc.spmv_mode == SPMV_F64_STORE64) ? 1 : 0;
  a.tile_d = h->tile_d;
  a.tile_kb = h->tile_kb;
  a.folded = (h->one_nu_per_tile && h->lc.spmv_mode == SPMV_F64) ? 1 : 0;
  // FP32 directions where CG runs to a relative residual: the converged
  // iterate does not depend on them beyond the tolerance, while the
  // fixed-count iterate of the reference mode would
  a.p32 = (P.tol_mode == 1 && h->lc.spmv_mode == SPMV_F64) ? 1 : 0;
  // and q = A p in FP32 once a body's relative residual is below 1e4 tol:
  // r -= alpha q then drifts from the true residual by ~6e-8 of the later
  // |alpha q|, far below tol (the FP64 oracle with q rounded this way: true
  // relative residuals at the stop as with FP64 q; rounded from the first
  // iteration, up to 9e-6 at tol 1e-6)
  const double q32_rel = 1e4 * P.tol;
  a.q32_rr = a.p32 ? q32_rel * q32_rel : 0.0;
  a.ordered = h->ordered ? 1 : 0;
  a.ord_begin = h->ord_begin;
  a.ord_ent = h->ord_ent;
  a.ordered_terms = h->ordered_terms;
  a.sub_hist = P.subspace > 0 && h->vring ? P.subspace : 0;
  a.vring = h->vring;
  a.vbuf = h->vbuf;
  a.sub_state = h->sub_state;
  a.sub_ring = h->ring_slots;
  a.sub_part = h->sub_part;
  a.body_tiles = h->body_tiles;
  a.sub_coef = h->sub_coef;
  a.prim_begin = h->prim_begin;
  a.prim_slot = h->prim_slot;
  a.sub_net = 0;
  a.vnet = nullptr;
  return a;
}

// Buffers of the subspace start for `hist` velocity differences (a ring of
// hist slots and the last two velocities; a larger request reallocates the
// ring and starts the history anew).
otfem_status sub_ensure(otfem_solver *h, int hist, int directions) {
  const int64_t N = h->L.N, nt = (int64_t)h->L.tiles.size(), B = h->L.B;
  if (!h->sub_part) HIPCHK(h, dalloc(&h->sub_part, kSubPart * std::max<int64_t>(nt, 1)));
  if (!h->sub_coef) {
    HIPCHK(h, dalloc(&h->sub_coef, kSubMax * B));
    HIPCHK(h, hipMemsetAsync(h->sub_coef, 0, kSubMax * B * sizeof(double), h->stream));
  }
  if (!h->sub_state) {
    HIPCHK(h, dalloc(&h->sub_state, 3));
    HIPCHK(h, hipMemsetAsync(h->sub_state, 0, 3 * sizeof(int32_t), h->stream));
  }
  if (!h->vbuf) HIPCHK(h, dalloc(&h->vbuf, 6 * N));
  if (!h->prim_begin) {
    // every tile's primary slots (one per tet), in slot order
    std::vector<int32_t> pb(nt + 1, 0), ps;
    ps.reserve(h->L.T);
    for (int64_t i = 0; i < nt; ++i) {
      const TileInfo &ti = h->L.tiles[i];
      for (int32_t sl = ti.slot_begin; sl < ti.slot_begin + ti.n_slots; ++sl)
        if (h->L.slot_tet[sl] < 0) ps.push_back(sl);
      pb[i + 1] = (int32_t)ps.size();
    }
    HIPCHK(h, dalloc(&h->prim_begin, nt + 1));
    HIPCHK(h, dalloc(&h->prim_slot, std::max<size_t>(ps.size(), 1)));
    HIPCHK(h, hipMemcpyAsync(h->prim_begin, pb.data(), (nt + 1) * sizeof(int32_t), hipMemcpyHostToDevice,
                             h->stream));
    if (!ps.empty())
      HIPCHK(h, hipMemcpyAsync(h->prim_slot, ps.data(), ps.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                               h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
  }
  if (h->ring_slots < hist) {
    // steps enqueued earlier may still read the old ring
    HIPCHK(h, hipStreamSynchronize(h->stream));
    drop_graphs(h);
    dfree(h->vring);
    h->ring_slots = 0;
    HIPCHK(h, dalloc(&h->vring, (size_t)hist * N));
    h->ring_slots = hist;
    HIPCHK(h, hipMemsetAsync(h->sub_state, 0, 3 * sizeof(int32_t), h->stream));
  }
  const size_t lds = sub_gram_lds_bytes(h->L.max_tile_nodes, h->L.max_tile_owned, directions);
  hipDeviceProp_t prop;
  HIPCHK(h, hipGetDeviceProperties(&prop, h->device));
  // a launch needs its dynamic and static LDS within the workgroup's share
  if (lds + sub_gram_static_lds_bytes() > (size_t)prop.sharedMemPerBlock)
    return fail(h, OTFEM_INVALID_ARG,
                "tile too large for the subspace start's LDS; use fewer directions or lower target_tile_nodes");
  h->lds_sub = std::max(h->lds_sub, lds);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return OTFEM_OK;
}

otfem_status resolve_timing(otfem_solver *h) {
  // five events per timed step: start, after the network, after the
  // element kernel, after the subspace start, after the CG kernels
  for (size_t k = 0; 5 * k + 4 < h->pool_used; ++k) {
    float e0 = 0.f, e1 = 0.f, e2 = 0.f, e3 = 0.f;
    HIPCHK(h, hipEventElapsedTime(&e0, h->pool[5 * k], h->pool[5 * k + 1]));
    HIPCHK(h, hipEventElapsedTime(&e1, h->pool[5 * k + 1], h->pool[5 * k + 2]));
    HIPCHK(h, hipEventElapsedTime(&e3, h->pool[5 * k + 2], h->pool[5 * k + 3]));
    HIPCHK(h, hipEventElapsedTime(&e2, h->pool[5 * k + 3], h->pool[5 * k + 4]));
    h->acc_ws_ms += e0;
    h->last_elem_ms = e1 + e3;
    h->last_cg_ms = e2;
    h->acc_elem_ms += e1;
    h->acc_sub_ms += e3;
    h->acc_cg_ms += e2;
    h->timed_steps++;
  }
  h->pool_used = 0;
  return OTFEM_OK;
}

otfem_status sync_and_check(otfem_solver *h, otfem_stats *stats) {
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->pending = false;
  otfem_status ts = resolve_timing(h);
  if (ts != OTFEM_OK) return ts;
  int ws_bad = 0;
  if (h->ws_bad) {
    HIPCHK(h, hipMemcpy(&ws_bad, h->ws_bad, sizeof(int), hipMemcpyDeviceToHost));
    if (ws_bad) HIPCHK(h, hipMemset(h->ws_bad, 0, sizeof(int)));
  }
  h->last_body.resize(h->L.B);
  HIPCHK(h, hipMemcpy(h->last_body.data(), h->body, h->L.B * sizeof(BodyState),
                      hipMemcpyDeviceToHost));
  otfem_status ret = OTFEM_OK;
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    for (const BodyState &b : h->last_body) {
      stats->iterations = std::max(stats->iterations, (int)b.iterations);
      stats->bodies_converged += b.converged ? 1 : 0;
      stats->flags |= b.flags;
    }
    stats->bodies_capped = (int)h->L.B - stats->bodies_converged;
    stats->rr0 = h->last_body[0].rr0;
    stats->rr = h->last_body[0].rr;
    stats->bb = h->last_body[0].bb;
    stats->elapsed_ms = h->last_elem_ms + h->last_cg_ms;
    if (ws_bad) stats->flags |= 2;
  }
  for (const BodyState &b : h->last_body)
    if (b.flags & 1) ret = fail(h, OTFEM_NONFINITE, "a CG scalar became non-finite");
  return ret;
}

}  // namespace

extern "C" {

const char *otfem_version(void) { return kVersion; }

const char *otfem_status_string(int status) {
  switch (status) {
    case OTFEM_OK: return "ok";
    case OTFEM_INVALID_ARG: return "invalid argument";
    case OTFEM_INVERTED_TET: return "inverted tetrahedron";
    case OTFEM_NONFINITE: return "non-finite value";
    case OTFEM_NOT_CONVERGED: return "not converged";
    case OTFEM_HIP_ERROR: return "HIP error";
    case OTFEM_NO_DEVICE: return "no device";
    case OTFEM_STATE: return "invalid state";
    case OTFEM_OUT_OF_MEMORY: return "out of host memory";
    default: return "unknown status";
  }
}

otfem_status otfem_default_params(otfem_params *p, int mode) {
  if (!p) return OTFEM_INVALID_ARG;
  std::memset(p, 0, sizeof(*p));
  if (mode == 0) {
    p->alpha = 2.0;
    p->beta = 0.0;
    p->cg_min_iterations = 20;
    p->cg_max_iterations = 20;
    p->tol_mode = OTFEM_TOL_ABS_R2;
    p->tol = 1e-3;
    p->tiny = 1e-10;
  } else if (mode == 1) {
    p->alpha = 0.1;
    p->beta = 0.001;
    p->cg_min_iterations = 0;
    p->cg_max_iterations = 200;
    p->tol_mode = OTFEM_TOL_REL_B;
    p->tol = 1e-6;
    p->tiny = 0.0;
  } else {
    return OTFEM_INVALID_ARG;
  }
  p->rotation_mode = OTFEM_ROT_GRAM_SCHMIDT;
  p->warm_start = OTFEM_WARM_PREVIOUS_VELOCITY;
  p->store_rotations = 0;
  return OTFEM_OK;
}

otfem_status otfem_device_count(int *count) {
  if (!count) return OTFEM_INVALID_ARG;
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
  *count = c;
  return OTFEM_OK;
}

// Why the calling thread's last otfem_create failed (no handle exists then).
static thread_local std::string g_create_err;

const char *otfem_create_last_error(void) { return g_create_err.c_str(); }

otfem_status otfem_create(const otfem_mesh_desc *d, otfem_handle *out) {
  g_create_err.clear();
  auto bad = [](const char *why) {
    try {
      g_create_err = why;
    } catch (...) {
    }
    return OTFEM_INVALID_ARG;
  };
  if (!d || !out) return bad("null descriptor or output handle");
  if (!d->coord || !d->tets) return bad("coordinates and tetrahedra are required");
  if (d->num_nodes <= 0 || d->num_tets < 0) return bad("need at least one node and no negative tet count");
  *out = nullptr;
  const otfem_status st = no_throw(nullptr, [&]() -> otfem_status {
    std::unique_ptr<otfem_solver> h(new otfem_solver());
    const int64_t N = d->num_nodes, T = d->num_tets;
    h->x_init.assign(d->coord, d->coord + 3 * N);
    h->x0_host.assign(d->model_coord ? d->model_coord : d->coord,
                      (d->model_coord ? d->model_coord : d->coord) + 3 * N);
    if (d->velocity)
      h->v_init.assign(d->velocity, d->velocity + 3 * N);
    else
      h->v_init.assign(3 * N, 0.0);
    if (d->fixed)
      h->fixed_host.assign(d->fixed, d->fixed + N);
    else
      h->fixed_host.assign(N, 0);
    for (auto &f : h->fixed_host) f = f ? 1 : 0;
    h->tets.assign(d->tets, d->tets + 4 * T);
    for (double c : h->x_init)
      if (!std::isfinite(c)) return bad("non-finite coordinate");
    int B = d->num_bodies > 1 ? d->num_bodies : 1;
    if (B > 1 && (!d->body_node_offsets || !d->body_tet_offsets))
      return bad("num_bodies > 1 needs body_node_offsets and body_tet_offsets");
    std::string e = build_layout(N, T, h->x0_host.data(), h->tets.data(), B, d->body_node_offsets,
                                 d->body_tet_offsets, d->target_tile_nodes, kMaxTileNodes, h->L);
    if (!e.empty()) return bad(e.c_str());
    *out = h.release();
    return OTFEM_OK;
  });
  if (st != OTFEM_OK && g_create_err.empty())
    bad(st == OTFEM_OUT_OF_MEMORY ? "out of host memory" : "mesh layout failed");
  return st;
}

otfem_status otfem_destroy(otfem_handle h) {
  if (!h) return OTFEM_INVALID_ARG;
  if (h->device_ready) {
    (void)hipStreamSynchronize(h->stream);
    free_device(h);
  }
  delete h;
  return OTFEM_OK;
}

const char *otfem_last_error(otfem_handle h) { return h ? h->err.c_str() : "null handle"; }

static otfem_status otfem_init_body(otfem_handle h, double young, double poisson, double density,
                                    double c_yield, double c_creep, double c_max) {
  if (!h) return OTFEM_INVALID_ARG;
  return init_common(h, &young, &poisson, &density, true, c_yield, c_creep, c_max);
}

otfem_status otfem_init(otfem_handle h, double young, double poisson, double density,
                        double c_yield, double c_creep, double c_max) {
  return no_throw(h, [&] {
    return otfem_init_body(h, young, poisson, density, c_yield, c_creep, c_max);
  });
}

static otfem_status otfem_init_materials_body(otfem_handle h, const double *young, const double *poisson,
                                              const double *density, double c_yield, double c_creep,
                                              double c_max) {
  if (!h || !young || !poisson || !density) return OTFEM_INVALID_ARG;
  return init_common(h, young, poisson, density, false, c_yield, c_creep, c_max);
}

otfem_status otfem_init_materials(otfem_handle h, const double *young, const double *poisson,
                                  const double *density, double c_yield, double c_creep,
                                  double c_max) {
  return no_throw(h, [&] {
    return otfem_init_materials_body(h, young, poisson, density, c_yield, c_creep, c_max);
  });
}

otfem_status otfem_set_external_forces(otfem_handle h, const double *f) {
  if (!h || !f) return OTFEM_INVALID_ARG;
  otfem_status st = ensure_device(h);
  if (st) return st;
  std::vector<double> &tmp = h->stage;
  to_internal3(h->L, f, tmp);
  HIPCHK(h, hipMemcpyAsync(h->fext, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice,
                           h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return OTFEM_OK;
}

otfem_status otfem_get_external_forces(otfem_handle h, double *f) {
  if (!h || !f) return OTFEM_INVALID_ARG;
  otfem_status st = ensure_device(h);
  if (st) return st;
  return read_nodes3(h, h->fext, f);
}

otfem_status otfem_set_state(otfem_handle h, const double *x, const double *v) {
  if (!h) return OTFEM_INVALID_ARG;
  otfem_status st = ensure_device(h);
  if (st) return st;
  std::vector<double> &tmp = h->stage;
  if (x) {
    to_internal3(h->L, x, tmp);
    HIPCHK(h, hipMemcpyAsync(h->x, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice,
                             h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
  }
  if (v) {
    to_internal3(h->L, v, tmp);
    HIPCHK(h, hipMemcpyAsync(h->v, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice,
                             h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return ws_reset_history(h);  // new velocities: the old history no longer leads here
  }
  return OTFEM_OK;
}

otfem_status otfem_get_state(otfem_handle h, double *x, double *v) {
  if (!h) return OTFEM_INVALID_ARG;
  otfem_status st = ensure_device(h);
  if (st) return st;
  if (x && (st = read_nodes3(h, h->x, x)) != OTFEM_OK) return st;
  if (v && (st = read_nodes3(h, h->v, v)) != OTFEM_OK) return st;
  return OTFEM_OK;
}

otfem_status otfem_set_fixed(otfem_handle h, const unsigned char *fixed) {
  if (!h || !fixed) return OTFEM_INVALID_ARG;
  for (int64_t i = 0; i < h->L.N; ++i) h->fixed_host[i] = fixed[i] ? 1 : 0;
  if (!h->device_ready) return OTFEM_OK;
  std::vector<uint8_t> fx(h->L.N);
  for (int64_t g = 0; g < h->L.N; ++g) fx[g] = h->fixed_host[h->L.node_of_internal[g]];
  HIPCHK(h, hipMemcpyAsync(h->fixed, fx.data(), fx.size(), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (h->free_rank) return ws_upload_free(h);
  return OTFEM_OK;
}

otfem_status otfem_get_mass(otfem_handle h, double *mass) {
  if (!h || !mass) return OTFEM_INVALID_ARG;
  if (!h->initialized) return fail(h, OTFEM_STATE, "call otfem_init first");
  std::copy(h->mass_host.begin(), h->mass_host.end(), mass);
  return OTFEM_OK;
}

otfem_status otfem_get_volumes(otfem_handle h, double *V) {
  if (!h || !V) return OTFEM_INVALID_ARG;
  if (!h->initialized) return fail(h, OTFEM_STATE, "call otfem_init first");
  std::copy(h->vol_host.begin(), h->vol_host.end(), V);
  return OTFEM_OK;
}

otfem_status otfem_get_rotations(otfem_handle h, double *Re) {
  if (!h || !Re) return OTFEM_INVALID_ARG;
  if (!h->initialized) return fail(h, OTFEM_STATE, "call otfem_init first");
  if (!h->rotations_valid) {
    // fem_initialize_stiffness_elements.h:50: Re = I until a warped step stores it
    for (int64_t t = 0; t < h->L.T; ++t)
      for (int i = 0; i < 9; ++i) Re[9 * t + i] = (i % 4 == 0) ? 1.0 : 0.0;
    return OTFEM_OK;
  }
  return pinned_read(h, h->Re, 9 * (size_t)h->L.T, [&](size_t off, const double *p, size_t len) {
    std::memcpy(Re + off, p, len * sizeof(double));
  });
}

otfem_status otfem_get_plastic(otfem_handle h, double *plastic) {
  if (!h || !plastic) return OTFEM_INVALID_ARG;
  if (!h->initialized) return fail(h, OTFEM_STATE, "call otfem_init first");
  HIPCHK(h, hipMemcpyAsync(plastic, h->plastic, 6 * h->L.T * sizeof(double), hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return OTFEM_OK;
}

otfem_status otfem_step(otfem_handle h, double dt, int use_stiffness_warping,
                        const otfem_params *params, otfem_stats *stats) {
  if (!h) return OTFEM_INVALID_ARG;
  if (!h->initialized) return fail(h, OTFEM_STATE, "call otfem_init before otfem_step");
  if (!(dt > 0) || !std::isfinite(dt)) return fail(h, OTFEM_INVALID_ARG, "dt must be positive");
  otfem_params P;
  if (params)
    P = *params;
  else
    otfem_default_params(&P, 0);
  if (P.cg_max_iterations < 0 || P.cg_min_iterations < 0 || !(P.tol >= 0) || !(P.tiny >= 0) ||
      (P.tol_mode != 0 && P.tol_mode != 1) || (P.rotation_mode != 0 && P.rotation_mode != 1) ||
      !std::isfinite(P.alpha) || !std::isfinite(P.beta))
    return fail(h, OTFEM_INVALID_ARG, "bad solver parameters");
  if (P.warm_start != OTFEM_WARM_PREVIOUS_VELOCITY && P.warm_start != OTFEM_WARM_LEARNED)
    return fail(h, OTFEM_INVALID_ARG, "unknown warm start");
  if (P.subspace < 0 || P.subspace > kSubMax - 2)
    return fail(h, OTFEM_INVALID_ARG, "subspace must be 0..10");
  if (P.subspace > 0) {
    otfem_status ss =
        sub_ensure(h, P.subspace, 1 + P.subspace + (P.warm_start == OTFEM_WARM_LEARNED ? 1 : 0));
    if (ss != OTFEM_OK) return ss;
  }
  const bool ws_on = h->model && !h->ws_suspended;
  const bool learned = P.warm_start == OTFEM_WARM_LEARNED;
  if (learned && !ws_on) return fail(h, OTFEM_STATE, "no warm-start model attached");
  WsArgs w{};
  ModelView mv{};
  if (ws_on) {
    model_view(h->model, &mv);
    if (h->recording && mv.capacity - mv.count < h->n_free)
      return fail(h, OTFEM_STATE, "sample store full (otfem_model_reserve)");
    w.v = h->v;
    w.vw = h->vw;
    w.hist1 = h->hist1;
    w.hist2 = h->hist2;
    w.kd = h->kd;
    w.mass = h->mass;
    w.fixed = h->fixed;
    w.free_rank = h->free_rank;
    w.N = h->L.N;
    w.cM = 1.0 + P.alpha * dt;
    w.cK = dt * dt + P.beta * dt;
    w.predict = learned ? 1 : 0;
    w.record = h->recording ? 1 : 0;
    w.rec_base = mv.count;
    w.rec_X = mv.X;
    w.rec_S = mv.S;
    w.rec_T = mv.T;
    w.wf = mv.wf;
    w.theta = mv.theta;
    w.bad = h->ws_bad;
  }
  if (h->lc.spmv_mode == SPMV_F64_STORE64 && !h->bt8)
    HIPCHK(h, dalloc(&h->bt8, 3 * (size_t)h->L.slot_tet.size()));
  StepArgs a = make_args(h, dt, use_stiffness_warping, P);
  if (learned && a.sub_hist > 0) {
    // the network's start is one direction of the subspace start; the
    // element kernel forms the plain residual
    a.sub_net = 1;
    a.vnet = h->vw;
  } else if (learned) {
    a.vw = h->vw;
  }
  hipStream_t s = h->stream;
  const bool timed = (stats != nullptr) || h->timing;
  hipEvent_t e[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  if (timed) {
    while (h->pool.size() < h->pool_used + 5) {
      hipEvent_t ev;
      HIPCHK(h, hipEventCreate(&ev));
      h->pool.push_back(ev);
    }
    for (int k = 0; k < 5; ++k) e[k] = h->pool[h->pool_used + k];
    h->pool_used += 5;
    HIPCHK(h, hipEventRecord(e[0], s));
  }
  if (ws_on) HIPCHK(h, enqueue_ws_predict(w, s));
  if (timed) HIPCHK(h, hipEventRecord(e[1], s));
  const int cap = std::max(0, P.cg_max_iterations);
  // capture `enqueue` (on the capture-only stream) into an executable graph
  auto capture = [&](auto &&enqueue, hipGraphExec_t *out, const char *what) -> otfem_status {
    hipGraph_t g;
    HIPCHK(h, hipStreamBeginCapture(h->cap_stream, hipStreamCaptureModeThreadLocal));
    hipError_t ce = enqueue(h->cap_stream);
    hipError_t ee = hipStreamEndCapture(h->cap_stream, &g);
    if (ce != hipSuccess) return hip_fail(h, ce, what);
    HIPCHK(h, ee);
    hipError_t ie = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHK(h, ie);
    return OTFEM_OK;
  };
  if (h->use_graphs) {
    // the graphs bake in every kernel argument: rebuild when any changes
    std::vector<unsigned char> key(sizeof(StepArgs) + sizeof(LaunchConfig));
    std::memcpy(key.data(), &a, sizeof(StepArgs));
    std::memcpy(key.data() + sizeof(StepArgs), &h->lc, sizeof(LaunchConfig));
    if (key != h->graph_key || !h->g_elem || !h->g_fin) {
      // earlier steps may still be running the graphs about to be destroyed
      if (h->g_elem || h->g_fin || !h->g_chunks.empty()) HIPCHK(h, hipStreamSynchronize(s));
      drop_graphs(h);
      otfem_status cs = capture([&](hipStream_t cs_) { return enqueue_element(a, h->lds_elem, cs_); },
                                &h->g_elem, "enqueue_element");
      if (cs == OTFEM_OK)
        cs = capture([&](hipStream_t cs_) { return enqueue_finish(a, h->lc, cs_); }, &h->g_fin,
                     "enqueue_finish");
      if (cs == OTFEM_OK && a.sub_hist > 0)
        cs = capture([&](hipStream_t cs_) { return enqueue_subspace(a, h->lds_sub, h->lds_cg, cs_); },
                     &h->g_sub, "enqueue_subspace");
      if (cs != OTFEM_OK) {
        drop_graphs(h);
        return cs;
      }
      h->graph_key = key;
    }
    HIPCHK(h, hipGraphLaunch(h->g_elem, s));
    if (timed) HIPCHK(h, hipEventRecord(e[2], s));
    if (h->g_sub) HIPCHK(h, hipGraphLaunch(h->g_sub, s));
  } else {
    HIPCHK(h, enqueue_element(a, h->lds_elem, s));
    if (timed) HIPCHK(h, hipEventRecord(e[2], s));
    HIPCHK(h, enqueue_subspace(a, h->lds_sub, h->lds_cg, s));
  }
  if (timed) HIPCHK(h, hipEventRecord(e[3], s));
  // CG in chunks: cg_chunk(a) iterations, and kCgTail from 16 iterations before
  // the iteration at which the previous step was seen to stop (so the chunks
  // that run after the stop are short).  The host keeps one chunk queued
  // ahead of the one whose stopped-body count it reads back, and issues no
  // more once every body has stopped; kernels of a chunk that runs after the
  // stop exit at once.
  // Batched meshes also shrink the CG grid to the active tiles: the count
  // read back (one chunk behind) only falls within a step, so a grid sized
  // from it covers the list the next chunk builds; grids are powers of two
  // times 256, capped at the tile count.
  const int ntiles = (int)h->L.tiles.size();
  const size_t lds_cg = a.p32 ? h->lds_cg32 : h->lds_cg;
  int known_active = ntiles;
  const int tail_from = h->cg_stop_seen > 0 ? std::max(0, h->cg_stop_seen - 16) : cap;
  int prev_end = 0;
  const int chunk = cg_chunk(a);
  for (int j = 0, it0 = 0; it0 < cap; ++j) {
    const int len = (it0 % chunk == 0 && it0 + chunk <= tail_from) ? chunk : kCgTail;
    const int it1 = std::min(cap, it0 + len);
    int grid = ntiles;
    if (a.compact) {
      grid = 256;
      while (grid < known_active) grid *= 2;
      grid = std::min(grid, ntiles);
    }
    if (h->use_graphs) {
      hipGraphExec_t &ge = h->g_chunks[std::make_tuple(it0, it1 - it0, grid)];
      if (!ge) {
        otfem_status cs = capture(
            [&](hipStream_t cs_) { return enqueue_cg(a, h->lc, lds_cg, cs_, it0, it1, grid); },
            &ge, "enqueue_cg");
        if (cs != OTFEM_OK) return cs;
      }
      HIPCHK(h, hipGraphLaunch(ge, s));
    } else {
      HIPCHK(h, enqueue_cg(a, h->lc, lds_cg, s, it0, it1, grid));
    }
    if (it1 < cap) {
      HIPCHK(h, hipMemcpyAsync(&h->done_host[2 * (j & 1)], h->done_count, 2 * sizeof(int),
                               hipMemcpyDeviceToHost, s));
      HIPCHK(h, hipEventRecord(h->chunk_ev[j & 1], s));
      if (j >= 1) {
        HIPCHK(h, hipEventSynchronize(h->chunk_ev[(j - 1) & 1]));
        const volatile int *dh = &h->done_host[2 * ((j - 1) & 1)];
        if (dh[0] >= (int)h->L.B) {
          h->cg_stop_seen = prev_end;
          break;
        }
        const int act_now = dh[1];
        if (a.compact) known_active = std::min(known_active, std::max(act_now, 1));
      }
    }
    prev_end = it1;
    it0 = it1;
  }
  if (h->use_graphs)
    HIPCHK(h, hipGraphLaunch(h->g_fin, s));
  else
    HIPCHK(h, enqueue_finish(a, h->lc, s));
  if (ws_on && h->recording) {
    HIPCHK(h, enqueue_ws_target(w, s));
    model_commit(h->model, h->n_free);
  }
  if (timed) HIPCHK(h, hipEventRecord(e[4], s));
  h->rotations_valid = (P.store_rotations != 0) && use_stiffness_warping;
  h->pending = true;
  if (!stats) return OTFEM_OK;
  return sync_and_check(h, stats);
}

}  // extern "C"

// ---- the step without allocation (otfem_step_noalloc) ----
namespace {

// Scratch of otfem_step_noalloc, for at most kScratchTiles tiles: the tile
// partial sums of the element kernel (b.b, r.r), of the CG kernels
// ([2][tile][4]) and of the subspace start's Gram ([tile][kSubPart]), the
// stopped-body / active-tile counters and the active-tile list.  Offsets in
// bytes, each 256-aligned.
constexpr int64_t kScratchTiles = 32768;
constexpr size_t kScrPartBB = 0;
constexpr size_t kScrPartRR = kScrPartBB + sizeof(double) * kScratchTiles;
constexpr size_t kScrPart4 = kScrPartRR + sizeof(double) * kScratchTiles;
constexpr size_t kScrDone = kScrPart4 + sizeof(double) * 8 * kScratchTiles;
constexpr size_t kScrAct = kScrDone + 256;
constexpr size_t kScrSubPart = kScrAct + sizeof(int32_t) * kScratchTiles;
constexpr size_t kScratchBytes = kScrSubPart + sizeof(double) * kSubPart * kScratchTiles;  // 25.1 MiB

// Each GPU that loads the library holds its own copy until the process ends.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kScratchDevices] = {};

// The device address of `scratch` on the solver's device (looked up once per
// device, with that device current; host code never uses the symbol itself).
otfem_status scratch_address(otfem_solver *h, unsigned char **out) {
  int cur = -1;
  HIPCHK(h, hipGetDevice(&cur));
  if (cur != h->device) return fail(h, OTFEM_STATE, "make the solver's device current");
  if (cur < 0 || cur >= kScratchDevices) return fail(h, OTFEM_NO_DEVICE, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[cur]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(h, e, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[cur] = static_cast<unsigned char *>(p);
  }
  *out = g_scratch_ptr[cur];
  return OTFEM_OK;
}

}  // namespace

extern "C" {

otfem_status otfem_step_noalloc(otfem_handle h, double dt, int use_stiffness_warping,
                                const otfem_params *params) {
  if (!h) return OTFEM_INVALID_ARG;
  if (!h->initialized) return fail(h, OTFEM_STATE, "call otfem_init before otfem_step_noalloc");
  if (!(dt > 0) || !std::isfinite(dt)) return fail(h, OTFEM_INVALID_ARG, "dt must be positive");
  otfem_params P;
  if (params)
    P = *params;
  else
    otfem_default_params(&P, 0);
  if (P.cg_max_iterations < 0 || P.cg_min_iterations < 0 || !(P.tol >= 0) || !(P.tiny >= 0) ||
      (P.tol_mode != 0 && P.tol_mode != 1) || (P.rotation_mode != 0 && P.rotation_mode != 1) ||
      !std::isfinite(P.alpha) || !std::isfinite(P.beta) || P.subspace < 0 || P.subspace > kSubMax - 2 ||
      (P.warm_start != OTFEM_WARM_PREVIOUS_VELOCITY && P.warm_start != OTFEM_WARM_LEARNED))
    return fail(h, OTFEM_INVALID_ARG, "bad solver parameters");
  const int64_t ntiles = (int64_t)h->L.tiles.size();
  if (ntiles > kScratchTiles)
    return fail(h, OTFEM_INVALID_ARG,
                "the mesh has " + std::to_string(ntiles) + " tiles; otfem_step_noalloc takes at most " +
                    std::to_string(kScratchTiles));
  // everything but the scratch must exist already
  const bool learned = P.warm_start == OTFEM_WARM_LEARNED;
  const bool ws_on = h->model && !h->ws_suspended;
  if (learned && !ws_on) return fail(h, OTFEM_STATE, "no warm-start model attached");
  if (ws_on && h->recording) return fail(h, OTFEM_STATE, "recording needs otfem_step");
  if (h->lc.spmv_mode == SPMV_F64_STORE64 && !h->bt8)
    return fail(h, OTFEM_STATE, "take one otfem_step in this SpMV mode first (its buffers)");
  if (P.subspace > 0) {
    const int dirs = 1 + P.subspace + (learned ? 1 : 0);
    if (!h->vring || h->ring_slots < P.subspace || !h->prim_begin ||
        sub_gram_lds_bytes(h->L.max_tile_nodes, h->L.max_tile_owned, dirs) > h->lds_sub)
      return fail(h, OTFEM_STATE,
                  "take one otfem_step with these subspace settings first (its history buffers)");
  }
  unsigned char *scr = nullptr;
  otfem_status st = scratch_address(h, &scr);
  if (st != OTFEM_OK) return st;
  StepArgs a = make_args(h, dt, use_stiffness_warping, P);
  a.part_bb = reinterpret_cast<double *>(scr + kScrPartBB);
  a.part_rr = reinterpret_cast<double *>(scr + kScrPartRR);
  a.part4 = reinterpret_cast<double *>(scr + kScrPart4);
  a.done_count = reinterpret_cast<int *>(scr + kScrDone);
  a.act = reinterpret_cast<int32_t *>(scr + kScrAct);
  a.sub_part = reinterpret_cast<double *>(scr + kScrSubPart);
  if (learned && a.sub_hist > 0) {
    a.sub_net = 1;
    a.vnet = h->vw;
  } else if (learned) {
    a.vw = h->vw;
  }
  hipStream_t s = h->stream;
  if (ws_on) {
    ModelView mv{};
    model_view(h->model, &mv);
    WsArgs w{};
    w.v = h->v;
    w.vw = h->vw;
    w.hist1 = h->hist1;
    w.hist2 = h->hist2;
    w.kd = h->kd;
    w.mass = h->mass;
    w.fixed = h->fixed;
    w.free_rank = h->free_rank;
    w.N = h->L.N;
    w.cM = 1.0 + P.alpha * dt;
    w.cK = dt * dt + P.beta * dt;
    w.predict = learned ? 1 : 0;
    w.record = 0;
    w.wf = mv.wf;
    w.theta = mv.theta;
    w.bad = h->ws_bad;
    HIPCHK(h, enqueue_ws_predict(w, s));
  }
  // the kernels write every partial sum and counter they read later in the
  // step (k_element clears the stopped-body count), so nothing is cleared
  HIPCHK(h, enqueue_element(a, h->lds_elem, s));
  HIPCHK(h, enqueue_subspace(a, h->lds_sub, h->lds_cg, s));
  const int cap = std::max(0, P.cg_max_iterations);
  const int nt = (int)ntiles;
  int known_active = nt;
  const int tail_from = h->cg_stop_seen > 0 ? std::max(0, h->cg_stop_seen - 16) : cap;
  int prev_end = 0;
  const int chunk = cg_chunk(a);
  for (int j = 0, it0 = 0; it0 < cap; ++j) {
    const int len = (it0 % chunk == 0 && it0 + chunk <= tail_from) ? chunk : kCgTail;
    const int it1 = std::min(cap, it0 + len);
    int grid = nt;
    if (a.compact) {
      grid = 256;
      while (grid < known_active) grid *= 2;
      grid = std::min(grid, nt);
    }
    HIPCHK(h, enqueue_cg(a, h->lc, a.p32 ? h->lds_cg32 : h->lds_cg, s, it0, it1, grid));
    if (it1 < cap) {
      HIPCHK(h, hipMemcpyAsync(&h->done_host[2 * (j & 1)], a.done_count, 2 * sizeof(int),
                               hipMemcpyDeviceToHost, s));
      HIPCHK(h, hipEventRecord(h->chunk_ev[j & 1], s));
      if (j >= 1) {
        HIPCHK(h, hipEventSynchronize(h->chunk_ev[(j - 1) & 1]));
        const volatile int *dh = &h->done_host[2 * ((j - 1) & 1)];
        if (dh[0] >= (int)h->L.B) {
          h->cg_stop_seen = prev_end;
          break;
        }
        if (a.compact) known_active = std::min(known_active, std::max((int)dh[1], 1));
      }
    }
    prev_end = it1;
    it0 = it1;
  }
  HIPCHK(h, enqueue_finish(a, h->lc, s));
  h->rotations_valid = (P.store_rotations != 0) && use_stiffness_warping;
  h->pending = true;
  return OTFEM_OK;
}

otfem_status otfem_synchronize(otfem_handle h, otfem_stats *stats) {
  if (!h) return OTFEM_INVALID_ARG;
  if (!h->device_ready) return OTFEM_OK;
  return sync_and_check(h, stats);
}

otfem_status otfem_set_timing(otfem_handle h, int enable) {
  if (!h) return OTFEM_INVALID_ARG;
  if (h->device_ready && h->pool_used) {
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->pool_used = 0;
  }
  h->timing = enable != 0;
  h->acc_elem_ms = h->acc_cg_ms = h->acc_ws_ms = h->acc_sub_ms = 0.0;
  h->timed_steps = 0;
  return OTFEM_OK;
}

otfem_status otfem_get_timing(otfem_handle h, double *element_ms, double *cg_ms,
                              double *network_ms, long long *steps) {
  if (!h) return OTFEM_INVALID_ARG;
  if (h->device_ready && h->pool_used) {
    HIPCHK(h, hipStreamSynchronize(h->stream));
    otfem_status ts = resolve_timing(h);
    if (ts != OTFEM_OK) return ts;
  }
  if (element_ms) *element_ms = h->acc_elem_ms + h->acc_sub_ms;
  if (network_ms) *network_ms = h->acc_ws_ms;
  if (cg_ms) *cg_ms = h->acc_cg_ms;
  if (steps) *steps = h->timed_steps;
  return OTFEM_OK;
}

otfem_status otfem_get_timing_detail(otfem_handle h, double *ms, long long *steps) {
  if (!h || !ms) return OTFEM_INVALID_ARG;
  otfem_status st = otfem_get_timing(h, nullptr, nullptr, nullptr, steps);
  if (st != OTFEM_OK) return st;
  ms[0] = h->acc_ws_ms;
  ms[1] = h->acc_elem_ms;
  ms[2] = h->acc_sub_ms;
  ms[3] = h->acc_cg_ms;
  return OTFEM_OK;
}

otfem_status otfem_iteration_totals(otfem_handle h, int64_t *totals, int reset) {
  if (!h) return OTFEM_INVALID_ARG;
  if (!h->device_ready) return fail(h, OTFEM_STATE, "no device state yet");
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (totals)
    HIPCHK(h, hipMemcpy(totals, h->iter_acc, h->L.B * sizeof(int64_t), hipMemcpyDeviceToHost));
  if (reset) HIPCHK(h, hipMemset(h->iter_acc, 0, h->L.B * sizeof(unsigned long long)));
  return OTFEM_OK;
}

otfem_status otfem_unconverged_totals(otfem_handle h, int64_t *totals, int reset) {
  if (!h) return OTFEM_INVALID_ARG;
  if (!h->device_ready) return fail(h, OTFEM_STATE, "no device state yet");
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (totals)
    HIPCHK(h, hipMemcpy(totals, h->unconv_acc, h->L.B * sizeof(int64_t), hipMemcpyDeviceToHost));
  if (reset) HIPCHK(h, hipMemset(h->unconv_acc, 0, h->L.B * sizeof(unsigned long long)));
  return OTFEM_OK;
}

otfem_status otfem_set_kernel_config(otfem_handle h, int cg_block, int spmv_mode, int use_graphs) {
  if (!h) return OTFEM_INVALID_ARG;
  if (cg_block != 256 && cg_block != 512 && cg_block != 1024)
    return fail(h, OTFEM_INVALID_ARG, "cg_block must be 256, 512 or 1024");
  if (spmv_mode < SPMV_F32 || spmv_mode > SPMV_F64_STORE64)
    return fail(h, OTFEM_INVALID_ARG, "unknown SpMV precision mode");
  if (spmv_mode == SPMV_F64_STORE64 && !h->bt8 && h->device_ready) {
    HIPCHK(h, dalloc(&h->bt8, 3 * (size_t)h->L.slot_tet.size()));
  }
  h->lc.cg_block = cg_block;
  h->lc.spmv_mode = spmv_mode;
  h->use_graphs = use_graphs != 0;
  return OTFEM_OK;
}

otfem_status otfem_set_deterministic(otfem_handle h, int on) {
  if (!h) return OTFEM_INVALID_ARG;
  if (!on) {
    // steps enqueued without stats may still use the buffers
    if (h->device_ready) HIPCHK(h, hipStreamSynchronize(h->stream));
    h->ordered = false;
    dfree(h->ord_begin); dfree(h->ord_ent); dfree(h->ordered_terms);
    return OTFEM_OK;
  }
  if (h->ordered) return OTFEM_OK;
  otfem_status st = ensure_device(h);
  if (st) return st;
  return no_throw(h, [&]() -> otfem_status {
    // per internal node, the (slot, corner) pairs it owns, in slot order:
    // every node's terms are then added in the same order on every run
    const Layout &L = h->L;
    const int64_t N = L.N, S = (int64_t)L.slot_tet.size();
    if (S >= (int64_t)1 << 29) return fail(h, OTFEM_INVALID_ARG, "too many slots for the deterministic mode");
    std::vector<int32_t> begin(N + 1, 0);
    for (const TileInfo &ti : L.tiles)
      for (int64_t sl = ti.slot_begin; sl < ti.slot_begin + ti.n_slots; ++sl)
        for (int c = 0; c < 4; ++c) {
          const int32_t li = L.slot_conn[4 * sl + c];
          if (li < ti.n_owned) ++begin[ti.node_begin + li + 1];
        }
    for (int64_t g = 0; g < N; ++g) begin[g + 1] += begin[g];
    std::vector<int32_t> ent(std::max<int32_t>(begin[N], 1)), fill(begin.begin(), begin.end() - 1);
    for (const TileInfo &ti : L.tiles)
      for (int64_t sl = ti.slot_begin; sl < ti.slot_begin + ti.n_slots; ++sl)
        for (int c = 0; c < 4; ++c) {
          const int32_t li = L.slot_conn[4 * sl + c];
          if (li < ti.n_owned) ent[fill[ti.node_begin + li]++] = (int32_t)(4 * sl + c);
        }
    hipError_t e = dalloc(&h->ord_begin, N + 1);
    if (e == hipSuccess) e = dalloc(&h->ord_ent, ent.size());
    if (e == hipSuccess) e = dalloc(&h->ordered_terms, 36 * (size_t)std::max<int64_t>(S, 1));
    if (e == hipSuccess)
      e = hipMemcpyAsync(h->ord_begin, begin.data(), (N + 1) * sizeof(int32_t), hipMemcpyHostToDevice,
                         h->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(h->ord_ent, ent.data(), ent.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                         h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
      dfree(h->ord_begin); dfree(h->ord_ent); dfree(h->ordered_terms);
      return hip_fail(h, e, "deterministic mode buffers");
    }
    h->ordered = true;
    return OTFEM_OK;
  });
}

// Block-size sweep of otfem_autotune; *best_ms is the element + CG device
// time of the `steps` timed steps at the chosen block size.  The trial steps
// leave the solver as it was: node state, plastic strain, the per-body
// iteration and unconverged counters, the last step's body records, the
// rotation flag, the chunking hint and the timing totals are restored on
// every exit; the trials store no rotations and run the plain start (so the
// learned and subspace histories are not touched).
static otfem_status tune_blocks(otfem_handle h, double dt, int use_stiffness_warping,
                                const otfem_params *params, int steps, int *best_block,
                                double *best_ms) {
  if (!h) return OTFEM_INVALID_ARG;
  if (!h->initialized) return fail(h, OTFEM_STATE, "call otfem_init before otfem_autotune");
  if (steps <= 0) steps = 2;
  const int64_t N = h->L.N, T = h->L.T, B = h->L.B;
  double *sx = nullptr, *sv = nullptr, *sf = nullptr, *spl = nullptr;
  unsigned long long *sit = nullptr, *sun = nullptr;
  hipError_t ae = dalloc(&sx, 3 * N);
  if (ae == hipSuccess) ae = dalloc(&sv, 3 * N);
  if (ae == hipSuccess) ae = dalloc(&sf, 3 * N);
  if (ae == hipSuccess) ae = dalloc(&spl, 6 * T);
  if (ae == hipSuccess) ae = dalloc(&sit, B);
  if (ae == hipSuccess) ae = dalloc(&sun, B);
  auto release = [&]() {
    dfree(sx); dfree(sv); dfree(sf); dfree(spl); dfree(sit); dfree(sun);
  };
  if (ae != hipSuccess) {
    release();
    return hip_fail(h, ae, "autotune: state copies");
  }
  hipStream_t s = h->stream;
  auto copy_state = [&](bool restore) -> hipError_t {
    auto copy = [&](auto *live, auto *kept, int64_t count) {
      return hipMemcpyAsync(restore ? live : kept, restore ? kept : live, count * sizeof(*live),
                            hipMemcpyDeviceToDevice, s);
    };
    hipError_t e;
    if ((e = copy(h->x, sx, 3 * N)) || (e = copy(h->v, sv, 3 * N)) || (e = copy(h->fext, sf, 3 * N)) ||
        (e = copy(h->plastic, spl, 6 * T)) || (e = copy(h->iter_acc, sit, B)) ||
        (e = copy(h->unconv_acc, sun, B)))
      return e;
    return hipStreamSynchronize(s);
  };
  // steps enqueued before: finish them so their records and timings are kept
  if (h->pending) {
    otfem_status ps = sync_and_check(h, nullptr);
    if (ps != OTFEM_OK) {
      release();
      return ps;
    }
  }
  // host-side records of the solver
  const std::vector<BodyState> kept_body = h->last_body;
  const bool kept_rot = h->rotations_valid;
  const int kept_stop = h->cg_stop_seen;
  const double kept_ms[4] = {h->acc_elem_ms, h->acc_cg_ms, h->acc_ws_ms, h->acc_sub_ms};
  const long long kept_steps = h->timed_steps;
  const bool timing_was = h->timing;
  otfem_params P;
  if (params)
    P = *params;
  else
    otfem_default_params(&P, 0);
  P.warm_start = OTFEM_WARM_PREVIOUS_VELOCITY;
  P.subspace = 0;
  P.store_rotations = 0;
  h->ws_suspended = true;
  otfem_status ret = OTFEM_OK;
  const int blocks[3] = {256, 512, 1024};
  double best = 1e300;
  int best_b = h->lc.cg_block;
  if (copy_state(false) != hipSuccess) ret = fail(h, OTFEM_HIP_ERROR, "autotune: state copy failed");
  for (int bi = 0; bi < 3 && ret == OTFEM_OK; ++bi) {
    h->lc.cg_block = blocks[bi];
    otfem_stats st;
    ret = otfem_step(h, dt, use_stiffness_warping, &P, &st);  // warm-up, builds the graphs
    if (ret == OTFEM_OK && copy_state(true) != hipSuccess)
      ret = fail(h, OTFEM_HIP_ERROR, "autotune: restore failed");
    if (ret != OTFEM_OK) break;
    h->timing = true;
    const double e0 = h->acc_elem_ms + h->acc_cg_ms;
    for (int k = 0; k < steps && ret == OTFEM_OK; ++k) ret = otfem_step(h, dt, use_stiffness_warping, &P, nullptr);
    if (ret == OTFEM_OK) ret = sync_and_check(h, nullptr);
    const double ms = h->acc_elem_ms + h->acc_cg_ms - e0;
    h->timing = timing_was;
    if (ret == OTFEM_OK && copy_state(true) != hipSuccess)
      ret = fail(h, OTFEM_HIP_ERROR, "autotune: restore failed");
    if (ret == OTFEM_OK && ms < best) {
      best = ms;
      best_b = blocks[bi];
    }
  }
  // restore on every path (a failed trial may have left partial state)
  if (ret != OTFEM_OK) {
    (void)hipStreamSynchronize(s);
    (void)copy_state(true);
  }
  h->timing = timing_was;
  h->ws_suspended = false;
  h->lc.cg_block = best_b;
  h->last_body = kept_body;
  h->rotations_valid = kept_rot;
  h->cg_stop_seen = kept_stop;
  h->acc_elem_ms = kept_ms[0];
  h->acc_cg_ms = kept_ms[1];
  h->acc_ws_ms = kept_ms[2];
  h->acc_sub_ms = kept_ms[3];
  h->timed_steps = kept_steps;
  release();
  if (best_block) *best_block = best_b;
  if (best_ms) *best_ms = best;
  return ret;
}

otfem_status otfem_autotune(otfem_handle h, double dt, int use_stiffness_warping,
                            const otfem_params *params, int steps, int *best_block) {
  return tune_blocks(h, dt, use_stiffness_warping, params, steps, best_block, nullptr);
}

static thread_local std::string g_tune_err;

const char *otfem_autotune_layout_last_error(void) { return g_tune_err.c_str(); }

static otfem_status autotune_layout_body(const otfem_mesh_desc *desc, const double *young,
                                        const double *poisson, const double *density,
                                        const double *f_ext, double dt, int use_stiffness_warping,
                                        const otfem_params *params, int steps,
                                        const int32_t *candidates, int num_candidates,
                                        int32_t *best_tile_nodes, int *best_block, double *best_ms) {
  if (!desc || !young || !poisson || !density || !best_tile_nodes || num_candidates < 0 ||
      (num_candidates > 0 && !candidates) || desc->num_nodes <= 0) {
    g_tune_err = "invalid argument";
    return OTFEM_INVALID_ARG;
  }
  std::vector<int32_t> cand;
  if (num_candidates > 0) {
    cand.assign(candidates, candidates + num_candidates);
  } else if (desc->num_bodies <= 1) {
    // one body: the default tiling, 512 owned nodes per tile, and 120, 256,
    // 768 and 1,024 tiles where that leaves 16 (120 tiles) or 64 owned nodes
    // per tile on average
    const int64_t N = desc->num_nodes;
    cand = {0, 512};
    if (N >= 16 * 120) cand.push_back((int32_t)((N + 119) / 120));
    for (int64_t tiles : {256, 768, 1024})
      if (N >= 64 * tiles) cand.push_back((int32_t)((N + tiles - 1) / tiles));
  } else {
    cand = {256, 512, 1024};
  }
  double best = 1e300;
  int32_t best_c = 0;
  int best_b = 0;
  std::string why;
  for (int32_t c : cand) {
    otfem_mesh_desc d = *desc;
    d.target_tile_nodes = c;
    otfem_handle h = nullptr;
    otfem_status st = otfem_create(&d, &h);
    if (st == OTFEM_OK && params && params->subspace > 0) {
      // a layout whose tiles cannot hold the subspace start's directions in
      // LDS cannot run these settings
      int dev = 0;
      hipDeviceProp_t prop;
      if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess &&
          sub_gram_lds_bytes(h->L.max_tile_nodes, h->L.max_tile_owned, 2 + params->subspace) +
                  sub_gram_static_lds_bytes() > (size_t)prop.sharedMemPerBlock) {
        why += "target " + std::to_string(c) + ": tiles too large for the subspace start; ";
        otfem_destroy(h);
        continue;
      }
    }
    if (st == OTFEM_OK) st = otfem_init_materials(h, young, poisson, density, 0.0, 0.0, 0.0);
    if (st == OTFEM_OK && f_ext) st = otfem_set_external_forces(h, f_ext);
    int b = 0;
    double ms = 1e300;
    if (st == OTFEM_OK) st = tune_blocks(h, dt, use_stiffness_warping, params, steps, &b, &ms);
    if (st != OTFEM_OK)
      why += "target " + std::to_string(c) + ": " + (h ? h->err : std::string("otfem_create failed")) + "; ";
    if (h) otfem_destroy(h);
    if (st == OTFEM_NO_DEVICE || st == OTFEM_HIP_ERROR) {
      g_tune_err = why;
      return st;
    }
    if (st == OTFEM_OK && ms < best) {
      best = ms;
      best_c = c;
      best_b = b;
    }
  }
  if (best == 1e300) {
    g_tune_err = "no candidate layout could be stepped: " + why;
    return OTFEM_INVALID_ARG;
  }
  *best_tile_nodes = best_c;
  if (best_block) *best_block = best_b;
  if (best_ms) *best_ms = best;
  return OTFEM_OK;
}

otfem_status otfem_autotune_layout(const otfem_mesh_desc *desc, const double *young,
                                   const double *poisson, const double *density,
                                   const double *f_ext, double dt, int use_stiffness_warping,
                                   const otfem_params *params, int steps,
                                   const int32_t *candidates, int num_candidates,
                                   int32_t *best_tile_nodes, int *best_block, double *best_ms) {
  const otfem_status st = no_throw(nullptr, [&] {
    return autotune_layout_body(desc, young, poisson, density, f_ext, dt, use_stiffness_warping,
                                params, steps, candidates, num_candidates, best_tile_nodes,
                                best_block, best_ms);
  });
  if (st == OTFEM_OUT_OF_MEMORY) {
    try {
      g_tune_err = "out of host memory";
    } catch (...) {
    }
  }
  return st;
}

otfem_status otfem_body_stats(otfem_handle h, int *iterations, double *rel_residual) {
  if (!h) return OTFEM_INVALID_ARG;
  if (h->last_body.size() != (size_t)h->L.B) return fail(h, OTFEM_STATE, "no completed step");
  for (int64_t b = 0; b < h->L.B; ++b) {
    const BodyState &s = h->last_body[b];
    if (iterations) iterations[b] = s.iterations;
    if (rel_residual) rel_residual[b] = s.bb > 0 ? std::sqrt(s.rr / s.bb) : 0.0;
  }
  return OTFEM_OK;
}

otfem_status otfem_body_start_residuals(otfem_handle h, double *rel_residual) {
  if (!h || !rel_residual) return OTFEM_INVALID_ARG;
  if (h->last_body.size() != (size_t)h->L.B) return fail(h, OTFEM_STATE, "no completed step");
  for (int64_t b = 0; b < h->L.B; ++b) {
    const BodyState &s = h->last_body[b];
    rel_residual[b] = s.bb > 0 ? std::sqrt(s.rr0 / s.bb) : 0.0;
  }
  return OTFEM_OK;
}

otfem_status otfem_subspace_coefficients(otfem_handle h, double *coef) {
  if (!h || !coef) return OTFEM_INVALID_ARG;
  if (!h->sub_coef) return fail(h, OTFEM_STATE, "no step with the subspace start yet");
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipMemcpy(coef, h->sub_coef, kSubMax * h->L.B * sizeof(double), hipMemcpyDeviceToHost));
  return OTFEM_OK;
}

otfem_status otfem_update_original_coord(otfem_handle h) {
  if (!h) return OTFEM_INVALID_ARG;
  otfem_status st = ensure_device(h);
  if (st) return st;
  std::vector<double> tmp(3 * h->L.N);
  HIPCHK(h, hipMemcpyAsync(h->x0, h->x, tmp.size() * sizeof(double), hipMemcpyDeviceToDevice,
                           h->stream));
  HIPCHK(h, hipMemcpyAsync(tmp.data(), h->x, tmp.size() * sizeof(double), hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  to_caller3(h->L, tmp, h->x0_host.data());
  // fem_update_original_coord.h:19-23: the force offsets are now stale and the
  // caller must initialise again before the next step.
  h->initialized = false;
  return OTFEM_OK;
}

static otfem_status otfem_export_K_body(otfem_handle h, int64_t *nnz_blocks, int64_t *row_ptr, int32_t *cols,
                                        double *blocks) {
  if (!h || !nnz_blocks) return OTFEM_INVALID_ARG;
  if (!h->initialized) return fail(h, OTFEM_STATE, "call otfem_init first");
  const int64_t N = h->L.N, T = h->L.T;
  std::vector<double> R(9 * T);
  if (otfem_get_rotations(h, R.data()) != OTFEM_OK) return OTFEM_HIP_ERROR;
  // K'_ij = sum_e Re Ke_ij Re^T (fem_stiffness_assembly.h:179-203), rows as
  // ordered maps like the reference's m_K_row.
  std::vector<std::map<int32_t, std::array<double, 9>>> rows(N);
  const double *x0 = h->x0_host.data();
  for (int64_t t = 0; t < T; ++t) {
    const int32_t *tt = &h->tets[4 * t];
    ElemData e = element_data(x0 + 3 * tt[0], x0 + 3 * tt[1], x0 + 3 * tt[2], x0 + 3 * tt[3],
                              h->young[t], h->poisson[t]);
    double b[4][3];
    for (int c = 0; c < 3; ++c) {
      b[1][c] = e.grad[c];
      b[2][c] = e.grad[3 + c];
      b[3][c] = e.grad[6 + c];
      b[0][c] = -(b[1][c] + b[2][c] + b[3][c]);
    }
    const double *Rt = &R[9 * t];
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) {
        // Ke_ij = D1 b_i b_j^T + D2 b_j b_i^T + (D2 b_i.b_j) I + (D0-D1-2D2) diag(b_i o b_j)
        double K[9];
        double bij = b[i][0] * b[j][0] + b[i][1] * b[j][1] + b[i][2] * b[j][2];
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) {
            double v = e.vd[1] * b[i][r] * b[j][c] + e.vd[2] * b[i][c] * b[j][r];
            if (r == c) v += e.vd[2] * bij + (e.vd[0] - e.vd[1] - 2 * e.vd[2]) * b[i][r] * b[j][r];
            K[3 * r + c] = v;
          }
        double RK[9], W[9];
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c)
            RK[3 * r + c] = Rt[3 * r] * K[c] + Rt[3 * r + 1] * K[3 + c] + Rt[3 * r + 2] * K[6 + c];
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c)
            W[3 * r + c] = RK[3 * r] * Rt[3 * c] + RK[3 * r + 1] * Rt[3 * c + 1] +
                           RK[3 * r + 2] * Rt[3 * c + 2];
        auto &blk = rows[tt[i]][tt[j]];
        for (int k = 0; k < 9; ++k) blk[k] += W[k];
      }
  }
  int64_t nnz = 0;
  for (int64_t i = 0; i < N; ++i) {
    if (h->fixed_host[i]) {
      nnz += 1;
      continue;
    }
    for (auto &kv : rows[i])
      if (!h->fixed_host[kv.first]) ++nnz;
  }
  *nnz_blocks = nnz;
  if (!blocks) return OTFEM_OK;
  if (!row_ptr || !cols) return OTFEM_INVALID_ARG;
  int64_t k = 0;
  row_ptr[0] = 0;
  for (int64_t i = 0; i < N; ++i) {
    if (h->fixed_host[i]) {
      cols[k] = (int32_t)i;
      for (int q = 0; q < 9; ++q) blocks[9 * k + q] = (q % 4 == 0) ? 1.0 : 0.0;
      ++k;
    } else {
      for (auto &kv : rows[i]) {
        if (h->fixed_host[kv.first]) continue;
        cols[k] = kv.first;
        for (int q = 0; q < 9; ++q) blocks[9 * k + q] = kv.second[q];
        ++k;
      }
    }
    row_ptr[i + 1] = k;
  }
  return OTFEM_OK;
}

otfem_status otfem_export_K(otfem_handle h, int64_t *nnz_blocks, int64_t *row_ptr, int32_t *cols,
                            double *blocks) {
  return no_throw(h, [&] { return otfem_export_K_body(h, nnz_blocks, row_ptr, cols, blocks); });
}

otfem_status otfem_layout_info(otfem_handle h, int64_t *num_tiles, int64_t *num_slots,
                               int64_t *num_halo, int64_t *max_tile_nodes) {
  if (!h) return OTFEM_INVALID_ARG;
  if (num_tiles) *num_tiles = (int64_t)h->L.tiles.size();
  if (num_slots) *num_slots = (int64_t)h->L.slot_tet.size();
  if (num_halo) *num_halo = (int64_t)h->L.halo.size();
  if (max_tile_nodes) *max_tile_nodes = h->L.max_tile_nodes;
  return OTFEM_OK;
}

// ---- checkpoint / resume ----
//
// File: "OTFEMSV1", uint32 version 1, uint32 flags (bit 0: velocity history
// of the learned start present), int64 N, T, B, uint64 connectivity hash,
// then in the caller's order: x0, x, v, f_ext (double[3N] each), fixed
// (uint8[N]), E, nu, rho (double[T] each), plastic yield / creep / max
// (double[3]), plastic strain (double[6T]), with bit 0 the two history
// velocities (float[4N] each), and the per-body iteration and unconverged
// step counters (uint64[B] each).

namespace {

uint64_t connectivity_hash(const otfem_solver *h) {
  uint64_t x = 1469598103934665603ull;
  auto mix = [&](const void *p, size_t n) {
    const unsigned char *b = static_cast<const unsigned char *>(p);
    for (size_t i = 0; i < n; ++i) {
      x ^= b[i];
      x *= 1099511628211ull;
    }
  };
  const int64_t dims[3] = {h->L.N, h->L.T, h->L.B};
  mix(dims, sizeof(dims));
  mix(h->tets.data(), h->tets.size() * sizeof(int32_t));
  mix(h->L.body_node_begin.data(), h->L.body_node_begin.size() * sizeof(int64_t));
  mix(h->L.body_tet_begin.data(), h->L.body_tet_begin.size() * sizeof(int64_t));
  return x;
}

// float4 per node: internal order <-> caller order
void hist_to_caller(const Layout &L, const std::vector<float> &in, std::vector<float> &out) {
  out.resize(in.size());
  for (int64_t g = 0; g < L.N; ++g)
    for (int c = 0; c < 4; ++c) out[4 * L.node_of_internal[g] + c] = in[4 * g + c];
}
void hist_to_internal(const Layout &L, const std::vector<float> &in, std::vector<float> &out) {
  out.resize(in.size());
  for (int64_t g = 0; g < L.N; ++g)
    for (int c = 0; c < 4; ++c) out[4 * g + c] = in[4 * L.node_of_internal[g] + c];
}

}  // namespace

static otfem_status otfem_save_state_body(otfem_handle h, const char *path) {
  if (!h || !path) return OTFEM_INVALID_ARG;
  if (!h->initialized) return fail(h, OTFEM_STATE, "call otfem_init first");
  otfem_status st = ensure_device(h);
  if (st) return st;
  const int64_t N = h->L.N, T = h->L.T, B = h->L.B;
  std::vector<double> x(3 * N), v(3 * N), f(3 * N), tmp(3 * N), pl(6 * T);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (auto [dst, src] : {std::pair<double *, double *>{x.data(), h->x}, {v.data(), h->v},
                          {f.data(), h->fext}}) {
    HIPCHK(h, hipMemcpy(tmp.data(), src, 3 * N * sizeof(double), hipMemcpyDeviceToHost));
    to_caller3(h->L, tmp, dst);
  }
  if (T > 0) HIPCHK(h, hipMemcpy(pl.data(), h->plastic, 6 * T * sizeof(double), hipMemcpyDeviceToHost));
  const bool hist = h->hist1 != nullptr;
  std::vector<float> h1, h2, raw(4 * N);
  if (hist) {
    HIPCHK(h, hipMemcpy(raw.data(), h->hist1, N * sizeof(float4), hipMemcpyDeviceToHost));
    hist_to_caller(h->L, raw, h1);
    HIPCHK(h, hipMemcpy(raw.data(), h->hist2, N * sizeof(float4), hipMemcpyDeviceToHost));
    hist_to_caller(h->L, raw, h2);
  }
  std::vector<unsigned long long> it(B), un(B);
  HIPCHK(h, hipMemcpy(it.data(), h->iter_acc, B * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(un.data(), h->unconv_acc, B * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  // the subspace start's history: ring of velocity differences, v^n-1 / v^n
  const bool sub = h->vring && h->sub_state && h->vbuf;
  int32_t sub_head[3] = {0, 0, 0};
  std::vector<std::vector<float>> ring;
  std::vector<double> vb0(3 * N), vb1(3 * N);
  if (sub) {
    HIPCHK(h, hipMemcpy(sub_head, h->sub_state, sizeof(sub_head), hipMemcpyDeviceToHost));
    ring.resize(h->ring_slots);
    for (int k = 0; k < h->ring_slots; ++k) {
      HIPCHK(h, hipMemcpy(raw.data(), h->vring + (int64_t)k * N, N * sizeof(float4), hipMemcpyDeviceToHost));
      hist_to_caller(h->L, raw, ring[k]);
    }
    for (auto [dst, half] : {std::pair<double *, int>{vb0.data(), 0}, {vb1.data(), 1}}) {
      HIPCHK(h, hipMemcpy(tmp.data(), h->vbuf + (int64_t)half * 3 * N, 3 * N * sizeof(double),
                          hipMemcpyDeviceToHost));
      to_caller3(h->L, tmp, dst);
    }
  }
  std::ofstream o(path, std::ios::binary);
  if (!o) return fail(h, OTFEM_INVALID_ARG, std::string("cannot write ") + path);
  auto put = [&](const void *p, size_t n) { o.write(static_cast<const char *>(p), (std::streamsize)n); };
  const uint32_t version = 1, flags = (hist ? 1u : 0u) | (sub ? 2u : 0u);
  const int64_t dims[3] = {N, T, B};
  const uint64_t hash = connectivity_hash(h);
  put("OTFEMSV1", 8);
  put(&version, 4);
  put(&flags, 4);
  put(dims, sizeof(dims));
  put(&hash, 8);
  put(h->x0_host.data(), 3 * N * 8);
  put(x.data(), 3 * N * 8);
  put(v.data(), 3 * N * 8);
  put(f.data(), 3 * N * 8);
  put(h->fixed_host.data(), N);
  put(h->young.data(), T * 8);
  put(h->poisson.data(), T * 8);
  put(h->density.data(), T * 8);
  put(h->plastic_c, 3 * 8);
  put(pl.data(), 6 * T * 8);
  if (hist) {
    put(h1.data(), 4 * N * 4);
    put(h2.data(), 4 * N * 4);
  }
  put(it.data(), B * 8);
  put(un.data(), B * 8);
  if (sub) {
    const int32_t slots = h->ring_slots;
    put(&slots, 4);
    put(sub_head, sizeof(sub_head));
    for (const auto &r : ring) put(r.data(), 4 * N * 4);
    put(vb0.data(), 3 * N * 8);
    put(vb1.data(), 3 * N * 8);
  }
  if (!o) return fail(h, OTFEM_INVALID_ARG, std::string("write failed: ") + path);
  return OTFEM_OK;
}

otfem_status otfem_save_state(otfem_handle h, const char *path) {
  return no_throw(h, [&] { return otfem_save_state_body(h, path); });
}

static otfem_status otfem_load_state_body(otfem_handle h, const char *path) {
  if (!h || !path) return OTFEM_INVALID_ARG;
  std::ifstream in(path, std::ios::binary);
  if (!in) return fail(h, OTFEM_INVALID_ARG, std::string("cannot read ") + path);
  auto get = [&](void *p, size_t n) { in.read(static_cast<char *>(p), (std::streamsize)n); };
  char magic[8];
  uint32_t version = 0, flags = 0;
  int64_t dims[3] = {0, 0, 0};
  uint64_t hash = 0;
  get(magic, 8);
  get(&version, 4);
  get(&flags, 4);
  get(dims, sizeof(dims));
  get(&hash, 8);
  if (!in || std::memcmp(magic, "OTFEMSV1", 8) != 0 || version != 1)
    return fail(h, OTFEM_INVALID_ARG, std::string("not a solver state file: ") + path);
  const int64_t N = h->L.N, T = h->L.T, B = h->L.B;
  if (dims[0] != N || dims[1] != T || dims[2] != B || hash != connectivity_hash(h))
    return fail(h, OTFEM_INVALID_ARG, "state file is of another mesh (node / tet counts, "
                                      "connectivity or body offsets differ)");
  std::vector<double> x0(3 * N), x(3 * N), v(3 * N), f(3 * N), E(T), nu(T), rho(T), pl(6 * T);
  std::vector<uint8_t> fixed(N);
  double pc[3];
  std::vector<float> h1(4 * N), h2(4 * N);
  std::vector<unsigned long long> it(B), un(B);
  get(x0.data(), 3 * N * 8);
  get(x.data(), 3 * N * 8);
  get(v.data(), 3 * N * 8);
  get(f.data(), 3 * N * 8);
  get(fixed.data(), N);
  get(E.data(), T * 8);
  get(nu.data(), T * 8);
  get(rho.data(), T * 8);
  get(pc, 3 * 8);
  get(pl.data(), 6 * T * 8);
  if (flags & 1u) {
    get(h1.data(), 4 * N * 4);
    get(h2.data(), 4 * N * 4);
  }
  get(it.data(), B * 8);
  get(un.data(), B * 8);
  int32_t slots = 0, sub_head[3] = {0, 0, 0};
  std::vector<std::vector<float>> ring;
  std::vector<double> vb0, vb1;
  if (flags & 2u) {
    get(&slots, 4);
    get(sub_head, sizeof(sub_head));
    if (!in || slots < 1 || slots > kSubMax - 2 || sub_head[0] < 0 || sub_head[0] >= slots ||
        sub_head[1] < 0 || sub_head[1] > slots + 1 || (sub_head[2] & ~1))
      return fail(h, OTFEM_INVALID_ARG, std::string("bad subspace history in ") + path);
    ring.assign(slots, std::vector<float>(4 * N));
    for (auto &r : ring) get(r.data(), 4 * N * 4);
    vb0.resize(3 * N);
    vb1.resize(3 * N);
    get(vb0.data(), 3 * N * 8);
    get(vb1.data(), 3 * N * 8);
  }
  if (!in) return fail(h, OTFEM_INVALID_ARG, std::string("state file is truncated: ") + path);
  otfem_status st = ensure_device(h);
  if (st) return st;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  // rest shape and fixed flags first: the element data and masses follow
  // them.  init_common checks everything before it changes the solver; if it
  // rejects the file, the previous rest shape and flags are put back.
  std::vector<double> tmp;
  std::vector<uint8_t> fx(N);
  auto put_shape = [&]() -> hipError_t {
    for (int64_t g = 0; g < N; ++g) fx[g] = h->fixed_host[h->L.node_of_internal[g]];
    to_internal3(h->L, h->x0_host.data(), tmp);
    hipError_t e = hipMemcpy(h->x0, tmp.data(), 3 * N * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h->fixed, fx.data(), N, hipMemcpyHostToDevice);
    return e;
  };
  std::vector<double> old_x0 = h->x0_host;
  std::vector<uint8_t> old_fixed = h->fixed_host;
  h->x0_host = x0;
  for (int64_t i = 0; i < N; ++i) h->fixed_host[i] = fixed[i] ? 1 : 0;
  HIPCHK(h, put_shape());
  st = init_common(h, E.data(), nu.data(), rho.data(), false, pc[0], pc[1], pc[2]);
  if (st) {
    const std::string why = h->err;
    h->x0_host.swap(old_x0);
    h->fixed_host.swap(old_fixed);
    HIPCHK(h, put_shape());
    h->err = why;
    return st;
  }
  if (h->free_rank) {
    st = ws_upload_free(h);
    if (st) return st;
  }
  for (auto [src, dst] : {std::pair<const double *, double *>{x.data(), h->x}, {v.data(), h->v},
                          {f.data(), h->fext}}) {
    to_internal3(h->L, src, tmp);
    HIPCHK(h, hipMemcpy(dst, tmp.data(), 3 * N * sizeof(double), hipMemcpyHostToDevice));
  }
  if (T > 0) HIPCHK(h, hipMemcpy(h->plastic, pl.data(), 6 * T * sizeof(double), hipMemcpyHostToDevice));
  if ((flags & 1u) && h->hist1) {
    std::vector<float> raw;
    hist_to_internal(h->L, h1, raw);
    HIPCHK(h, hipMemcpy(h->hist1, raw.data(), N * sizeof(float4), hipMemcpyHostToDevice));
    hist_to_internal(h->L, h2, raw);
    HIPCHK(h, hipMemcpy(h->hist2, raw.data(), N * sizeof(float4), hipMemcpyHostToDevice));
  }
  HIPCHK(h, hipMemcpy(h->iter_acc, it.data(), B * sizeof(unsigned long long), hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(h->unconv_acc, un.data(), B * sizeof(unsigned long long), hipMemcpyHostToDevice));
  if (flags & 2u) {
    // the subspace start continues from the saved history (init_common above
    // cleared the solver's own); the ring keeps the file's slot count
    if (h->vring && h->ring_slots != slots) {
      HIPCHK(h, hipStreamSynchronize(h->stream));
      drop_graphs(h);
      dfree(h->vring);
      h->ring_slots = 0;
    }
    st = sub_ensure(h, slots, 2);
    if (st) return st;
    std::vector<float> raw;
    for (int k = 0; k < slots; ++k) {
      hist_to_internal(h->L, ring[k], raw);
      HIPCHK(h, hipMemcpy(h->vring + (int64_t)k * N, raw.data(), N * sizeof(float4), hipMemcpyHostToDevice));
    }
    to_internal3(h->L, vb0.data(), tmp);
    HIPCHK(h, hipMemcpy(h->vbuf, tmp.data(), 3 * N * sizeof(double), hipMemcpyHostToDevice));
    to_internal3(h->L, vb1.data(), tmp);
    HIPCHK(h, hipMemcpy(h->vbuf + 3 * N, tmp.data(), 3 * N * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->sub_state, sub_head, sizeof(sub_head), hipMemcpyHostToDevice));
  }
  h->rotations_valid = false;
  drop_graphs(h);
  return OTFEM_OK;
}

otfem_status otfem_load_state(otfem_handle h, const char *path) {
  return no_throw(h, [&] { return otfem_load_state_body(h, path); });
}

// ---- surface coupling ----

static otfem_status otfem_surface_bind_body(otfem_handle h, int64_t num_vertices, const double *vertices,
                                            double *barycentric, int32_t *tet_index) {
  if (!h) return OTFEM_INVALID_ARG;
  if (num_vertices < 0 || (num_vertices > 0 && (!vertices || !barycentric || !tet_index)))
    return fail(h, OTFEM_INVALID_ARG, "surface arrays missing or negative vertex count");
  otfem_status st = ensure_device(h);
  if (st) return st;
  // current node positions (the reference binds at node->m_coord), caller order
  std::vector<double> tmp(3 * h->L.N), x(3 * h->L.N);
  HIPCHK(h, hipMemcpyAsync(tmp.data(), h->x, tmp.size() * sizeof(double), hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  to_caller3(h->L, tmp, x.data());
  const int threads = (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
  coupling::bind_points(h->L.N, x.data(), h->L.T, h->tets.data(), num_vertices, vertices,
                        barycentric, tet_index, threads);
  return OTFEM_OK;
}

otfem_status otfem_surface_bind(otfem_handle h, int64_t num_vertices, const double *vertices,
                                double *barycentric, int32_t *tet_index) {
  return no_throw(h, [&] {
    return otfem_surface_bind_body(h, num_vertices, vertices, barycentric, tet_index);
  });
}

struct otfem_surface {
  otfem_solver *solver = nullptr;
  int64_t n = 0;
  int4 *ids = nullptr;      // [n] internal node ids of the tetrahedron; x = -1: unbound
  double4 *w = nullptr;     // [n] w0..w3
  double *rest = nullptr;   // [3n] positions kept for unbound vertices
  double *out = nullptr;    // [3n]
  int32_t *tet = nullptr;   // [n] caller tetrahedron index, -1: unbound
  double *n0 = nullptr;     // [3n] rest normals (otfem_surface_set_normals)
  double *nout = nullptr;   // [3n]
  bool has_normals = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;
};

namespace {
otfem_status surface_fail(otfem_surface *s, otfem_status st, const std::string &msg) {
  s->err = msg;
  return st;
}
void surface_free(otfem_surface *s) {
  if (s->ids) (void)hipFree(s->ids);
  if (s->w) (void)hipFree(s->w);
  if (s->rest) (void)hipFree(s->rest);
  if (s->out) (void)hipFree(s->out);
  if (s->tet) (void)hipFree(s->tet);
  if (s->n0) (void)hipFree(s->n0);
  if (s->nout) (void)hipFree(s->nout);
  if (s->ev0) (void)hipEventDestroy(s->ev0);
  if (s->ev1) (void)hipEventDestroy(s->ev1);
}
}  // namespace

#define SCHK(s, call)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return surface_fail((s), OTFEM_HIP_ERROR, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

static otfem_status otfem_surface_create_body(otfem_handle h, int64_t num_vertices, const double *vertices,
                                              const double *barycentric, const int32_t *tet_index,
                                              otfem_surface_handle *out) {
  if (!h || !out) return OTFEM_INVALID_ARG;
  *out = nullptr;
  if (num_vertices < 0 || (num_vertices > 0 && (!vertices || !barycentric || !tet_index)))
    return fail(h, OTFEM_INVALID_ARG, "surface arrays missing or negative vertex count");
  otfem_status st = ensure_device(h);
  if (st) return st;
  const int64_t n = num_vertices;
  std::vector<int4> ids((size_t)n);
  std::vector<double4> w((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    const int32_t t = tet_index[i];
    if (t < -1 || t >= h->L.T)
      return fail(h, OTFEM_INVALID_ARG, "surface vertex " + std::to_string(i) +
                                            ": tetrahedron index out of range");
    if (t < 0) {
      ids[(size_t)i] = make_int4(-1, -1, -1, -1);
      w[(size_t)i] = make_double4(0.0, 0.0, 0.0, 0.0);
      continue;
    }
    const int32_t *tt = &h->tets[4 * (int64_t)t];
    ids[(size_t)i] = make_int4(h->L.internal_of_node[tt[0]], h->L.internal_of_node[tt[1]],
                               h->L.internal_of_node[tt[2]], h->L.internal_of_node[tt[3]]);
    const double w1 = barycentric[3 * i], w2 = barycentric[3 * i + 1], w3 = barycentric[3 * i + 2];
    w[(size_t)i] = make_double4(1.0 - w1 - w2 - w3, w1, w2, w3);
  }
  auto *s = new otfem_surface();
  s->solver = h;
  s->n = n;
  auto bail = [&](otfem_status e) {
    fail(h, e, s->err);
    surface_free(s);
    delete s;
    return e;
  };
  otfem_status e = OTFEM_OK;
  auto step = [&](hipError_t r, const char *what) {
    if (r != hipSuccess && e == OTFEM_OK) {
      s->err = std::string(what) + ": " + hipGetErrorString(r);
      e = OTFEM_HIP_ERROR;
    }
  };
  if (n > 0) {
    step(hipMalloc(&s->ids, n * sizeof(int4)), "hipMalloc");
    step(hipMalloc(&s->w, n * sizeof(double4)), "hipMalloc");
    step(hipMalloc(&s->rest, 3 * n * sizeof(double)), "hipMalloc");
    step(hipMalloc(&s->out, 3 * n * sizeof(double)), "hipMalloc");
    step(hipMalloc(&s->tet, n * sizeof(int32_t)), "hipMalloc");
    if (e == OTFEM_OK) {
      step(hipMemcpy(s->tet, tet_index, n * sizeof(int32_t), hipMemcpyHostToDevice), "hipMemcpy");
      step(hipMemcpy(s->ids, ids.data(), n * sizeof(int4), hipMemcpyHostToDevice), "hipMemcpy");
      step(hipMemcpy(s->w, w.data(), n * sizeof(double4), hipMemcpyHostToDevice), "hipMemcpy");
      step(hipMemcpy(s->rest, vertices, 3 * n * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
    }
  }
  step(hipEventCreate(&s->ev0), "hipEventCreate");
  step(hipEventCreate(&s->ev1), "hipEventCreate");
  if (e != OTFEM_OK) return bail(e);
  *out = s;
  return OTFEM_OK;
}

otfem_status otfem_surface_create(otfem_handle h, int64_t num_vertices, const double *vertices,
                                  const double *barycentric, const int32_t *tet_index,
                                  otfem_surface_handle *out) {
  return no_throw(h, [&] {
    return otfem_surface_create_body(h, num_vertices, vertices, barycentric, tet_index, out);
  });
}

otfem_status otfem_surface_update(otfem_surface_handle s, double *vertices_out, double *kernel_ms) {
  if (!s) return OTFEM_INVALID_ARG;
  if (s->n > 0 && !vertices_out)
    return surface_fail(s, OTFEM_INVALID_ARG, "vertices_out is NULL");
  otfem_solver *h = s->solver;
  hipStream_t st = h->stream;
  SCHK(s, hipEventRecord(s->ev0, st));
  SCHK(s, enqueue_surface_update(s->n, s->ids, s->w, s->rest, h->x, s->out, st));
  SCHK(s, hipEventRecord(s->ev1, st));
  if (s->n > 0)
    SCHK(s, hipMemcpyAsync(vertices_out, s->out, 3 * s->n * sizeof(double), hipMemcpyDeviceToHost, st));
  SCHK(s, hipStreamSynchronize(st));
  if (kernel_ms) {
    float ms = 0.0f;
    SCHK(s, hipEventElapsedTime(&ms, s->ev0, s->ev1));
    *kernel_ms = ms;
  }
  return OTFEM_OK;
}

otfem_status otfem_surface_set_normals(otfem_surface_handle s, const double *normals) {
  if (!s) return OTFEM_INVALID_ARG;
  if (s->n > 0 && !normals) return surface_fail(s, OTFEM_INVALID_ARG, "normals is NULL");
  if (s->n > 0) {
    if (!s->n0) {
      SCHK(s, hipMalloc(&s->n0, 3 * s->n * sizeof(double)));
      SCHK(s, hipMalloc(&s->nout, 3 * s->n * sizeof(double)));
    }
    SCHK(s, hipMemcpyAsync(s->n0, normals, 3 * s->n * sizeof(double), hipMemcpyHostToDevice,
                           s->solver->stream));
    SCHK(s, hipStreamSynchronize(s->solver->stream));
  }
  s->has_normals = true;
  return OTFEM_OK;
}

otfem_status otfem_surface_update_normals(otfem_surface_handle s, double *normals_out) {
  if (!s) return OTFEM_INVALID_ARG;
  if (!s->has_normals) return surface_fail(s, OTFEM_STATE, "call otfem_surface_set_normals first");
  if (s->n > 0 && !normals_out) return surface_fail(s, OTFEM_INVALID_ARG, "normals_out is NULL");
  otfem_solver *h = s->solver;
  hipStream_t st = h->stream;
  SCHK(s, enqueue_surface_normals(s->n, s->tet, h->Re, h->rotations_valid ? 1 : 0, s->n0, s->nout, st));
  if (s->n > 0)
    SCHK(s, hipMemcpyAsync(normals_out, s->nout, 3 * s->n * sizeof(double), hipMemcpyDeviceToHost, st));
  SCHK(s, hipStreamSynchronize(st));
  return OTFEM_OK;
}

const char *otfem_surface_last_error(otfem_surface_handle s) { return s ? s->err.c_str() : ""; }

void otfem_surface_destroy(otfem_surface_handle s) {
  if (!s) return;
  if (s->solver && s->solver->stream) (void)hipStreamSynchronize(s->solver->stream);
  surface_free(s);
  delete s;
}

otfem_status otfem_set_warmstart_model(otfem_handle h, otfem_model_handle m) {
  if (!h) return OTFEM_INVALID_ARG;
  if (!m) {
    h->model = nullptr;
    h->recording = false;
    return OTFEM_OK;
  }
  otfem_status st = ensure_device(h);
  if (st) return st;
  ModelView mv{};
  model_view(m, &mv);
  if (mv.device != h->device) return fail(h, OTFEM_INVALID_ARG, "model and solver are on different devices");
  const int64_t N = h->L.N;
  if (!h->vw) {
    HIPCHK(h, dalloc(&h->vw, 3 * N));
    HIPCHK(h, dalloc(&h->hist1, N));
    HIPCHK(h, dalloc(&h->hist2, N));
    HIPCHK(h, dalloc(&h->kd, N));
    HIPCHK(h, dalloc(&h->free_rank, N));
    HIPCHK(h, dalloc(&h->ws_bad, 1));
    HIPCHK(h, hipMemsetAsync(h->kd, 0, N * sizeof(float4), h->stream));
    HIPCHK(h, hipMemsetAsync(h->ws_bad, 0, sizeof(int), h->stream));
    st = ws_upload_kd(h);
    if (st) return st;
  }
  st = ws_upload_free(h);
  if (st) return st;
  st = ws_reset_history(h);
  if (st) return st;
  h->model = m;
  return OTFEM_OK;
}

otfem_status otfem_set_recording(otfem_handle h, int on) {
  if (!h) return OTFEM_INVALID_ARG;
  if (on && !h->model) return fail(h, OTFEM_STATE, "attach a model before recording");
  h->recording = on != 0;
  return OTFEM_OK;
}

// A = cM M + cK K_rest over the free nodes (caller order), K_rest from the
// rest element stiffness Ke_ab = V [d17 g_a g_b^T + d18 g_b g_a^T + d18 (g_a.g_b) I]
// (fem_compute_ke.h:78-99 with d16 = d17 + 2 d18), fixed rows and columns dropped.
static otfem_status residual_operator(otfem_solver *h, double dt, const otfem_params *params,
                                      ResidualOp &op) {
  if (!h->initialized) return fail(h, OTFEM_STATE, "call otfem_init first");
  if (!(dt > 0)) return fail(h, OTFEM_INVALID_ARG, "dt must be positive");
  otfem_params P;
  if (params)
    P = *params;
  else
    otfem_default_params(&P, 1);
  const double cM = 1.0 + P.alpha * dt, cK = dt * dt + P.beta * dt;
  const int64_t N = h->L.N, T = h->L.T;
  std::vector<int32_t> rank(N);
  int64_t n = 0;
  for (int64_t u = 0; u < N; ++u) rank[u] = h->fixed_host[u] ? -1 : (int32_t)n++;
  std::vector<std::vector<std::pair<int32_t, std::array<double, 9>>>> rows(n);
  auto add_block = [&](int32_t r, int32_t c, const double *b) {
    for (auto &e : rows[r])
      if (e.first == c) {
        for (int q = 0; q < 9; ++q) e.second[q] += b[q];
        return;
      }
    std::array<double, 9> a;
    for (int q = 0; q < 9; ++q) a[q] = b[q];
    rows[r].emplace_back(c, a);
  };
  const double *x0 = h->x0_host.data();
  for (int64_t t = 0; t < T; ++t) {
    const int32_t *tt = &h->tets[4 * t];
    const ElemData ed = element_data(x0 + 3 * tt[0], x0 + 3 * tt[1], x0 + 3 * tt[2],
                                     x0 + 3 * tt[3], h->young[t], h->poisson[t]);
    double g[4][3];
    for (int d = 0; d < 3; ++d) {
      g[1][d] = ed.grad[d];
      g[2][d] = ed.grad[3 + d];
      g[3][d] = ed.grad[6 + d];
      g[0][d] = -(g[1][d] + g[2][d] + g[3][d]);
    }
    for (int a = 0; a < 4; ++a) {
      const int32_t ra = rank[tt[a]];
      if (ra < 0) continue;
      for (int b = 0; b < 4; ++b) {
        const int32_t rb = rank[tt[b]];
        if (rb < 0) continue;
        const double dot = g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2];
        double blk[9];
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c)
            blk[3 * r + c] = cK * (ed.vd[1] * g[a][r] * g[b][c] + ed.vd[2] * g[b][r] * g[a][c] +
                                   (r == c ? ed.vd[2] * dot : 0.0));
        add_block(ra, rb, blk);
      }
    }
  }
  for (int64_t u = 0; u < N; ++u) {
    if (rank[u] < 0) continue;
    const double mm = cM * h->mass_host[u];
    const double blk[9] = {mm, 0, 0, 0, mm, 0, 0, 0, mm};
    add_block(rank[u], rank[u], blk);
  }
  op.n = n;
  op.row_ptr.assign(n + 1, 0);
  op.cols.clear();
  op.blocks.clear();
  for (int64_t r = 0; r < n; ++r) {
    std::sort(rows[r].begin(), rows[r].end(),
              [](const auto &a, const auto &b) { return a.first < b.first; });
    for (auto &e : rows[r]) {
      op.cols.push_back(e.first);
      for (int q = 0; q < 9; ++q) op.blocks.push_back((float)e.second[q]);
    }
    op.row_ptr[r + 1] = (int64_t)op.cols.size();
  }
  return OTFEM_OK;
}

static otfem_status otfem_model_train_residual_body(otfem_model_handle m, otfem_handle h, double dt,
                                                    const otfem_params *params, int steps,
                                                    int steps_per_batch, double lr, uint64_t seed,
                                                    float *loss) {
  if (!m || !h) return OTFEM_INVALID_ARG;
  ResidualOp op;
  otfem_status st = residual_operator(h, dt, params, op);
  if (st) return st;
  std::string err;
  st = (otfem_status)model_train_residual(m, op, steps, steps_per_batch, lr, seed, loss, &err);
  if (st) h->err = err;
  return st;
}

otfem_status otfem_model_train_residual(otfem_model_handle m, otfem_handle h, double dt,
                                        const otfem_params *params, int steps,
                                        int steps_per_batch, double lr, uint64_t seed,
                                        float *loss) {
  return no_throw(h, [&] {
    return otfem_model_train_residual_body(m, h, dt, params, steps, steps_per_batch, lr, seed, loss);
  });
}

static otfem_status otfem_model_residual_gradient_body(otfem_model_handle m, otfem_handle h, double dt,
                                                       const otfem_params *params, const int32_t *steps,
                                                       int num_steps, double *loss, float *grad) {
  if (!m || !h || !steps || num_steps <= 0) return OTFEM_INVALID_ARG;
  ResidualOp op;
  otfem_status st = residual_operator(h, dt, params, op);
  if (st) return st;
  std::string err;
  st = (otfem_status)model_residual(m, op, 1, num_steps, 1.0, 0, nullptr, steps, grad, loss, &err);
  if (st) h->err = err;
  return st;
}

otfem_status otfem_model_residual_gradient(otfem_model_handle m, otfem_handle h, double dt,
                                           const otfem_params *params, const int32_t *steps,
                                           int num_steps, double *loss, float *grad) {
  return no_throw(h, [&] {
    return otfem_model_residual_gradient_body(m, h, dt, params, steps, num_steps, loss, grad);
  });
}

}  // extern "C"
