// Host-side device layout of a mesh (or CSR-of-bodies batch) for the otfem
// kernels.  No HIP calls here: built on the CPU at otfem_create.
//
// Layout ("owner-computes tiles"):
//   * Nodes are split per body into tiles by recursive coordinate bisection of
//     the rest coordinates; a tile OWNS a contiguous range of internal node ids.
//   * Every tetrahedron is copied into each tile that owns at least one of its
//     corners; such a copy is a SLOT.  A slot carries the four corners as
//     tile-local indices into the tile's node table [owned..., halo...].
//   * A tile computes element terms for all its slots but accumulates only
//     into its owned nodes, so every node sum is complete inside one
//     workgroup: no global atomics, and per-node epilogues (dot products,
//     axpys) fuse into the same pass.  The node sums are LDS atomics, so
//     their rounding depends on wave order (not bitwise reproducible).
#pragma once
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <exception>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace otfem {

// Runs f(i) for i in [0, n) on up to `threads` threads (the caller's thread
// is one of them); items are handed out one at a time, so uneven bodies and
// tiles balance.  Each f(i) must write only to data that item owns.  An
// exception in any item (a failed allocation) stops the hand-out and is
// rethrown on the calling thread once every worker has joined.
template <class F>
void parallel_for(int64_t n, int threads, F &&f) {
  const int64_t nt = std::min<int64_t>(threads, n);
  if (nt <= 1) {
    for (int64_t i = 0; i < n; ++i) f(i);
    return;
  }
  std::atomic<int64_t> next{0};
  std::exception_ptr error;
  std::mutex error_lock;
  auto work = [&] {
    try {
      for (int64_t i; (i = next.fetch_add(1, std::memory_order_relaxed)) < n;) f(i);
    } catch (...) {
      std::lock_guard<std::mutex> g(error_lock);
      if (!error) error = std::current_exception();
      next.store(n);
    }
  };
  std::vector<std::thread> pool;
  pool.reserve(nt - 1);
  for (int64_t t = 1; t < nt; ++t) pool.emplace_back(work);
  work();
  for (auto &t : pool) t.join();
  if (error) std::rethrow_exception(error);
}

// Host threads for set-up work: one per hardware thread, at most 16.
inline int host_threads() {
  return (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
}


struct TileInfo {
  int32_t node_begin;  // first owned internal node id
  int32_t n_owned;
  int32_t halo_begin;  // offset into halo[]
  int32_t n_halo;
  int32_t slot_begin;
  int32_t n_slots;
  int32_t body;
  int32_t body_rank;   // index of this tile within its body
  int32_t body_tile_begin;  // tiles [body_tile_begin, body_tile_end) are the body's
  int32_t body_tile_end;
};

struct Layout {
  int64_t N = 0, T = 0, B = 0;
  std::vector<int32_t> node_of_internal;  // internal -> user node id
  std::vector<int32_t> internal_of_node;  // user -> internal node id
  std::vector<TileInfo> tiles;
  std::vector<int32_t> halo;              // internal node ids
  std::vector<uint16_t> slot_conn;        // 4 per slot, tile-local indices
  std::vector<int32_t> slot_tet;          // user tet id; bit 31 = primary copy
  std::vector<int32_t> body_tile_begin;   // B+1
  std::vector<int64_t> body_node_begin;   // B+1, user node offsets
  std::vector<int64_t> body_tet_begin;    // B+1, user tet offsets
  int32_t max_tile_nodes = 0;             // max n_owned + n_halo
  int32_t max_tile_owned = 0;
  int32_t max_tile_slots = 0;
};

// Builds the layout.  `coords` are rest coordinates [3N] (used only for the
// spatial partition), `tets` [4T].  Bodies: B == 0/1 means one body spanning
// everything.  Returns an empty string on success, else an error message.
// The work runs on `threads` host threads (0: one per hardware thread, at
// most 16), split by body and by tile; every body and tile is processed in
// the same order whatever the thread count, so the layout does not depend on
// it.
std::string build_layout(int64_t N, int64_t T, const double *coords, const int32_t *tets,
                         int64_t B, const int64_t *body_node_off, const int64_t *body_tet_off,
                         int target_tile_nodes, int max_tile_nodes, Layout &out,
                         int threads = 0);

}  // namespace otfem
