// Owner-computes tile layout (see layout.h).
#include "layout.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <exception>
#include <mutex>
#include <numeric>
#include <thread>
#include <unordered_map>

namespace otfem {

namespace {

// Recursive coordinate bisection of `idx[lo, hi)` into `parts` parts of
// about equal total weight `w` (every part keeps at least one node); the part
// sizes are appended to `part_sizes` in order.
void rcb(std::vector<int32_t> &idx, int64_t lo, int64_t hi, int64_t parts, const double *x,
         const double *w, std::vector<int32_t> &part_sizes) {
  if (parts <= 1 || hi - lo <= 1) {
    part_sizes.push_back((int32_t)(hi - lo));
    return;
  }
  double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
  for (int64_t i = lo; i < hi; ++i)
    for (int c = 0; c < 3; ++c) {
      double v = x[3 * (int64_t)idx[i] + c];
      mn[c] = std::min(mn[c], v);
      mx[c] = std::max(mx[c], v);
    }
  int ax = 0;
  for (int c = 1; c < 3; ++c)
    if (mx[c] - mn[c] > mx[ax] - mn[ax]) ax = c;
  int64_t left_parts = parts / 2;
  std::sort(idx.begin() + lo, idx.begin() + hi, [&](int32_t a, int32_t b) {
    double va = x[3 * (int64_t)a + ax], vb = x[3 * (int64_t)b + ax];
    if (va != vb) return va < vb;
    return a < b;
  });
  double total = 0.0;
  for (int64_t i = lo; i < hi; ++i) total += w[idx[i]];
  const double target = total * (double)left_parts / (double)parts;
  double acc = 0.0;
  int64_t mid = lo;
  while (mid < hi && acc + 0.5 * w[idx[mid]] < target) acc += w[idx[mid++]];
  mid = std::min(std::max(mid, lo + left_parts), hi - (parts - left_parts));
  rcb(idx, lo, mid, left_parts, x, w, part_sizes);
  rcb(idx, mid, hi, parts - left_parts, x, w, part_sizes);
}

// Picks an order for the slots of one tile, in windows of 64 (one wave
// instruction) in which no two slots share an owned corner, so the per-corner
// LDS atomics of a wave never hit the same address.  Each quarter window (16
// lanes) also spreads each corner's node over the 16 residues of its local
// index mod 16: a tile node's three FP64 values start at 8-byte word 3 l, an
// 8-byte access of 16 lanes is served in one pass when their words fall in
// distinct bank pairs (word mod 16), so lanes of a quarter wave that read
// different nodes of one residue wait for each other.  (Measured on the 64^3
// cube: 32.6-33.1 us per CG iteration with quarter windows mod 16, 33.1 us
// with half windows mod 32 -- equal within the spread between runs -- and
// 36.0 us without spreading.)  Quarter windows are filled in
// passes, each taking compatible slots in order: at most 1, 2, 3 and then 4
// distinct nodes per corner and residue (looking at most kLook slots past the
// first remaining one), then any slot without a shared owned corner, then any
// slot, so later windows stay aligned.
constexpr int kLook = 512;
constexpr int kCaps = 4;
constexpr int kGroup = 16;  // lanes served by the LDS in one pass
constexpr int kRes = 16;    // bank residues of a node's local index

void pick_slots(const uint16_t *sc, int32_t n, int32_t n_owned, std::vector<int32_t> &order) {
  order.clear();
  order.reserve(n);
  std::vector<char> taken(n, 0);
  std::vector<int32_t> mark(n_owned, -1);
  int32_t first = 0;
  for (int32_t window = 0; (int32_t)order.size() < n; ++window) {
    for (int part = 0; part < 64 / kGroup && (int32_t)order.size() < n; ++part) {
      // per corner and residue: distinct nodes taken in this part (first kCaps kept)
      int count[4][kRes] = {};
      int32_t seen[4][kRes][kCaps];
      auto fresh = [&](int c, int32_t li) {  // li not yet taken at corner c
        const int r = li & (kRes - 1);
        for (int q = 0; q < std::min(count[c][r], kCaps); ++q)
          if (seen[c][r][q] == li) return false;
        return true;
      };
      int got = 0;
      for (int pass = 0; pass < kCaps + 2 && got < kGroup; ++pass) {
        const int cap = pass + 1;  // passes 0..kCaps-1: residue cap 1..kCaps
        int looked = 0;
        for (int32_t k = first; k < n && got < kGroup; ++k) {
          if (taken[k]) continue;
          if (pass < kCaps && ++looked > kLook) break;
          bool ok = true;
          if (pass < kCaps + 1)
            for (int c = 0; c < 4 && ok; ++c) {
              const int32_t li = sc[4 * k + c];
              if (li < n_owned && mark[li] == window) ok = false;
            }
          if (pass < kCaps)
            for (int c = 0; c < 4 && ok; ++c) {
              const int32_t li = sc[4 * k + c];
              if (fresh(c, li) && count[c][li & (kRes - 1)] >= cap) ok = false;
            }
          if (!ok) continue;
          for (int c = 0; c < 4; ++c) {
            const int32_t li = sc[4 * k + c];
            if (li < n_owned) mark[li] = window;
            const int r = li & (kRes - 1);
            if (fresh(c, li)) {
              if (count[c][r] < kCaps) seen[c][r][count[c][r]] = li;
              ++count[c][r];
            }
          }
          taken[k] = 1;
          order.push_back(k);
          ++got;
        }
      }
      while (first < n && taken[first]) ++first;
    }
  }
}

// The slot order of one tile: two rounds of pick_slots, the second over the
// first's sequence, whose windows already hold compatible slots side by side
// (the second round spreads the residues further).
void order_slots(const TileInfo &ti, uint16_t *slot_conn, int32_t *slot_tet) {
  const int32_t n = ti.n_slots;
  const uint16_t *sc = slot_conn + 4 * (int64_t)ti.slot_begin;
  std::vector<int32_t> o1, o2;
  pick_slots(sc, n, ti.n_owned, o1);
  std::vector<uint16_t> c1(4 * (size_t)n);
  for (int32_t k = 0; k < n; ++k)
    for (int c = 0; c < 4; ++c) c1[4 * k + c] = sc[4 * o1[k] + c];
  pick_slots(c1.data(), n, ti.n_owned, o2);
  std::vector<int32_t> order(n);
  for (int32_t k = 0; k < n; ++k) order[k] = o1[o2[k]];
  std::vector<uint16_t> conn(4 * (size_t)n);
  std::vector<int32_t> tet(n);
  int32_t *st = slot_tet + ti.slot_begin;
  for (int32_t k = 0; k < n; ++k) {
    for (int c = 0; c < 4; ++c) conn[4 * k + c] = sc[4 * order[k] + c];
    tet[k] = st[order[k]];
  }
  std::copy(conn.begin(), conn.end(), slot_conn + 4 * (int64_t)ti.slot_begin);
  std::copy(tet.begin(), tet.end(), st);
}

}  // namespace

std::string build_layout(int64_t N, int64_t T, const double *coords, const int32_t *tets,
                         int64_t B, const int64_t *body_node_off, const int64_t *body_tet_off,
                         int target_tile_nodes, int max_tile_nodes, Layout &L, int threads) {
  L = Layout();
  if (N <= 0 || T < 0) return "num_nodes must be positive and num_tets non-negative";
  if (N >= (int64_t)INT32_MAX || T >= (int64_t)INT32_MAX) return "mesh too large for 32-bit ids";
  if (threads <= 0) threads = (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
  // Default tiling: about 512 owned nodes per tile; a single body large
  // enough to fill the GPU gets a tile count that is a multiple of the 256
  // compute units of an MI355X, so every CU runs the same number of tiles,
  // and a smaller single body gets up to 248 smaller tiles (see below).
  const bool auto_tiles = target_tile_nodes <= 0;
  if (auto_tiles) target_tile_nodes = 512;
  if (max_tile_nodes <= 0) max_tile_nodes = 8192;
  L.N = N;
  L.T = T;
  if (B <= 1) {
    L.B = 1;
    L.body_node_begin = {0, N};
    L.body_tet_begin = {0, T};
  } else {
    L.B = B;
    L.body_node_begin.assign(body_node_off, body_node_off + B + 1);
    L.body_tet_begin.assign(body_tet_off, body_tet_off + B + 1);
    if (L.body_node_begin[0] != 0 || L.body_node_begin[B] != N || L.body_tet_begin[0] != 0 ||
        L.body_tet_begin[B] != T)
      return "body offsets must start at 0 and end at num_nodes / num_tets";
    for (int64_t b = 0; b < B; ++b)
      if (L.body_node_begin[b + 1] <= L.body_node_begin[b] ||
          L.body_tet_begin[b + 1] < L.body_tet_begin[b])
        return "body offsets must be increasing (every body needs a node)";
  }
  const int64_t nb = L.B;
  const int64_t *bn = L.body_node_begin.data();
  const int64_t *bt = L.body_tet_begin.data();
  // Every tet must index a node of its own body (the whole mesh when there
  // is one body).  After this check the bodies share no node and no tet, so
  // the per-body passes below write disjoint data.
  {
    std::atomic<int> bad{0};
    parallel_for(nb, threads, [&](int64_t b) {
      for (int64_t t = bt[b]; t < bt[b + 1]; ++t)
        for (int k = 0; k < 4; ++k) {
          int64_t v = tets[4 * t + k];
          if (v < 0 || v >= N) {
            bad.store(1);
            return;
          }
          if (v < bn[b] || v >= bn[b + 1]) {  // reported only if no index is out of range
            int expected = 0;
            bad.compare_exchange_strong(expected, 2);
          }
        }
    });
    if (bad.load() == 1) return "tetrahedron node index out of range";
    if (bad.load() == 2) return "a tetrahedron references a node of another body";
  }

  // A batch whose 512-node tiling leaves CUs idle (fewer than 248 tiles in
  // all) is latency-bound like a small single body: its bodies get smaller
  // tiles, the largest node target (16 at least) that keeps the batch within
  // 248 tiles.
  if (auto_tiles && nb > 1) {
    auto tiles_at = [&](int64_t target) {
      int64_t total = 0;
      for (int64_t b = 0; b < nb; ++b) total += (bn[b + 1] - bn[b] + target - 1) / target;
      return total;
    };
    if (tiles_at(512) < 248) {
      int64_t target = std::max<int64_t>(16, (N + 247) / 248);
      while (target < 512 && tiles_at(target) > 248) target += target / 8 + 1;
      target_tile_nodes = (int)std::min<int64_t>(target, 512);
    }
  }

  // 1. Partition nodes of each body into tiles (internal ids = tile-major).
  // A tile's CG work is mostly its slot count.  Nodes inside a body touch
  // more tets than nodes on its surface, so tiles of equal node count hold
  // up to ~10% more slots in the interior than at the surface; weighting
  // each node by (tets at the node)^3 + 1 in the bisection brings the
  // largest tile to ~2-3% above the mean slot count (64^3 cube: 3,523 vs
  // 3,805 slots at a mean of 3,450).
  // The bisection runs on smoothed coordinates (each pass moves a node to the
  // mean centroid of its tets): on a jittered lattice, cuts through the raw
  // coordinates are ragged and cut more tets (64^3 cube, jitter 0.2 h: 1.43
  // slots per tet; after 8 passes 1.37, against 1.35 without jitter).  Nodes
  // of no tet keep their coordinates.
  std::vector<double> weight(N, 0.0);
  std::vector<double> xs(coords, coords + 3 * N);
  std::vector<int32_t> order(N);
  std::vector<std::vector<int32_t>> body_sizes(nb);
  {
    std::vector<double> acc(3 * N);
    std::vector<int32_t> cnt(N);
    parallel_for(nb, threads, [&](int64_t b) {
      const int64_t n0 = bn[b], n1 = bn[b + 1];
      for (int64_t q = 4 * bt[b]; q < 4 * bt[b + 1]; ++q) weight[tets[q]] += 1.0;
      for (int64_t i = n0; i < n1; ++i) weight[i] = weight[i] * weight[i] * weight[i] + 1.0;
      for (int pass = 0; pass < 8 && bt[b + 1] > bt[b]; ++pass) {
        std::fill(acc.begin() + 3 * n0, acc.begin() + 3 * n1, 0.0);
        std::fill(cnt.begin() + n0, cnt.begin() + n1, 0);
        for (int64_t t = bt[b]; t < bt[b + 1]; ++t) {
          const int32_t *tt = tets + 4 * t;
          double c[3] = {0.0, 0.0, 0.0};
          for (int k = 0; k < 4; ++k)
            for (int d = 0; d < 3; ++d) c[d] += xs[3 * (int64_t)tt[k] + d];
          for (int k = 0; k < 4; ++k) {
            for (int d = 0; d < 3; ++d) acc[3 * (int64_t)tt[k] + d] += 0.25 * c[d];
            ++cnt[tt[k]];
          }
        }
        for (int64_t i = n0; i < n1; ++i)
          if (cnt[i] > 0)
            for (int d = 0; d < 3; ++d) xs[3 * i + d] = acc[3 * i + d] / cnt[i];
      }
      // bisection of the body's nodes; the body's internal ids are
      // [n0, n1) because bodies keep their order
      std::iota(order.begin() + n0, order.begin() + n1, (int32_t)n0);
      int64_t parts = std::max<int64_t>(1, (n1 - n0 + target_tile_nodes - 1) / target_tile_nodes);
      if (auto_tiles && nb == 1) {
        const int64_t cus = 256;
        const int64_t k = (int64_t)std::llround((double)(n1 - n0) / (512.0 * cus));
        if (k >= 1) {
          parts = k * cus;
        } else {
          // A body too small to give every CU a 512-node tile is
          // latency-bound: more, smaller tiles shorten an iteration although
          // more tets are copied into several tiles.  About one tile per CU,
          // up to 248 so that none waits for a second round, and 16 owned
          // nodes per tile on average (measured, 3 steps: a 4,961-node cantilever
          // 6.6 ms on 10 tiles, 3.7 ms on 249; a 30^3 cube 14.4 ms on 59
          // tiles, 9.4 ms on 233, 11.9 ms on 257).
          parts = std::max<int64_t>(1, std::min<int64_t>(248, (n1 - n0) / 16));
        }
      }
      rcb(order, n0, n1, parts, xs.data(), weight.data(), body_sizes[b]);
      // keep user order inside a tile (locality of generator order)
      int64_t pos = n0;
      for (int32_t s : body_sizes[b]) {
        std::sort(order.begin() + pos, order.begin() + pos + s);
        pos += s;
      }
    });
  }
  L.body_tile_begin.assign(nb + 1, 0);
  for (int64_t b = 0; b < nb; ++b)
    L.body_tile_begin[b + 1] = L.body_tile_begin[b] + (int32_t)body_sizes[b].size();
  const int64_t ntiles = L.body_tile_begin[nb];
  L.tiles.assign(ntiles, TileInfo{});
  L.node_of_internal = std::move(order);
  L.internal_of_node.assign(N, -1);
  std::vector<int32_t> tile_of_internal(N);
  parallel_for(nb, threads, [&](int64_t b) {
    int64_t next_internal = bn[b];
    int32_t tile = L.body_tile_begin[b], rank = 0;
    for (int32_t s : body_sizes[b]) {
      TileInfo &ti = L.tiles[tile];
      ti.node_begin = (int32_t)next_internal;
      ti.n_owned = s;
      ti.body = (int32_t)b;
      ti.body_rank = rank++;
      ti.body_tile_begin = L.body_tile_begin[b];
      ti.body_tile_end = L.body_tile_begin[b + 1];
      for (int32_t k = 0; k < s; ++k) {
        L.internal_of_node[L.node_of_internal[next_internal]] = (int32_t)next_internal;
        tile_of_internal[next_internal] = tile;
        ++next_internal;
      }
      ++tile;
    }
  });
  body_sizes.clear();
  body_sizes.shrink_to_fit();

  // 2. Slots: each tet goes into every distinct tile owning one of its corners.
  // A body's tets only reach the body's own tiles.
  std::vector<int64_t> tile_nslots(ntiles, 0);
  parallel_for(nb, threads, [&](int64_t b) {
    for (int64_t t = bt[b]; t < bt[b + 1]; ++t) {
      int32_t tl[4];
      for (int k = 0; k < 4; ++k) tl[k] = tile_of_internal[L.internal_of_node[tets[4 * t + k]]];
      for (int k = 0; k < 4; ++k) {
        bool dup = false;
        for (int q = 0; q < k; ++q) dup |= (tl[q] == tl[k]);
        if (!dup) tile_nslots[tl[k]]++;
      }
    }
  });
  int64_t S = 0;
  for (int64_t i = 0; i < ntiles; ++i) {
    if (S + tile_nslots[i] >= (int64_t)INT32_MAX) return "too many slots for 32-bit offsets";
    L.tiles[i].slot_begin = (int32_t)S;
    L.tiles[i].n_slots = 0;
    S += tile_nslots[i];
  }
  L.slot_conn.assign(4 * S, 0);
  L.slot_tet.assign(S, 0);
  std::vector<int32_t> slot_corner_internal(4 * S);
  parallel_for(nb, threads, [&](int64_t b) {
    for (int64_t t = bt[b]; t < bt[b + 1]; ++t) {
      int32_t in[4], tl[4];
      for (int k = 0; k < 4; ++k) {
        in[k] = L.internal_of_node[tets[4 * t + k]];
        tl[k] = tile_of_internal[in[k]];
      }
      for (int k = 0; k < 4; ++k) {
        bool dup = false;
        for (int q = 0; q < k; ++q) dup |= (tl[q] == tl[k]);
        if (dup) continue;
        TileInfo &ti = L.tiles[tl[k]];
        int64_t s = ti.slot_begin + ti.n_slots++;
        L.slot_tet[s] = (int32_t)t | (k == 0 ? (int32_t)0x80000000u : 0);
        for (int c = 0; c < 4; ++c) slot_corner_internal[4 * s + c] = in[c];
      }
    }
  });

  // 3. Halo tables and tile-local corner indices, per tile; the halo lists
  // are then concatenated in tile order.
  std::vector<std::vector<int32_t>> tile_halo(ntiles);
  std::atomic<bool> too_many{false};
  parallel_for(ntiles, threads, [&](int64_t i) {
    TileInfo &ti = L.tiles[i];
    std::vector<int32_t> &halo = tile_halo[i];
    std::unordered_map<int32_t, int32_t> local;
    for (int64_t s = ti.slot_begin; s < ti.slot_begin + ti.n_slots; ++s)
      for (int c = 0; c < 4; ++c) {
        int32_t g = slot_corner_internal[4 * s + c];
        int32_t li;
        if (g >= ti.node_begin && g < ti.node_begin + ti.n_owned) {
          li = g - ti.node_begin;
        } else {
          auto it = local.find(g);
          if (it == local.end()) {
            li = ti.n_owned + (int32_t)halo.size();
            local.emplace(g, li);
            halo.push_back(g);
          } else {
            li = it->second;
          }
        }
        if (li > 65535) {
          too_many.store(true);
          return;
        }
        L.slot_conn[4 * s + c] = (uint16_t)li;
      }
    ti.n_halo = (int32_t)halo.size();
    order_slots(ti, L.slot_conn.data(), L.slot_tet.data());
  });
  if (too_many.load()) return "tile node table exceeds 16-bit local indices";
  int64_t H = 0;
  for (int64_t i = 0; i < ntiles; ++i) {
    L.tiles[i].halo_begin = (int32_t)H;
    H += L.tiles[i].n_halo;
  }
  L.halo.resize(H);
  parallel_for(ntiles, threads, [&](int64_t i) {
    std::copy(tile_halo[i].begin(), tile_halo[i].end(), L.halo.begin() + L.tiles[i].halo_begin);
  });
  for (const TileInfo &ti : L.tiles) {
    L.max_tile_nodes = std::max(L.max_tile_nodes, ti.n_owned + ti.n_halo);
    L.max_tile_owned = std::max(L.max_tile_owned, ti.n_owned);
    L.max_tile_slots = std::max(L.max_tile_slots, ti.n_slots);
  }
  if (L.max_tile_nodes > max_tile_nodes) return "tile node table exceeds the LDS budget";
  return std::string();
}

}  // namespace otfem
