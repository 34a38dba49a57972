// Host check of the tile layout (opentissue_amd/csrc/layout.cpp), built with
// g++ from the layout source alone:
//   * the layout is the same for 1, 3 and 8 host threads (batched bodies and
//     one large body with the CU-multiple default tiling), and a small body
//     gets the small-body default (min(248, N / 16) tiles);
//   * structure: node ids form a permutation, tiles never mix bodies, every
//     tet has exactly one primary slot and one slot per distinct owning tile
//     of its corners, local corner indices resolve to the right nodes, and
//     the first window of 64 slots of a tile with at least 256 owned nodes
//     repeats no owned corner, and quarter windows of 16 slots spread their
//     corner nodes over the residues mod 16 (LDS banks);
//   * errors: an out-of-range index is reported before a tet that reaches
//     into another body.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>
#include <vector>

#include "layout.h"

using otfem::Layout;
using otfem::TileInfo;

static int failures = 0;
#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #c); \
      ++failures;                                                  \
    }                                                              \
  } while (0)

// Lattice of I x J x K cubes, five tets per cube (alternating split), nodes
// jittered by a fixed-seed LCG; appended to coords / tets.
static void lattice(int I, int J, int K, double h, unsigned &seed, std::vector<double> &x,
                    std::vector<int32_t> &t) {
  const int32_t base = (int32_t)(x.size() / 3);
  auto id = [&](int i, int j, int k) { return base + (i * (J + 1) + j) * (K + 1) + k; };
  for (int i = 0; i <= I; ++i)
    for (int j = 0; j <= J; ++j)
      for (int k = 0; k <= K; ++k) {
        const int c[3] = {i, j, k};
        for (int d = 0; d < 3; ++d) {
          seed = seed * 1664525u + 1013904223u;
          x.push_back(h * c[d] + 0.2 * h * ((seed >> 8) * (1.0 / 16777216.0) - 0.5));
        }
      }
  for (int i = 0; i < I; ++i)
    for (int j = 0; j < J; ++j)
      for (int k = 0; k < K; ++k) {
        int p[8] = {id(i, j, k),         id(i + 1, j, k),         id(i + 1, j + 1, k),
                    id(i, j + 1, k),     id(i, j, k + 1),         id(i + 1, j, k + 1),
                    id(i + 1, j + 1, k + 1), id(i, j + 1, k + 1)};
        static const int odd[5][4] = {{1, 2, 3, 6}, {3, 6, 7, 4}, {1, 4, 5, 6}, {1, 3, 0, 4}, {1, 6, 3, 4}};
        static const int even[5][4] = {{2, 0, 1, 5}, {2, 7, 3, 0}, {2, 5, 6, 7}, {0, 7, 4, 5}, {2, 0, 5, 7}};
        const int(*q)[4] = ((i + j + k) & 1) ? odd : even;
        for (int e = 0; e < 5; ++e)
          for (int c = 0; c < 4; ++c) t.push_back(p[q[e][c]]);
      }
}

static bool same(const Layout &a, const Layout &b) {
  return a.tiles.size() == b.tiles.size() &&
         std::memcmp(a.tiles.data(), b.tiles.data(), a.tiles.size() * sizeof(TileInfo)) == 0 &&
         a.node_of_internal == b.node_of_internal && a.internal_of_node == b.internal_of_node &&
         a.halo == b.halo && a.slot_conn == b.slot_conn && a.slot_tet == b.slot_tet &&
         a.body_tile_begin == b.body_tile_begin && a.max_tile_nodes == b.max_tile_nodes &&
         a.max_tile_owned == b.max_tile_owned && a.max_tile_slots == b.max_tile_slots;
}

// Mean over full quarter windows (16 slots) and corners of the largest number
// of distinct tile nodes that share a residue of the local index mod 16: the
// lanes of a quarter wave that read such nodes wait on the same LDS banks.
static double bank_spread(const Layout &L) {
  double sum = 0.0;
  int64_t n = 0;
  for (const TileInfo &ti : L.tiles)
    for (int32_t w = 0; w + 16 <= ti.n_slots; w += 16)
      for (int c = 0; c < 4; ++c) {
        std::set<int32_t> by[16];
        for (int32_t s = w; s < w + 16; ++s) {
          const int32_t li = L.slot_conn[4 * (ti.slot_begin + s) + c];
          by[li & 15].insert(li);
        }
        size_t most = 0;
        for (const auto &b : by) most = std::max(most, b.size());
        sum += (double)most;
        ++n;
      }
  return n ? sum / (double)n : 0.0;
}

static void structure(const Layout &L, const std::vector<int32_t> &tets) {
  const int64_t N = L.N, T = L.T;
  std::vector<int> seen(N, 0);
  for (int64_t i = 0; i < N; ++i) {
    const int32_t u = L.node_of_internal[i];
    CHECK(u >= 0 && u < N);
    if (u >= 0 && u < N) {
      ++seen[u];
      CHECK(L.internal_of_node[u] == i);
    }
  }
  for (int64_t u = 0; u < N; ++u) CHECK(seen[u] == 1);
  std::vector<int> primary(T, 0), copies(T, 0);
  int64_t next_node = 0, next_slot = 0;
  for (size_t ti_i = 0; ti_i < L.tiles.size(); ++ti_i) {
    const TileInfo &ti = L.tiles[ti_i];
    CHECK(ti.node_begin == next_node);
    CHECK(ti.slot_begin == next_slot);
    next_node += ti.n_owned;
    next_slot += ti.n_slots;
    CHECK(ti_i >= (size_t)ti.body_tile_begin && ti_i < (size_t)ti.body_tile_end);
    // owned nodes lie in the tile's body
    for (int32_t k = 0; k < ti.n_owned; ++k) {
      const int32_t u = L.node_of_internal[ti.node_begin + k];
      CHECK(u >= L.body_node_begin[ti.body] && u < L.body_node_begin[ti.body + 1]);
    }
    for (int32_t k = 0; k < ti.n_halo; ++k) {
      const int32_t g = L.halo[ti.halo_begin + k];
      CHECK(g < ti.node_begin || g >= ti.node_begin + ti.n_owned);
    }
    for (int32_t s = 0; s < ti.n_slots; ++s) {
      const int64_t gs = ti.slot_begin + s;
      const int32_t t = L.slot_tet[gs] & 0x7fffffff;
      CHECK(t >= L.body_tet_begin[ti.body] && t < L.body_tet_begin[ti.body + 1]);
      if (L.slot_tet[gs] < 0) ++primary[t];
      ++copies[t];
      bool owns = false;
      for (int c = 0; c < 4; ++c) {
        const int32_t li = L.slot_conn[4 * gs + c];
        CHECK(li < ti.n_owned + ti.n_halo);
        const int32_t g = li < ti.n_owned ? ti.node_begin + li : L.halo[ti.halo_begin + li - ti.n_owned];
        CHECK(L.node_of_internal[g] == tets[4 * t + c]);
        owns |= li < ti.n_owned;
      }
      CHECK(owns);
    }
    // the slot order fills each window of 64 with slots whose owned corners
    // differ; 64 slots mark at most 256 owned corners, so a tile with that
    // many owned nodes fills its first window without repeats on these meshes
    if (ti.n_slots >= 64 && ti.n_owned >= 64 * 4) {
      std::set<int32_t> owned;
      for (int32_t s = 0; s < 64; ++s)
        for (int c = 0; c < 4; ++c) {
          const int32_t li = L.slot_conn[4 * (ti.slot_begin + s) + c];
          if (li < ti.n_owned) CHECK(owned.insert(li).second);
        }
    }
  }
  CHECK(next_node == N);
  CHECK(next_slot == (int64_t)L.slot_tet.size());
  std::vector<int32_t> tile_of(N, -1);
  for (size_t k = 0; k < L.tiles.size(); ++k)
    for (int32_t i = 0; i < L.tiles[k].n_owned; ++i) tile_of[L.tiles[k].node_begin + i] = (int32_t)k;
  for (int64_t t = 0; t < T; ++t) {
    CHECK(primary[t] == 1);
    std::set<int32_t> tiles_of_corners;
    for (int c = 0; c < 4; ++c) tiles_of_corners.insert(tile_of[L.internal_of_node[tets[4 * t + c]]]);
    CHECK(copies[t] == (int)tiles_of_corners.size());
  }
}

int main() {
  unsigned seed = 12345u;
  // batched: 40 uneven bodies
  std::vector<double> x;
  std::vector<int32_t> t;
  std::vector<int64_t> no{0}, to{0};
  for (int b = 0; b < 40; ++b) {
    lattice(2 + b % 5, 3 + (b * 7) % 6, 2 + (b * 3) % 7, 0.01, seed, x, t);
    no.push_back((int64_t)x.size() / 3);
    to.push_back((int64_t)t.size() / 4);
  }
  const int64_t N = no.back(), T = to.back(), B = 40;
  Layout ref;
  CHECK(otfem::build_layout(N, T, x.data(), t.data(), B, no.data(), to.data(), 48, 16384, ref, 1).empty());
  structure(ref, t);
  for (int th : {3, 8}) {
    Layout L;
    CHECK(otfem::build_layout(N, T, x.data(), t.data(), B, no.data(), to.data(), 48, 16384, L, th).empty());
    CHECK(same(ref, L));
  }
  // one body large enough for the CU-multiple default tiling (256 tiles)
  std::vector<double> x1;
  std::vector<int32_t> t1;
  lattice(40, 40, 40, 0.01, seed, x1, t1);
  const int64_t N1 = (int64_t)x1.size() / 3, T1 = (int64_t)t1.size() / 4;
  Layout one_ref, one;
  CHECK(otfem::build_layout(N1, T1, x1.data(), t1.data(), 1, nullptr, nullptr, 0, 16384, one_ref, 1).empty());
  CHECK(one_ref.tiles.size() == 256);
  CHECK(otfem::build_layout(N1, T1, x1.data(), t1.data(), 1, nullptr, nullptr, 0, 16384, one, 8).empty());
  CHECK(same(one_ref, one));
  structure(one, t1);
  // slots spread over the LDS banks (2.2 here; the order that only keeps
  // owned corners apart gives 2.8)
  const double spread = bank_spread(one);
  std::printf("bank spread of the 40^3 body: %.3f\n", spread);
  CHECK(spread > 1.0 && spread < 2.5);
  // a small single body: min(248, N / 16) tiles (16 owned nodes on average;
  // the bisection balances slot work, so interior tiles hold fewer)
  {
    std::vector<double> xs;
    std::vector<int32_t> ts;
    lattice(10, 10, 10, 0.01, seed, xs, ts);
    const int64_t Ns = (int64_t)xs.size() / 3, Ts = (int64_t)ts.size() / 4;
    Layout small;
    CHECK(otfem::build_layout(Ns, Ts, xs.data(), ts.data(), 1, nullptr, nullptr, 0, 16384, small, 8).empty());
    CHECK((int64_t)small.tiles.size() == std::min<int64_t>(248, Ns / 16));
    int32_t lo = 1 << 30, hi = 0;
    for (const TileInfo &ti : small.tiles) {
      CHECK(ti.n_owned >= 1);
      lo = std::min(lo, ti.n_owned);
      hi = std::max(hi, ti.n_owned);
    }
    std::printf("small body: %lld nodes, %zu tiles of %d..%d owned nodes\n", (long long)Ns,
                small.tiles.size(), lo, hi);
    structure(small, ts);
  }
  // a small batch: smaller tiles, at most 248 of them
  {
    std::vector<double> xb;
    std::vector<int32_t> tb;
    std::vector<int64_t> nob{0}, tob{0};
    for (int b = 0; b < 8; ++b) {
      lattice(6 + b % 3, 7, 5, 0.01, seed, xb, tb);
      nob.push_back((int64_t)xb.size() / 3);
      tob.push_back((int64_t)tb.size() / 4);
    }
    Layout lb;
    CHECK(otfem::build_layout(nob.back(), tob.back(), xb.data(), tb.data(), 8, nob.data(), tob.data(),
                              0, 16384, lb, 8).empty());
    CHECK(lb.tiles.size() > 100 && lb.tiles.size() <= 248);
    std::printf("small batch: %lld nodes, %zu tiles\n", (long long)nob.back(), lb.tiles.size());
    structure(lb, tb);
  }
  // errors: out of range wins over a cross-body index wherever they are
  {
    std::vector<int32_t> bad = t;
    bad[0] = (int32_t)no[1];  // body 0 reaches into body 1
    Layout L;
    CHECK(otfem::build_layout(N, T, x.data(), bad.data(), B, no.data(), to.data(), 48, 16384, L, 8) ==
          "a tetrahedron references a node of another body");
    bad[4 * (T - 1) + 2] = (int32_t)N;  // last tet out of range
    CHECK(otfem::build_layout(N, T, x.data(), bad.data(), B, no.data(), to.data(), 48, 16384, L, 8) ==
          "tetrahedron node index out of range");
  }
  if (failures) {
    std::fprintf(stderr, "layout_check: %d failure(s)\n", failures);
    return 1;
  }
  std::printf("layout_check ok: %lld bodies, %zu tiles, %zu slots; one body %zu tiles, %zu slots\n",
              (long long)B, ref.tiles.size(), ref.slot_tet.size(), one.tiles.size(), one.slot_tet.size());
  return 0;
}
