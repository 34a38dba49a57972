// Surface-to-volume mesh coupling (skinning a render surface onto the
// tetrahedral mesh the FEM step deforms).
//
// Reference interfaces mirrored (paths relative to the OpenTissue tree):
//   t4mesh::mesh_coupling::bind_surface    core/containers/t4mesh/util/t4mesh_mesh_coupling.h:100-200
//   t4mesh::mesh_coupling::update_surface  core/containers/t4mesh/util/t4mesh_mesh_coupling.h:202-238
//   barycentric coordinates of a point     core/geometry/geometry_barycentric.h:339-372
//
// bind_surface finds for every surface vertex a tetrahedron that contains it
// -- all four barycentric coordinates in (-1e-4, 1 + 1e-4), the reference's
// tolerance -- and stores the coordinates w1, w2, w3 of corners j, k, m and
// the tetrahedron's index; update_surface places each vertex at
// w0 p_i + w1 p_j + w2 p_k + w3 p_m (w0 = 1 - w1 - w2 - w3) of the current node
// positions.  The reference takes the first containing tetrahedron its spatial
// hash reports; this one takes the containing tetrahedron of lowest index (a
// uniform grid over the tetrahedra's bounding boxes), so the binding is
// reproducible.  A vertex inside no tetrahedron gets index -1 here (the
// reference leaves its entries default-initialised, which binds it to
// tetrahedron 0) and update_surface leaves it where it is.
//
// The array form (coupling::bind_points) also backs otfem_surface_bind in the
// C ABI; the device update is otfem_surface_create / otfem_surface_update.
#ifndef OTFEM_MESH_COUPLING_HPP
#define OTFEM_MESH_COUPLING_HPP

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <thread>
#include <vector>

namespace otfem {
namespace coupling {

// Barycentric coordinates of p in the tetrahedron (x1, x2, x3, x4):
// w = P^{-1} (p - x4) with P = [x1 - x4, x2 - x4, x3 - x4] (columns),
// w4 = 1 - w1 - w2 - w3.  Returns false for a degenerate tetrahedron.
inline bool barycentric(const double *x1, const double *x2, const double *x3, const double *x4,
                        const double *p, double w[4]) {
  double P[3][3], r[3];
  for (int d = 0; d < 3; ++d) {
    P[d][0] = x1[d] - x4[d];
    P[d][1] = x2[d] - x4[d];
    P[d][2] = x3[d] - x4[d];
    r[d] = p[d] - x4[d];
  }
  const double c00 = P[1][1] * P[2][2] - P[1][2] * P[2][1];
  const double c01 = P[1][2] * P[2][0] - P[1][0] * P[2][2];
  const double c02 = P[1][0] * P[2][1] - P[1][1] * P[2][0];
  const double det = P[0][0] * c00 + P[0][1] * c01 + P[0][2] * c02;
  if (!(det != 0.0) || !std::isfinite(det)) return false;
  const double id = 1.0 / det;
  // rows of P^{-1} = adj(P) / det
  const double inv[3][3] = {
      {c00 * id, (P[0][2] * P[2][1] - P[0][1] * P[2][2]) * id, (P[0][1] * P[1][2] - P[0][2] * P[1][1]) * id},
      {c01 * id, (P[0][0] * P[2][2] - P[0][2] * P[2][0]) * id, (P[0][2] * P[1][0] - P[0][0] * P[1][2]) * id},
      {c02 * id, (P[0][1] * P[2][0] - P[0][0] * P[2][1]) * id, (P[0][0] * P[1][1] - P[0][1] * P[1][0]) * id}};
  for (int i = 0; i < 3; ++i) w[i] = inv[i][0] * r[0] + inv[i][1] * r[1] + inv[i][2] * r[2];
  w[3] = 1.0 - w[0] - w[1] - w[2];
  return true;
}

// For each of the n points pts[3n], the lowest-index tetrahedron (tets[4T]
// over nodes[3N]) containing it within the tolerance, and the barycentric
// coordinates of corners 1, 2, 3 (w123[3n]); tet[i] = -1 when none does.
// The point queries run on `threads` threads (each point's result is its
// own, so it does not depend on the thread count).
inline void bind_points(int64_t N, const double *nodes, int64_t T, const int32_t *tets, int64_t n,
                        const double *pts, double *w123, int32_t *tet, int threads = 1) {
  const double lower = -1e-4, upper = 1.0 + 1e-4;
  for (int64_t i = 0; i < n; ++i) {
    tet[i] = -1;
    w123[3 * i] = w123[3 * i + 1] = w123[3 * i + 2] = 0.0;
  }
  if (T <= 0 || N <= 0 || n <= 0) return;
  // uniform grid: cell edge = mean tetrahedron bounding-box edge
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  double mean_edge = 0.0;
  std::vector<double> box(6 * (size_t)T);
  for (int64_t t = 0; t < T; ++t) {
    double *b = &box[6 * (size_t)t];
    for (int d = 0; d < 3; ++d) {
      b[d] = 1e300;
      b[3 + d] = -1e300;
    }
    for (int k = 0; k < 4; ++k)
      for (int d = 0; d < 3; ++d) {
        const double v = nodes[3 * (int64_t)tets[4 * t + k] + d];
        b[d] = std::min(b[d], v);
        b[3 + d] = std::max(b[3 + d], v);
      }
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::min(lo[d], b[d]);
      hi[d] = std::max(hi[d], b[3 + d]);
      mean_edge += (b[3 + d] - b[d]) / 3.0;
    }
  }
  mean_edge /= (double)T;
  double cell = mean_edge > 0.0 ? mean_edge : 1.0;
  int64_t dim[3];
  for (int d = 0; d < 3; ++d) {
    // at most ~2^21 cells per axis and ~64 M in all
    double ext = hi[d] - lo[d];
    dim[d] = std::max<int64_t>(1, std::min<int64_t>((int64_t)(ext / cell) + 1, 1 << 21));
  }
  while ((double)dim[0] * dim[1] * dim[2] > 64.0 * 1024 * 1024) {
    cell *= 2.0;
    for (int d = 0; d < 3; ++d)
      dim[d] = std::max<int64_t>(1, (int64_t)((hi[d] - lo[d]) / cell) + 1);
  }
  auto cell_of = [&](int d, double v) {
    int64_t c = (int64_t)std::floor((v - lo[d]) / cell);
    return std::min<int64_t>(std::max<int64_t>(c, 0), dim[d] - 1);
  };
  // cell -> tetrahedra (CSR, each list in increasing tetrahedron order)
  const int64_t ncell = dim[0] * dim[1] * dim[2];
  std::vector<int64_t> start(ncell + 1, 0);
  std::vector<int64_t> span(6 * (size_t)T);  // cell ranges of each box
  for (int64_t t = 0; t < T; ++t) {
    const double *b = &box[6 * (size_t)t];
    int64_t *sp = &span[6 * (size_t)t];
    for (int d = 0; d < 3; ++d) {
      const double pad = 1e-4 * (b[3 + d] - b[d]) + 1e-12 * cell;
      sp[d] = cell_of(d, b[d] - pad);
      sp[3 + d] = cell_of(d, b[3 + d] + pad);
    }
    for (int64_t x = sp[0]; x <= sp[3]; ++x)
      for (int64_t y = sp[1]; y <= sp[4]; ++y)
        for (int64_t z = sp[2]; z <= sp[5]; ++z) ++start[(x * dim[1] + y) * dim[2] + z + 1];
  }
  for (int64_t c = 0; c < ncell; ++c) start[c + 1] += start[c];
  std::vector<int64_t> fill(start.begin(), start.end() - 1);
  std::vector<int32_t> list((size_t)start[ncell]);
  for (int64_t t = 0; t < T; ++t) {
    const int64_t *sp = &span[6 * (size_t)t];
    for (int64_t x = sp[0]; x <= sp[3]; ++x)
      for (int64_t y = sp[1]; y <= sp[4]; ++y)
        for (int64_t z = sp[2]; z <= sp[5]; ++z)
          list[(size_t)fill[(x * dim[1] + y) * dim[2] + z]++] = (int32_t)t;
  }
  auto query = [&](int64_t i) {
    const double *p = pts + 3 * i;
    if (!std::isfinite(p[0]) || !std::isfinite(p[1]) || !std::isfinite(p[2])) return;
    bool inside_box = true;
    for (int d = 0; d < 3; ++d) {
      const double pad = 1e-4 * (hi[d] - lo[d]) + 1e-12;
      inside_box = inside_box && p[d] >= lo[d] - pad && p[d] <= hi[d] + pad;
    }
    if (!inside_box) return;
    const int64_t c = (cell_of(0, p[0]) * dim[1] + cell_of(1, p[1])) * dim[2] + cell_of(2, p[2]);
    for (int64_t k = start[c]; k < start[c + 1]; ++k) {
      const int32_t t = list[(size_t)k];
      const int32_t *tt = tets + 4 * (int64_t)t;
      double w[4];
      if (!barycentric(nodes + 3 * (int64_t)tt[0], nodes + 3 * (int64_t)tt[1],
                       nodes + 3 * (int64_t)tt[2], nodes + 3 * (int64_t)tt[3], p, w))
        continue;
      if (w[0] > lower && w[0] < upper && w[1] > lower && w[1] < upper && w[2] > lower &&
          w[2] < upper && w[3] > lower && w[3] < upper) {
        // the reference's barycentric_algebraic(p_i, p_j, p_k, p_m) order:
        // stored are the coordinates of corners j, k, m
        w123[3 * i] = w[1];
        w123[3 * i + 1] = w[2];
        w123[3 * i + 2] = w[3];
        tet[i] = t;
        return;
      }
    }
  };
  const int64_t chunk = 4096;
  const int64_t nchunks = (n + chunk - 1) / chunk;
  const int64_t nt = std::max<int64_t>(1, std::min<int64_t>(threads, nchunks));
  std::atomic<int64_t> next{0};
  auto work = [&] {
    for (int64_t k; (k = next.fetch_add(1, std::memory_order_relaxed)) < nchunks;)
      for (int64_t i = k * chunk, e = std::min(n, i + chunk); i < e; ++i) query(i);
  };
  std::vector<std::thread> pool;
  for (int64_t t = 1; t < nt; ++t) pool.emplace_back(work);
  work();
  for (auto &th : pool) th.join();
}

}  // namespace coupling

namespace t4mesh {
namespace mesh_coupling {

// t4mesh_mesh_coupling.h:100-200 for a surface mesh with vertex iteration
// (vertex_begin/end, size_vertices, v->m_coord, v->get_handle().get_idx())
// and a tetrahedral mesh with the T4Mesh interface of <otfem/fem.hpp>.
// `barycentric` gets one 3-vector and `bind_indices` one index per vertex
// (-1: the vertex lies in no tetrahedron).
template <typename surface_mesh, typename volume_mesh, typename point_container,
          typename index_container>
void bind_surface(surface_mesh &M, volume_mesh &V, point_container &barycentric,
                  index_container &bind_indices) {
  const int64_t N = (int64_t)V.size_nodes(), T = (int64_t)V.size_tetrahedra();
  std::vector<double> nodes(3 * (size_t)N);
  for (auto nd = V.node_begin(); nd != V.node_end(); ++nd)
    for (int d = 0; d < 3; ++d) nodes[3 * nd->idx() + d] = nd->m_coord(d);
  std::vector<int32_t> tets(4 * (size_t)T);
  for (auto t = V.tetrahedron_begin(); t != V.tetrahedron_end(); ++t)
    for (int c = 0; c < 4; ++c) tets[4 * t->idx() + c] = (int32_t)t->node(c)->idx();
  const int64_t n = (int64_t)M.size_vertices();
  std::vector<double> pts(3 * (size_t)n), w(3 * (size_t)n);
  std::vector<int32_t> idx((size_t)n);
  for (auto v = M.vertex_begin(); v != M.vertex_end(); ++v) {
    const size_t i = (size_t)v->get_handle().get_idx();
    for (int d = 0; d < 3; ++d) pts[3 * i + d] = v->m_coord(d);
  }
  coupling::bind_points(N, nodes.data(), T, tets.data(), n, pts.data(), w.data(), idx.data());
  barycentric.resize((size_t)n);
  bind_indices.resize((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    for (int d = 0; d < 3; ++d) barycentric[(size_t)i](d) = w[3 * (size_t)i + d];
    bind_indices[(size_t)i] = idx[(size_t)i];
  }
}

// t4mesh_mesh_coupling.h:202-238: every bound vertex at the barycentric
// combination of its tetrahedron's current node positions.
template <typename surface_mesh, typename volume_mesh, typename point_container,
          typename index_container>
void update_surface(surface_mesh &M, volume_mesh &V, point_container const &barycentric,
                    index_container const &bind_info) {
  for (auto v = M.vertex_begin(); v != M.vertex_end(); ++v) {
    const size_t i = (size_t)v->get_handle().get_idx();
    const long long t = (long long)bind_info[i];
    if (t < 0) continue;
    auto T = V.tetrahedron((size_t)t);
    const double w1 = barycentric[i](0), w2 = barycentric[i](1), w3 = barycentric[i](2);
    const double w0 = 1.0 - w1 - w2 - w3;
    for (int d = 0; d < 3; ++d)
      v->m_coord(d) = T->i()->m_coord(d) * w0 + T->j()->m_coord(d) * w1 +
                      T->k()->m_coord(d) * w2 + T->m()->m_coord(d) * w3;
  }
}

}  // namespace mesh_coupling
}  // namespace t4mesh
}  // namespace otfem

#endif  // OTFEM_MESH_COUPLING_HPP
