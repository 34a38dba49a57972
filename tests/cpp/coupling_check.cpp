// Binds points to a T4MESH XML mesh with <otfem/mesh_coupling.hpp>
// (bind_surface on a minimal surface-mesh type), then moves the mesh nodes by
// an affine map and runs update_surface.  Used by tests/test_coupling.py.
//
//   coupling_check MESH.xml POINTS.txt OUT.txt
// POINTS.txt: n, then n lines "x y z".  OUT.txt: n lines
// "tet w1 w2 w3 x' y' z'" (x' after the update; %.17g).
#include <otfem/fem.hpp>
#include <otfem/mesh_coupling.hpp>
#include <otfem/t4mesh_io.hpp>

#include <cstdio>
#include <fstream>
#include <vector>

typedef OpenTissue::math::BasicMathTypes<double, size_t> math_types;
typedef OpenTissue::fem::Mesh<math_types> mesh_type;
typedef OpenTissue::math::Vector3<double> vec3;

struct SurfaceMesh {  // the vertex interface bind_surface / update_surface use
  struct Handle {
    size_t i;
    size_t get_idx() const { return i; }
  };
  struct Vertex {
    Handle h;
    vec3 m_coord;
    Handle get_handle() const { return h; }
  };
  std::vector<Vertex> v;
  std::vector<Vertex>::iterator vertex_begin() { return v.begin(); }
  std::vector<Vertex>::iterator vertex_end() { return v.end(); }
  size_t size_vertices() const { return v.size(); }
};

int main(int argc, char **argv) {
  if (argc < 4) return 2;
  mesh_type V;
  if (!OpenTissue::t4mesh::xml_read(argv[1], V)) return 1;
  std::ifstream in(argv[2]);
  size_t n = 0;
  in >> n;
  SurfaceMesh M;
  M.v.resize(n);
  for (size_t i = 0; i < n; ++i) {
    M.v[i].h.i = i;
    in >> M.v[i].m_coord(0) >> M.v[i].m_coord(1) >> M.v[i].m_coord(2);
  }
  if (!in) return 1;
  std::vector<vec3> bary;
  std::vector<int> idx;
  OpenTissue::t4mesh::mesh_coupling::bind_surface(M, V, bary, idx);
  {  // the array form gives the same binding on 1 and 4 threads
    std::vector<double> nodes, pts(3 * n);
    std::vector<int32_t> tets;
    for (auto nd = V.node_begin(); nd != V.node_end(); ++nd)
      for (int d = 0; d < 3; ++d) nodes.push_back(nd->m_coord(d));
    for (auto t = V.tetrahedron_begin(); t != V.tetrahedron_end(); ++t) {
      for (int k = 0; k < 4; ++k) tets.push_back((int32_t)t->node(k)->idx());
    }
    for (size_t i = 0; i < n; ++i)
      for (int d = 0; d < 3; ++d) pts[3 * i + d] = M.v[i].m_coord(d);
    std::vector<double> w1(3 * n), w4(3 * n);
    std::vector<int32_t> t1(n), t4(n);
    const int64_t N = (int64_t)nodes.size() / 3, T = (int64_t)tets.size() / 4;
    otfem::coupling::bind_points(N, nodes.data(), T, tets.data(), (int64_t)n, pts.data(), w1.data(), t1.data(), 1);
    otfem::coupling::bind_points(N, nodes.data(), T, tets.data(), (int64_t)n, pts.data(), w4.data(), t4.data(), 4);
    if (w1 != w4 || t1 != t4) {
      std::fprintf(stderr, "bind_points differs between 1 and 4 threads\n");
      return 3;
    }
  }
  // affine motion x -> A x + t of every node
  const double A[3][3] = {{1.1, 0.2, -0.1}, {0.05, 0.9, 0.3}, {-0.2, 0.1, 1.3}};
  const double t[3] = {0.3, -0.7, 0.25};
  for (auto nd = V.node_begin(); nd != V.node_end(); ++nd) {
    const vec3 p = nd->m_coord;
    for (int r = 0; r < 3; ++r) nd->m_coord(r) = A[r][0] * p(0) + A[r][1] * p(1) + A[r][2] * p(2) + t[r];
  }
  OpenTissue::t4mesh::mesh_coupling::update_surface(M, V, bary, idx);
  FILE *f = std::fopen(argv[3], "w");
  if (!f) return 1;
  for (size_t i = 0; i < n; ++i)
    std::fprintf(f, "%d %.17g %.17g %.17g %.17g %.17g %.17g\n", idx[i], bary[i](0), bary[i](1),
                 bary[i](2), M.v[i].m_coord(0), M.v[i].m_coord(1), M.v[i].m_coord(2));
  std::fclose(f);
  return 0;
}
