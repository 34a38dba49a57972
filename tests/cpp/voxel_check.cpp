// Runs t4mesh::generate_blocks on a voxel grid read from a file and prints the
// mesh.  Used by tests/test_mesh_io.py.
//
//   voxel_check GRID.txt
// GRID.txt: I J K, min x y z, max x y z, then I*J*K 0/1 values in storage
// order (i fastest).  Output: "N T", N lines "x y z", T lines "a b c d".
#include <otfem/fem.hpp>

#include <cstdio>
#include <fstream>
#include <vector>

typedef OpenTissue::math::BasicMathTypes<double, size_t> math_types;
typedef OpenTissue::fem::Mesh<math_types> mesh_type;

struct VoxelGrid {  // the part of the reference's grid::Grid the generator uses
  size_t n[3];
  double lo[3], hi[3];
  std::vector<int> v;
  size_t I() const { return n[0]; }
  size_t J() const { return n[1]; }
  size_t K() const { return n[2]; }
  double dx() const { return (hi[0] - lo[0]) / (n[0] - 1); }
  double dy() const { return (hi[1] - lo[1]) / (n[1] - 1); }
  double dz() const { return (hi[2] - lo[2]) / (n[2] - 1); }
  double min_coord(int d) const { return lo[d]; }
  int operator()(size_t p) const { return v[p]; }
};

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  std::ifstream in(argv[1]);
  VoxelGrid g;
  in >> g.n[0] >> g.n[1] >> g.n[2] >> g.lo[0] >> g.lo[1] >> g.lo[2] >> g.hi[0] >> g.hi[1] >> g.hi[2];
  g.v.resize(g.n[0] * g.n[1] * g.n[2]);
  for (auto &x : g.v) in >> x;
  if (!in) return 1;
  mesh_type mesh;
  OpenTissue::t4mesh::generate_blocks(g, mesh);
  std::printf("%zu %zu\n", mesh.size_nodes(), mesh.size_tetrahedra());
  for (auto n = mesh.node_begin(); n != mesh.node_end(); ++n)
    std::printf("%.17g %.17g %.17g\n", n->m_coord(0), n->m_coord(1), n->m_coord(2));
  for (auto t = mesh.tetrahedron_begin(); t != mesh.tetrahedron_end(); ++t)
    std::printf("%zu %zu %zu %zu\n", t->i()->idx(), t->j()->idx(), t->k()->idx(), t->m()->idx());
  return 0;
}
