// Reads a T4MESH XML file (and optionally a TetGen file pair) with
// <otfem/t4mesh_io.hpp> into the drop-in fem mesh type, prints counts and
// checksums, and writes the XML mesh back out.  Used by tests/test_mesh_io.py.
//
//   mesh_io_check IN.xml OUT.xml [TETGEN_BASENAME]
#include <otfem/fem.hpp>
#include <otfem/t4mesh_io.hpp>

#include <cstdio>

typedef OpenTissue::math::BasicMathTypes<double, size_t> math_types;
typedef OpenTissue::fem::Mesh<math_types> mesh_type;

static void report(const char *what, mesh_type &mesh) {
  double sc = 0.0, wc = 0.0;
  for (auto n = mesh.node_begin(); n != mesh.node_end(); ++n)
    for (int d = 0; d < 3; ++d) {
      sc += n->m_coord(d);
      wc += (double)(n->idx() + 1) * (d + 1) * n->m_coord(d);
    }
  unsigned long long si = 0, wi = 0;
  for (auto t = mesh.tetrahedron_begin(); t != mesh.tetrahedron_end(); ++t)
    for (int c = 0; c < 4; ++c) {
      si += t->node(c)->idx();
      wi += (t->idx() + 1) * (unsigned long long)(c + 1) * t->node(c)->idx();
    }
  std::printf("%s nodes %zu tets %zu coord_sum %.17g coord_wsum %.17g index_sum %llu index_wsum %llu\n",
              what, mesh.size_nodes(), mesh.size_tetrahedra(), sc, wc, si, wi);
}

int main(int argc, char **argv) {
  if (argc < 3) return 2;
  mesh_type mesh;
  if (!OpenTissue::t4mesh::xml_read(argv[1], mesh)) return 1;
  report("xml", mesh);
  if (!OpenTissue::t4mesh::xml_write(argv[2], mesh)) return 1;
  if (argc > 3) {
    mesh_type tg;
    if (!OpenTissue::t4mesh::tetgen_read(argv[3], tg)) return 1;
    report("tetgen", tg);
  }
  return 0;
}
