// Caller code written against the reference's fem API (as in a program that
// includes <OpenTissue/dynamics/fem/fem.h>), compiled against <otfem/fem.hpp>.
// Usage: dropin_c1 I J K steps mode out.bin
//   builds generate_blocks(I, J, K, 0.01, ...), fixes nodes at z = 0,
//   E 5e5, nu 0.33, rho 1000, gravity m g, and steps fem::simulate(mesh,
//   0.01, true) with the reference's settings (mode 0), the relative
//   residual settings (mode 1), or those with plasticity (yield 1e-4, creep
//   20, max 0.05) and 50 g (mode 2).  Writes m_coord and m_velocity (3N
//   doubles each) before the first step and after every step, then the
//   export_K triplet count, then m_f_external (3N doubles) after every step:
//   plasticity adds its forces into it, as in the reference.
#include <otfem/fem.hpp>

#include <cstdio>
#include <cstdlib>
#include <vector>

struct Triplets {  // the three uBLAS calls fem::export_K makes
  size_t rows = 0, cols = 0;
  std::vector<size_t> r, c;
  std::vector<double> v;
  void resize(size_t a, size_t b, bool) { rows = a; cols = b; }
  void clear() { r.clear(); c.clear(); v.clear(); }
  void push_back(size_t i, size_t j, double x) { r.push_back(i); c.push_back(j); v.push_back(x); }
};

int main(int argc, char **argv) {
  if (argc < 7) {
    std::fprintf(stderr, "usage: %s I J K steps mode out.bin\n", argv[0]);
    return 2;
  }
  typedef OpenTissue::math::BasicMathTypes<double, size_t> math_types;
  typedef OpenTissue::fem::Mesh<math_types> mesh_type;
  typedef math_types::vector3_type vector3_type;
  unsigned int I = std::atoi(argv[1]), J = std::atoi(argv[2]), K = std::atoi(argv[3]);
  int steps = std::atoi(argv[4]);
  int mode = std::atoi(argv[5]);
  FILE *f = std::fopen(argv[6], "wb");
  if (!f) return 3;
  mesh_type mesh;
  OpenTissue::t4mesh::generate_blocks(I, J, K, 0.01, 0.01, 0.01, mesh);
  for (auto n = mesh.node_begin(); n != mesh.node_end(); ++n) {
    n->m_model_coord = n->m_coord;
    n->m_fixed = (n->m_coord(2) == 0.0);
    n->m_velocity.clear();
  }
  if (mode == 2)
    OpenTissue::fem::init(mesh, 5e5, 0.33, 1000.0, 1e-4, 20.0, 0.05);
  else
    OpenTissue::fem::init(mesh, 5e5, 0.33, 1000.0, 0.0, 0.0, 0.0);
  const double g = mode == 2 ? 50.0 * 9.81 : 9.81;
  std::vector<double> fext_log;
  if (mode >= 1) {
    otfem_params p;
    otfem_default_params(&p, 1);
    OpenTissue::fem::set_params(mesh, p);
  }
  auto dump = [&]() {
    for (auto n = mesh.node_begin(); n != mesh.node_end(); ++n)
      std::fwrite(&n->m_coord.x, sizeof(double), 3, f);
    for (auto n = mesh.node_begin(); n != mesh.node_end(); ++n)
      std::fwrite(&n->m_velocity.x, sizeof(double), 3, f);
  };
  dump();
  for (int s = 0; s < steps; ++s) {
    for (auto n = mesh.node_begin(); n != mesh.node_end(); ++n)
      n->m_f_external = n->m_fixed ? vector3_type(0, 0, 0) : vector3_type(0, 0, -g * n->m_mass);
    OpenTissue::fem::simulate(mesh, 0.01, true);
    dump();
    for (auto n = mesh.node_begin(); n != mesh.node_end(); ++n)
      fext_log.insert(fext_log.end(), {n->m_f_external(0), n->m_f_external(1), n->m_f_external(2)});
  }
  Triplets K3;
  OpenTissue::fem::export_K(mesh, K3);
  double nk = (double)K3.v.size();
  std::fwrite(&nk, sizeof(double), 1, f);
  std::fwrite(fext_log.data(), sizeof(double), fext_log.size(), f);
  std::fclose(f);
  std::printf("dropin_c1: %zu nodes, %zu tets, %d steps, export_K %zu entries\n", mesh.size_nodes(),
              mesh.size_tetrahedra(), steps, K3.v.size());
  return 0;
}
