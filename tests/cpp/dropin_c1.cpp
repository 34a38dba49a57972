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
//   plasticity adds its forces into it, as in the reference.  Mode 3 is
//   mode 1 solved to rel-res 1e-10 with the rotation read-back off and a
//   caller that edits the mesh between steps: before every odd step s it
//   shifts free node j's m_coord by 1e-4 sin(j) along x, before step 2 it
//   halves every m_velocity (the recorded states are those before the edits).
// Usage: dropin_c1 --time I J K steps warmup rotations
//   the same block mesh and gravity with the relative-residual settings;
//   after `warmup` untimed steps times `steps` fem::simulate calls (caller
//   work included: m_f_external reset every step, node and rotation
//   transfers; rotations 0 turns the m_Re read-back off), then as many
//   otfem_step calls on the same solver without the mesh's node arrays,
//   and prints one JSON line with both rates and the device kernel time.
#include <otfem/fem.hpp>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

struct Triplets {  // the three uBLAS calls fem::export_K makes
  size_t rows = 0, cols = 0;
  std::vector<size_t> r, c;
  std::vector<double> v;
  void resize(size_t a, size_t b, bool) { rows = a; cols = b; }
  void clear() { r.clear(); c.clear(); v.clear(); }
  void push_back(size_t i, size_t j, double x) { r.push_back(i); c.push_back(j); v.push_back(x); }
};

static int time_steps(int argc, char **argv) {
  if (argc < 8) {
    std::fprintf(stderr, "usage: %s --time I J K steps warmup rotations\n", argv[0]);
    return 2;
  }
  typedef OpenTissue::math::BasicMathTypes<double, size_t> math_types;
  typedef OpenTissue::fem::Mesh<math_types> mesh_type;
  typedef math_types::vector3_type vector3_type;
  typedef std::chrono::steady_clock clk;
  unsigned int I = std::atoi(argv[2]), J = std::atoi(argv[3]), K = std::atoi(argv[4]);
  const int steps = std::atoi(argv[5]), warmup = std::atoi(argv[6]);
  const bool rotations = std::atoi(argv[7]) != 0;
  if (steps < 1 || warmup < 0) return 2;
  mesh_type mesh;
  OpenTissue::t4mesh::generate_blocks(I, J, K, 0.01, 0.01, 0.01, mesh);
  for (auto n = mesh.node_begin(); n != mesh.node_end(); ++n) {
    n->m_model_coord = n->m_coord;
    n->m_fixed = (n->m_coord(2) == 0.0);
    n->m_velocity.clear();
  }
  OpenTissue::fem::init(mesh, 5e5, 0.33, 1000.0, 0.0, 0.0, 0.0);
  otfem_params p;
  otfem_default_params(&p, 1);
  OpenTissue::fem::set_params(mesh, p);
  OpenTissue::fem::set_rotation_readback(mesh, rotations);
  otfem_handle h = mesh.m_solver.get();
  const double g = 9.81;
  auto step = [&]() {
    for (auto n = mesh.node_begin(); n != mesh.node_end(); ++n)
      n->m_f_external = n->m_fixed ? vector3_type(0, 0, 0) : vector3_type(0, 0, -g * n->m_mass);
    OpenTissue::fem::simulate(mesh, 0.01, true);
  };
  for (int s = 0; s < warmup; ++s) step();
  int64_t its = 0;
  otfem_iteration_totals(h, &its, 1);
  otfem_set_timing(h, 1);
  auto t0 = clk::now();
  for (int s = 0; s < steps; ++s) step();
  const double wall_dropin = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
  double ms[4] = {0, 0, 0, 0};
  long long ts = 0;
  otfem_get_timing_detail(h, ms, &ts);
  otfem_iteration_totals(h, &its, 1);
  const int64_t its_dropin = its;
  // the same solver stepped through the C ABI alone: no node arrays cross
  otfem_set_timing(h, 1);
  p.store_rotations = 0;
  otfem_stats st;
  t0 = clk::now();
  for (int s = 0; s < steps; ++s)
    if (otfem_step(h, 0.01, 1, &p, &st) != OTFEM_OK) return 4;
  const double wall_direct = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
  double ms2[4] = {0, 0, 0, 0};
  otfem_get_timing_detail(h, ms2, &ts);
  otfem_iteration_totals(h, &its, 1);
  const double T = (double)mesh.size_tetrahedra();
  std::printf("{\"nodes\": %zu, \"tets\": %zu, \"steps\": %d, \"warmup\": %d, "
              "\"rotations\": %d, \"dropin_ms_per_step\": %.4f, \"dropin_tet_steps_per_s\": %.6g, "
              "\"dropin_kernel_ms_per_step\": %.4f, \"dropin_cg_iterations_per_step\": %.2f, "
              "\"direct_ms_per_step\": %.4f, \"direct_tet_steps_per_s\": %.6g, "
              "\"direct_kernel_ms_per_step\": %.4f, \"direct_cg_iterations_per_step\": %.2f}\n",
              mesh.size_nodes(), mesh.size_tetrahedra(), steps, warmup, rotations ? 1 : 0,
              wall_dropin / steps, T * steps / (wall_dropin * 1e-3),
              (ms[0] + ms[1] + ms[2] + ms[3]) / steps, (double)its_dropin / steps,
              wall_direct / steps, T * steps / (wall_direct * 1e-3),
              (ms2[0] + ms2[1] + ms2[2] + ms2[3]) / steps, (double)its / steps);
  return 0;
}

int main(int argc, char **argv) {
  if (argc >= 2 && std::strcmp(argv[1], "--time") == 0) return time_steps(argc, argv);
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
    if (mode == 3) {  // solved tightly: the edits make steps far from the last one
      p.tol = 1e-10;
      p.cg_max_iterations = 5000;
    }
    OpenTissue::fem::set_params(mesh, p);
  }
  if (mode == 3) OpenTissue::fem::set_rotation_readback(mesh, false);
  auto dump = [&]() {
    for (auto n = mesh.node_begin(); n != mesh.node_end(); ++n)
      std::fwrite(&n->m_coord.x, sizeof(double), 3, f);
    for (auto n = mesh.node_begin(); n != mesh.node_end(); ++n)
      std::fwrite(&n->m_velocity.x, sizeof(double), 3, f);
  };
  dump();
  for (int s = 0; s < steps; ++s) {
    if (mode == 3 && s % 2 == 1) {
      size_t j = 0;
      for (auto n = mesh.node_begin(); n != mesh.node_end(); ++n, ++j)
        if (!n->m_fixed) n->m_coord(0) += 1e-4 * std::sin((double)j);
    }
    if (mode == 3 && s == 2)
      for (auto n = mesh.node_begin(); n != mesh.node_end(); ++n) n->m_velocity *= 0.5;
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
