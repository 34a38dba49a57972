// Drop-in C++ front end for the OpenTissue::fem path, backed by the otfem C ABI.
//
// Replace
//     #include <OpenTissue/dynamics/fem/fem.h>
//     #include <OpenTissue/core/containers/t4mesh/util/t4mesh_block_generator.h>
// with
//     #include <otfem/fem.hpp>
// and link libotfem.so.  Caller code written against the reference keeps
// compiling:
//     typedef OpenTissue::math::BasicMathTypes<double, size_t> math_types;
//     typedef OpenTissue::fem::Mesh<math_types>                mesh_type;
//     OpenTissue::t4mesh::generate_blocks(10, 40, 10, h, h, h, mesh);
//     for each node: n->m_model_coord = n->m_coord; n->m_fixed = ...;
//     OpenTissue::fem::init(mesh, E, nu, rho, c_yield, c_creep, c_max);
//     loop { set n->m_f_external; OpenTissue::fem::simulate(mesh, dt, true); }
//
// Reference interfaces mirrored (paths relative to the OpenTissue tree):
//   fem::Mesh                      dynamics/fem/fem_mesh.h:21-34
//   node / tet trait fields        dynamics/fem/fem_node_traits.h:37-52,
//                                  dynamics/fem/fem_tetrahedron_traits.h:29-35
//   T4Mesh insert / node / iterate core/containers/t4mesh/t4mesh_t4mesh.h:93-252
//   fem::init                      dynamics/fem/fem_init.h:45-65
//   fem::simulate                  dynamics/fem/fem_simulate.h:49-113
//   fem::update_original_coord     dynamics/fem/fem_update_original_coord.h:28-33
//   fem::export_K                  dynamics/fem/fem_export_k.h:23-68
//   t4mesh::generate_blocks        core/containers/t4mesh/util/t4mesh_block_generator.h:32-110
//                                  (voxel grid: :119-243)
//
// Each fem::simulate call marshals the caller's node fields (m_coord,
// m_velocity, m_f_external, m_fixed) to the device, runs one step there, and
// reads m_coord, m_velocity, m_f_external and m_Re back, so the host-side
// mesh behaves as the reference's does.  Callers that keep state on the device
// use the C ABI (<otfem/otfem.h>) directly.
#ifndef OTFEM_FEM_HPP
#define OTFEM_FEM_HPP

#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "otfem/otfem.h"

namespace otfem {

namespace math {

template <typename T>
class Vector3 {
 public:
  typedef T value_type;
  T x, y, z;
  Vector3() : x(0), y(0), z(0) {}  // zero-initialised (math_vector3.h:63-67)
  Vector3(T a, T b, T c) : x(a), y(b), z(c) {}
  T &operator()(size_t i) { return i == 0 ? x : (i == 1 ? y : z); }
  const T &operator()(size_t i) const { return i == 0 ? x : (i == 1 ? y : z); }
  T &operator[](size_t i) { return (*this)(i); }
  const T &operator[](size_t i) const { return (*this)(i); }
  void clear() { x = y = z = 0; }
  Vector3 &operator+=(const Vector3 &v) { x += v.x; y += v.y; z += v.z; return *this; }
  Vector3 &operator-=(const Vector3 &v) { x -= v.x; y -= v.y; z -= v.z; return *this; }
  Vector3 &operator*=(T s) { x *= s; y *= s; z *= s; return *this; }
  Vector3 operator+(const Vector3 &v) const { return Vector3(x + v.x, y + v.y, z + v.z); }
  Vector3 operator-(const Vector3 &v) const { return Vector3(x - v.x, y - v.y, z - v.z); }
  Vector3 operator-() const { return Vector3(-x, -y, -z); }
  T operator*(const Vector3 &v) const { return x * v.x + y * v.y + z * v.z; }
  Vector3 operator%(const Vector3 &v) const {
    return Vector3(y * v.z - v.y * z, v.x * z - x * v.z, x * v.y - v.x * y);
  }
};
template <typename T>
Vector3<T> operator*(const Vector3<T> &v, T s) { return Vector3<T>(v.x * s, v.y * s, v.z * s); }
template <typename T>
Vector3<T> operator*(T s, const Vector3<T> &v) { return Vector3<T>(v.x * s, v.y * s, v.z * s); }

template <typename T>
class Matrix3x3 {
 public:
  typedef T value_type;
  typedef Vector3<T> vector3_type;
  Vector3<T> r0, r1, r2;
  Matrix3x3() {}
  Matrix3x3(T a00, T a01, T a02, T a10, T a11, T a12, T a20, T a21, T a22)
      : r0(a00, a01, a02), r1(a10, a11, a12), r2(a20, a21, a22) {}
  T &operator()(size_t i, size_t j) { return (i == 0 ? r0 : (i == 1 ? r1 : r2))(j); }
  const T &operator()(size_t i, size_t j) const { return (i == 0 ? r0 : (i == 1 ? r1 : r2))(j); }
  Vector3<T> operator*(const Vector3<T> &v) const { return Vector3<T>(r0 * v, r1 * v, r2 * v); }
  size_t size1() const { return 3u; }
  size_t size2() const { return 3u; }
};

template <typename T>
Matrix3x3<T> diag(T d) { return Matrix3x3<T>(d, 0, 0, 0, d, 0, 0, 0, d); }

// math_basic_types.h:26-44
template <typename real, typename index = size_t>
struct BasicMathTypes {
  typedef real real_type;
  typedef index index_type;
  typedef Vector3<real> vector3_type;
  typedef Matrix3x3<real> matrix3x3_type;
};

}  // namespace math

namespace detail {

inline void check(otfem_status st, otfem_handle h, const char *what) {
  if (st == OTFEM_OK) return;
  std::string msg = std::string(what) + ": " + otfem_status_string(st);
  if (h)
    msg += std::string(" (") + otfem_last_error(h) + ")";
  else if (*otfem_create_last_error())
    msg += std::string(" (") + otfem_create_last_error() + ")";
  if (st == OTFEM_INVALID_ARG) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

struct HandleDeleter {
  void operator()(otfem_solver *h) const {
    if (h) otfem_destroy(h);
  }
};

}  // namespace detail

namespace t4mesh {

// Minimal T4Mesh (t4mesh_t4mesh.h:49-506): compact 0..N-1 node and 0..T-1
// tetrahedron indices, insert(), node(idx) with a range check, iterators.
template <typename math_types, typename node_base, typename tet_base>
class T4Mesh {
 public:
  typedef typename math_types::real_type real_type;
  typedef typename math_types::vector3_type vector3_type;
  typedef typename math_types::matrix3x3_type matrix3x3_type;

  class Node : public node_base {
   public:
    size_t m_idx = 0;
    vector3_type m_coord;
    size_t idx() const { return m_idx; }
  };
  class Tetrahedron;
  typedef Node node_type;
  typedef Tetrahedron tetrahedron_type;
  typedef typename std::vector<Node>::iterator node_iterator;
  typedef typename std::vector<Tetrahedron>::iterator tetrahedron_iterator;

  class Tetrahedron : public tet_base {
   public:
    size_t m_idx = 0;
    size_t m_nodes[4] = {0, 0, 0, 0};
    T4Mesh *m_owner = nullptr;
    size_t idx() const { return m_idx; }
    node_iterator node(int n) const { return m_owner->node(m_nodes[n]); }
    node_iterator i() const { return node(0); }
    node_iterator j() const { return node(1); }
    node_iterator k() const { return node(2); }
    node_iterator m() const { return node(3); }
  };

  T4Mesh() {}
  T4Mesh(const T4Mesh &) = delete;
  T4Mesh &operator=(const T4Mesh &) = delete;

  void clear() {
    m_nodes.clear();
    m_tets.clear();
    ++m_topology_version;
  }
  node_iterator insert() {
    Node n;
    n.m_idx = m_nodes.size();
    m_nodes.push_back(n);
    ++m_topology_version;
    return m_nodes.end() - 1;
  }
  tetrahedron_iterator insert(size_t a, size_t b, size_t c, size_t d) {
    if (a >= m_nodes.size() || b >= m_nodes.size() || c >= m_nodes.size() || d >= m_nodes.size())
      throw std::out_of_range("T4Mesh::insert(): node index out of range");
    Tetrahedron t;
    t.m_idx = m_tets.size();
    t.m_nodes[0] = a; t.m_nodes[1] = b; t.m_nodes[2] = c; t.m_nodes[3] = d;
    t.m_owner = this;
    m_tets.push_back(t);
    ++m_topology_version;
    return m_tets.end() - 1;
  }
  node_iterator node(size_t idx) {
    if (idx >= m_nodes.size()) throw std::out_of_range("T4Mesh::node(): index out of range");
    return m_nodes.begin() + idx;
  }
  tetrahedron_iterator tetrahedron(size_t idx) {
    if (idx >= m_tets.size()) throw std::out_of_range("T4Mesh::tetrahedron(): index out of range");
    return m_tets.begin() + idx;
  }
  node_iterator node_begin() { return m_nodes.begin(); }
  node_iterator node_end() { return m_nodes.end(); }
  tetrahedron_iterator tetrahedron_begin() { return m_tets.begin(); }
  tetrahedron_iterator tetrahedron_end() { return m_tets.end(); }
  size_t size_nodes() const { return m_nodes.size(); }
  size_t size_tetrahedra() const { return m_tets.size(); }

  // device solver bound by fem::init
  std::unique_ptr<otfem_solver, detail::HandleDeleter> m_solver;
  size_t m_solver_version = ~size_t(0);
  size_t m_topology_version = 0;
  otfem_params m_params{};
  bool m_params_set = false;
  // What the device holds, as last written or read back by fem::init /
  // fem::simulate: fem::simulate uploads only what the caller changed since,
  // so a caller that leaves a node array alone pays no transfer for it.
  std::vector<double> m_dev_x, m_dev_v, m_dev_f;
  std::vector<unsigned char> m_dev_fixed;
  // fem::simulate's gather and read-back buffers, kept between steps
  std::vector<double> m_host_x, m_host_v, m_host_f, m_host_Re;
  std::vector<unsigned char> m_host_fixed;
  bool m_dev_valid = false;
  bool m_plastic = false;         // plasticity adds into m_f_external on the device
  bool m_read_rotations = true;   // fem::simulate refreshes every m_Re

 protected:
  std::vector<Node> m_nodes;
  std::vector<Tetrahedron> m_tets;
};

// t4mesh_block_generator.h:32-110: (I+1)(J+1)(K+1) nodes in x, J, K order with
// x = w*i, coord(2) = d*j, coord(1) = h*k; five tetrahedra per cube.
template <typename real_type, typename t4mesh_type>
void generate_blocks(unsigned int const &I, unsigned int const &J, unsigned int const &K,
                     real_type const &block_width, real_type const &block_height,
                     real_type const &block_depth, t4mesh_type &mesh) {
  mesh.clear();
  const unsigned int nv = (I + 1) * (J + 1) * (K + 1);
  for (unsigned int i = 0; i < nv; ++i) mesh.insert();
  auto n = mesh.node_begin();
  for (unsigned int x = 0; x <= I; ++x)
    for (unsigned int y = 0; y <= J; ++y)
      for (unsigned int z = 0; z <= K; ++z, ++n) {
        n->m_coord(0) = block_width * x;
        n->m_coord(2) = block_depth * y;
        n->m_coord(1) = block_height * z;
      }
  for (unsigned int i = 0; i < I; ++i)
    for (unsigned int j = 0; j < J; ++j)
      for (unsigned int k = 0; k < K; ++k) {
        size_t p0 = (i * (J + 1) + j) * (K + 1) + k, p1 = p0 + 1;
        size_t p3 = ((i + 1) * (J + 1) + j) * (K + 1) + k, p2 = p3 + 1;
        size_t p7 = ((i + 1) * (J + 1) + (j + 1)) * (K + 1) + k, p6 = p7 + 1;
        size_t p4 = (i * (J + 1) + (j + 1)) * (K + 1) + k, p5 = p4 + 1;
        if ((i + j + k) % 2 == 1) {
          mesh.insert(p1, p2, p6, p3);
          mesh.insert(p3, p6, p4, p7);
          mesh.insert(p1, p4, p6, p5);
          mesh.insert(p1, p3, p4, p0);
          mesh.insert(p1, p6, p4, p3);
        } else {
          mesh.insert(p2, p0, p5, p1);
          mesh.insert(p2, p7, p0, p3);
          mesh.insert(p2, p5, p7, p6);
          mesh.insert(p0, p7, p5, p4);
          mesh.insert(p2, p0, p7, p5);
        }
      }
}

// t4mesh_block_generator.h:119-243: nodes at the set voxels of a grid (in its
// storage order, i fastest: p = (k J + j) I + i), placed at min + (i dx, j dy,
// k dz) (grid_idx2coord.h:29-46); each cell contributes those of its five
// tetrahedra whose four corners are set.  `grid_type` needs I(), J(), K(),
// dx(), dy(), dz(), min_coord(int) and linear access voxels(p), as the
// reference's grid::Grid has.  Nodes in no tetrahedron are kept (the
// reference reports them).
template <typename t4mesh_type, typename grid_type>
void generate_blocks(grid_type &voxels, t4mesh_type &mesh) {
  mesh.clear();
  const size_t I = voxels.I(), J = voxels.J(), K = voxels.K();
  std::vector<long long> node_of(I * J * K, -1);
  for (size_t k = 0; k < K; ++k)
    for (size_t j = 0; j < J; ++j)
      for (size_t i = 0; i < I; ++i) {
        const size_t p = (k * J + j) * I + i;
        if (!voxels(p)) continue;
        auto n = mesh.insert();
        n->m_coord(0) = i * voxels.dx() + voxels.min_coord(0);
        n->m_coord(1) = j * voxels.dy() + voxels.min_coord(1);
        n->m_coord(2) = k * voxels.dz() + voxels.min_coord(2);
        node_of[p] = (long long)n->idx();
      }
  if (I < 2 || J < 2 || K < 2) return;
  auto add = [&](size_t a, size_t b, size_t c, size_t d) {
    if (voxels(a) && voxels(b) && voxels(c) && voxels(d))
      mesh.insert((size_t)node_of[a], (size_t)node_of[b], (size_t)node_of[c], (size_t)node_of[d]);
  };
  for (size_t i = 0; i + 1 < I; ++i)
    for (size_t j = 0; j + 1 < J; ++j)
      for (size_t k = 0; k + 1 < K; ++k) {
        const size_t p0 = (k * J + j) * I + i, p1 = (k * J + j) * I + (i + 1);
        const size_t p2 = (k * J + (j + 1)) * I + (i + 1), p3 = (k * J + (j + 1)) * I + i;
        const size_t p4 = ((k + 1) * J + j) * I + i, p5 = ((k + 1) * J + j) * I + (i + 1);
        const size_t p6 = ((k + 1) * J + (j + 1)) * I + (i + 1), p7 = ((k + 1) * J + (j + 1)) * I + i;
        if ((i + j + k) % 2 == 1) {
          add(p1, p2, p3, p6);
          add(p3, p6, p7, p4);
          add(p1, p4, p5, p6);
          add(p1, p3, p0, p4);
          add(p1, p6, p3, p4);
        } else {
          add(p2, p0, p1, p5);
          add(p2, p7, p3, p0);
          add(p2, p5, p6, p7);
          add(p0, p7, p4, p5);
          add(p2, p0, p5, p7);
        }
      }
}

}  // namespace t4mesh

namespace fem {

namespace detail {

template <typename math_types>
struct NodeTraits {  // fem_node_traits.h:37-52 (solver rows live on the device)
  typedef typename math_types::real_type real_type;
  typedef typename math_types::vector3_type vector3_type;
  bool m_fixed = false;
  vector3_type m_model_coord;
  vector3_type m_velocity;
  real_type m_mass = 0;
  vector3_type m_f_external;
};

template <typename math_types>
struct TetrahedronTraits {  // fem_tetrahedron_traits.h:29-35
  typedef typename math_types::real_type real_type;
  typedef typename math_types::matrix3x3_type matrix3x3_type;
  real_type m_young = 0;
  real_type m_poisson = 0;
  real_type m_density = 0;
  matrix3x3_type m_Re = math::diag<real_type>(1);
  real_type m_V = 0;
};

}  // namespace detail

// fem_mesh.h:21-34
template <typename math_types>
class Mesh : public t4mesh::T4Mesh<math_types, detail::NodeTraits<math_types>,
                                   detail::TetrahedronTraits<math_types>> {
 public:
  typedef typename math_types::real_type real_type;
  typedef typename math_types::vector3_type vector3_type;
  typedef typename math_types::matrix3x3_type matrix3x3_type;
};

namespace detail {

template <typename fem_mesh>
void create_solver(fem_mesh &mesh) {
  mesh.m_dev_valid = false;  // until finish_init records the new solver's arrays
  const size_t N = mesh.size_nodes(), T = mesh.size_tetrahedra();
  std::vector<double> x(3 * N), x0(3 * N), v(3 * N);
  std::vector<unsigned char> fixed(N);
  std::vector<int32_t> tets(4 * T);
  size_t i = 0;
  for (auto n = mesh.node_begin(); n != mesh.node_end(); ++n, ++i) {
    for (int c = 0; c < 3; ++c) {
      x[3 * i + c] = n->m_coord(c);
      x0[3 * i + c] = n->m_model_coord(c);
      v[3 * i + c] = n->m_velocity(c);
    }
    fixed[i] = n->m_fixed ? 1 : 0;
  }
  size_t t = 0;
  for (auto T4 = mesh.tetrahedron_begin(); T4 != mesh.tetrahedron_end(); ++T4, ++t)
    for (int k = 0; k < 4; ++k) tets[4 * t + k] = (int32_t)T4->node(k)->idx();
  otfem_mesh_desc d{};
  d.num_nodes = (int)N;
  d.num_tets = (int)T;
  d.coord = x.data();
  d.model_coord = x0.data();
  d.velocity = v.data();
  d.fixed = fixed.data();
  d.tets = tets.data();
  otfem_handle h = nullptr;
  ::otfem::detail::check(otfem_create(&d, &h), nullptr, "otfem_create");
  mesh.m_solver.reset(h);
  mesh.m_solver_version = mesh.m_topology_version;
}

template <typename fem_mesh>
void finish_init(fem_mesh &mesh, bool plastic) {
  otfem_handle h = mesh.m_solver.get();
  const size_t N = mesh.size_nodes(), T = mesh.size_tetrahedra();
  // the device now holds the mesh's coordinates, velocities and fixed flags
  // (otfem_create) and zero external forces
  mesh.m_dev_x.assign(3 * N, 0.0);
  mesh.m_dev_v.assign(3 * N, 0.0);
  mesh.m_dev_f.assign(3 * N, 0.0);
  mesh.m_dev_fixed.assign(N, 0);
  {
    size_t j = 0;
    for (auto n = mesh.node_begin(); n != mesh.node_end(); ++n, ++j) {
      for (int c = 0; c < 3; ++c) {
        mesh.m_dev_x[3 * j + c] = n->m_coord(c);
        mesh.m_dev_v[3 * j + c] = n->m_velocity(c);
      }
      mesh.m_dev_fixed[j] = n->m_fixed ? 1 : 0;
    }
  }
  mesh.m_dev_valid = true;
  mesh.m_plastic = plastic;
  std::vector<double> m(N), V(T);
  ::otfem::detail::check(otfem_get_mass(h, m.data()), h, "otfem_get_mass");
  ::otfem::detail::check(otfem_get_volumes(h, V.data()), h, "otfem_get_volumes");
  size_t i = 0;
  for (auto n = mesh.node_begin(); n != mesh.node_end(); ++n, ++i) n->m_mass = m[i];
  size_t t = 0;
  typedef typename fem_mesh::real_type real_type;
  for (auto T4 = mesh.tetrahedron_begin(); T4 != mesh.tetrahedron_end(); ++T4, ++t) {
    T4->m_V = V[t];
    T4->m_Re = math::diag<real_type>(1);
  }
}

}  // namespace detail

// fem_init.h:45-65
template <typename fem_mesh, typename real_type>
inline void init(fem_mesh &mesh, real_type const &young, real_type const &poisson,
                 real_type const &density, real_type const &c_yield, real_type const &c_creep,
                 real_type const &c_max) {
  for (auto T = mesh.tetrahedron_begin(); T != mesh.tetrahedron_end(); ++T) {
    T->m_young = young;
    T->m_poisson = poisson;
    T->m_density = density;
  }
  detail::create_solver(mesh);
  otfem_handle h = mesh.m_solver.get();
  ::otfem::detail::check(otfem_init(h, young, poisson, density, c_yield, c_creep, c_max), h,
                         "fem::init");
  detail::finish_init(mesh, c_creep > 0 && c_max > 0);
}

// The user-written init of fem_init.h:32-35: materials already set per tet
// in m_young / m_poisson / m_density.
template <typename fem_mesh, typename real_type>
inline void init_materials(fem_mesh &mesh, real_type const &c_yield, real_type const &c_creep,
                           real_type const &c_max) {
  std::vector<double> E, nu, rho;
  for (auto T = mesh.tetrahedron_begin(); T != mesh.tetrahedron_end(); ++T) {
    E.push_back(T->m_young);
    nu.push_back(T->m_poisson);
    rho.push_back(T->m_density);
  }
  detail::create_solver(mesh);
  otfem_handle h = mesh.m_solver.get();
  ::otfem::detail::check(otfem_init_materials(h, E.data(), nu.data(), rho.data(), c_yield,
                                              c_creep, c_max), h, "fem::init_materials");
  detail::finish_init(mesh, c_creep > 0 && c_max > 0);
}

// Solver settings for subsequent fem::simulate calls on this mesh (the
// reference hard-codes them; default = the reference's values).
template <typename fem_mesh>
inline void set_params(fem_mesh &mesh, const otfem_params &p) {
  mesh.m_params = p;
  mesh.m_params_set = true;
}

// Whether fem::simulate refreshes every tetrahedron's m_Re (default, as the
// reference's update_orientation does).  A caller that never reads m_Re can
// turn it off: the device then stores no rotations and none cross the bus.
template <typename fem_mesh>
inline void set_rotation_readback(fem_mesh &mesh, bool on) {
  mesh.m_read_rotations = on;
}

// fem_simulate.h:49-113.  The caller's node arrays are compared with what the
// device already holds and only changed ones are uploaded (a velocity upload
// restarts the learned and subspace starts' histories, a position or force
// upload does not); positions and velocities come back every step, external
// forces only when plasticity added into them, rotations unless turned off.
template <typename fem_mesh, typename real_type>
inline void simulate(fem_mesh &mesh, real_type const &time_step, bool use_stiffness_warping) {
  otfem_handle h = mesh.m_solver.get();
  if (!h || mesh.m_solver_version != mesh.m_topology_version || !mesh.m_dev_valid)
    throw std::logic_error("fem::simulate(): fem::init must be called after the mesh is built");
  const size_t N = mesh.size_nodes(), T = mesh.size_tetrahedra();
  std::vector<double> &x = mesh.m_host_x, &v = mesh.m_host_v, &f = mesh.m_host_f;
  std::vector<unsigned char> &fixed = mesh.m_host_fixed;
  x.resize(3 * N);
  v.resize(3 * N);
  f.resize(3 * N);
  fixed.resize(N);
  size_t i = 0;
  for (auto n = mesh.node_begin(); n != mesh.node_end(); ++n, ++i) {
    for (int c = 0; c < 3; ++c) {
      x[3 * i + c] = n->m_coord(c);
      v[3 * i + c] = n->m_velocity(c);
      f[3 * i + c] = n->m_f_external(c);
    }
    fixed[i] = n->m_fixed ? 1 : 0;
  }
  if (fixed != mesh.m_dev_fixed) {
    ::otfem::detail::check(otfem_set_fixed(h, fixed.data()), h, "otfem_set_fixed");
    mesh.m_dev_fixed = fixed;
  }
  const bool x_new = x != mesh.m_dev_x, v_new = v != mesh.m_dev_v;
  if (x_new || v_new)
    ::otfem::detail::check(otfem_set_state(h, x_new ? x.data() : nullptr, v_new ? v.data() : nullptr),
                           h, "otfem_set_state");
  if (f != mesh.m_dev_f) {
    ::otfem::detail::check(otfem_set_external_forces(h, f.data()), h, "otfem_set_external_forces");
    mesh.m_dev_f = f;
  }
  otfem_params P;
  if (mesh.m_params_set)
    P = mesh.m_params;
  else
    otfem_default_params(&P, 0);
  P.store_rotations = mesh.m_read_rotations ? 1 : 0;
  otfem_stats st;
  ::otfem::detail::check(otfem_step(h, (double)time_step, use_stiffness_warping ? 1 : 0, &P, &st),
                         h, "fem::simulate");
  ::otfem::detail::check(otfem_get_state(h, x.data(), v.data()), h, "otfem_get_state");
  if (mesh.m_plastic) {
    ::otfem::detail::check(otfem_get_external_forces(h, f.data()), h, "otfem_get_external_forces");
    mesh.m_dev_f = f;
  }
  i = 0;
  for (auto n = mesh.node_begin(); n != mesh.node_end(); ++n, ++i)
    for (int c = 0; c < 3; ++c) {
      n->m_coord(c) = x[3 * i + c];
      n->m_velocity(c) = v[3 * i + c];
      if (mesh.m_plastic) n->m_f_external(c) = f[3 * i + c];
    }
  mesh.m_dev_x.swap(x);
  mesh.m_dev_v.swap(v);
  if (mesh.m_read_rotations) {
    std::vector<double> &Re = mesh.m_host_Re;
    Re.resize(9 * T);
    ::otfem::detail::check(otfem_get_rotations(h, Re.data()), h, "otfem_get_rotations");
    size_t t = 0;
    for (auto T4 = mesh.tetrahedron_begin(); T4 != mesh.tetrahedron_end(); ++T4, ++t)
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) T4->m_Re(r, c) = Re[9 * t + 3 * r + c];
  }
}

// fem_update_original_coord.h:28-33 (host-side; fem::init must follow, as in
// the reference)
template <typename node_iterator>
inline void update_original_coord(node_iterator const &begin, node_iterator const &end) {
  for (node_iterator N = begin; N != end; ++N) N->m_model_coord = N->m_coord;
}

// fem_export_k.h:23-68.  matrix_type needs resize(rows, cols, bool), clear()
// and push_back(row, col, value) -- the uBLAS calls the reference makes.
template <typename fem_mesh, typename matrix_type>
inline void export_K(fem_mesh &mesh, matrix_type &bigK) {
  otfem_handle h = mesh.m_solver.get();
  if (!h) throw std::logic_error("fem::export_K(): fem::init has not been called");
  const size_t N = mesh.size_nodes();
  int64_t nnz = 0;
  ::otfem::detail::check(otfem_export_K(h, &nnz, nullptr, nullptr, nullptr), h, "fem::export_K");
  std::vector<int64_t> rp(N + 1);
  std::vector<int32_t> cols((size_t)nnz);
  std::vector<double> blocks(9 * (size_t)nnz);
  ::otfem::detail::check(otfem_export_K(h, &nnz, rp.data(), cols.data(), blocks.data()), h,
                         "fem::export_K");
  bigK.resize(N * 3, N * 3, false);
  bigK.clear();
  for (size_t i = 0; i < N; ++i)
    for (unsigned int r = 0; r < 3; ++r)
      for (int64_t k = rp[i]; k < rp[i + 1]; ++k) {
        bool fixed_row = mesh.node(i)->m_fixed;
        for (unsigned int c = 0; c < 3; ++c) {
          double val = blocks[9 * k + 3 * r + c];
          if (fixed_row) {
            if (c == r) bigK.push_back(3 * i + r, 3 * i + r, 1.0);
          } else {
            bigK.push_back(3 * i + r, 3 * (size_t)cols[k] + c, val);
          }
        }
      }
}

}  // namespace fem
}  // namespace otfem

#ifndef OTFEM_NO_OPENTISSUE_ALIAS
namespace OpenTissue {
namespace math = ::otfem::math;
namespace t4mesh = ::otfem::t4mesh;
namespace fem = ::otfem::fem;
}  // namespace OpenTissue
#endif

#endif  // OTFEM_FEM_HPP
