// FP64 CPU oracle for the corotated tetrahedral FEM step of OpenTissue::fem.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP library is
// compared against, and the single-core CPU baseline that is timed beside it.
// Nothing in the product library links, loads or calls it.
//
// It restates the reference algorithm with the reference's own data structure
// (array-of-structs nodes and tetrahedra, one std::map<int, 3x3> row per node
// for K' and A, the same loop order and the same operation order inside the
// 3-vector / 3x3 arithmetic), so that it is both a faithful parity oracle and a
// faithful timing stand-in for the reference (which cannot be compiled here:
// its include closure needs Boost, absent from this image).
//
// Reference files followed (paths relative to the OpenTissue tree):
//   OpenTissue/dynamics/fem/fem_init.h:45-65                  init
//   OpenTissue/dynamics/fem/fem_simulate.h:49-113             simulate
//   OpenTissue/dynamics/fem/fem_compute_ke.h:29-100,106-121,132-292  compute_Ke
//   OpenTissue/dynamics/fem/fem_compute_volume.h:27-39        compute_volume
//   OpenTissue/dynamics/fem/fem_initialize_stiffness_elements.h:30-62
//   OpenTissue/dynamics/fem/fem_compute_mass.h:30-172         lumped mass
//   OpenTissue/dynamics/fem/fem_update_orientation.h:22-215   Re (Gram-Schmidt)
//   OpenTissue/dynamics/fem/fem_reset_orientation.h:22-27     Re = I
//   OpenTissue/dynamics/fem/fem_clear_stiffness_assembly.h:23-33
//   OpenTissue/dynamics/fem/fem_stiffness_assembly.h:171-209  K', f0'
//   OpenTissue/dynamics/fem/fem_add_plasticity_force.h:26-204 plasticity
//   OpenTissue/dynamics/fem/fem_initialize_plastic.h:33-57
//   OpenTissue/dynamics/fem/fem_compute_b.h:28-396            B vectors
//   OpenTissue/dynamics/fem/fem_compute_isotropic_elasticity.h:25-51
//   OpenTissue/dynamics/fem/fem_dynamics_assembly.h:47-93     A, b
//   OpenTissue/dynamics/fem/fem_conjugate_gradients.h:25-133  CG
//   OpenTissue/dynamics/fem/fem_position_update.h:25-39       x += dt v
//   OpenTissue/dynamics/fem/fem_update_original_coord.h:28-33
//   OpenTissue/dynamics/fem/fem_export_k.h:23-68
//   OpenTissue/core/math/math_vector3.h:63-67,190-194,374-391 vector ops
//   OpenTissue/core/math/math_matrix3x3.h:151,253-266,535-587 matrix ops
//   OpenTissue/core/math/math_polar_decomposition.h:49-91     polar (POLAR mode)
//   OpenTissue/core/math/math_eigen_system_decomposition.h:38-170 eigen (POLAR)
//
// Extensions beyond the reference (each reduces to the reference exactly for
// the reference's hard-coded values, mass_damping=2, beta=0, min=max=20):
//   * stiffness damping beta: A_ij = K'_ij*(dt*dt + beta*dt)  (beta=0 adds an
//     exact 0.0, so the product is bit-identical to K'_ij*(dt*dt));
//   * a relative-residual stop sqrt(r.r) <= tol*sqrt(b.b) over free dofs;
//   * POLAR rotation mode via the reference's own (commented-out) eigen-based
//     polar decomposition, fem_update_orientation.h:210-213.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <vector>

namespace orc {

struct V3 {
  double x = 0, y = 0, z = 0;  // zero-initialised like math_vector3.h:63-67
  V3() = default;
  V3(double a, double b, double c) : x(a), y(b), z(c) {}
  double &operator()(int i) { return i == 0 ? x : (i == 1 ? y : z); }
  double operator()(int i) const { return i == 0 ? x : (i == 1 ? y : z); }
  void clear() { x = y = z = 0; }
  V3 &operator+=(const V3 &v) { x += v.x; y += v.y; z += v.z; return *this; }
  V3 &operator-=(const V3 &v) { x -= v.x; y -= v.y; z -= v.z; return *this; }
  V3 &operator*=(double s) { x *= s; y *= s; z *= s; return *this; }
  V3 &operator/=(double s) { x /= s; y /= s; z /= s; return *this; }
  V3 operator+(const V3 &v) const { return V3(x + v.x, y + v.y, z + v.z); }
  V3 operator-(const V3 &v) const { return V3(x - v.x, y - v.y, z - v.z); }
  // math_vector3.h:193  operator%  (cross product, this operand order)
  V3 operator%(const V3 &v) const {
    return V3(y * v.z - v.y * z, v.x * z - x * v.z, x * v.y - v.x * y);
  }
  // math_vector3.h:194  operator*  (dot product)
  double operator*(const V3 &v) const { return x * v.x + y * v.y + z * v.z; }
};
inline V3 operator*(const V3 &v, double s) { return V3(v.x * s, v.y * s, v.z * s); }
inline V3 operator*(double s, const V3 &v) { return V3(v.x * s, v.y * s, v.z * s); }
// math_vector3.h:374-377 cross() (different operand order from operator%)
inline V3 cross(const V3 &a, const V3 &b) {
  return V3(a.y * b.z - b.y * a.z, -a.x * b.z + b.x * a.z, a.x * b.y - b.x * a.y);
}
inline double dot(const V3 &a, const V3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double length(const V3 &v) { return std::sqrt(dot(v, v)); }

struct M3 {  // three row vectors, math_matrix3x3.h:50-52
  V3 r[3];
  M3() = default;
  M3(double a00, double a01, double a02, double a10, double a11, double a12,
     double a20, double a21, double a22) {
    r[0] = V3(a00, a01, a02); r[1] = V3(a10, a11, a12); r[2] = V3(a20, a21, a22);
  }
  double &operator()(int i, int j) { return r[i](j); }
  double operator()(int i, int j) const { return r[i](j); }
  void clear() { r[0].clear(); r[1].clear(); r[2].clear(); }
  M3 &operator+=(const M3 &m) { r[0] += m.r[0]; r[1] += m.r[1]; r[2] += m.r[2]; return *this; }
  M3 &operator*=(double s) { r[0] *= s; r[1] *= s; r[2] *= s; return *this; }
  // math_matrix3x3.h:151  M*v as three row dots
  V3 operator*(const V3 &v) const { return V3(r[0] * v, r[1] * v, r[2] * v); }
};
inline M3 operator*(const M3 &m, double s) {
  M3 c; c.r[0] = m.r[0] * s; c.r[1] = m.r[1] * s; c.r[2] = m.r[2] * s; return c;
}
// math_matrix3x3.h:253-266  column-outer loop, accumulate from zero
inline M3 operator*(const M3 &A, const M3 &B) {
  M3 C;
  for (int c = 0; c < 3; ++c)
    for (int rr = 0; rr < 3; ++rr) {
      C(rr, c) = 0.0;
      for (int i = 0; i < 3; ++i) C(rr, c) += A(rr, i) * B(i, c);
    }
  return C;
}
inline M3 trans(const M3 &M) {  // math_matrix3x3.h:560-567
  return M3(M(0, 0), M(1, 0), M(2, 0), M(0, 1), M(1, 1), M(2, 1), M(0, 2), M(1, 2), M(2, 2));
}
inline M3 diag1() { return M3(1, 0, 0, 0, 1, 0, 0, 0, 1); }

// math_matrix3x3.h:535-557  Gram-Schmidt by rows
inline M3 ortonormalize(const M3 &A) {
  V3 row0(A(0, 0), A(0, 1), A(0, 2));
  V3 row1(A(1, 0), A(1, 1), A(1, 2));
  V3 row2(A(2, 0), A(2, 1), A(2, 2));
  const double l0 = length(row0);
  if (l0) row0 /= l0;
  row1 -= row0 * dot(row0, row1);
  const double l1 = length(row1);
  if (l1) row1 /= l1;
  row2 = cross(row0, row1);
  return M3(row0.x, row0.y, row0.z, row1.x, row1.y, row1.z, row2.x, row2.y, row2.z);
}

inline M3 outer_prod(const V3 &a, const V3 &b) {
  return M3(a.x * b.x, a.x * b.y, a.x * b.z, a.y * b.x, a.y * b.y, a.y * b.z, a.z * b.x,
            a.z * b.y, a.z * b.z);
}

inline M3 inverse3(const M3 &A) {
  // plain cofactor inverse (used only by the optional POLAR mode)
  double a = A(0, 0), b = A(0, 1), c = A(0, 2), d = A(1, 0), e = A(1, 1), f = A(1, 2),
         g = A(2, 0), h = A(2, 1), k = A(2, 2);
  double det = a * (e * k - f * h) - b * (d * k - f * g) + c * (d * h - e * g);
  double id = 1.0 / det;
  return M3((e * k - f * h) * id, (c * h - b * k) * id, (b * f - c * e) * id,
            (f * g - d * k) * id, (a * k - c * g) * id, (c * d - a * f) * id,
            (d * h - e * g) * id, (b * g - a * h) * id, (a * e - b * d) * id);
}

// math_eigen_system_decomposition.h:38-170 (tridiagonal QL, 32 iterations max)
inline void eigen_sym(const M3 &A, M3 &V, V3 &diag) {
  V3 sub;
  V = A;
  double fM00 = V(0, 0), fM01 = V(0, 1), fM02 = V(0, 2), fM11 = V(1, 1), fM12 = V(1, 2),
         fM22 = V(2, 2);
  diag(0) = fM00;
  sub(2) = 0.0;
  if (fM02 != 0.0) {
    double fLength = std::sqrt(fM01 * fM01 + fM02 * fM02);
    double fInv = 1.0 / fLength;
    fM01 *= fInv;
    fM02 *= fInv;
    double fQ = 2.0 * fM01 * fM12 + fM02 * (fM22 - fM11);
    diag(1) = fM11 + fM02 * fQ;
    diag(2) = fM22 - fM02 * fQ;
    sub(0) = fLength;
    sub(1) = fM12 - fM01 * fQ;
    V = M3(1, 0, 0, 0, fM01, fM02, 0, fM02, -fM01);
  } else {
    diag(1) = fM11;
    diag(2) = fM22;
    sub(0) = fM01;
    sub(1) = fM12;
    V = diag1();
  }
  const int max_it = 32, dim = 3;
  for (int i0 = 0; i0 < dim; ++i0) {
    int i1;
    for (i1 = 0; i1 < max_it; ++i1) {
      int i2;
      for (i2 = i0; i2 <= dim - 2; ++i2) {
        double fTmp = std::fabs(diag(i2)) + std::fabs(diag(i2 + 1));
        if (std::fabs(sub(i2)) + fTmp == fTmp) break;
      }
      if (i2 == i0) break;
      double fG = (diag(i0 + 1) - diag(i0)) / (2.0 * sub(i0));
      double fR = std::sqrt(fG * fG + 1.0);
      if (fG < 0.0)
        fG = diag(i2) - diag(i0) + sub(i0) / (fG - fR);
      else
        fG = diag(i2) - diag(i0) + sub(i0) / (fG + fR);
      double fSin = 1.0, fCos = 1.0, fP = 0.0;
      for (int i3 = i2 - 1; i3 >= i0; --i3) {
        double fF = fSin * sub(i3);
        double fB = fCos * sub(i3);
        if (std::fabs(fF) >= std::fabs(fG)) {
          fCos = fG / fF;
          fR = std::sqrt(fCos * fCos + 1.0);
          sub(i3 + 1) = fF * fR;
          fSin = 1.0 / fR;
          fCos *= fSin;
        } else {
          fSin = fF / fG;
          fR = std::sqrt(fSin * fSin + 1.0);
          sub(i3 + 1) = fG * fR;
          fCos = 1.0 / fR;
          fSin *= fCos;
        }
        fG = diag(i3 + 1) - fP;
        fR = (diag(i3) - fG) * fSin + 2.0 * fB * fCos;
        fP = fSin * fR;
        diag(i3 + 1) = fG + fP;
        fG = fCos * fR - fB;
        for (int i4 = 0; i4 < dim; ++i4) {
          fF = V(i4, i3 + 1);
          V(i4, i3 + 1) = fSin * V(i4, i3) + fCos * fF;
          V(i4, i3) = fCos * V(i4, i3) - fSin * fF;
        }
      }
      diag(i0) -= fP;
      sub(i0) = fG;
      sub(i2) = 0.0;
    }
    if (i1 == max_it) break;
  }
}

// math_polar_decomposition.h:49-91  A = R S via eigen of A^T A
inline bool polar_eigen(const M3 &A, M3 &R, M3 &S) {
  M3 V;
  V3 d;
  M3 S2 = trans(A) * A;
  eigen_sym(S2, V, d);
  if (d(0) <= 0.0 || d(1) <= 0.0 || d(2) <= 0.0) return false;
  V3 v0(V(0, 0), V(1, 0), V(2, 0)), v1(V(0, 1), V(1, 1), V(2, 1)), v2(V(0, 2), V(1, 2), V(2, 2));
  M3 t0 = outer_prod(v0, v0) * std::sqrt(d(0));
  M3 t1 = outer_prod(v1, v1) * std::sqrt(d(1));
  M3 t2 = outer_prod(v2, v2) * std::sqrt(d(2));
  S = t0;
  S += t1;
  S += t2;
  R = A * inverse3(S);
  return true;
}

struct Node {  // fem_node_traits.h:21-70
  std::map<int, M3> K_row, A_row;
  V3 f0, b;
  bool fixed = false;
  V3 model_coord, coord, velocity;
  double mass = 0;
  V3 f_external;
  V3 update, prev, residual;
};

struct Tet {  // fem_tetrahedron_traits.h:18-49
  int n[4];
  double young = 0, poisson = 0, density = 0;
  M3 Ke[4][4];
  M3 Re;
  double V = 0;
  V3 e10, e20, e30;
  V3 B[4];
  V3 D;
  double plastic[6] = {0, 0, 0, 0, 0, 0};
  double yield = 0, creep = 0, cmax = 0;
};

struct Mesh {
  std::vector<Node> nodes;
  std::vector<Tet> tets;
};

// fem_compute_ke.h:29-100
inline void compute_Ke(const V3 &p0, const V3 &p1, const V3 &p2, const V3 &p3, double E,
                       double nu, M3 Ke[4][4]) {
  double d = p0(0), d4 = p0(1), d8 = p0(2);
  double d1 = p1(0) - d, d5 = p1(1) - d4, d9 = p1(2) - d8;
  double d2 = p2(0) - d, d6 = p2(1) - d4, d10 = p2(2) - d8;
  double d3 = p3(0) - d, d7 = p3(1) - d4, d11 = p3(2) - d8;
  double d12 = (d1 * d6 * d11 + d2 * d7 * d9 + d3 * d5 * d10) - d1 * d7 * d10 - d2 * d5 * d11 -
               d3 * d6 * d9;
  double d13 = 1.0 / d12;
  double d14 = std::fabs(d12 / 6.0);
  V3 B[4];
  B[1](0) = (d10 * d7 - d6 * d11) * d13;
  B[2](0) = (d5 * d11 - d9 * d7) * d13;
  B[3](0) = (d9 * d6 - d5 * d10) * d13;
  B[0](0) = -B[1](0) - B[2](0) - B[3](0);
  B[1](1) = (d2 * d11 - d10 * d3) * d13;
  B[2](1) = (d9 * d3 - d1 * d11) * d13;
  B[3](1) = (d1 * d10 - d9 * d2) * d13;
  B[0](1) = -B[1](1) - B[2](1) - B[3](1);
  B[1](2) = (d6 * d3 - d2 * d7) * d13;
  B[2](2) = (d1 * d7 - d5 * d3) * d13;
  B[3](2) = (d5 * d2 - d1 * d6) * d13;
  B[0](2) = -B[1](2) - B[2](2) - B[3](2);
  double d15 = E / (1.0 + nu) / (1.0 - 2 * nu);
  double d16 = (1.0 - nu) * d15;
  double d17 = nu * d15;
  double d18 = E / 2 / (1.0 + nu);
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      double d19 = B[i](0), d20 = B[i](1), d21 = B[i](2);
      double d22 = B[j](0), d23 = B[j](1), d24 = B[j](2);
      Ke[i][j](0, 0) = d16 * d19 * d22 + d18 * (d20 * d23 + d21 * d24);
      Ke[i][j](0, 1) = d17 * d19 * d23 + d18 * (d20 * d22);
      Ke[i][j](0, 2) = d17 * d19 * d24 + d18 * (d21 * d22);
      Ke[i][j](1, 0) = d17 * d20 * d22 + d18 * (d19 * d23);
      Ke[i][j](1, 1) = d16 * d20 * d23 + d18 * (d19 * d22 + d21 * d24);
      Ke[i][j](1, 2) = d17 * d20 * d24 + d18 * (d21 * d23);
      Ke[i][j](2, 0) = d17 * d21 * d22 + d18 * (d19 * d24);
      Ke[i][j](2, 1) = d17 * d21 * d23 + d18 * (d20 * d24);
      Ke[i][j](2, 2) = d16 * d21 * d24 + d18 * (d20 * d23 + d19 * d22);
      Ke[i][j] *= d14;
    }
}

// fem_compute_ke.h:132-292  Ke_ij from (B_i, D, B_j, V)
inline void compute_Ke_ij(const V3 &Bi, const V3 &D, const V3 &Bj, double volume, M3 &Ke) {
  double bi = Bi(0), ci = Bi(1), di = Bi(2);
  double bj = Bj(0), cj = Bj(1), dj = Bj(2);
  double D0 = D(0) * volume, D1 = D(1) * volume, D2 = D(2) * volume;
  Ke(0, 0) = D0 * bi * bj + D2 * (ci * cj + di * dj);
  Ke(0, 1) = D1 * bi * cj + D2 * (ci * bj);
  Ke(0, 2) = D1 * bi * dj + D2 * (di * bj);
  Ke(1, 0) = D1 * ci * bj + D2 * (bi * cj);
  Ke(1, 1) = D0 * ci * cj + D2 * (bi * bj + di * dj);
  Ke(1, 2) = D1 * ci * dj + D2 * (di * cj);
  Ke(2, 0) = D1 * di * bj + D2 * (bi * dj);
  Ke(2, 1) = D1 * di * cj + D2 * (ci * dj);
  Ke(2, 2) = D0 * di * dj + D2 * (ci * cj + bi * bj);
}

// fem_compute_ke.h:106-121
inline void compute_Ke_BD(const V3 *B, const V3 &D, double volume, M3 Ke[4][4]) {
  for (int i = 0; i < 4; ++i)
    for (int j = i; j < 4; ++j) compute_Ke_ij(B[i], D, B[j], volume, Ke[i][j]);
  for (int i = 1; i < 4; ++i)
    for (int j = 0; j < i; ++j) Ke[i][j] = trans(Ke[j][i]);
}

// fem_compute_volume.h:27-39
inline double compute_volume(const V3 &e10, const V3 &e20, const V3 &e30) {
  double sixV = e10 * (e20 % e30);
  return sixV / 6.0;
}

// fem_compute_b.h:28-396 (note: the reference's sign convention, B = -grad N)
inline void compute_B(const V3 &e10, const V3 &e20, const V3 &e30, double volume, V3 *B) {
  double div6V = 1 / (6.0 * volume);
  B[1](0) = (e20(2) * e30(1) - e20(1) * e30(2)) * div6V;
  B[2](0) = (e10(1) * e30(2) - e10(2) * e30(1)) * div6V;
  B[3](0) = (e10(2) * e20(1) - e10(1) * e20(2)) * div6V;
  B[0](0) = -B[1](0) - B[2](0) - B[3](0);
  B[1](1) = (e20(0) * e30(2) - e20(2) * e30(0)) * div6V;
  B[2](1) = (e10(2) * e30(0) - e10(0) * e30(2)) * div6V;
  B[3](1) = (e10(0) * e20(2) - e10(2) * e20(0)) * div6V;
  B[0](1) = -B[1](1) - B[2](1) - B[3](1);
  B[1](2) = (e20(1) * e30(0) - e20(0) * e30(1)) * div6V;
  B[2](2) = (e10(0) * e30(1) - e10(1) * e30(0)) * div6V;
  B[3](2) = (e10(1) * e20(0) - e10(0) * e20(1)) * div6V;
  B[0](2) = -B[1](2) - B[2](2) - B[3](2);
}

// fem_compute_isotropic_elasticity.h:25-51
inline void compute_isotropic_elasticity_vector(double young, double poisson, V3 &D) {
  double poisson2 = 2.0 * poisson;
  double scale = young / ((1.0 + poisson) * (1.0 - poisson2));
  D(0) = (1.0 - poisson) * scale;
  D(1) = poisson * scale;
  D(2) = young / (2.0 + poisson2);
}

}  // namespace orc

using namespace orc;

extern "C" {

// Solver parameters.  The reference's hard-coded values are
//   alpha=2 (fem_simulate.h:104), beta=0 (fem_dynamics_assembly.h:39-41),
//   min=max=20 (fem_simulate.h:108-109), tol_mode=0 with tol=1e-3 and
//   tiny=1e-10 (fem_conjugate_gradients.h:43-44), rotation Gram-Schmidt.
typedef struct {
  double alpha;        // mass damping coefficient, C = alpha*M + beta*K
  double beta;         // stiffness damping coefficient
  int cg_min_iterations;
  int cg_max_iterations;
  int tol_mode;        // 0: break if iteration>=min && r.r < tol (reference)
                       // 1: stop when sqrt(r.r) <= tol*sqrt(b.b), free dofs
  double tol;
  double tiny;         // |d2|,|d| clamp of fem_conjugate_gradients.h:101-102,121-122
  int rotation_mode;   // 0 Gram-Schmidt (reference), 1 polar decomposition
} orc_params;

typedef struct {
  int iterations;      // CG iterations carried out
  int converged;       // 1 if the stopping rule fired (always 0 in reference mode)
  double rr0;          // r.r after the warm start
  double rr;           // final (recursive) r.r
  double bb;           // b.b over free dofs
} orc_stats;

void *orc_create(int N, int T, const double *coord, const double *model_coord,
                 const double *velocity, const unsigned char *fixed, const int *tets) {
  Mesh *m = new Mesh();
  m->nodes.resize(N);
  m->tets.resize(T);
  for (int i = 0; i < N; ++i) {
    Node &n = m->nodes[i];
    n.coord = V3(coord[3 * i], coord[3 * i + 1], coord[3 * i + 2]);
    n.model_coord = V3(model_coord[3 * i], model_coord[3 * i + 1], model_coord[3 * i + 2]);
    if (velocity) n.velocity = V3(velocity[3 * i], velocity[3 * i + 1], velocity[3 * i + 2]);
    n.fixed = fixed ? (fixed[i] != 0) : false;
  }
  for (int t = 0; t < T; ++t)
    for (int k = 0; k < 4; ++k) m->tets[t].n[k] = tets[4 * t + k];
  return m;
}

void orc_destroy(void *h) { delete static_cast<Mesh *>(h); }

static void initialize_stiffness_elements(Mesh &m) {  // :30-62
  for (Tet &T : m.tets) {
    const V3 &p0 = m.nodes[T.n[0]].model_coord;
    const V3 &p1 = m.nodes[T.n[1]].model_coord;
    const V3 &p2 = m.nodes[T.n[2]].model_coord;
    const V3 &p3 = m.nodes[T.n[3]].model_coord;
    T.e10 = p1 - p0;
    T.e20 = p2 - p0;
    T.e30 = p3 - p0;
    T.V = compute_volume(T.e10, T.e20, T.e30);
    T.Re = diag1();
    compute_Ke(p0, p1, p2, p3, T.young, T.poisson, T.Ke);
  }
}

static void clear_stiffness_assembly(Mesh &m) {  // fem_clear_stiffness_assembly.h:23-33
  for (Node &n : m.nodes) {
    n.f0.clear();
    for (auto &kv : n.K_row) kv.second.clear();
  }
}

static void compute_mass(Mesh &m) {  // fem_compute_mass.h:154-171
  for (Node &n : m.nodes) n.mass = n.fixed ? std::numeric_limits<double>::max() : 0.0;
  for (Tet &T : m.tets) {
    double amount = T.density * T.V * .25;
    for (int k = 0; k < 4; ++k) m.nodes[T.n[k]].mass += amount;
  }
}

static void initialize_plastic(Mesh &m, double c_yield, double c_creep, double c_max) {
  for (Tet &T : m.tets) {  // fem_initialize_plastic.h:45-55
    T.yield = c_yield;
    T.creep = c_creep;
    T.cmax = c_max;
    for (int i = 0; i < 6; ++i) T.plastic[i] = 0;
    compute_B(T.e10, T.e20, T.e30, T.V, T.B);
    compute_isotropic_elasticity_vector(T.young, T.poisson, T.D);
  }
}

static int finish_init(Mesh &m, double c_yield, double c_creep, double c_max) {
  initialize_stiffness_elements(m);
  clear_stiffness_assembly(m);
  compute_mass(m);
  initialize_plastic(m, c_yield, c_creep, c_max);
  int inverted = 0;
  for (const Tet &T : m.tets)
    if (!(T.V > 0)) ++inverted;  // the reference asserts V>0 (:48)
  return inverted;
}

// fem_init.h:45-65.  Returns the number of tetrahedra with V <= 0.
int orc_init(void *h, double young, double poisson, double density, double c_yield,
             double c_creep, double c_max) {
  Mesh &m = *static_cast<Mesh *>(h);
  for (Tet &T : m.tets) {
    T.young = young;
    T.poisson = poisson;
    T.density = density;
  }
  return finish_init(m, c_yield, c_creep, c_max);
}

// The user-written init that fem_init.h:32-35 invites: per-tet materials.
int orc_init_materials(void *h, const double *young, const double *poisson, const double *density,
                       double c_yield, double c_creep, double c_max) {
  Mesh &m = *static_cast<Mesh *>(h);
  for (size_t t = 0; t < m.tets.size(); ++t) {
    m.tets[t].young = young[t];
    m.tets[t].poisson = poisson[t];
    m.tets[t].density = density[t];
  }
  return finish_init(m, c_yield, c_creep, c_max);
}

static void update_orientation(Mesh &m, int rotation_mode) {  // :22-215
  for (Tet &T : m.tets) {
    double div6V = 1.0 / T.V * 6.0;  // sic: (1/V)*6, fem_update_orientation.h:31
    double e1x = T.e10(0), e1y = T.e10(1), e1z = T.e10(2);
    double e2x = T.e20(0), e2y = T.e20(1), e2z = T.e20(2);
    double e3x = T.e30(0), e3y = T.e30(1), e3z = T.e30(2);
    double n1x = (e2y * e3z - e3y * e2z) * div6V;
    double n1y = (e3x * e2z - e2x * e3z) * div6V;
    double n1z = (e2x * e3y - e3x * e2y) * div6V;
    double n2x = (e1z * e3y - e1y * e3z) * div6V;
    double n2y = (e1x * e3z - e1z * e3x) * div6V;
    double n2z = (e1y * e3x - e1x * e3y) * div6V;
    double n3x = (e1y * e2z - e1z * e2y) * div6V;
    double n3y = (e1z * e2x - e1x * e2z) * div6V;
    double n3z = (e1x * e2y - e1y * e2x) * div6V;
    const V3 &p0 = m.nodes[T.n[0]].coord;
    const V3 &p1 = m.nodes[T.n[1]].coord;
    const V3 &p2 = m.nodes[T.n[2]].coord;
    const V3 &p3 = m.nodes[T.n[3]].coord;
    e1x = p1(0) - p0(0); e1y = p1(1) - p0(1); e1z = p1(2) - p0(2);
    e2x = p2(0) - p0(0); e2y = p2(1) - p0(1); e2z = p2(2) - p0(2);
    e3x = p3(0) - p0(0); e3y = p3(1) - p0(1); e3z = p3(2) - p0(2);
    M3 R;
    R(0, 0) = e1x * n1x + e2x * n2x + e3x * n3x;
    R(0, 1) = e1x * n1y + e2x * n2y + e3x * n3y;
    R(0, 2) = e1x * n1z + e2x * n2z + e3x * n3z;
    R(1, 0) = e1y * n1x + e2y * n2x + e3y * n3x;
    R(1, 1) = e1y * n1y + e2y * n2y + e3y * n3y;
    R(1, 2) = e1y * n1z + e2y * n2z + e3y * n3z;
    R(2, 0) = e1z * n1x + e2z * n2x + e3z * n3x;
    R(2, 1) = e1z * n1y + e2z * n2y + e3z * n3y;
    R(2, 2) = e1z * n1z + e2z * n2z + e3z * n3z;
    if (rotation_mode == 1) {
      M3 Rp, S;
      if (polar_eigen(R, Rp, S))
        T.Re = Rp;
      else
        T.Re = ortonormalize(R);  // degenerate: fall back to Gram-Schmidt
    } else {
      T.Re = ortonormalize(R);
    }
  }
}

static void reset_orientation(Mesh &m) {  // fem_reset_orientation.h:22-27
  for (Tet &T : m.tets) T.Re = diag1();
}

static void stiffness_assembly(Mesh &m) {  // fem_stiffness_assembly.h:171-209
  for (Tet &T : m.tets) {
    const M3 &Re = T.Re;
    for (int i = 0; i < 4; ++i) {
      Node &p_i = m.nodes[T.n[i]];
      V3 f;
      for (int j = 0; j < 4; ++j) {
        Node &p_j = m.nodes[T.n[j]];
        const M3 &Ke_ij = T.Ke[i][j];
        f += Ke_ij * p_j.model_coord;
        if (j >= i) {
          M3 tmp = Re * Ke_ij * trans(Re);
          p_i.K_row[T.n[j]] += tmp;
          if (j > i) p_j.K_row[T.n[i]] += trans(tmp);
        }
      }
      p_i.f0 -= Re * f;
    }
  }
}

static void add_plasticity_force(Mesh &m, double dt) {  // fem_add_plasticity_force.h:26-204
  for (Tet &T : m.tets) {
    double e_total[6] = {0, 0, 0, 0, 0, 0}, e_elastic[6];
    for (int j = 0; j < 4; ++j) {
      const V3 &x_j = m.nodes[T.n[j]].coord;
      const V3 &x0_j = m.nodes[T.n[j]].model_coord;
      V3 tmp = (trans(T.Re) * x_j) - x0_j;
      double bj = T.B[j](0), cj = T.B[j](1), dj = T.B[j](2);
      e_total[0] += bj * tmp(0);
      e_total[1] += cj * tmp(1);
      e_total[2] += dj * tmp(2);
      e_total[3] += cj * tmp(0) + bj * tmp(1);
      e_total[4] += dj * tmp(0) + bj * tmp(2);
      e_total[5] += dj * tmp(1) + cj * tmp(2);
    }
    for (int i = 0; i < 6; ++i) e_elastic[i] = e_total[i] - T.plastic[i];
    double norm_elastic = 0;
    for (int i = 0; i < 6; ++i) norm_elastic += e_elastic[i] * e_elastic[i];
    norm_elastic = std::sqrt(norm_elastic);
    if (norm_elastic > T.yield) {
      double amount = dt * std::min(T.creep, (1.0 / dt));
      for (int i = 0; i < 6; ++i) T.plastic[i] += amount * e_elastic[i];
    }
    double norm_plastic = 0;
    for (int i = 0; i < 6; ++i) norm_plastic += T.plastic[i] * T.plastic[i];
    norm_plastic = std::sqrt(norm_plastic);
    if (norm_plastic > T.cmax) {
      double scale = T.cmax / norm_plastic;
      for (int i = 0; i < 6; ++i) T.plastic[i] *= scale;
    }
    for (int j = 0; j < 4; ++j) {
      const double *plastic = T.plastic;
      double bj = T.B[j](0), cj = T.B[j](1), dj = T.B[j](2);
      double E0 = T.D(0), E1 = T.D(1), E2 = T.D(2);
      double bjE0 = bj * E0, bjE1 = bj * E1, bjE2 = bj * E2;
      double cjE0 = cj * E0, cjE1 = cj * E1, cjE2 = cj * E2;
      double djE0 = dj * E0, djE1 = dj * E1, djE2 = dj * E2;
      V3 f;
      f(0) = bjE0 * plastic[0] + bjE1 * plastic[1] + bjE1 * plastic[2] + cjE2 * plastic[3] +
             djE2 * plastic[4];
      f(1) = cjE1 * plastic[0] + cjE0 * plastic[1] + cjE1 * plastic[2] + bjE2 * plastic[3] +
             +djE2 * plastic[5];
      f(2) = djE1 * plastic[0] + djE1 * plastic[1] + djE0 * plastic[2] + bjE2 * plastic[4] +
             cjE2 * plastic[5];
      f *= T.V;
      m.nodes[T.n[j]].f_external += T.Re * f;
    }
  }
}

// fem_dynamics_assembly.h:47-93, with the beta extension described above.
static void dynamics_assembly(Mesh &m, double mass_damping, double beta, double dt) {
  const int N = (int)m.nodes.size();
  const double kscale = dt * dt + beta * dt;
  for (int i = 0; i < N; ++i) {
    Node &n_i = m.nodes[i];
    V3 &b_i = n_i.b;
    double m_i = n_i.mass;
    b_i.clear();
    for (auto &kv : n_i.K_row) {
      int j = kv.first;
      const M3 &K_ij = kv.second;
      const V3 &x_j = m.nodes[j].coord;
      M3 &A_ij = n_i.A_row[j];
      A_ij = K_ij * kscale;
      b_i -= K_ij * x_j;
      if (i == j) {
        double c_i = mass_damping * m_i;
        double tmp = m_i + dt * c_i;
        A_ij(0, 0) += tmp;
        A_ij(1, 1) += tmp;
        A_ij(2, 2) += tmp;
      }
    }
    b_i -= n_i.f0;
    b_i += n_i.f_external;
    b_i *= dt;
    b_i += n_i.velocity * m_i;
  }
}

// fem_conjugate_gradients.h:25-133, with the relative-residual extension.
static void conjugate_gradients(Mesh &m, const orc_params &P, orc_stats *st) {
  const double tiny = P.tiny;
  const int N = (int)m.nodes.size();
  double bb = 0.0;
  for (int i = 0; i < N; ++i) {
    Node &n_i = m.nodes[i];
    if (n_i.fixed) continue;
    n_i.residual = n_i.b;
    for (auto &kv : n_i.A_row) n_i.residual -= kv.second * m.nodes[kv.first].velocity;
    n_i.prev = n_i.residual;
    bb += n_i.b * n_i.b;
  }
  double rr0 = 0.0;
  for (int i = 0; i < N; ++i)
    if (!m.nodes[i].fixed) rr0 += m.nodes[i].residual * m.nodes[i].residual;
  const double rel_stop = P.tol * std::sqrt(bb);
  int done_iterations = 0;
  int converged = 0;
  double last_rr = rr0;
  if (P.tol_mode == 1 && P.cg_min_iterations <= 0 && std::sqrt(rr0) <= rel_stop) {
    converged = 1;
  } else {
    for (int iteration = 0; iteration < P.cg_max_iterations; ++iteration) {
      double d = 0.0, d2 = 0.0;
      for (int i = 0; i < N; ++i) {
        Node &n_i = m.nodes[i];
        if (n_i.fixed) continue;
        n_i.update.clear();
        for (auto &kv : n_i.A_row) n_i.update += kv.second * m.nodes[kv.first].prev;
        d += n_i.residual * n_i.residual;
        d2 += n_i.prev * n_i.update;
      }
      if (std::fabs(d2) < tiny) d2 = tiny;
      double d3 = d / d2;
      double d1 = 0.0;
      for (int i = 0; i < N; ++i) {
        Node &n_i = m.nodes[i];
        if (n_i.fixed) continue;
        n_i.velocity += n_i.prev * d3;
        n_i.residual -= n_i.update * d3;
        d1 += n_i.residual * n_i.residual;
      }
      done_iterations = iteration + 1;
      last_rr = d1;
      if (P.tol_mode == 0) {
        if (iteration >= P.cg_min_iterations && d1 < P.tol) { converged = 1; break; }
      } else {
        if (iteration + 1 >= P.cg_min_iterations && std::sqrt(d1) <= rel_stop) {
          converged = 1;
          break;
        }
      }
      if (std::fabs(d) < tiny) d = tiny;
      double d4 = d1 / d;
      for (int i = 0; i < N; ++i) {
        Node &n_i = m.nodes[i];
        if (n_i.fixed) continue;
        n_i.prev = n_i.residual + n_i.prev * d4;
      }
    }
  }
  if (st) {
    st->iterations = done_iterations;
    st->converged = converged;
    st->rr0 = rr0;
    st->rr = last_rr;
    st->bb = bb;
  }
}

static void position_update(Mesh &m, double dt) {  // fem_position_update.h:25-39
  for (Node &n : m.nodes) {
    if (n.fixed) continue;
    n.coord += dt * n.velocity;
  }
}

// fem_simulate.h:49-113
int orc_simulate(void *h, double dt, int use_stiffness_warping, const orc_params *P,
                 orc_stats *st) {
  Mesh &m = *static_cast<Mesh *>(h);
  clear_stiffness_assembly(m);
  if (use_stiffness_warping)
    update_orientation(m, P->rotation_mode);
  else
    reset_orientation(m);
  stiffness_assembly(m);
  add_plasticity_force(m, dt);
  dynamics_assembly(m, P->alpha, P->beta, dt);
  conjugate_gradients(m, *P, st);
  position_update(m, dt);
  return 0;
}

void orc_default_params(orc_params *P) {
  P->alpha = 2.0;
  P->beta = 0.0;
  P->cg_min_iterations = 20;
  P->cg_max_iterations = 20;
  P->tol_mode = 0;
  P->tol = 0.001;
  P->tiny = 1e-10;
  P->rotation_mode = 0;
}

// fem_update_original_coord.h:28-33
void orc_update_original_coord(void *h) {
  Mesh &m = *static_cast<Mesh *>(h);
  for (Node &n : m.nodes) n.model_coord = n.coord;
}

void orc_set_state(void *h, const double *x, const double *v) {
  Mesh &m = *static_cast<Mesh *>(h);
  for (size_t i = 0; i < m.nodes.size(); ++i) {
    if (x) m.nodes[i].coord = V3(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
    if (v) m.nodes[i].velocity = V3(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
  }
}

void orc_get_state(void *h, double *x, double *v) {
  Mesh &m = *static_cast<Mesh *>(h);
  for (size_t i = 0; i < m.nodes.size(); ++i) {
    const Node &n = m.nodes[i];
    if (x) { x[3 * i] = n.coord.x; x[3 * i + 1] = n.coord.y; x[3 * i + 2] = n.coord.z; }
    if (v) { v[3 * i] = n.velocity.x; v[3 * i + 1] = n.velocity.y; v[3 * i + 2] = n.velocity.z; }
  }
}

void orc_set_fixed(void *h, const unsigned char *fixed) {
  Mesh &m = *static_cast<Mesh *>(h);
  for (size_t i = 0; i < m.nodes.size(); ++i) m.nodes[i].fixed = fixed[i] != 0;
}

void orc_set_external_forces(void *h, const double *f) {
  Mesh &m = *static_cast<Mesh *>(h);
  for (size_t i = 0; i < m.nodes.size(); ++i)
    m.nodes[i].f_external = V3(f[3 * i], f[3 * i + 1], f[3 * i + 2]);
}

void orc_get_external_forces(void *h, double *f) {
  Mesh &m = *static_cast<Mesh *>(h);
  for (size_t i = 0; i < m.nodes.size(); ++i) {
    const V3 &g = m.nodes[i].f_external;
    f[3 * i] = g.x; f[3 * i + 1] = g.y; f[3 * i + 2] = g.z;
  }
}

void orc_get_mass(void *h, double *mass) {
  Mesh &m = *static_cast<Mesh *>(h);
  for (size_t i = 0; i < m.nodes.size(); ++i) mass[i] = m.nodes[i].mass;
}

void orc_get_volume(void *h, double *V) {
  Mesh &m = *static_cast<Mesh *>(h);
  for (size_t t = 0; t < m.tets.size(); ++t) V[t] = m.tets[t].V;
}

void orc_get_rotations(void *h, double *Re) {  // 9 per tet, row-major
  Mesh &m = *static_cast<Mesh *>(h);
  for (size_t t = 0; t < m.tets.size(); ++t)
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) Re[9 * t + 3 * r + c] = m.tets[t].Re(r, c);
}

void orc_get_Ke(void *h, int t, double *Ke) {  // Ke[i][j](r,c) at ((i*4+j)*3+r)*3+c
  Mesh &m = *static_cast<Mesh *>(h);
  const Tet &T = m.tets[t];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) Ke[((i * 4 + j) * 3 + r) * 3 + c] = T.Ke[i][j](r, c);
}

void orc_get_plastic(void *h, double *plastic) {
  Mesh &m = *static_cast<Mesh *>(h);
  for (size_t t = 0; t < m.tets.size(); ++t)
    for (int i = 0; i < 6; ++i) plastic[6 * t + i] = m.tets[t].plastic[i];
}

void orc_get_rhs(void *h, double *b) {
  Mesh &m = *static_cast<Mesh *>(h);
  for (size_t i = 0; i < m.nodes.size(); ++i) {
    b[3 * i] = m.nodes[i].b.x; b[3 * i + 1] = m.nodes[i].b.y; b[3 * i + 2] = m.nodes[i].b.z;
  }
}

void orc_get_f0(void *h, double *f0) {
  Mesh &m = *static_cast<Mesh *>(h);
  for (size_t i = 0; i < m.nodes.size(); ++i) {
    f0[3 * i] = m.nodes[i].f0.x; f0[3 * i + 1] = m.nodes[i].f0.y; f0[3 * i + 2] = m.nodes[i].f0.z;
  }
}

// Number of stored 3x3 blocks in the K' (which=0) or A (which=1) rows.
long orc_row_blocks(void *h, int which) {
  Mesh &m = *static_cast<Mesh *>(h);
  long n = 0;
  for (const Node &nd : m.nodes) n += (long)(which ? nd.A_row.size() : nd.K_row.size());
  return n;
}

// Block-CSR copy of the stored K' (which=0) or A (which=1) rows, all rows and
// all columns (no fixed-node masking); blocks are 9 doubles row-major.
void orc_export_rows(void *h, int which, long *row_ptr, int *cols, double *blocks) {
  Mesh &m = *static_cast<Mesh *>(h);
  long k = 0;
  row_ptr[0] = 0;
  for (size_t i = 0; i < m.nodes.size(); ++i) {
    const auto &row = which ? m.nodes[i].A_row : m.nodes[i].K_row;
    for (const auto &kv : row) {
      cols[k] = kv.first;
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) blocks[9 * k + 3 * r + c] = kv.second(r, c);
      ++k;
    }
    row_ptr[i + 1] = k;
  }
}

// fem_export_k.h:23-68 into a dense 3N x 3N row-major matrix (small meshes):
// fixed rows become identity rows, fixed columns are skipped.
void orc_export_K_dense(void *h, double *K) {
  Mesh &m = *static_cast<Mesh *>(h);
  const long N = (long)m.nodes.size();
  std::memset(K, 0, sizeof(double) * 9 * N * N);
  for (long i = 0; i < N; ++i) {
    const Node &n_i = m.nodes[i];
    for (int r = 0; r < 3; ++r) {
      long row = 3 * i + r;
      if (n_i.fixed) {
        K[row * 3 * N + row] = 1.0;
        continue;
      }
      for (const auto &kv : n_i.K_row) {
        if (m.nodes[kv.first].fixed) continue;
        for (int c = 0; c < 3; ++c) K[row * 3 * N + 3 * kv.first + c] += kv.second(r, c);
      }
    }
  }
}

// --- building blocks exposed for unit tests -------------------------------
void orc_compute_Ke(const double *p, double E, double nu, double *Ke) {
  V3 p0(p[0], p[1], p[2]), p1(p[3], p[4], p[5]), p2(p[6], p[7], p[8]), p3(p[9], p[10], p[11]);
  M3 K[4][4];
  compute_Ke(p0, p1, p2, p3, E, nu, K);
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) Ke[((i * 4 + j) * 3 + r) * 3 + c] = K[i][j](r, c);
}

void orc_compute_Ke_BD(const double *p, double E, double nu, double *Ke) {
  V3 p0(p[0], p[1], p[2]), p1(p[3], p[4], p[5]), p2(p[6], p[7], p[8]), p3(p[9], p[10], p[11]);
  V3 e10 = p1 - p0, e20 = p2 - p0, e30 = p3 - p0;
  double V = compute_volume(e10, e20, e30);
  V3 B[4], D;
  compute_B(e10, e20, e30, V, B);
  compute_isotropic_elasticity_vector(E, nu, D);
  M3 K[4][4];
  compute_Ke_BD(B, D, V, K);
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) Ke[((i * 4 + j) * 3 + r) * 3 + c] = K[i][j](r, c);
}

void orc_ortonormalize(const double *A, double *R) {
  M3 M(A[0], A[1], A[2], A[3], A[4], A[5], A[6], A[7], A[8]);
  M3 Q = ortonormalize(M);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) R[3 * r + c] = Q(r, c);
}

int orc_polar_decomposition(const double *A, double *R, double *S) {
  M3 M(A[0], A[1], A[2], A[3], A[4], A[5], A[6], A[7], A[8]);
  M3 Rm, Sm;
  int ok = polar_eigen(M, Rm, Sm) ? 1 : 0;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      R[3 * r + c] = Rm(r, c);
      S[3 * r + c] = Sm(r, c);
    }
  return ok;
}

}  // extern "C"
