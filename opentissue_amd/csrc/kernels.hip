// otfem device kernels for gfx950 (MI355X).
//
// Two kernels make one implicit step of the corotated linear-elastic FEM of
// OpenTissue::fem (fem_simulate.h:49-113):
//
//   k_element  one workgroup per tile, FP64 arithmetic.  Per tet copy (slot):
//              deformation gradient F = I + sum_k (u_k - u_0) gradN_k^T from
//              displacements, rotation R (Gram-Schmidt of F's rows as
//              fem_update_orientation.h:191-208 / math_matrix3x3.h:535-557, or
//              the polar factor), corotated elastic force -R Ke (R^T x - x0)
//              (fem_stiffness_assembly.h:179-205 + fem_dynamics_assembly.h:
//              80-91), K'v for the warm-start residual, optional plasticity
//              (fem_add_plasticity_force.h:85-199), and the rotated gradients
//              R gradN_k (FP32) that the SpMV uses.  Owned-node epilogue: right
//              hand side b, initial residual r = b - A v, p = r
//              (fem_conjugate_gradients.h:48-67), and per-body b.b, r.r.
//
//   k_cg       persistent CG (fem_conjugate_gradients.h:69-132) for every body
//              at once, two grid barriers per iteration:
//                phase S  p = r + beta p (owned + halo), q = A p matrix-free in
//                         FP32 from FP64 differences, p.q        -> alpha
//                phase U  v += alpha p, r -= alpha q, r.r          -> beta, stop
//              then x += dt v (fem_position_update.h:25-39).
//
// Matrix-free operator.  For isotropic Ke_ij = V [D1 b_i b_j^T + D2 b_j b_i^T
// + D2 (b_i.b_j) I] (fem_compute_ke.h:78-99 rewritten), R Ke(b) R^T equals
// Ke(R b) for orthonormal R, so the warped stiffness K' = sum R Ke R^T is
// applied per element with the rotated gradients bt_k = R gradN_k:
//   G = sum_k (p_k - p_0) bt_k^T,  sigma = D(G),  f_k = V sigma bt_k,
//   f_0 = -(f_1 + f_2 + f_3).
// Accumulation is into the tile's owned nodes only (LDS atomics), so node
// sums never leave a workgroup.
#include "kernels.h"

#include <cmath>

namespace otfem {

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

__device__ __forceinline__ int xcd_remap(int bid, int nb) {
  // Workgroups are dispatched round-robin over the 8 XCDs.  Give each XCD a
  // contiguous range of logical ids, so spatially adjacent tiles (which share
  // halo nodes) run behind the same L2.
  const int nx = 8;
  int q = nb / nx, r = nb % nx;
  int xcd = bid % nx, k = bid / nx;
  return xcd * q + (xcd < r ? xcd : r) + k;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Deterministic block sum of two doubles; the result is valid in thread 0.
__device__ __forceinline__ void block_sum2(double &a, double &b, double *red) {
  a = wave_sum(a);
  b = wave_sum(b);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) {
    red[2 * w] = a;
    red[2 * w + 1] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s0 = 0.0, s1 = 0.0;
    for (int i = 0; i < kWaves; ++i) {
      s0 += red[2 * i];
      s1 += red[2 * i + 1];
    }
    a = s0;
    b = s1;
  }
  __syncthreads();
}

__device__ __forceinline__ double ld_agent(const double *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid barrier over the co-resident workgroups of the persistent kernel.
// Bounded: a wait longer than spin_limit ticks raises ctrl->abort, and every
// workgroup leaves at its next barrier, so the kernel always terminates.
__device__ bool grid_barrier(Control *c, unsigned int nblocks, long long spin_limit) {
  __shared__ int s_ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned int g = __hip_atomic_load(&c->bar_gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    unsigned int arrived =
        __hip_atomic_fetch_add(&c->bar_count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (arrived == nblocks - 1) {
      __hip_atomic_store(&c->bar_count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&c->bar_gen, g + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      long long t0 = wall_clock64();
      while (__hip_atomic_load(&c->bar_gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) {
        if (__hip_atomic_load(&c->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
        if (wall_clock64() - t0 > spin_limit) {
          __hip_atomic_store(&c->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
    }
    __threadfence();
    s_ok = __hip_atomic_load(&c->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u;
  }
  __syncthreads();
  return s_ok != 0;
}

// Publishes this tile's two partial sums and tells (uniformly across the
// workgroup) whether it is the last tile of its body to do so in `phase`.
__device__ bool publish_partials(const StepArgs &A, int tile, int body, int phase, int ntiles_body,
                                 double a, double b) {
  __shared__ int s_last;
  if (threadIdx.x == 0) {
    __hip_atomic_store(&A.tile_partial[2 * tile], a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&A.tile_partial[2 * tile + 1], b, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    unsigned int prev = __hip_atomic_fetch_add(&A.body_cnt[2 * body + phase], 1u,
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int last = (prev == (unsigned int)(ntiles_body - 1));
    if (last) {
      __hip_atomic_store(&A.body_cnt[2 * body + phase], 0u, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
      __threadfence();
    }
    s_last = last;
  }
  __syncthreads();
  return s_last != 0;
}

// Sum of the body's tile partials in tile order (fixed reduction tree), valid
// in thread 0.
__device__ void body_total(const StepArgs &A, int body, double &a, double &b, double *red) {
  const int t0 = A.body_tile_begin[body], t1 = A.body_tile_begin[body + 1];
  double sa = 0.0, sb = 0.0;
  for (int t = t0 + (int)threadIdx.x; t < t1; t += kBlock) {
    sa += ld_agent(&A.tile_partial[2 * t]);
    sb += ld_agent(&A.tile_partial[2 * t + 1]);
  }
  block_sum2(sa, sb, red);
  a = sa;
  b = sb;
}

__device__ __forceinline__ bool finite(double x) { return isfinite(x); }

// ---- 3x3 helpers (row-major double[9]) -----------------------------------

__device__ __forceinline__ void gram_schmidt_rows(const double *F, double *R) {
  // math_matrix3x3.h:535-557 on the rows of F
  double r0x = F[0], r0y = F[1], r0z = F[2];
  double r1x = F[3], r1y = F[4], r1z = F[5];
  double l0 = sqrt(r0x * r0x + r0y * r0y + r0z * r0z);
  if (l0 != 0.0) {
    r0x /= l0;
    r0y /= l0;
    r0z /= l0;
  }
  double d = r0x * r1x + r0y * r1y + r0z * r1z;
  r1x -= r0x * d;
  r1y -= r0y * d;
  r1z -= r0z * d;
  double l1 = sqrt(r1x * r1x + r1y * r1y + r1z * r1z);
  if (l1 != 0.0) {
    r1x /= l1;
    r1y /= l1;
    r1z /= l1;
  }
  R[0] = r0x; R[1] = r0y; R[2] = r0z;
  R[3] = r1x; R[4] = r1y; R[5] = r1z;
  // cross(row0,row1) in math_vector3.h:374-377 operand order
  R[6] = r0y * r1z - r1y * r0z;
  R[7] = -r0x * r1z + r1x * r0z;
  R[8] = r0x * r1y - r1x * r0y;
}

// Orthogonal polar factor of F by scaled Newton iteration (Higham 1986):
// X <- (g X + X^{-T} / g) / 2, g = sqrt(||X^{-1}||_F / ||X||_F).  Same limit
// as the eigen-based polar decomposition of math_polar_decomposition.h:49-91
// for non-singular F; singular F falls back to Gram-Schmidt, as the oracle
// does when that decomposition reports failure.
__device__ void polar_rotation(const double *F, double *R) {
  double X[9];
  for (int i = 0; i < 9; ++i) X[i] = F[i];
  for (int it = 0; it < 24; ++it) {
    double c00 = X[4] * X[8] - X[5] * X[7];
    double c01 = X[5] * X[6] - X[3] * X[8];
    double c02 = X[3] * X[7] - X[4] * X[6];
    double c10 = X[2] * X[7] - X[1] * X[8];
    double c11 = X[0] * X[8] - X[2] * X[6];
    double c12 = X[1] * X[6] - X[0] * X[7];
    double c20 = X[1] * X[5] - X[2] * X[4];
    double c21 = X[2] * X[3] - X[0] * X[5];
    double c22 = X[0] * X[4] - X[1] * X[3];
    double det = X[0] * c00 + X[1] * c01 + X[2] * c02;
    double nx = 0.0;
    for (int i = 0; i < 9; ++i) nx += X[i] * X[i];
    if (!(fabs(det) > 1e-300 * nx) || !finite(det)) {
      gram_schmidt_rows(F, R);
      return;
    }
    // X^{-T} = cofactor(X) / det
    double id = 1.0 / det;
    double Y[9] = {c00 * id, c01 * id, c02 * id, c10 * id, c11 * id,
                   c12 * id, c20 * id, c21 * id, c22 * id};
    double ny = 0.0;
    for (int i = 0; i < 9; ++i) ny += Y[i] * Y[i];
    double g = sqrt(sqrt(ny / nx));
    double diff = 0.0;
    for (int i = 0; i < 9; ++i) {
      double xn = 0.5 * (g * X[i] + Y[i] / g);
      diff += (xn - X[i]) * (xn - X[i]);
      X[i] = xn;
    }
    if (diff <= 1e-30) break;
  }
  for (int i = 0; i < 9; ++i) R[i] = X[i];
}

// sigma = D(G) for the reference's Ke structure (fem_compute_ke.h:78-99):
// s_rr = D0 G_rr + D1 (sum_{c != r} G_cc),  s_rc = D2 (G_rc + G_cr).
template <typename T>
__device__ __forceinline__ void stress(const T *G, T D0, T D1, T D2, T *S) {
  S[0] = D0 * G[0] + D1 * (G[4] + G[8]);
  S[4] = D0 * G[4] + D1 * (G[0] + G[8]);
  S[8] = D0 * G[8] + D1 * (G[0] + G[4]);
  T s01 = D2 * (G[1] + G[3]);
  T s02 = D2 * (G[2] + G[6]);
  T s12 = D2 * (G[5] + G[7]);
  S[1] = s01; S[3] = s01;
  S[2] = s02; S[6] = s02;
  S[5] = s12; S[7] = s12;
}

template <typename T>
__device__ __forceinline__ void matvec(const T *M, const T *x, T *y) {
  y[0] = M[0] * x[0] + M[1] * x[1] + M[2] * x[2];
  y[1] = M[3] * x[0] + M[4] * x[1] + M[5] * x[2];
  y[2] = M[6] * x[0] + M[7] * x[1] + M[8] * x[2];
}

// lds layout helpers for the element kernel
struct ElemLds {
  double *u, *v, *f, *kv, *fp;
};

}  // namespace

__global__ __launch_bounds__(kBlock) void k_element(StepArgs A) {
  extern __shared__ double lds[];
  __shared__ double red[2 * kWaves];
  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  const TileInfo ti = A.tiles[tile];
  const int nn = ti.n_owned + ti.n_halo;
  const int no = ti.n_owned;
  double *su = lds;
  double *sv = su + 3 * nn;
  double *sf = sv + 3 * nn;
  double *skv = sf + 3 * no;
  double *sfp = skv + 3 * no;

  for (int l = threadIdx.x; l < nn; l += kBlock) {
    int g = l < no ? ti.node_begin + l : A.halo[ti.halo_begin + l - no];
    for (int d = 0; d < 3; ++d) {
      su[3 * l + d] = A.x[3 * (int64_t)g + d] - A.x0[3 * (int64_t)g + d];
      sv[3 * l + d] = A.v[3 * (int64_t)g + d];
    }
    if (l < no) {
      for (int d = 0; d < 3; ++d) {
        sf[3 * l + d] = 0.0;
        skv[3 * l + d] = 0.0;
        if (A.plastic_on) sfp[3 * l + d] = 0.0;
      }
    }
  }
  __syncthreads();

  const int64_t S = A.S;
  for (int s = ti.slot_begin + threadIdx.x; s < ti.slot_begin + ti.n_slots; s += kBlock) {
    const ushort4 c = A.conn[s];
    const int cn[4] = {c.x, c.y, c.z, c.w};
    double b[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int d = 0; d < 3; ++d) b[k][d] = A.rest64[(3 * k + d) * S + s];
    const double D0 = A.vd64[s], D1 = A.vd64[S + s], D2 = A.vd64[2 * S + s];

    // F = I + sum_k (u_k - u_0) b_k^T  (displacement form: no cancellation
    // between absolute coordinates)
    double du[3][3], dv[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        du[k][d] = su[3 * cn[k + 1] + d] - su[3 * cn[0] + d];
        dv[k][d] = sv[3 * cn[k + 1] + d] - sv[3 * cn[0] + d];
      }
    double F[9], Hv[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        F[3 * r + q] = du[0][r] * b[0][q] + du[1][r] * b[1][q] + du[2][r] * b[2][q];
        Hv[3 * r + q] = dv[0][r] * b[0][q] + dv[1][r] * b[1][q] + dv[2][r] * b[2][q];
      }
    double H[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) H[i] = F[i];
    F[0] += 1.0;
    F[4] += 1.0;
    F[8] += 1.0;

    double R[9];
    if (A.warp) {
      if (A.rotation_mode == 1)
        polar_rotation(F, R);
      else
        gram_schmidt_rows(F, R);
    } else {
      R[0] = 1; R[1] = 0; R[2] = 0; R[3] = 0; R[4] = 1; R[5] = 0; R[6] = 0; R[7] = 0; R[8] = 1;
    }

    // G = R^T F - I = (R^T - I) + R^T H ; Gv = R^T Hv
    double G[9], Gv[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        double rt = R[r] * H[q] + R[3 + r] * H[3 + q] + R[6 + r] * H[6 + q];
        double rtI = R[3 * q + r] - (r == q ? 1.0 : 0.0);
        G[3 * r + q] = rtI + rt;
        Gv[3 * r + q] = R[r] * Hv[q] + R[3 + r] * Hv[3 + q] + R[6 + r] * Hv[6 + q];
      }
    double Sg[9], Sv[9];
    stress(G, D0, D1, D2, Sg);
    stress(Gv, D0, D1, D2, Sv);

    double fe[4][3], kv[4][3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      fe[0][d] = 0.0;
      kv[0][d] = 0.0;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double t[3], w[3], rt[3], rw[3];
      matvec(Sg, b[k], t);
      matvec(Sv, b[k], w);
      matvec(R, t, rt);
      matvec(R, w, rw);
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        fe[k + 1][d] = -rt[d];
        kv[k + 1][d] = rw[d];
        fe[0][d] += rt[d];
        kv[0][d] -= rw[d];
      }
      // rotated gradient for the per-iteration SpMV
      double bt[3];
      matvec(R, b[k], bt);
#pragma unroll
      for (int d = 0; d < 3; ++d) A.bt[(3 * k + d) * S + s] = (float)bt[d];
    }

    const int32_t st = A.slot_tet[s];
    const bool primary = st < 0;
    const int64_t tet = st & 0x7fffffff;
    if (primary && A.store_rotations && A.Re) {
#pragma unroll
      for (int i = 0; i < 9; ++i) A.Re[9 * tet + i] = R[i];
    }

    double fpl[4][3];
    if (A.plastic_on) {
      // fem_add_plasticity_force.h:85-199, with the reference's sign
      // convention B = -grad N (fem_compute_b.h:80-95).
      const double yield = A.plastic_param[3 * tet], creep = A.plastic_param[3 * tet + 1],
                   cmax = A.plastic_param[3 * tet + 2];
      double pl[6];
#pragma unroll
      for (int i = 0; i < 6; ++i) pl[i] = A.plastic[6 * tet + i];
      double et[6] = {-G[0], -G[4], -G[8], -(G[1] + G[3]), -(G[2] + G[6]), -(G[5] + G[7])};
      double ee[6], ne = 0.0;
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        ee[i] = et[i] - pl[i];
        ne += ee[i] * ee[i];
      }
      ne = sqrt(ne);
      if (ne > yield) {
        double amount = A.dt * fmin(creep, 1.0 / A.dt);
#pragma unroll
        for (int i = 0; i < 6; ++i) pl[i] += amount * ee[i];
      }
      double np = 0.0;
#pragma unroll
      for (int i = 0; i < 6; ++i) np += pl[i] * pl[i];
      np = sqrt(np);
      if (np > cmax) {
        double sc = cmax / np;
#pragma unroll
        for (int i = 0; i < 6; ++i) pl[i] *= sc;
      }
      if (primary) {
#pragma unroll
        for (int i = 0; i < 6; ++i) A.plastic[6 * tet + i] = pl[i];
      }
      // |V| D is stored; the reference multiplies D by the signed V, which is
      // positive for every accepted mesh.
      double Bref[4][3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        Bref[1][d] = -b[0][d];
        Bref[2][d] = -b[1][d];
        Bref[3][d] = -b[2][d];
        Bref[0][d] = -Bref[1][d] - Bref[2][d] - Bref[3][d];
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        double bj = Bref[j][0], cj = Bref[j][1], dj = Bref[j][2];
        double f[3];
        f[0] = bj * D0 * pl[0] + bj * D1 * pl[1] + bj * D1 * pl[2] + cj * D2 * pl[3] + dj * D2 * pl[4];
        f[1] = cj * D1 * pl[0] + cj * D0 * pl[1] + cj * D1 * pl[2] + bj * D2 * pl[3] + dj * D2 * pl[5];
        f[2] = dj * D1 * pl[0] + dj * D1 * pl[1] + dj * D0 * pl[2] + bj * D2 * pl[4] + cj * D2 * pl[5];
        matvec(R, f, fpl[j]);
      }
    }

#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int li = cn[j];
      if (li < no) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          atomicAdd(&sf[3 * li + d], fe[j][d]);
          atomicAdd(&skv[3 * li + d], kv[j][d]);
          if (A.plastic_on) atomicAdd(&sfp[3 * li + d], fpl[j][d]);
        }
      }
    }
  }
  __syncthreads();

  double bb = 0.0, rr = 0.0;
  for (int l = threadIdx.x; l < no; l += kBlock) {
    const int64_t g = ti.node_begin + l;
    double fx[3];
    for (int d = 0; d < 3; ++d) {
      fx[d] = A.fext[3 * g + d];
      if (A.plastic_on) {
        fx[d] += sfp[3 * l + d];
        A.fext[3 * g + d] = fx[d];
      }
    }
    if (A.fixed[g]) {
      for (int d = 0; d < 3; ++d) {
        A.rhs[3 * g + d] = 0.0;
        A.r[3 * g + d] = 0.0;
        A.p0[3 * g + d] = 0.0;
      }
      continue;
    }
    const double m = A.mass[g];
    for (int d = 0; d < 3; ++d) {
      double vel = sv[3 * l + d];
      double bd = (sf[3 * l + d] + fx[d]) * A.dt + vel * m;
      double rd = bd - (A.cM * m * vel + A.cK * skv[3 * l + d]);
      A.rhs[3 * g + d] = bd;
      A.r[3 * g + d] = rd;
      A.p0[3 * g + d] = rd;
      bb += bd * bd;
      rr += rd * rd;
    }
  }
  block_sum2(bb, rr, red);
  const int nt_body = A.body_tile_begin[ti.body + 1] - A.body_tile_begin[ti.body];
  if (publish_partials(A, tile, ti.body, 0, nt_body, bb, rr)) {
    double tb, tr;
    body_total(A, ti.body, tb, tr, red);
    if (threadIdx.x == 0) {
      BodyState st;
      st.bb = tb;
      st.rr0 = tr;
      st.d = tr;
      st.d2 = 0.0;
      st.alpha = 0.0;
      st.beta = 0.0;
      st.rr = tr;
      st.iterations = 0;
      st.flags = (finite(tb) && finite(tr)) ? 0 : 1;
      st.converged = 0;
      st.active = 1;
      if (st.flags) st.active = 0;
      if (A.tol_mode == 1 && A.cg_min <= 0 && sqrt(tr) <= A.tol * sqrt(tb)) {
        st.active = 0;
        st.converged = 1;
      }
      if (A.cg_max <= 0) st.active = 0;
      A.body[ti.body] = st;
      __threadfence();
      if (st.active)
        __hip_atomic_fetch_add(&A.ctrl->active_bodies, 1, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

namespace {

// Reads a body's active flag and one scalar once (thread 0) and broadcasts
// both, so every thread of the workgroup takes the same branch.
__device__ bool body_gate(const BodyState *bs, const double *scalar, double &value) {
  __shared__ int s_act;
  __shared__ double s_val;
  if (threadIdx.x == 0) {
    s_act = __hip_atomic_load(&bs->active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_val = ld_agent(scalar);
  }
  __syncthreads();
  const int act = s_act;
  value = s_val;
  __syncthreads();
  return act != 0;
}

__device__ void cg_phase_s(const StepArgs &A, int tile, int it, double *lds, double *red) {
  const TileInfo ti = A.tiles[tile];
  const BodyState *bs = &A.body[ti.body];
  double beta;
  if (!body_gate(bs, &bs->beta, beta)) return;
  const double *pin = (it & 1) ? A.p1 : A.p0;
  double *pout = (it & 1) ? A.p0 : A.p1;
  const int nn = ti.n_owned + ti.n_halo, no = ti.n_owned;
  double *sp = lds;
  float *sy = reinterpret_cast<float *>(sp + 3 * nn);

  for (int l = threadIdx.x; l < nn; l += kBlock) {
    const int64_t g = l < no ? ti.node_begin + l : A.halo[ti.halo_begin + l - no];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double pn = A.r[3 * g + d] + beta * pin[3 * g + d];
      sp[3 * l + d] = pn;
      if (l < no) {
        pout[3 * g + d] = pn;
        sy[3 * l + d] = 0.0f;
      }
    }
  }
  __syncthreads();

  const int64_t S = A.S;
  for (int s = ti.slot_begin + threadIdx.x; s < ti.slot_begin + ti.n_slots; s += kBlock) {
    const ushort4 c = A.conn[s];
    float bt[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int d = 0; d < 3; ++d) bt[k][d] = A.bt[(3 * k + d) * S + s];
    const float D0 = A.vd32[s], D1 = A.vd32[S + s], D2 = A.vd32[2 * S + s];
    const double p0x = sp[3 * c.x], p0y = sp[3 * c.x + 1], p0z = sp[3 * c.x + 2];
    float dp[3][3];
    const int cn[3] = {c.y, c.z, c.w};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      dp[k][0] = (float)(sp[3 * cn[k]] - p0x);
      dp[k][1] = (float)(sp[3 * cn[k] + 1] - p0y);
      dp[k][2] = (float)(sp[3 * cn[k] + 2] - p0z);
    }
    float G[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int q = 0; q < 3; ++q)
        G[3 * r + q] = dp[0][r] * bt[0][q] + dp[1][r] * bt[1][q] + dp[2][r] * bt[2][q];
    float Sg[9];
    stress(G, D0, D1, D2, Sg);
    float f[4][3];
    f[0][0] = f[0][1] = f[0][2] = 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      matvec(Sg, bt[k], f[k + 1]);
#pragma unroll
      for (int d = 0; d < 3; ++d) f[0][d] -= f[k + 1][d];
    }
    const int cc[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (cc[j] < no) {
#pragma unroll
        for (int d = 0; d < 3; ++d) atomicAdd(&sy[3 * cc[j] + d], f[j][d]);
      }
    }
  }
  __syncthreads();

  double pq = 0.0;
  for (int l = threadIdx.x; l < no; l += kBlock) {
    const int64_t g = ti.node_begin + l;
    if (A.fixed[g]) continue;
    const double m = A.mass[g];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      float yd = sy[3 * l + d];
      A.y[3 * g + d] = yd;
      double p = sp[3 * l + d];
      double q = A.cM * m * p + A.cK * (double)yd;
      pq += p * q;
    }
  }
  double unused = 0.0;
  block_sum2(pq, unused, red);
  const int nt_body = A.body_tile_begin[ti.body + 1] - A.body_tile_begin[ti.body];
  if (publish_partials(A, tile, ti.body, 0, nt_body, pq, 0.0)) {
    double d2, z;
    body_total(A, ti.body, d2, z, red);
    if (threadIdx.x == 0) {
      BodyState *w = &A.body[ti.body];
      if (fabs(d2) < A.tiny) d2 = A.tiny;  // fem_conjugate_gradients.h:101-102
      double alpha = w->d / d2;
      w->d2 = d2;
      w->alpha = alpha;
      if (!finite(alpha)) {
        w->flags |= 1;
        w->active = 0;
        __threadfence();
        __hip_atomic_fetch_add(&A.ctrl->active_bodies, -1, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }
}

__device__ void cg_phase_u(const StepArgs &A, int tile, int it, double *red) {
  const TileInfo ti = A.tiles[tile];
  BodyState *bs = &A.body[ti.body];
  double alpha;
  if (!body_gate(bs, &bs->alpha, alpha)) return;
  const double *pn = (it & 1) ? A.p0 : A.p1;
  double rr = 0.0;
  for (int l = threadIdx.x; l < ti.n_owned; l += kBlock) {
    const int64_t g = ti.node_begin + l;
    if (A.fixed[g]) continue;
    const double m = A.mass[g];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double p = pn[3 * g + d];
      double q = A.cM * m * p + A.cK * (double)A.y[3 * g + d];
      A.v[3 * g + d] += alpha * p;
      double rd = A.r[3 * g + d] - alpha * q;
      A.r[3 * g + d] = rd;
      rr += rd * rd;
    }
  }
  double unused = 0.0;
  block_sum2(rr, unused, red);
  const int nt_body = A.body_tile_begin[ti.body + 1] - A.body_tile_begin[ti.body];
  if (publish_partials(A, tile, ti.body, 1, nt_body, rr, 0.0)) {
    double d1, z;
    body_total(A, ti.body, d1, z, red);
    if (threadIdx.x == 0) {
      BodyState *w = bs;
      const int it1 = w->iterations + 1;
      w->iterations = it1;
      w->rr = d1;
      bool stop = false;
      if (!finite(d1)) {
        w->flags |= 1;
        stop = true;
      } else if (A.tol_mode == 0) {
        // fem_conjugate_gradients.h:118 (iteration index is it1 - 1)
        if (it1 - 1 >= A.cg_min && d1 < A.tol) {
          w->converged = 1;
          stop = true;
        }
      } else if (it1 >= A.cg_min && sqrt(d1) <= A.tol * sqrt(w->bb)) {
        w->converged = 1;
        stop = true;
      }
      if (it1 >= A.cg_max) stop = true;
      if (!stop) {
        double d = w->d;
        if (fabs(d) < A.tiny) d = A.tiny;  // fem_conjugate_gradients.h:121-122
        w->beta = d1 / d;
        w->d = d1;
      } else {
        w->active = 0;
        __threadfence();
        __hip_atomic_fetch_add(&A.ctrl->active_bodies, -1, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }
}

}  // namespace

__global__ __launch_bounds__(kBlock) void k_cg(StepArgs A) {
  extern __shared__ double lds[];
  __shared__ double red[2 * kWaves];
  __shared__ int s_active;
  const int lb = xcd_remap(blockIdx.x, gridDim.x);
  const int nb = gridDim.x;
  for (int it = 0; it < A.cg_max; ++it) {
    if (threadIdx.x == 0)
      s_active = __hip_atomic_load(&A.ctrl->active_bodies, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    if (s_active <= 0) break;
    for (int t = lb; t < A.ntiles; t += nb) cg_phase_s(A, t, it, lds, red);
    if (!grid_barrier(A.ctrl, nb, A.spin_limit)) return;
    for (int t = lb; t < A.ntiles; t += nb) cg_phase_u(A, t, it, red);
    if (!grid_barrier(A.ctrl, nb, A.spin_limit)) return;
  }
  // fem_position_update.h:33-38
  for (int t = lb; t < A.ntiles; t += nb) {
    const TileInfo ti = A.tiles[t];
    for (int l = threadIdx.x; l < ti.n_owned; l += kBlock) {
      const int64_t g = ti.node_begin + l;
      if (A.fixed[g]) continue;
#pragma unroll
      for (int d = 0; d < 3; ++d) A.x[3 * g + d] += A.dt * A.v[3 * g + d];
    }
  }
}

size_t element_lds_bytes(int max_nodes, int max_owned, int plastic_on) {
  return sizeof(double) * (6 * (size_t)max_nodes + (plastic_on ? 9 : 6) * (size_t)max_owned);
}

size_t cg_lds_bytes(int max_nodes, int max_owned) {
  return sizeof(double) * 3 * (size_t)max_nodes + sizeof(float) * 3 * (size_t)max_owned;
}

hipError_t launch_element(const StepArgs &a, int block, size_t lds, hipStream_t s) {
  if (block != kBlock) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_element, dim3(a.ntiles), dim3(kBlock), lds, s, a);
  return hipGetLastError();
}

hipError_t launch_cg(const StepArgs &a, int grid, int block, size_t lds, bool cooperative,
                     hipStream_t s) {
  if (block != kBlock) return hipErrorInvalidValue;
  if (cooperative) {
    StepArgs copy = a;
    void *args[] = {&copy};
    return hipLaunchCooperativeKernel(reinterpret_cast<void *>(k_cg), dim3(grid), dim3(kBlock),
                                      args, (unsigned int)lds, s);
  }
  hipLaunchKernelGGL(k_cg, dim3(grid), dim3(kBlock), lds, s, a);
  return hipGetLastError();
}

hipError_t cg_occupancy(int block, size_t lds, int *blocks_per_cu) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_cg, block, lds);
}

}  // namespace otfem
