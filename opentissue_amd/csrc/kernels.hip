// otfem device kernels for gfx950 (MI355X).
//
// One implicit step of the corotated linear-elastic FEM of OpenTissue::fem
// (fem_simulate.h:49-113) is a chain of kernels on one stream (captured in a
// hipGraph by the host):
//
//   k_element      one workgroup per tile, FP64 arithmetic.  Per tet copy
//                  (slot): deformation gradient F = I + sum_k (u_k - u_0)
//                  gradN_k^T from displacements, rotation R (Gram-Schmidt of
//                  F's rows as fem_update_orientation.h:191-208 /
//                  math_matrix3x3.h:535-557, or the polar factor), corotated
//                  elastic force -R Ke (R^T x - x0) (fem_stiffness_assembly.h:
//                  179-205 + fem_dynamics_assembly.h:80-91), K'v for the
//                  warm-start residual, optional plasticity
//                  (fem_add_plasticity_force.h:85-199), and the rotated
//                  gradients R gradN_k (FP32) that the SpMV uses.  Owned-node
//                  epilogue: right-hand side b, r = b - A v, p = r
//                  (fem_conjugate_gradients.h:48-67), tile partials of b.b, r.r.
//
//   k_cg_iter(it)  one CG iteration per kernel: completes iteration it-1
//                  (alpha, v += alpha p, r -= alpha q), p = r + beta p for owned
//                  and halo nodes, q = A p matrix-free, tile partials of p.q,
//                  r.q, q.q, r.r (fem_conjugate_gradients.h:69-132).
//   k_finish       completes the last iteration of bodies that reached the
//                  cap, then x += dt v (fem_position_update.h:25-39).
//
// No workgroup waits for another inside a kernel: the kernel boundary is the
// grid barrier (cheaper on this GPU than any in-kernel barrier over hundreds
// of workgroups).  Every workgroup sums its body's tile partials itself, in
// tile order, so all workgroups derive identical alpha / beta / stop
// decisions; device atomics are used only to count stopped bodies.
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

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace otfem {

namespace {

constexpr int kElemBlock = 256;

__device__ __forceinline__ int xcd_remap(int bid, int nb) {
  // Workgroups are dispatched round-robin over the 8 XCDs.  Give each XCD a
  // contiguous range of logical ids, so spatially adjacent tiles (which share
  // halo nodes) run behind the same L2.
  const int nx = 8;
  int q = nb / nx, r = nb % nx;
  int xcd = bid % nx, k = bid / nx;
  return xcd * q + (xcd < r ? xcd : r) + k;
}

// One DPP move of a double (two 32-bit halves); lanes the row mask leaves
// out receive 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROW_MASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over the 64 lanes (called with every lane active), in a fixed order,
// returned to every lane.  Cross-lane steps are DPP moves (no LDS): pairs
// and quads (quad_perm), half rows and rows (mirrors), then row 1 += row 0
// and row 3 += row 2 (row_bcast:15), row 3 += rows 0 + 1 (row_bcast:31);
// lane 63 ends with the total.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_f64<0xb1, 0xf>(v);   // quad_perm [1, 0, 3, 2]
  v += dpp_f64<0x4e, 0xf>(v);   // quad_perm [2, 3, 0, 1]
  v += dpp_f64<0x141, 0xf>(v);  // row_half_mirror
  v += dpp_f64<0x140, 0xf>(v);  // row_mirror
  v += dpp_f64<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_f64<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, 63);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Deterministic block sum (fixed tree), result broadcast to every thread:
// after one barrier every thread adds the waves' sums itself, in wave order.
// `red` holds BLOCK/64 + 1 doubles of LDS.  kReuse adds a trailing barrier,
// so that `red` may be written again right after the call.
template <int BLOCK, bool kReuse = true>
__device__ __forceinline__ double block_sum(double a, double *red) {
  constexpr int W = BLOCK / 64;
  a = wave_sum(a);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) red[w] = a;
  __syncthreads();
  double r = 0.0;
  for (int i = 0; i < W; ++i) r += red[i];
  if (kReuse) __syncthreads();
  return r;
}

// Four sums at once (one tree, one barrier); result broadcast.  `red4`
// holds 4 * (BLOCK/64 + 1) doubles of LDS; kReuse as for block_sum.
template <int BLOCK, bool kReuse = true>
__device__ __forceinline__ double4 block_sum4(double4 a, double *red4) {
  constexpr int W = BLOCK / 64;
  a.x = wave_sum(a.x);
  a.y = wave_sum(a.y);
  a.z = wave_sum(a.z);
  a.w = wave_sum(a.w);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) {
    red4[4 * w] = a.x;
    red4[4 * w + 1] = a.y;
    red4[4 * w + 2] = a.z;
    red4[4 * w + 3] = a.w;
  }
  __syncthreads();
  double4 r = make_double4(0.0, 0.0, 0.0, 0.0);
  for (int i = 0; i < W; ++i) {
    r.x += red4[4 * i];
    r.y += red4[4 * i + 1];
    r.z += red4[4 * i + 2];
    r.w += red4[4 * i + 3];
  }
  if (kReuse) __syncthreads();
  return r;
}

// All four fields of one body's [tile][4] partial records, in tile order.
template <int BLOCK>
__device__ __forceinline__ double4 body_sum4(const double4 *part4, int t0, int t1, double *red4) {
  double4 s = make_double4(0.0, 0.0, 0.0, 0.0);
  for (int t = t0 + (int)threadIdx.x; t < t1; t += BLOCK) {
    const double4 v = part4[t];
    s.x += v.x;
    s.y += v.y;
    s.z += v.z;
    s.w += v.w;
  }
  return block_sum4<BLOCK>(s, red4);
}

__device__ __forceinline__ bool finite(double x) { return isfinite(x); }

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

}  // namespace

// The four tile-local corners of slot s (12 bits each, packed in 6 bytes).
__device__ __forceinline__ ushort4 slot_corners(const StepArgs &A, int s) {
  const uint32_t lo = A.conn_lo[s], hi = A.conn_hi[s];
  ushort4 c;
  c.x = (unsigned short)(lo & 0xfffu);
  c.y = (unsigned short)((lo >> 12) & 0xfffu);
  c.z = (unsigned short)((lo >> 24) | ((hi & 0xfu) << 8));
  c.w = (unsigned short)(hi >> 4);
  return c;
}

__global__ __launch_bounds__(kElemBlock) void k_element(StepArgs A) {
  extern __shared__ double lds[];
  __shared__ double red[kElemBlock / 64 + 1];
  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  const TileInfo ti = A.tiles[tile];
  const int nn = ti.n_owned + ti.n_halo;
  const int no = ti.n_owned;
  double *su = lds;
  double *sv = su + 3 * nn;
  double *sf = sv + 3 * nn;
  double *skv = sf + 3 * no;
  double *sfp = skv + 3 * no;

  if (ti.body_rank == 0 && threadIdx.x == 0) {
    BodyState *w = &A.body[ti.body];
    w->iterations = 0;
    w->stopped = 0;
    w->converged = 0;
    w->flags = 0;
    w->q32_from = INT32_MAX;
  }
  if (tile == 0 && threadIdx.x == 0) *A.done_count = 0;

  for (int l = threadIdx.x; l < nn; l += kElemBlock) {
    int g = l < no ? ti.node_begin + l : A.halo[ti.halo_begin + l - no];
    for (int d = 0; d < 3; ++d) {
      su[3 * l + d] = A.x[3 * (int64_t)g + d] - A.x0[3 * (int64_t)g + d];
      sv[3 * l + d] = A.vw[3 * (int64_t)g + d];
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
  // a slot's corners, rest gradients and D; the next slot's are in flight
  // while one is used (a thread's last load repeats the tile's last slot)
  struct SlotIn {
    ushort4 c;
    double b[3][3], D[3];
  };
  auto load_in = [&](int s, SlotIn &in) {
    in.c = slot_corners(A, s);
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int d = 0; d < 3; ++d) in.b[k][d] = A.rest64[(3 * k + d) * S + s];
#pragma unroll
    for (int k = 0; k < 3; ++k) in.D[k] = A.vd64[k * S + s];
  };
  const int se = ti.slot_begin + ti.n_slots;
  SlotIn cur;
  if (ti.slot_begin + (int)threadIdx.x < se) load_in(ti.slot_begin + threadIdx.x, cur);
  for (int s = ti.slot_begin + threadIdx.x; s < se; s += kElemBlock) {
    SlotIn nxt;
    load_in(min(s + kElemBlock, se - 1), nxt);
    const ushort4 c = cur.c;
    const int cn[4] = {c.x, c.y, c.z, c.w};
    double b[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int d = 0; d < 3; ++d) b[k][d] = cur.b[k][d];
    const double D0 = cur.D[0], D1 = cur.D[1], D2 = cur.D[2];
    cur = nxt;

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
    double btr[3][3];
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
      // rotated gradient and |V| D_k for the per-iteration SpMV
      matvec(R, b[k], btr[k]);
      const double vdk = k == 0 ? D0 : (k == 1 ? D1 : D2);
      if (!A.folded)
        A.bt4[k * S + s] = make_float4((float)btr[k][0], (float)btr[k][1], (float)btr[k][2],
                                       (float)vdk);
      if (A.store64) A.bt8[k * S + s] = make_double4(btr[k][0], btr[k][1], btr[k][2], vdk);
    }
    if (A.folded) {
      // sqrt(E |V|) R grad N_k: K' p = |V| sigma_E(G) bt_k is quadratic in
      // the gradients and D = E |V| c(nu), so E |V| = D0 / c0 folds into
      // them and c(nu) comes from the tile.
      const double sq = sqrt(D0 / A.tile_d[tile].x);
      float4 *bc4 = A.bt4;
      float *bc1 = reinterpret_cast<float *>(A.bt4 + 2 * S);
      bc4[s] = make_float4((float)(sq * btr[0][0]), (float)(sq * btr[0][1]),
                           (float)(sq * btr[0][2]), (float)(sq * btr[1][0]));
      bc4[S + s] = make_float4((float)(sq * btr[1][1]), (float)(sq * btr[1][2]),
                               (float)(sq * btr[2][0]), (float)(sq * btr[2][1]));
      bc1[s] = (float)(sq * btr[2][2]);
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
      // every copy of the tet reads the strain of the start of the step and
      // updates it the same way; the primary copy stores the result apart, so
      // no copy can read a strain another tile has already updated
      if (primary) {
#pragma unroll
        for (int i = 0; i < 6; ++i) A.plastic_next[6 * tet + i] = pl[i];
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

    if (A.ordered) {
      // deterministic mode: the slot's terms go to memory (one contiguous
      // record) and are summed per node in slot order below
      double *o = A.ordered_terms + 36 * (int64_t)s;
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          o[3 * j + d] = fe[j][d];
          o[12 + 3 * j + d] = kv[j][d];
          o[24 + 3 * j + d] = A.plastic_on ? fpl[j][d] : 0.0;
        }
    } else {
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
  }
  __syncthreads();
  if (A.ordered) {
    for (int l = threadIdx.x; l < no; l += kElemBlock) {
      const int64_t g = ti.node_begin + l;
      double acc[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      for (int e = A.ord_begin[g]; e < A.ord_begin[g + 1]; ++e) {
        const int32_t code = A.ord_ent[e];
        const double *o = A.ordered_terms + 36 * (int64_t)(code >> 2) + 3 * (code & 3);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          acc[d] += o[d];
          acc[3 + d] += o[12 + d];
          acc[6 + d] += o[24 + d];
        }
      }
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        sf[3 * l + d] = acc[d];
        skv[3 * l + d] = acc[3 + d];
        if (A.plastic_on) sfp[3 * l + d] = acc[6 + d];
      }
    }
    __syncthreads();
  }

  double bb = 0.0, rr = 0.0;
  for (int l = threadIdx.x; l < no; l += kElemBlock) {
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
      }
      continue;
    }
    const double m = A.mass[g];
    for (int d = 0; d < 3; ++d) {
      // b uses v^n; the residual uses the CG start (fem_conjugate_gradients.h:52-66).
      // One code path for both starts: v^n is read from memory and the start
      // is always written back as CG's iterate (with the plain start vw == v
      // and the store rewrites the value just read), so the arithmetic does
      // not depend on where the start came from.
      const double vs = sv[3 * l + d];
      const double vel = A.v[3 * g + d];
      double bd = (sf[3 * l + d] + fx[d]) * A.dt + vel * m;
      double rd = bd - (A.cM * m * vs + A.cK * skv[3 * l + d]);
      A.rhs[3 * g + d] = bd;
      A.r[3 * g + d] = rd;
      A.v[3 * g + d] = vs;
      bb += bd * bd;
      rr += rd * rd;
    }
  }
  bb = block_sum<kElemBlock>(bb, red);
  rr = block_sum<kElemBlock>(rr, red);
  if (threadIdx.x == 0) {
    A.part_bb[tile] = bb;
    A.part_rr[tile] = rr;  // r_0.r_0 lives at [0 & 1]
  }
}

namespace {

// Stop test for a body entering iteration `it` (it >= 1: after the update of
// iteration it-1 produced r.r = rr).  fem_conjugate_gradients.h:118 for the
// reference's rule (tol_mode 0); relative residual for tol_mode 1.
__device__ __forceinline__ bool stop_rule(const StepArgs &A, int it, double rr, double bb,
                                          bool &converged, bool &nonfinite) {
  converged = false;
  nonfinite = !finite(rr) || !finite(bb);
  if (nonfinite) return true;
  if (it == 0) {
    if (A.tol_mode == 1 && A.cg_min <= 0 && sqrt(rr) <= A.tol * sqrt(bb)) {
      converged = true;
      return true;
    }
    return A.cg_max <= 0;
  }
  if (A.tol_mode == 0) {
    if (it - 1 >= A.cg_min && rr < A.tol) converged = true;
  } else if (it >= A.cg_min && sqrt(rr) <= A.tol * sqrt(bb)) {
    converged = true;
  }
  return converged || it >= A.cg_max;
}

// `kernel` is the CG kernel that records the stop; `stopped` keeps kernel + 1
// so that workgroups of that same kernel still run their prologue.
__device__ __forceinline__ void record_stop(const StepArgs &A, BodyState *w, int it, double rr,
                                            bool converged, bool nonfinite, int kernel) {
  w->iterations = it;
  w->rr = rr;
  w->converged = converged ? 1 : 0;
  w->flags |= nonfinite ? 1 : 0;
  w->stopped = kernel + 1;
  __threadfence();
  atomicAdd(A.done_count, 1);
}

}  // namespace

namespace {

// Raw slot record as loaded (conversion to the math type happens at use, so
// several records can be in flight per thread at little register cost).
template <typename ST>
struct SlotRaw;
template <>
struct SlotRaw<float> {
  ushort4 c;
  float4 q[3];
};
template <>
struct SlotRaw<double> {
  ushort4 c;
  double4 q[3];
};
// Tag of the folded records: nine floats sqrt(E |V|) R grad N_k (k = 1..3)
// per slot, 44 bytes with the corners; D / (E |V|) comes from the tile's one
// Poisson ratio.
struct Folded {};
template <>
struct SlotRaw<Folded> {
  ushort4 c;
  float4 q[2];
  float z;
};

template <typename ST>
__device__ __forceinline__ void load_raw(const StepArgs &A, int64_t S, int s, SlotRaw<ST> &r);
template <>
__device__ __forceinline__ void load_raw<float>(const StepArgs &A, int64_t S, int s,
                                                SlotRaw<float> &r) {
  r.c = slot_corners(A, s);
#pragma unroll
  for (int k = 0; k < 3; ++k) r.q[k] = A.bt4[k * S + s];
}
template <>
__device__ __forceinline__ void load_raw<double>(const StepArgs &A, int64_t S, int s,
                                                 SlotRaw<double> &r) {
  r.c = slot_corners(A, s);
#pragma unroll
  for (int k = 0; k < 3; ++k) r.q[k] = A.bt8[k * S + s];
}
template <>
__device__ __forceinline__ void load_raw<Folded>(const StepArgs &A, int64_t S, int s,
                                                 SlotRaw<Folded> &r) {
  r.c = slot_corners(A, s);
  r.q[0] = A.bt4[s];
  r.q[1] = A.bt4[S + s];
  r.z = reinterpret_cast<const float *>(A.bt4 + 2 * S)[s];
}

// Gradients and D of one slot in the math type; `Dt` is the tile's D (used
// by folded records only).
template <typename MT, typename ST>
__device__ __forceinline__ void unpack_raw(const SlotRaw<ST> &r, const MT Dt[3], MT bt[3][3],
                                           MT D[3]) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    bt[k][0] = (MT)r.q[k].x;
    bt[k][1] = (MT)r.q[k].y;
    bt[k][2] = (MT)r.q[k].z;
    D[k] = (MT)r.q[k].w;
  }
}
template <>
__device__ __forceinline__ void unpack_raw<double, Folded>(const SlotRaw<Folded> &r,
                                                           const double Dt[3], double bt[3][3],
                                                           double D[3]) {
  bt[0][0] = r.q[0].x; bt[0][1] = r.q[0].y; bt[0][2] = r.q[0].z;
  bt[1][0] = r.q[0].w; bt[1][1] = r.q[1].x; bt[1][2] = r.q[1].y;
  bt[2][0] = r.q[1].z; bt[2][1] = r.q[1].w; bt[2][2] = r.z;
  D[0] = Dt[0];
  D[1] = Dt[1];
  D[2] = Dt[2];
}

// (K' p)_e for one slot, per corner: G = sum_k (p_k - p_0) bt_k^T,
// sigma = D(G), f_k = sigma bt_k, f_0 = -(f_1 + f_2 + f_3).
template <typename MT, typename PS>
__device__ __forceinline__ void slot_forces(ushort4 c, const MT bt[3][3], const MT D[3],
                                            const PS *sp, MT f[4][3]) {
  const double p0x = sp[3 * c.x], p0y = sp[3 * c.x + 1], p0z = sp[3 * c.x + 2];
  MT dp[3][3];
  const int cn[3] = {c.y, c.z, c.w};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    dp[k][0] = (MT)((double)sp[3 * cn[k]] - p0x);
    dp[k][1] = (MT)((double)sp[3 * cn[k] + 1] - p0y);
    dp[k][2] = (MT)((double)sp[3 * cn[k] + 2] - p0z);
  }
  MT G[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int q = 0; q < 3; ++q)
      G[3 * r + q] = dp[0][r] * bt[0][q] + dp[1][r] * bt[1][q] + dp[2][r] * bt[2][q];
  MT Sg[9];
  stress(G, D[0], D[1], D[2], Sg);
  f[0][0] = f[0][1] = f[0][2] = (MT)0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    matvec(Sg, bt[k], f[k + 1]);
#pragma unroll
    for (int d = 0; d < 3; ++d) f[0][d] -= f[k + 1][d];
  }
}

// y_owned += (K' p)_e for one slot (LDS atomics).
template <typename MT, typename AT, typename PS>
__device__ __forceinline__ void slot_spmv(ushort4 c, const MT bt[3][3], const MT D[3],
                                          const PS *sp, AT *sy, int no) {
  MT f[4][3];
  slot_forces<MT, PS>(c, bt, D, sp, f);
  const int cc[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (cc[j] < no) {
#pragma unroll
      for (int d = 0; d < 3; ++d) atomicAdd(&sy[3 * cc[j] + d], (AT)f[j][d]);
    }
  }
}

// Deterministic mode: the slot's four corner terms to out[12] (one
// contiguous record in memory); each owned node adds its terms in slot order
// after the slot loop.
template <typename MT, typename PS>
__device__ __forceinline__ void slot_spmv_store(ushort4 c, const MT bt[3][3], const MT D[3],
                                                const PS *sp, double *out) {
  MT f[4][3];
  slot_forces<MT, PS>(c, bt, D, sp, f);
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int d = 0; d < 3; ++d) out[3 * j + d] = (double)f[j][d];
}

// Accumulator tag of the deterministic mode (K'p summed per node in slot
// order, in FP64); AccOf gives the LDS type of the node sums.
struct Ordered {};
// Accumulator tag of the deterministic CG: K'p summed per node in LDS as
// 64-bit fixed point.  Integer addition is exact and associative, so the sum
// does not depend on the order of the LDS atomics.  The tile's scale 2^k is
// chosen from a bound on every node sum, |(K'p)_i| <= kb * max|p| with
// kb = max over owned nodes of sqrt(12) sum trace(Ke) (||K'_e||_2 = ||Ke||_2 <=
// trace(Ke), ||p_e||_2 <= sqrt(12) max|p|), so that terms and sums stay below
// 2^50; a term is rounded to a multiple of 2^-k, an absolute error of
// 2^-51 of that bound.
struct Fixed {};
template <typename AT>
struct AccOf {
  using type = AT;
};
template <>
struct AccOf<Ordered> {
  using type = double;
};
template <>
struct AccOf<Fixed> {
  using type = unsigned long long;
};

// Node sum in LDS -> FP64 (inv = 2^-k of the fixed-point sums).
template <typename T>
__device__ __forceinline__ double acc_value(T v, double) {
  return (double)v;
}
template <>
__device__ __forceinline__ double acc_value<unsigned long long>(unsigned long long v, double inv) {
  return (double)(long long)v * inv;
}

// round(x sc) as a two's-complement integer, for |x sc| < 2^51: adding
// 1.5 * 2^52 leaves the rounded value in the low mantissa bits (the unit in
// the last place is 1 there), with one rounding in the fused multiply-add.
__device__ __forceinline__ unsigned long long to_fixed(double x, double sc) {
  constexpr double kMagic = 6755399441055744.0;  // 1.5 * 2^52
  return (unsigned long long)(__double_as_longlong(__fma_rn(x, sc, kMagic)) -
                              __double_as_longlong(kMagic));
}

// y_owned += (K' p)_e for one slot as fixed point (LDS integer atomics).
template <typename MT, typename PS>
__device__ __forceinline__ void slot_spmv_fixed(ushort4 c, const MT bt[3][3], const MT D[3],
                                                const PS *sp, unsigned long long *sy, int no,
                                                double sc) {
  MT f[4][3];
  slot_forces<MT, PS>(c, bt, D, sp, f);
  const int cc[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (cc[j] < no) {
#pragma unroll
      for (int d = 0; d < 3; ++d) atomicAdd(&sy[3 * cc[j] + d], to_fixed((double)f[j][d], sc));
    }
  }
}

}  // namespace

// Per-thread register budget of k_cg_iter, by block size: tile nodes per
// thread whose loads are all issued before any is used (one batch), slot
// records per buffer of the double-buffered slot loop (FP64 records take
// twice the registers), and the waves per SIMD the register allocation must
// allow.  512 threads at 4 waves per SIMD run two tiles per CU; 256 threads
// at 3 waves per SIMD three (2 node batches: 4 batches at 2 waves measured
// 2.6% slower on the 2,048-body set).
template <int BLOCK>
struct CgShape {
  static constexpr int kNodes = BLOCK <= 512 ? 2 : 1;
  static constexpr int kSlots = BLOCK <= 256 ? 2 : 1;
  static constexpr int kWaves = BLOCK <= 256 ? 3 : 4;
  // the stop test waits for nothing else and the body's tile partials are
  // loaded with the other early loads (with 128 registers per thread: stop
  // test first, partials after the node inputs)
  static constexpr bool kEarlyPartials = BLOCK <= 256;
  // even kernels leave v alone and the next (odd) kernel adds both
  // iterations' updates: 48 B less per owned node in every other kernel for
  // 12 B more in the odd ones (-4.2% CG time on the 2,048-body set).  Not at
  // 512 and 1024 threads, whose 128 registers would spill the extra loads
  // (the 64^3 cube, latency-bound at 512 threads, measured 7.6% slower).
  static constexpr bool kDeferV = BLOCK <= 256;
};

// One tile node's inputs to the CG prologue, as loaded (PT: the storage
// type of the direction p).
template <typename PT>
struct NodeIn {
  int32_t g;  // internal node id
  int32_t fx;
  double m;  // owned nodes only
  double v[3];
  // r and q = A p as the previous kernel stored them: FP64 bits, or (q_in32)
  // the bits of an FP32 value in the low half (one register pair either way)
  uint64_t rb[3], yb[3];
  PT p[3];
  PT pp[3];  // p of two iterations back (odd kernels from 3 on, owned nodes)
};

// Loads the inputs of tile nodes l = base + threadIdx.x + k*BLOCK (k < NB)
// without using any of them, so every load of the batch is in flight at
// once.  Lanes past the node table read node_begin (in bounds) and are never
// used.  With kDefer, v (and, from kernel 3 on, the p of two iterations
// back, still in the buffer this kernel writes p to) is loaded by odd
// kernels only: even kernels defer their update of v to the next kernel.
template <int BLOCK, int NB, typename PT, bool kDefer>
__device__ __forceinline__ void load_nodes(const StepArgs &A, const TileInfo &ti, int it,
                                           int base, const PT *Pp, const PT *Pq,
                                           NodeIn<PT> (&nd)[NB]) {
  const int nn = ti.n_owned + ti.n_halo, no = ti.n_owned;
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    const int l = base + (int)threadIdx.x + k * BLOCK;
    int g = ti.node_begin;
    if (l < no)
      g = ti.node_begin + l;
    else if (l < nn)
      g = __ldg(&A.halo[ti.halo_begin + l - no]);
    nd[k].g = g;
  }
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    const int64_t g = nd[k].g;
    const int l = base + (int)threadIdx.x + k * BLOCK;
    nd[k].fx = A.fixed[g];
    // the mass enters q only, which the previous kernel stored: owned nodes
    // need it for this kernel's q, halo nodes not at all
    nd[k].m = l < no ? A.mass[g] : 0.0;
    if (it > 0) {
#pragma unroll
      for (int d = 0; d < 3; ++d) nd[k].p[d] = Pp[3 * g + d];
      if (!kDefer || (it & 1)) {
#pragma unroll
        for (int d = 0; d < 3; ++d) nd[k].v[d] = l < no ? A.v[3 * g + d] : 0.0;
      }
      if (kDefer && (it & 1) && it >= 3) {
#pragma unroll
        for (int d = 0; d < 3; ++d) nd[k].pp[d] = l < no ? Pq[3 * g + d] : (PT)0;
      }
    }
  }
}

// Loads r and q = A p of the previous kernel for the nodes load_nodes set up,
// from the FP32 copies when that kernel stored them so (q_in32: the same for
// the whole workgroup; the first batch issues these after the other loads,
// once the body's record says which copies they are).  Kernel 0 reads the
// start's residual.
template <int NB, typename PT>
__device__ __forceinline__ void load_rq(const StepArgs &A, int it, const double *Rp,
                                        const float *Rfp, const double *Yp, const float *Yfp,
                                        bool q_in32, NodeIn<PT> (&nd)[NB]) {
  if (q_in32) {  // it >= 1
#pragma unroll
    for (int k = 0; k < NB; ++k)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const int64_t i = 3 * (int64_t)nd[k].g + d;
        nd[k].rb[d] = (uint64_t)__float_as_uint(Rfp[i]);
        nd[k].yb[d] = (uint64_t)__float_as_uint(Yfp[i]);
      }
  } else {
    const double *R = it == 0 ? A.r : Rp;
#pragma unroll
    for (int k = 0; k < NB; ++k)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const int64_t i = 3 * (int64_t)nd[k].g + d;
        nd[k].rb[d] = (uint64_t)__double_as_longlong(R[i]);
        if (it > 0) nd[k].yb[d] = (uint64_t)__double_as_longlong(Yp[i]);
      }
  }
}

// Completes iteration it-1 for one tile node and forms the new direction:
// r = r - alpha q, p = r + beta p (q = A p as the previous kernel stored
// it), v += alpha p for owned nodes; p goes to LDS for the slot loop, r and
// p of owned nodes to LDS and memory, and the mass of owned nodes to LDS (-1
// marks a fixed node).  Fixed nodes carry zero r and p.  p is rounded to its storage type PT before any
// use: the SpMV, the dot products and (next kernel) v += alpha p and
// r -= alpha A p all see the same vector, so r stays the residual of v.
// With kDefer, v is read and written by odd kernels only: kernel it (odd,
// from 3 on) adds alpha_prev p(it-2) and then alpha p(it-1), in the order
// single updates would; an even kernel updates v only when it records the
// body's stop.
template <typename AT, typename PT, bool kDefer>
__device__ __forceinline__ void update_node(const StepArgs &A, int it, bool stop, double alpha,
                                            double alpha_prev, double beta, bool q_in32,
                                            bool q_out32, int l, int no, const NodeIn<PT> &n,
                                            double *Rn, float *Rfn, PT *Pn, PT *sp, double *sr,
                                            AT *sy, double *sm, double &pmax) {
  const int64_t g = n.g;
  double rn[3], pn[3];
  if (n.fx) {
#pragma unroll
    for (int d = 0; d < 3; ++d) rn[d] = pn[d] = 0.0;
  } else {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double r = q_in32 ? (double)__uint_as_float((uint32_t)n.rb[d])
                              : __longlong_as_double((long long)n.rb[d]);
      if (it == 0) {
        rn[d] = r;
      } else {
        const double q = q_in32 ? (double)__uint_as_float((uint32_t)n.yb[d])
                                : __longlong_as_double((long long)n.yb[d]);
        rn[d] = r - alpha * q;  // q = A p of the previous kernel
      }
      // once the body's residual is small r is kept in FP32 like q, rounded
      // before any use
      if (q_out32) rn[d] = (double)(float)rn[d];
      pn[d] = it == 0 ? rn[d] : rn[d] + beta * n.p[d];
    }
    if (it > 0 && l < no) {
      if (!kDefer || (it & 1)) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          double v = n.v[d];
          if (kDefer && it >= 3) v = v + alpha_prev * n.pp[d];
          A.v[3 * g + d] = v + alpha * n.p[d];
        }
      } else if (stop) {
#pragma unroll
        for (int d = 0; d < 3; ++d) A.v[3 * g + d] = A.v[3 * g + d] + alpha * n.p[d];
      }
    }
  }
  PT pf[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    pf[d] = (PT)pn[d];
    pn[d] = (double)pf[d];
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    sp[3 * l + d] = pf[d];
    // a non-finite p makes the bound infinite (max ignores NaN)
    pmax = isfinite(pn[d]) ? fmax(pmax, fabs(pn[d])) : HUGE_VAL;
  }
  if (l < no) {
    // kernel 0 leaves the start's residual where it is, unless it already
    // goes to the FP32 copy
    const bool store_r = !stop && (it > 0 || q_out32);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (store_r) {
        if (q_out32)
          Rfn[3 * g + d] = (float)rn[d];
        else
          Rn[3 * g + d] = rn[d];
      }
      Pn[3 * g + d] = pf[d];
      sr[3 * l + d] = rn[d];
      sy[3 * l + d] = (AT)0;
    }
    sm[l] = n.fx ? -1.0 : n.m;
  }
}

// Owned node l's contribution to the tile partials p.q, r.q, q.q, r.r, and
// its q = A p = cM m p + cK K'p to memory, in FP64 or (q_out32) rounded to
// FP32 first: the partials and the next kernel's r -= alpha q use the same
// value, so the recurrence for r.r and the update of r agree.
template <typename AT, typename PS>
__device__ __forceinline__ void node_partials(const StepArgs &A, int64_t g, int l, double m,
                                              const PS *sp, const double *sr, const AT *sy,
                                              double *Yn, float *Yfn, bool q_out32,
                                              double4 &acc, double inv) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double yd = acc_value<AT>(sy[3 * l + d], inv);
    const double p = sp[3 * l + d], r = sr[3 * l + d];
    double q = A.cM * m * p + A.cK * yd;
    if (q_out32) {
      const float qf = (float)q;
      q = (double)qf;
      Yfn[3 * g + d] = qf;
    } else {
      Yn[3 * g + d] = q;
    }
    acc.x += p * q;
    acc.y += r * q;
    acc.z += q * q;
    acc.w += r * r;
  }
}

// One CG iteration per kernel (fem_conjugate_gradients.h:69-132).  Kernel j
// first completes iteration j-1 -- alpha = r.r / p.Ap from the previous
// kernel's partials (:101-105), v += alpha p (:113), r -= alpha q (:114) --
// then forms p = r + beta p (:124-131) for owned and halo nodes, applies
// q = A p matrix-free, and leaves partial sums of p.q, r.q, q.q and r.r.
// r.r after the update comes from the exact recurrence
// r'.r' = r.r - 2 alpha r.q + alpha^2 q.q, so the stop test and beta need no
// extra pass; alpha itself always uses the directly summed r.r.
// Buffers alternate with j: kernel j reads set (j-1)&1 and writes set j&1.
//
// Latency: the loads that do not depend on alpha and beta -- the first slot
// records and the node inputs of up to kNodes nodes per thread -- are issued
// before the body's partials are summed, so the kernel waits for memory about
// three times (tile record, node ids, everything else) rather than once per
// node loop trip.  The slot loop is double-buffered.
template <int BLOCK, typename MT, typename AT, typename ST, typename PT>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(CgShape<BLOCK>::kWaves)))
void k_cg_iter(StepArgs A, int it) {
  extern __shared__ double lds[];
  __shared__ double red[4 * (BLOCK / 64 + 1)];
  constexpr int NB = CgShape<BLOCK>::kNodes;
  constexpr int U = sizeof(ST) == 8 ? (CgShape<BLOCK>::kSlots + 1) / 2 : CgShape<BLOCK>::kSlots;
  constexpr bool kOrdered = std::is_same<AT, Ordered>::value;
  constexpr bool kFixed = std::is_same<AT, Fixed>::value;
  using Acc = typename AccOf<AT>::type;
  __shared__ unsigned long long s_pmax;  // max |p| over the tile (FP64 bits), fixed point only
  int tile;
  if (A.compact) {
    // workgroup -> active tile; the remap keeps neighbouring tiles on one XCD
    const int nact = *(volatile int *)(A.done_count + 1);
    if ((int)blockIdx.x >= nact) return;
    tile = A.act[xcd_remap(blockIdx.x, nact)];
  } else {
    tile = xcd_remap(blockIdx.x, gridDim.x);
  }
  const TileInfo ti = A.tiles[tile];
  MT Dt[3] = {(MT)0, (MT)0, (MT)0};
  if (std::is_same<ST, Folded>::value) {
    const double4 td = A.tile_d[tile];
    Dt[0] = (MT)td.x;
    Dt[1] = (MT)td.y;
    Dt[2] = (MT)td.z;
  }
  BodyState *bs = &A.body[ti.body];
  // Issued before any of them is used: the body's stop record, the first
  // tile partials of the body sums, the first slot records and the first
  // batch of node inputs (b.b is read after the sums: loading it early
  // measured slower).
  const int st = *(volatile int32_t *)&bs->stopped;
  // with an FP32 direction, q = A p goes to FP32 once the body's residual is
  // small (StepArgs::q32_rr); q32_from is the kernel that first stored it so,
  // and this kernel's workgroups may see it being set to `it` (not <= it - 1)
  constexpr bool kQ32 = std::is_same<PT, float>::value;
  const int q32_from = kQ32 ? bs->q32_from : INT32_MAX;
  if (!CgShape<BLOCK>::kEarlyPartials && st != 0 && st - 1 < it) return;
  const int t0 = ti.body_tile_begin, t1 = ti.body_tile_end;
  const int nt = A.ntiles;
  const int nn = ti.n_owned + ti.n_halo, no = ti.n_owned;
  const int cur = it & 1, prv = (it - 1) & 1;
  double *Rn = A.r + cur * 3 * A.N, *Yn = A.y + cur * 3 * A.N;
  float *Rfn = A.r32 + cur * 3 * A.N, *Yfn = A.y32 + cur * 3 * A.N;
  PT *Pn = reinterpret_cast<PT *>(A.p) + cur * 3 * A.N;
  const double *Rp = A.r + prv * 3 * A.N, *Yp = A.y + prv * 3 * A.N;
  const float *Rfp = A.r32 + prv * 3 * A.N, *Yfp = A.y32 + prv * 3 * A.N;
  const PT *Pp = reinterpret_cast<const PT *>(A.p) + prv * 3 * A.N;
  // alpha of iteration it-2, whose update of v the previous (even) kernel
  // deferred to this one
  constexpr bool kDefer = CgShape<BLOCK>::kDeferV;
  const double alpha_prev = (kDefer && (it & 1) && it >= 3) ? bs->alpha_hist[prv] : 0.0;
  // tile partials: r.r and b.b of the element kernel at it == 0, else
  // p.q, r.q, q.q, r.r of the previous kernel
  const double4 *prevp = reinterpret_cast<const double4 *>(A.part4 + prv * 4 * (int64_t)nt);
  auto partial = [&](int t) {
    return it == 0 ? make_double4(A.part_rr[t], A.part_bb[t], 0.0, 0.0) : prevp[t];
  };
  double4 part = make_double4(0.0, 0.0, 0.0, 0.0);
  if (CgShape<BLOCK>::kEarlyPartials && t0 + (int)threadIdx.x < t1)
    part = partial(t0 + (int)threadIdx.x);

  // first slot records; loads past the tile's last slot re-read that slot
  const int64_t S = A.S;
  const int sb = ti.slot_begin, send = sb + ti.n_slots;
  const bool has_slots = ti.n_slots > 0;
  SlotRaw<ST> ra[U], rb[U];
  if (has_slots) {
#pragma unroll
    for (int u = 0; u < U; ++u)
      load_raw<ST>(A, S, min(sb + (int)threadIdx.x + u * BLOCK, send - 1), ra[u]);
  }
  NodeIn<PT> nd[NB];
  load_nodes<BLOCK, NB, PT, kDefer>(A, ti, it, 0, Pp, Pn, nd);
  const bool q_in32 = kQ32 && __builtin_amdgcn_readfirstlane(q32_from) <= it - 1;  // uniform
  load_rq<NB, PT>(A, it, Rp, Rfp, Yp, Yfp, q_in32, nd);

  // exit if the body stopped in an earlier kernel (a stop recorded by this
  // kernel still needs this workgroup's prologue: v += alpha p)
  if (st != 0 && st - 1 < it) return;
  if (!CgShape<BLOCK>::kEarlyPartials && t0 + (int)threadIdx.x < t1)
    part = partial(t0 + (int)threadIdx.x);

  // body sums in tile order (bodies with more tiles than threads add the
  // rest here); every workgroup of the body gets the same values
  for (int t = t0 + (int)threadIdx.x + BLOCK; t < t1; t += BLOCK) {
    const double4 q = partial(t);
    part.x += q.x;
    part.y += q.y;
    part.z += q.z;
    part.w += q.w;
  }
  if (kFixed && threadIdx.x == 0) s_pmax = 0ull;  // ordered by block_sum4's barriers
  // no trailing barrier: `red` is next written after the barriers around
  // the slot loop
  const double4 s4 = block_sum4<BLOCK, false>(part, red);
  double alpha = 0.0, beta = 0.0, rr, bb;
  if (it == 0) {
    rr = s4.x;
    bb = s4.y;
  } else {
    double pq = s4.x;
    const double rq = s4.y, qq = s4.z, rrd = s4.w;
    bb = bs->bb;
    if (fabs(pq) < A.tiny) pq = A.tiny;  // fem_conjugate_gradients.h:101-102
    alpha = rrd / pq;
    if (!finite(alpha)) {
      if (ti.body_rank == 0 && threadIdx.x == 0) record_stop(A, bs, it - 1, rrd, false, true, it);
      // v keeps the iterate of iteration it-2: add the update deferred to
      // this kernel
      if (kDefer && (it & 1) && it >= 3) {
        for (int l = threadIdx.x; l < no; l += BLOCK) {
          const int64_t g = ti.node_begin + l;
          if (A.fixed[g]) continue;
#pragma unroll
          for (int d = 0; d < 3; ++d) A.v[3 * g + d] = A.v[3 * g + d] + alpha_prev * Pn[3 * g + d];
        }
      }
      return;
    }
    // the next (odd) kernel completes this kernel's update of v
    if (kDefer && !(it & 1) && ti.body_rank == 0 && threadIdx.x == 0) bs->alpha_hist[cur] = alpha;
    rr = fmax(rrd - 2.0 * alpha * rq + alpha * alpha * qq, 0.0);
    double d = rrd;
    if (fabs(d) < A.tiny) d = A.tiny;  // fem_conjugate_gradients.h:121-122
    beta = rr / d;
  }
  bool conv, nonfin;
  const bool stop = stop_rule(A, it, rr, bb, conv, nonfin);
  // every workgroup of the body derives the same rr and bb, so the same choice
  const bool q_out32 = kQ32 && (q_in32 || rr <= A.q32_rr * bb);
  if (ti.body_rank == 0 && threadIdx.x == 0) {
    if (it == 0) {
      bs->bb = bb;
      bs->rr0 = rr;
    }
    if (stop)
      record_stop(A, bs, it, rr, conv, nonfin, it);
    else
      bs->iterations = it;
    if (q_out32 && !q_in32) bs->q32_from = it;
  }

  double *sr = lds;                 // r of owned nodes
  double *sm = sr + 3 * no;         // mass of owned nodes, -1 if fixed
  Acc *sy = reinterpret_cast<Acc *>(sm + no);  // K'p of owned nodes
  // p of every tile node, in its storage type (after the 8-byte arrays)
  PT *sp = reinterpret_cast<PT *>(sm + no + (3 * no * sizeof(Acc) + 7) / 8);

  // prologue: finish iteration it-1 and form the new direction, NB nodes per
  // thread at a time (the first batch was loaded above)
  double pmax = 0.0;
  for (int base = 0;;) {
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      const int l = base + (int)threadIdx.x + k * BLOCK;
      if (l < nn)
        update_node<Acc, PT, kDefer>(A, it, stop, alpha, alpha_prev, beta, q_in32, q_out32, l,
                                     no, nd[k], Rn, Rfn, Pn, sp, sr, sy, sm, pmax);
    }
    base += NB * BLOCK;
    if (base >= nn) break;
    load_nodes<BLOCK, NB, PT, kDefer>(A, ti, it, base, Pp, Pn, nd);
    load_rq<NB, PT>(A, it, Rp, Rfp, Yp, Yfp, q_in32, nd);
  }
  if (stop) return;
  if (kFixed && pmax > 0.0) atomicMax(&s_pmax, (unsigned long long)__double_as_longlong(pmax));
  __syncthreads();
  // fixed-point scale 2^k with kb max|p| 2^k < 2^50 (inv = 2^-k; NaN when the
  // bound is not finite, so the node sums come out non-finite)
  double sc = 1.0, inv = 1.0;
  if constexpr (kFixed) {
    const double bound = A.tile_kb[tile] * __longlong_as_double((long long)s_pmax);
    if (!isfinite(bound)) {
      inv = __longlong_as_double(0x7ff8000000000000ll);
    } else if (bound > 0.0) {
      const int k = min(50 - (ilogb(bound) + 1), 1000);
      sc = ldexp(1.0, k);
      inv = ldexp(1.0, -k);
    }
  }

  // q = K' p over the tile's slots: buffer ra is in use while rb loads, and
  // the other way round
  if (has_slots) {
    for (int s = sb + (int)threadIdx.x; s < send; s += 2 * U * BLOCK) {
#pragma unroll
      for (int u = 0; u < U; ++u) load_raw<ST>(A, S, min(s + (U + u) * BLOCK, send - 1), rb[u]);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (s + u * BLOCK < send) {
          MT bt[3][3], D[3];
          unpack_raw<MT, ST>(ra[u], Dt, bt, D);
          if constexpr (kOrdered)
            slot_spmv_store<MT, PT>(ra[u].c, bt, D, sp, A.ordered_terms + 12 * (int64_t)(s + u * BLOCK));
          else if constexpr (kFixed)
            slot_spmv_fixed<MT, PT>(ra[u].c, bt, D, sp, sy, no, sc);
          else
            slot_spmv<MT, AT, PT>(ra[u].c, bt, D, sp, sy, no);
        }
      }
      if (s + U * BLOCK >= send) break;
#pragma unroll
      for (int u = 0; u < U; ++u)
        load_raw<ST>(A, S, min(s + (2 * U + u) * BLOCK, send - 1), ra[u]);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (s + (U + u) * BLOCK < send) {
          MT bt[3][3], D[3];
          unpack_raw<MT, ST>(rb[u], Dt, bt, D);
          if constexpr (kOrdered)
            slot_spmv_store<MT, PT>(rb[u].c, bt, D, sp,
                                A.ordered_terms + 12 * (int64_t)(s + (U + u) * BLOCK));
          else if constexpr (kFixed)
            slot_spmv_fixed<MT, PT>(rb[u].c, bt, D, sp, sy, no, sc);
          else
            slot_spmv<MT, AT, PT>(rb[u].c, bt, D, sp, sy, no);
        }
      }
    }
  }
  __syncthreads();

  double4 acc = make_double4(0.0, 0.0, 0.0, 0.0);
  for (int l = threadIdx.x; l < no; l += BLOCK) {
    const double m = sm[l];
    if constexpr (kOrdered) {
      // this node's K'p: its slot terms added in slot order (the thread that
      // writes sy[l] here is the one that reads it below)
      const int64_t g = ti.node_begin + l;
      double y0 = 0.0, y1 = 0.0, y2 = 0.0;
      for (int e = A.ord_begin[g]; e < A.ord_begin[g + 1]; ++e) {
        const int32_t code = A.ord_ent[e];
        const double *o = A.ordered_terms + 12 * (int64_t)(code >> 2) + 3 * (code & 3);
        y0 += o[0];
        y1 += o[1];
        y2 += o[2];
      }
      sy[3 * l] = y0;
      sy[3 * l + 1] = y1;
      sy[3 * l + 2] = y2;
    }
    if (m >= 0.0)
      node_partials<Acc, PT>(A, ti.node_begin + l, l, m, sp, sr, sy, Yn, Yfn, q_out32, acc, inv);
  }
  const double4 tot = block_sum4<BLOCK, false>(acc, red);
  if (threadIdx.x == 0) reinterpret_cast<double4 *>(A.part4 + cur * 4 * (int64_t)nt)[tile] = tot;
}

// ---------------------------------------------------------------------------
// Subspace (Galerkin) CG start.  The reference starts CG from v^n
// (fem_conjugate_gradients.h:48-67); here the start is v^n + D c with the
// directions D = [v^n, v^n - v^n-1, ..., v^n-nh+1 - v^n-nh (, v_net - v^n)]
// and c = (D^T A D)^-1 D^T (b - A v^n), the start of smallest A-norm error
// in v^n + span(D) (v^n itself is c = 0).  CG then runs to the same stopping
// rule.  Three kernels:
//   k_sub_gram   per tile, the partials of D^T A D and D^T r_plain: the
//                element part is the bilinear form p^T K'_e q =
//                eps(p)^T D_e eps(q), written as E~^T E~ with E~ = L_e eps
//                (L_e^T L_e = cK D_e), so a wave forms 4 slots x 16
//                directions of E~ per round and adds E~^T E~ with FP64 MFMA
//                (v_mfma_f64_16x16x4f64, 6 per round); the mass part
//                likewise from rows sqrt(cM m) d.
//   k_sub_solve  per body, the Gram sums in tile order and the coefficients.
//   k_sub_start  per tile, w = D c, K' w (one SpMV), v = v^n + w and
//                r = r_plain - A w, with the tile partials of r.r.
// Precision: the velocity differences are kept in FP32 (their rounding is
// relative to the difference, not to v), v^n in FP64, and every Gram product
// and sum is FP64, so nearly dependent directions still give usable
// coefficients; the start and its residual are FP64 and exactly consistent
// (every tile forms w of a node from the same values).

namespace {

constexpr int kSubGramBlock = 768;
constexpr int kSubStartBlock = 256;
constexpr int kSubSolveBlock = 1024;
constexpr int kSubRounds = 4;  // Gram rounds whose slot records are loaded ahead
typedef double v4d __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int sub_pair(int i, int j) { return j * (j + 1) / 2 + i; }

// The step's history: ring slot of this step's newest difference, earlier
// steps since the last reset, the half of vbuf with v^(n-1), the number of
// differences used and of directions.
struct SubState {
  int head, prev, par, nh, k;
};

__device__ __forceinline__ SubState sub_state(const StepArgs &A) {
  SubState s;
  s.head = A.sub_state[0];
  s.prev = A.sub_state[1];
  s.par = A.sub_state[2];
  s.nh = min(min(s.prev, A.sub_hist), A.sub_ring);
  s.k = 1 + s.nh + (A.sub_net ? 1 : 0);
  return s;
}

// The loaded directions of one node: ring entries h[j] for j = 2 .. kSubMax - 1
// (entries 0 and 1 unused) and the network's start vt.
struct SubDirs {
  float4 h[kSubMax];
  double vt[3];
};

// Issues every load of node g's directions at once, with no branch, so a
// node waits for memory once: ring slots past those the step uses repeat the
// last one used, and without the network vt reads v^n.
__device__ __forceinline__ void sub_load(const StepArgs &A, int64_t g, const SubState &q,
                                         SubDirs &s) {
  const int jmax = max(min(q.nh, q.k - 1), 2);
#pragma unroll
  for (int j = 2; j < kSubMax; ++j) {
    int slot = q.head - min(j, jmax) + 1;
    if (slot < 0) slot += A.sub_ring;
    s.h[j] = A.vring[(int64_t)slot * A.N + g];
  }
  const double *vp = A.sub_net ? A.vnet : A.v;
#pragma unroll
  for (int x = 0; x < 3; ++x) s.vt[x] = vp[3 * g + x];
}

// Component x of direction j >= 1 (j < k) at node g as FP32: d_1 = v^n -
// v^(n-1) (given), the older differences from the ring, the network's start
// - v^n.
__device__ __forceinline__ float sub_get(const SubDirs &s, const SubState &q, int j, int x,
                                         const float d1[3], const double vn[3]) {
  const float net = (float)(s.vt[x] - vn[x]);
  if (j == 1) return q.nh >= 1 ? d1[x] : net;
  const float hv = x == 0 ? s.h[j].x : (x == 1 ? s.h[j].y : s.h[j].z);
  return j <= q.nh ? hv : net;
}

// Scale factors of E~ = L eps with L^T L = cK D: eps^T D eps =
// (D0 + 2 D1) (tr e)^2 / 3 + (D0 - D1) |dev e|^2 + D2 |gamma|^2 (the stress
// of fem_compute_ke.h:78-99) as a sum of six squares.
__device__ __forceinline__ void sub_factors(const double D[3], double cK, double f[3]) {
  f[0] = sqrt(cK * (D[0] + 2.0 * D[1]) / 3.0);
  f[1] = sqrt(cK * (D[0] - D[1]));
  f[2] = sqrt(cK * D[2]);
}

// Per-tile partials of D^T A D (A = cM M + cK K') and D^T r_plain.  Owned
// nodes also store v^n (FP64) and v^n - v^(n-1) (FP32, ring slot `head`); no
// kernel of this step reads either before the kernel boundary.
template <typename ST>
__global__ __launch_bounds__(kSubGramBlock) void k_sub_gram(StepArgs A) {
  // LDS: directions [nn][k][3] (FP32); after the passes, the waves' 16 x 16
  // sums
  extern __shared__ double lds_g[];
  constexpr int W = kSubGramBlock / 64;
  __shared__ double gred[W][16];
  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  const TileInfo ti = A.tiles[tile];
  const SubState q = sub_state(A);
  const int k = q.k;
  const int nn = ti.n_owned + ti.n_halo, no = ti.n_owned;
  float *sdir = reinterpret_cast<float *>(lds_g);
  const double *vprev = A.vbuf + (int64_t)q.par * 3 * A.N;
  double *vcur = A.vbuf + (int64_t)(1 - q.par) * 3 * A.N;
  const double cK = A.cK, cM = A.cM;
  // two nodes per thread and trip (a tile has fewer than two nodes per
  // thread): both nodes' loads are issued before either is used; the second
  // node of a trip past the table repeats the first and stores nothing
  for (int l0 = threadIdx.x; l0 < nn; l0 += 2 * kSubGramBlock) {
    const int ln[2] = {l0, l0 + kSubGramBlock < nn ? l0 + kSubGramBlock : l0};
    int64_t g[2];
#pragma unroll
    for (int h = 0; h < 2; ++h)
      g[h] = ln[h] < no ? ti.node_begin + ln[h] : A.halo[ti.halo_begin + ln[h] - no];
    double vn[2][3], vp[2][3];
    bool fx[2];
    SubDirs sd[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
#pragma unroll
      for (int x = 0; x < 3; ++x) {
        vn[h][x] = A.v[3 * g[h] + x];
        vp[h][x] = vprev[3 * g[h] + x];
      }
      fx[h] = A.fixed[g[h]] != 0;
      sub_load(A, g[h], q, sd[h]);
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (h == 1 && ln[1] == ln[0]) break;
      const int l = ln[h];
      float d1[3] = {0.f, 0.f, 0.f};
      if (q.prev >= 1)
#pragma unroll
        for (int x = 0; x < 3; ++x) d1[x] = (float)(vn[h][x] - vp[h][x]);
      if (l < no) {
#pragma unroll
        for (int x = 0; x < 3; ++x) vcur[3 * g[h] + x] = vn[h][x];
        if (q.prev >= 1)
          A.vring[(int64_t)q.head * A.N + g[h]] = make_float4(d1[0], d1[1], d1[2], 0.f);
      }
      float *o = sdir + l * k * 3;
#pragma unroll
      for (int x = 0; x < 3; ++x) o[x] = fx[h] ? 0.f : (float)vn[h][x];
#pragma unroll
      for (int j = 1; j < kSubMax; ++j)
        if (j < k)
#pragma unroll
          for (int x = 0; x < 3; ++x) o[3 * j + x] = fx[h] ? 0.f : sub_get(sd[h], q, j, x, d1, vn[h]);
    }
  }
  __syncthreads();

  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int qs = lane >> 4, dir = lane & 15;  // lane = (tet of the round, direction)
  const bool live = dir < k;
  v4d C0 = {0.0, 0.0, 0.0, 0.0}, C1 = {0.0, 0.0, 0.0, 0.0};
  double Dt[3] = {0.0, 0.0, 0.0}, ft[3] = {0.0, 0.0, 0.0};
  if (std::is_same<ST, Folded>::value) {
    const double4 td = A.tile_d[tile];
    Dt[0] = td.x;
    Dt[1] = td.y;
    Dt[2] = td.z;
    sub_factors(Dt, cK, ft);
  }
  const int64_t S = A.S;
  // element part over the tile's primary slots (each tet once): E~ rows of 4
  // tets per wave and round, the next round's slot record in flight while
  // this one is used (wave-uniform loop)
  const int pb = A.prim_begin[tile], pe = A.prim_begin[tile + 1];
  auto fetch = [&](int i, SlotRaw<ST> &raw, bool &valid) {
    valid = i < pe && live;
    load_raw<ST>(A, S, A.prim_slot[min(i, pe - 1)], raw);
  };
  auto rows = [&](const SlotRaw<ST> &raw, bool valid) {
    double e[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (valid) {
      double bt[3][3], D[3], f[3];
      unpack_raw<double, ST>(raw, Dt, bt, D);
      if (std::is_same<ST, Folded>::value) {
        f[0] = ft[0];
        f[1] = ft[1];
        f[2] = ft[2];
      } else {
        sub_factors(D, cK, f);
      }
      const int cn[4] = {raw.c.x, raw.c.y, raw.c.z, raw.c.w};
      const float *p0 = sdir + ((int64_t)cn[0] * k + dir) * 3;
      double dd[3][3];
#pragma unroll
      for (int kk = 0; kk < 3; ++kk) {
        const float *pk = sdir + ((int64_t)cn[kk + 1] * k + dir) * 3;
#pragma unroll
        for (int x = 0; x < 3; ++x) dd[kk][x] = (double)pk[x] - (double)p0[x];
      }
      double G[9];
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
          G[3 * r + c] = dd[0][r] * bt[0][c] + dd[1][r] * bt[1][c] + dd[2][r] * bt[2][c];
      const double e0 = G[0], e1 = G[4], e2 = G[8];
      e[0] = f[0] * (e0 + e1 + e2);
      e[1] = f[1] * (e0 - e1) * 0.70710678118654752440;
      e[2] = f[1] * (e0 + e1 - 2.0 * e2) * 0.40824829046386301637;
      e[3] = f[2] * (G[1] + G[3]);
      e[4] = f[2] * (G[2] + G[6]);
      e[5] = f[2] * (G[5] + G[7]);
    }
    // two accumulators: consecutive MFMAs do not wait for each other
    C0 = __builtin_amdgcn_mfma_f64_16x16x4f64(e[0], e[0], C0, 0, 0, 0);
    C1 = __builtin_amdgcn_mfma_f64_16x16x4f64(e[1], e[1], C1, 0, 0, 0);
    C0 = __builtin_amdgcn_mfma_f64_16x16x4f64(e[2], e[2], C0, 0, 0, 0);
    C1 = __builtin_amdgcn_mfma_f64_16x16x4f64(e[3], e[3], C1, 0, 0, 0);
    C0 = __builtin_amdgcn_mfma_f64_16x16x4f64(e[4], e[4], C0, 0, 0, 0);
    C1 = __builtin_amdgcn_mfma_f64_16x16x4f64(e[5], e[5], C1, 0, 0, 0);
  };
  // rounds of 4 tets per wave; U rounds' slot records are in flight while
  // the previous U are used (wave-uniform loop; the larger unfolded records
  // keep 2 in flight, within the registers of 3 waves per SIMD)
  constexpr int U = std::is_same<ST, Folded>::value ? kSubRounds : 2;
  constexpr int R = 4 * W;  // tets per round over the workgroup
  int base = pb + 4 * wave + qs;
  if (base - qs < pe) {
    SlotRaw<ST> ra[U], rb[U];
    bool va[U], vb[U];
#pragma unroll
    for (int u = 0; u < U; ++u) fetch(base + u * R, ra[u], va[u]);
    for (;; base += 2 * U * R) {
#pragma unroll
      for (int u = 0; u < U; ++u) fetch(base + (U + u) * R, rb[u], vb[u]);
#pragma unroll
      for (int u = 0; u < U; ++u) rows(ra[u], va[u]);
      if (base - qs + U * R >= pe) break;
#pragma unroll
      for (int u = 0; u < U; ++u) fetch(base + (2 * U + u) * R, ra[u], va[u]);
#pragma unroll
      for (int u = 0; u < U; ++u) rows(rb[u], vb[u]);
      if (base - qs + 2 * U * R >= pe) break;
    }
  }
  // mass part (rows sqrt(cM m) d_x) and D^T r_plain over owned free nodes:
  // four rounds of 4 nodes per wave at a time, their loads issued together
  // (rows of lanes past the k directions or the owned nodes, and of fixed
  // nodes, are zero)
  double gacc = 0.0;
  const int dc = min(dir, k - 1);
  for (int nb = 4 * wave; nb < no; nb += 16 * W) {
    bool ok[4];
    double m[4], rv[4][3];
    int lc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int l = nb + 4 * W * t + qs;
      lc[t] = min(l, no - 1);
      const int64_t g = ti.node_begin + lc[t];
      const bool free_node = A.fixed[g] == 0;
      ok[t] = (l < no) & live & free_node;  // no short circuit: every load is issued
      m[t] = A.mass[g];
#pragma unroll
      for (int x = 0; x < 3; ++x) rv[t][x] = A.r[3 * g + x];
    }
    // the masses are used below only where a node is free; this keeps their
    // loads unconditional and issued with the others
    asm volatile("" : "+v"(m[0]), "+v"(m[1]), "+v"(m[2]), "+v"(m[3]));
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const double sq = sqrt(cM * (ok[t] ? m[t] : 0.0));  // fixed nodes hold DBL_MAX
      const float *d = sdir + (lc[t] * k + dc) * 3;
      double e[3];
#pragma unroll
      for (int x = 0; x < 3; ++x) {
        e[x] = sq * (double)d[x];
        gacc += ok[t] ? (double)d[x] * rv[t][x] : 0.0;
      }
      C0 = __builtin_amdgcn_mfma_f64_16x16x4f64(e[0], e[0], C0, 0, 0, 0);
      C1 = __builtin_amdgcn_mfma_f64_16x16x4f64(e[1], e[1], C1, 0, 0, 0);
      C0 = __builtin_amdgcn_mfma_f64_16x16x4f64(e[2], e[2], C0, 0, 0, 0);
    }
  }
  gacc += __shfl_xor(gacc, 16, 64);
  gacc += __shfl_xor(gacc, 32, 64);
  if (qs == 0) gred[wave][dir] = gacc;
  __syncthreads();  // every wave is done with the staged data: reuse its LDS
  double *cred = lds_g;  // [W][256]
  // v_mfma_f64_16x16x4f64 leaves row lane / 16 + 4 v, column lane % 16 of
  // the 16 x 16 sum in C[v] (checked on the device)
#pragma unroll
  for (int v = 0; v < 4; ++v) cred[wave * 256 + (qs + 4 * v) * 16 + dir] = C0[v] + C1[v];
  __syncthreads();
  for (int e = threadIdx.x; e < kSubPart; e += kSubGramBlock) {
    double s = 0.0;
    if (e < kSubPairs) {
      // pair (i, j), i <= j, at e = j (j + 1) / 2 + i
      int j = 0;
      while ((j + 1) * (j + 2) / 2 <= e) ++j;
      const int i = e - j * (j + 1) / 2;
      for (int w = 0; w < W; ++w) s += cred[w * 256 + i * 16 + j];
    } else {
      for (int w = 0; w < W; ++w) s += gred[w][e - kSubPairs];
    }
    A.sub_part[(int64_t)tile * kSubPart + e] = s;
  }
}

// Coefficients of the start on one wave: lane i holds row i of G = D^T A D
// scaled to unit diagonal; Cholesky by columns with directions dropped whose
// pivot falls below 1e-12 (numerically dependent on the earlier ones, or
// zero); dropped directions get c = 0.  out[kSubMax] receives c.
// Lane j's value of x to every lane (j wave-uniform: a scalar read, no LDS).
__device__ __forceinline__ double lane_f64(double x, int j) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)b, j);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), j);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ void sub_solve_wave(const double *gram, int k, double *out) {
  const int lane = threadIdx.x & 63;
  const bool in = lane < k;
  const double gii = in ? gram[sub_pair(lane, lane)] : 0.0;
  const bool kp = in && gii > 0.0 && isfinite(gii);
  const double si = kp ? 1.0 / sqrt(gii) : 0.0;
  double row[kSubMax];
#pragma unroll
  for (int j = 0; j < kSubMax; ++j) {
    const double sj = lane_f64(si, j);
    row[j] = (in && j < k) ? gram[sub_pair(min(lane, j), max(lane, j))] * si * sj : 0.0;
  }
  const double bi = in ? gram[kSubPairs + lane] * si : 0.0;
  unsigned keep = (unsigned)(__ballot(kp) & ((1ull << kSubMax) - 1));
  double ldiag[kSubMax];
#pragma unroll
  for (int j = 0; j < kSubMax; ++j) {
    ldiag[j] = 1.0;
    if ((keep >> j) & 1u) {
      double d = row[j];
#pragma unroll
      for (int m = 0; m < j; ++m)
        if ((keep >> m) & 1u) d -= row[m] * row[m];
      d = lane_f64(d, j);
      if (!(d > 1e-12)) {
        keep &= ~(1u << j);
      } else {
        const double ljj = sqrt(d);
        ldiag[j] = ljj;
        double v = row[j];
#pragma unroll
        for (int m = 0; m < j; ++m)
          if ((keep >> m) & 1u) v -= row[m] * lane_f64(row[m], j);
        if (lane > j)
          row[j] = v / ljj;
        else if (lane == j)
          row[j] = ljj;
      }
    }
  }
  double y[kSubMax];
#pragma unroll
  for (int j = 0; j < kSubMax; ++j) {
    y[j] = 0.0;
    if ((keep >> j) & 1u) {
      double v = bi;
#pragma unroll
      for (int m = 0; m < j; ++m)
        if ((keep >> m) & 1u) v -= row[m] * y[m];
      y[j] = lane_f64(v, j) / ldiag[j];
    }
  }
  const bool mine = lane < kSubMax && ((keep >> (lane < kSubMax ? lane : 0)) & 1u);
  double cl = 0.0;
#pragma unroll
  for (int j = kSubMax - 1; j >= 0; --j) {
    if ((keep >> j) & 1u) {
      const double t = wave_sum((mine && lane > j) ? row[j] * cl : 0.0);
      const double cj = (y[j] - t) / ldiag[j];
      if (lane == j) cl = cj;
    }
  }
  const bool ok = __ballot(!isfinite(cl)) == 0;
  if (lane < kSubMax) out[lane] = (ok && mine) ? cl * si : 0.0;
}

// One workgroup per body: its tiles' Gram partials summed in a fixed order
// (wave w takes tiles w, w + 4, ..., then the waves in order), then the
// coefficients into sub_coef[body].
__global__ __launch_bounds__(kSubSolveBlock) void k_sub_solve(StepArgs A) {
  constexpr int W = kSubSolveBlock / 64;
  __shared__ double part[W][kSubPart];
  __shared__ double gram[kSubPart];
  const int b = blockIdx.x;
  const SubState q = sub_state(A);
  const int t0 = A.body_tiles[b], t1 = A.body_tiles[b + 1];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double s0 = 0.0, s1 = 0.0;
  const bool two = lane + 64 < kSubPart;
#pragma unroll 4
  for (int t = t0 + wave; t < t1; t += W) {
    const double *p = A.sub_part + (int64_t)t * kSubPart;
    s0 += p[lane];
    if (two) s1 += p[lane + 64];
  }
  part[wave][lane] = s0;
  if (two) part[wave][lane + 64] = s1;
  __syncthreads();
  for (int e = threadIdx.x; e < kSubPart; e += kSubSolveBlock) {
    double s = 0.0;
    for (int w = 0; w < W; ++w) s += part[w][e];
    gram[e] = s;
  }
  __syncthreads();
  if (wave == 0) sub_solve_wave(gram, q.k, A.sub_coef + (int64_t)b * kSubMax);
}

template <int BLOCK, typename AT, typename ST>
__global__ __launch_bounds__(BLOCK) void k_sub_start(StepArgs A) {
  extern __shared__ double lds[];
  __shared__ double red[BLOCK / 64 + 1];
  constexpr bool kOrdered = std::is_same<AT, Ordered>::value;
  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  const TileInfo ti = A.tiles[tile];
  const SubState q = sub_state(A);
  const int k = q.k;
  const int nn = ti.n_owned + ti.n_halo, no = ti.n_owned;
  const double *vcur = A.vbuf + (int64_t)(1 - q.par) * 3 * A.N;  // v^n
  // the body's coefficients, loaded once (every entry is written, 0 past k)
  double cf[kSubMax];
#pragma unroll
  for (int j = 0; j < kSubMax; ++j) cf[j] = A.sub_coef[(int64_t)ti.body * kSubMax + j];
  double *sp = lds;         // w = D c at every tile node
  double *sy = sp + 3 * nn; // K' w at owned nodes
  for (int l = threadIdx.x; l < nn; l += BLOCK) {
    const int64_t g = l < no ? ti.node_begin + l : A.halo[ti.halo_begin + l - no];
    // every load of the node is issued before any value is used (fixed nodes
    // load too and then take w = 0)
    const double vn[3] = {vcur[3 * g], vcur[3 * g + 1], vcur[3 * g + 2]};
    const float4 h1 = A.vring[(int64_t)q.head * A.N + g];
    SubDirs sd;
    sub_load(A, g, q, sd);
    const bool fx = A.fixed[g] != 0;
    // d_1 is used only with a history (nh >= 1); this keeps its load with the
    // others rather than in a branch of its own
    asm volatile("" ::"v"(h1.x), "v"(h1.y), "v"(h1.z));
    const float d1[3] = {h1.x, h1.y, h1.z};
    double wv[3];
    const double c0 = cf[0];
#pragma unroll
    for (int x = 0; x < 3; ++x) wv[x] = c0 * vn[x];
#pragma unroll
    for (int j = 1; j < kSubMax; ++j) {
      if (j < k) {
        const double cj = cf[j];
#pragma unroll
        for (int x = 0; x < 3; ++x) wv[x] += cj * (double)sub_get(sd, q, j, x, d1, vn);
      }
    }
#pragma unroll
    for (int x = 0; x < 3; ++x) wv[x] = fx ? 0.0 : wv[x];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      sp[3 * l + x] = wv[x];
      if (l < no) sy[3 * l + x] = 0.0;
    }
  }
  __syncthreads();
  double Dt[3] = {0.0, 0.0, 0.0};
  if (std::is_same<ST, Folded>::value) {
    const double4 td = A.tile_d[tile];
    Dt[0] = td.x;
    Dt[1] = td.y;
    Dt[2] = td.z;
  }
  const int64_t S = A.S;
  // the next slot's record is in flight while one is applied (a thread's
  // last load repeats the tile's last slot)
  const int se = ti.slot_begin + ti.n_slots;
  int s = ti.slot_begin + (int)threadIdx.x;
  SlotRaw<ST> raw;
  if (s < se) load_raw<ST>(A, S, s, raw);
  for (; s < se; s += BLOCK) {
    SlotRaw<ST> next;
    load_raw<ST>(A, S, min(s + BLOCK, se - 1), next);
    double bt[3][3], D[3];
    unpack_raw<double, ST>(raw, Dt, bt, D);
    if constexpr (kOrdered)
      slot_spmv_store<double, double>(raw.c, bt, D, sp, A.ordered_terms + 12 * (int64_t)s);
    else
      slot_spmv<double, double, double>(raw.c, bt, D, sp, sy, no);
    raw = next;
  }
  __syncthreads();
  double rr = 0.0;
  for (int l = threadIdx.x; l < no; l += BLOCK) {
    const int64_t g = ti.node_begin + l;
    if constexpr (kOrdered) {
      double y0 = 0.0, y1 = 0.0, y2 = 0.0;
      for (int e = A.ord_begin[g]; e < A.ord_begin[g + 1]; ++e) {
        const int32_t code = A.ord_ent[e];
        const double *o = A.ordered_terms + 12 * (int64_t)(code >> 2) + 3 * (code & 3);
        y0 += o[0];
        y1 += o[1];
        y2 += o[2];
      }
      sy[3 * l] = y0;
      sy[3 * l + 1] = y1;
      sy[3 * l + 2] = y2;
    }
    if (A.fixed[g]) continue;
    const double m = A.mass[g];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      const double wv = sp[3 * l + x];
      const double r = A.r[3 * g + x] - (A.cM * m * wv + A.cK * sy[3 * l + x]);
      A.r[3 * g + x] = r;
      A.v[3 * g + x] += wv;
      rr += r * r;
    }
  }
  rr = block_sum<BLOCK>(rr, red);
  if (threadIdx.x == 0) A.part_rr[tile] = rr;
}

// The history moves on: the slot just written (if any) becomes the newest,
// v^n becomes v^(n-1).
__global__ void k_sub_advance(StepArgs A) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    if (A.sub_state[1] >= 1) A.sub_state[0] = (A.sub_state[0] + 1) % A.sub_ring;
    A.sub_state[1] = min(A.sub_state[1] + 1, A.sub_ring + 1);
    A.sub_state[2] = 1 - A.sub_state[2];
  }
}

template <typename AT, typename ST>
hipError_t enqueue_sub(const StepArgs &a, size_t lds_gram, size_t lds_start, hipStream_t s) {
  hipLaunchKernelGGL(k_sub_gram<ST>, dim3(a.ntiles), dim3(kSubGramBlock), lds_gram, s, a);
  hipLaunchKernelGGL(k_sub_solve, dim3(a.B), dim3(kSubSolveBlock), 0, s, a);
  hipLaunchKernelGGL((k_sub_start<kSubStartBlock, AT, ST>), dim3(a.ntiles), dim3(kSubStartBlock),
                     lds_start, s, a);
  hipLaunchKernelGGL(k_sub_advance, dim3(1), dim3(64), 0, s, a);
  return hipGetLastError();
}

}  // namespace

size_t sub_gram_lds_bytes(int max_nodes, int max_owned, int directions) {
  // the directions of every tile node, then (reused) the waves' 16 x 16 sums
  (void)max_owned;
  const size_t staged = sizeof(double) * ((3 * (size_t)max_nodes * directions + 1) / 2);
  return std::max(staged, sizeof(double) * 256 * (kSubGramBlock / 64));
}

size_t sub_gram_static_lds_bytes() { return sizeof(double) * 16 * (kSubGramBlock / 64); }

hipError_t enqueue_subspace(const StepArgs &a, size_t lds_gram, size_t lds_start, hipStream_t s) {
  if (a.sub_hist <= 0 || a.ntiles == 0) return hipSuccess;
  if (a.ordered) {
    if (a.store64) return enqueue_sub<Ordered, double>(a, lds_gram, lds_start, s);
    if (a.folded) return enqueue_sub<Ordered, Folded>(a, lds_gram, lds_start, s);
    return enqueue_sub<Ordered, float>(a, lds_gram, lds_start, s);
  }
  if (a.store64) return enqueue_sub<double, double>(a, lds_gram, lds_start, s);
  if (a.folded) return enqueue_sub<double, Folded>(a, lds_gram, lds_start, s);
  return enqueue_sub<double, float>(a, lds_gram, lds_start, s);
}

// Bodies still running after cg_max iterations complete their last update
// here (v += alpha p, and the update kernel cg_max-1 deferred when it was an
// even kernel of a deferring shape, CgShape::kDeferV) and record their stop; then x += dt v
// (fem_position_update.h:33-38).  This kernel counts as CG kernel cg_max: it
// records the stop as 1 + cg_max, so workgroups of the body that start after
// the body's first tile has recorded it still complete their nodes.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_finish(StepArgs A) {
  __shared__ double red[4 * (BLOCK / 64 + 1)];
  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  const TileInfo ti = A.tiles[tile];
  BodyState *bs = &A.body[ti.body];
  const int stopped = *(volatile int32_t *)&bs->stopped;
  const int it = A.cg_max;
  const bool running = stopped == 0 || stopped - 1 >= it;
  if (running && it > 0) {
    const int t0 = ti.body_tile_begin, t1 = ti.body_tile_end;
    const double *prevp = A.part4 + ((it - 1) & 1) * 4 * (int64_t)A.ntiles;
    const double4 s4 = body_sum4<BLOCK>(reinterpret_cast<const double4 *>(prevp), t0, t1, red);
    double pq = s4.x;
    const double rq = s4.y, qq = s4.z, rrd = s4.w;
    if (fabs(pq) < A.tiny) pq = A.tiny;
    const double alpha = rrd / pq;
    const bool ok = finite(alpha);
    // kernel it-1 was even (it odd, from 3 on): it deferred alpha(it-2) p(it-2)
    const bool two = CgShape<BLOCK>::kDeferV && (it & 1) && it >= 3;
    const double alpha_prev = two ? bs->alpha_hist[(it - 1) & 1] : 0.0;
    if (ok || two) {
      const int64_t off1 = ((it - 1) & 1) * 3 * A.N, off2 = (it & 1) * 3 * A.N;
      const double *P1 = A.p + off1, *P2 = A.p + off2;
      const float *F1 = reinterpret_cast<const float *>(A.p) + off1;
      const float *F2 = reinterpret_cast<const float *>(A.p) + off2;
      for (int l = threadIdx.x; l < ti.n_owned; l += BLOCK) {
        const int64_t g = ti.node_begin + l;
        if (A.fixed[g]) continue;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          double v = A.v[3 * g + d];
          if (two) v = v + alpha_prev * (A.p32 ? (double)F2[3 * g + d] : P2[3 * g + d]);
          if (ok) v = v + alpha * (A.p32 ? (double)F1[3 * g + d] : P1[3 * g + d]);
          A.v[3 * g + d] = v;
        }
      }
    }
    if (ti.body_rank == 0 && threadIdx.x == 0) {
      const double rr = ok ? fmax(rrd - 2.0 * alpha * rq + alpha * alpha * qq, 0.0) : rrd;
      bool conv, nonfin;
      stop_rule(A, it, rr, bs->bb, conv, nonfin);
      bs->iterations = it;
      bs->rr = rr;
      bs->converged = conv ? 1 : 0;
      bs->flags |= (nonfin || !ok) ? 1 : 0;
      bs->stopped = it + 1;
    }
    __syncthreads();
  } else if (running && ti.body_rank == 0 && threadIdx.x == 0) {
    bs->iterations = 0;
    bs->rr = bs->rr0;
    bs->stopped = 1;
  }
  if (ti.body_rank == 0 && threadIdx.x == 0) {
    A.iter_acc[ti.body] += (unsigned long long)(*(volatile int32_t *)&bs->iterations);
    if (!*(volatile int32_t *)&bs->converged) A.unconv_acc[ti.body] += 1ull;
  }
  for (int l = threadIdx.x; l < ti.n_owned; l += BLOCK) {
    const int64_t g = ti.node_begin + l;
    if (A.fixed[g]) continue;
#pragma unroll
    for (int d = 0; d < 3; ++d) A.x[3 * g + d] += A.dt * A.v[3 * g + d];
  }
}

size_t element_lds_bytes(int max_nodes, int max_owned, int plastic_on) {
  return sizeof(double) * (6 * (size_t)max_nodes + (plastic_on ? 9 : 6) * (size_t)max_owned);
}

size_t cg_lds_bytes(int max_nodes, int max_owned, int p_bytes) {
  // p of all tile nodes in its storage type; FP64 r and mass and the K'p
  // accumulators (at most 8 bytes) of owned nodes
  return (size_t)p_bytes * 3 * (size_t)max_nodes + sizeof(double) * 7 * (size_t)max_owned;
}

namespace {

template <int BLOCK, typename MT, typename AT, typename ST, typename PT = double>
hipError_t enqueue_iterations(const StepArgs &a, size_t lds_cg, hipStream_t s, int it0, int it1,
                              int grid) {
  for (int it = it0; it < it1; ++it)
    hipLaunchKernelGGL((k_cg_iter<BLOCK, MT, AT, ST, PT>), dim3(grid), dim3(BLOCK), lds_cg, s, a, it);
  return hipGetLastError();
}

// FP32 directions (a.p32) or FP64 ones.
template <int BLOCK, typename MT, typename AT, typename ST>
hipError_t enqueue_p(const StepArgs &a, size_t lds_cg, hipStream_t s, int it0, int it1, int grid) {
  if (a.p32) return enqueue_iterations<BLOCK, MT, AT, ST, float>(a, lds_cg, s, it0, it1, grid);
  return enqueue_iterations<BLOCK, MT, AT, ST, double>(a, lds_cg, s, it0, it1, grid);
}

// Active-tile list: tiles whose body has not stopped, in tile order, and
// their count in done_count[1].  One workgroup; each thread scans a
// contiguous range of tiles.
constexpr int kCompactBlock = 1024;
__global__ __launch_bounds__(kCompactBlock) void k_compact(StepArgs A) {
  // by body: a running body contributes its contiguous tile range, so each
  // thread reads one stop flag per body (not one per tile), and the list
  // comes out in tile order
  __shared__ int scan[kCompactBlock];
  const int nb = A.B;
  const int per = (nb + kCompactBlock - 1) / kCompactBlock;
  const int b0 = threadIdx.x * per, b1 = min(nb, b0 + per);
  int cnt = 0;
  for (int b = b0; b < b1; ++b)
    if (!*(volatile int32_t *)&A.body[b].stopped) cnt += A.body_tiles[b + 1] - A.body_tiles[b];
  scan[threadIdx.x] = cnt;
  __syncthreads();
  // inclusive Hillis-Steele scan
  for (int off = 1; off < kCompactBlock; off <<= 1) {
    const int v = threadIdx.x >= off ? scan[threadIdx.x - off] : 0;
    __syncthreads();
    scan[threadIdx.x] += v;
    __syncthreads();
  }
  int pos = scan[threadIdx.x] - cnt;
  for (int b = b0; b < b1; ++b)
    if (!*(volatile int32_t *)&A.body[b].stopped)
      for (int t = A.body_tiles[b]; t < A.body_tiles[b + 1]; ++t) A.act[pos++] = t;
  if (threadIdx.x == kCompactBlock - 1) A.done_count[1] = scan[kCompactBlock - 1];
}

template <int BLOCK>
hipError_t enqueue_mode(const StepArgs &a, int mode, size_t lds_cg, hipStream_t s, int it0,
                        int it1, int g) {
  if (a.ordered) {  // deterministic mode: FP64 arithmetic, fixed-point node sums
    if (mode == SPMV_F64_STORE64)
      return enqueue_iterations<BLOCK, double, Fixed, double>(a, lds_cg, s, it0, it1, g);
    if (a.folded) return enqueue_p<BLOCK, double, Fixed, Folded>(a, lds_cg, s, it0, it1, g);
    return enqueue_p<BLOCK, double, Fixed, float>(a, lds_cg, s, it0, it1, g);
  }
  switch (mode) {
    case SPMV_F32: return enqueue_iterations<BLOCK, float, float, float>(a, lds_cg, s, it0, it1, g);
    case SPMV_F32_ACC64:
      return enqueue_iterations<BLOCK, float, double, float>(a, lds_cg, s, it0, it1, g);
    case SPMV_F64:
      if (a.folded) return enqueue_p<BLOCK, double, double, Folded>(a, lds_cg, s, it0, it1, g);
      return enqueue_p<BLOCK, double, double, float>(a, lds_cg, s, it0, it1, g);
    case SPMV_F64_STORE64:
      return enqueue_iterations<BLOCK, double, double, double>(a, lds_cg, s, it0, it1, g);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

namespace {

// plastic := plastic_next once every tile has finished k_element.
__global__ __launch_bounds__(256) void k_plastic_commit(StepArgs A) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < 6 * A.T) A.plastic[i] = A.plastic_next[i];
}

}  // namespace

hipError_t enqueue_element(const StepArgs &a, size_t lds_elem, hipStream_t s) {
  hipLaunchKernelGGL(k_element, dim3(a.ntiles), dim3(kElemBlock), lds_elem, s, a);
  if (a.plastic_on && a.T > 0)
    hipLaunchKernelGGL(k_plastic_commit, dim3((unsigned)((6 * a.T + 255) / 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t enqueue_cg(const StepArgs &a, const LaunchConfig &lc, size_t lds_cg, hipStream_t s,
                      int it0, int it1, int grid) {
  if (a.compact) {
    hipLaunchKernelGGL(k_compact, dim3(1), dim3(kCompactBlock), 0, s, a);
  } else {
    grid = a.ntiles;
  }
  switch (lc.cg_block) {
    case 256: return enqueue_mode<256>(a, lc.spmv_mode, lds_cg, s, it0, it1, grid);
    case 512: return enqueue_mode<512>(a, lc.spmv_mode, lds_cg, s, it0, it1, grid);
    case 1024: return enqueue_mode<1024>(a, lc.spmv_mode, lds_cg, s, it0, it1, grid);
    default: return hipErrorInvalidValue;
  }
}

namespace {

// t4mesh_mesh_coupling.h:224-235 per surface vertex: the barycentric
// combination of its tetrahedron's four current node positions, summed in the
// reference's order (p0 w0 + p1 w1) + p2 w2) + p3 w3.  HBM-bound gather.
constexpr int kSurfaceBlock = 256;
__global__ __launch_bounds__(kSurfaceBlock) void k_surface_update(int64_t n, const int4 *ids,
                                                                  const double4 *w,
                                                                  const double *rest,
                                                                  const double *x, double *out) {
  const int64_t i = (int64_t)blockIdx.x * kSurfaceBlock + threadIdx.x;
  if (i >= n) return;
  const int4 id = ids[i];
  if (id.x < 0) {
#pragma unroll
    for (int d = 0; d < 3; ++d) out[3 * i + d] = rest[3 * i + d];
    return;
  }
  const double4 ww = w[i];
  const double *p0 = x + 3 * (int64_t)id.x, *p1 = x + 3 * (int64_t)id.y;
  const double *p2 = x + 3 * (int64_t)id.z, *p3 = x + 3 * (int64_t)id.w;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double v = p0[d] * ww.x;
    v = v + p1[d] * ww.y;
    v = v + p2[d] * ww.z;
    v = v + p3[d] * ww.w;
    out[3 * i + d] = v;
  }
}

}  // namespace

namespace {

// Vertex normals carried by the element rotations, as the coupling notes of
// t4mesh_mesh_coupling.h:171-176 describe (n = Re n0): Re of the vertex's
// tetrahedron (row-major, caller tetrahedron order) times the rest normal;
// unbound vertices (tet < 0), or all of them when no rotations are stored
// (Re = I), keep n0.
__global__ __launch_bounds__(kSurfaceBlock) void k_surface_normals(int64_t n, const int32_t *tet,
                                                                   const double *Re, int use_re,
                                                                   const double *n0, double *out) {
  const int64_t i = (int64_t)blockIdx.x * kSurfaceBlock + threadIdx.x;
  if (i >= n) return;
  const double a = n0[3 * i], b = n0[3 * i + 1], c = n0[3 * i + 2];
  const int32_t t = tet[i];
  if (t < 0 || !use_re) {
    out[3 * i] = a;
    out[3 * i + 1] = b;
    out[3 * i + 2] = c;
    return;
  }
  const double *R = Re + 9 * (int64_t)t;
#pragma unroll
  for (int r = 0; r < 3; ++r) out[3 * i + r] = (R[3 * r] * a + R[3 * r + 1] * b) + R[3 * r + 2] * c;
}

}  // namespace

hipError_t enqueue_surface_normals(int64_t n, const int32_t *tet, const double *Re, int use_re,
                                   const double *n0, double *out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int64_t blocks = (n + kSurfaceBlock - 1) / kSurfaceBlock;
  hipLaunchKernelGGL(k_surface_normals, dim3((unsigned)blocks), dim3(kSurfaceBlock), 0, s, n, tet, Re,
                     use_re, n0, out);
  return hipGetLastError();
}

hipError_t enqueue_surface_update(int64_t n, const int4 *ids, const double4 *w, const double *rest,
                                  const double *x, double *out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int64_t blocks = (n + kSurfaceBlock - 1) / kSurfaceBlock;
  hipLaunchKernelGGL(k_surface_update, dim3((unsigned)blocks), dim3(kSurfaceBlock), 0, s, n, ids, w,
                     rest, x, out);
  return hipGetLastError();
}

hipError_t enqueue_finish(const StepArgs &a, const LaunchConfig &lc, hipStream_t s) {
  switch (lc.cg_block) {
    case 256: hipLaunchKernelGGL((k_finish<256>), dim3(a.ntiles), dim3(256), 0, s, a); break;
    case 512: hipLaunchKernelGGL((k_finish<512>), dim3(a.ntiles), dim3(512), 0, s, a); break;
    case 1024: hipLaunchKernelGGL((k_finish<1024>), dim3(a.ntiles), dim3(1024), 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace otfem
