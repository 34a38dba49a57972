// Lattice bodies, fixed faces and traction loads built on the device
// (otfem_blocks_count / otfem_blocks_build in include/otfem/otfem.h).
//
// Reference behaviour followed (paths relative to the OpenTissue tree):
//   core/containers/t4mesh/util/t4mesh_block_generator.h:32-110
//       node (i, j, k) of an I x J x K block is number (i (J+1) + j)(K+1) + k
//       and sits at (width i, height k, depth j) (:53-65); cube (i, j, k)
//       gives five tetrahedra whose corner choice flips with the parity of
//       i + j + k (:67-106).
//   core/containers/t4mesh/t4mesh_t4boundary_faces.h:87-150
//       boundary faces (those of one tetrahedron); the faces on a bounding
//       plane of a block are exactly the tetrahedron faces with three
//       corners on that plane, which is what the load kernel gathers.
//
// Jitter (no reference counterpart; it makes the irregular test bodies):
// every interior node moves by a caller-supplied (or built-in) random
// displacement; then, round after round, each tetrahedron below a volume
// floor marks its interior corners, and exactly those nodes are moved again
// from their lattice position.  The marked nodes are listed in ascending id
// (an ordered stream compaction), so a caller drawing from a sequential
// random stream gets the same result as a host loop that lists
// unique(nodes of bad tets) in order.
//
// Kernels are one thread per node, cube or tetrahedron; all of this is
// integer work plus a few flops per item, bound by HBM bandwidth and done
// once before otfem_create.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "otfem/otfem.h"

namespace {

thread_local std::string g_err;

otfem_status fail(otfem_status s, const std::string &m) {
  g_err = m;
  return s;
}

struct Body {
  int32_t I, J, K, pad;
  int64_t node0;  // first global node id
  int64_t cube0;  // first global cube id (tets 5 cube0 ..)
  double tx, ty, tz;  // traction on the load face
};

// Body of global item `g` given the bodies' first ids (node0 or cube0).
template <bool kByCube>
__device__ __forceinline__ int find_body(const Body *bd, int nb, int64_t g) {
  int lo = 0, hi = nb - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const int64_t first = kByCube ? bd[mid].cube0 : bd[mid].node0;
    if (first <= g)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Lattice index along world axis a: x -> i, y -> k, z -> j.
__device__ __forceinline__ int axis_index(int a, int i, int j, int k) {
  return a == 0 ? i : (a == 1 ? k : j);
}
__device__ __forceinline__ int axis_extent(int a, const Body &b) {
  return a == 0 ? b.I : (a == 1 ? b.K : b.J);
}

__device__ __forceinline__ bool on_face(int axis, int side, const Body &b, int i, int j, int k) {
  if (axis < 0) return false;
  return axis_index(axis, i, j, k) == (side ? axis_extent(axis, b) : 0);
}

// Corner c (0..7) of cube (i, j, k) as the generator numbers them:
// p0 (i,j,k), p1 (i,j,k+1), p2 (i+1,j,k+1), p3 (i+1,j,k), p4 (i,j+1,k),
// p5 (i,j+1,k+1), p6 (i+1,j+1,k+1), p7 (i+1,j+1,k).
__device__ __forceinline__ void cube_corner(int c, int i, int j, int k, int &ci, int &cj, int &ck) {
  const int di[8] = {0, 0, 1, 1, 0, 0, 1, 1};
  const int dj[8] = {0, 0, 0, 0, 1, 1, 1, 1};
  const int dk[8] = {0, 1, 1, 0, 0, 1, 1, 0};
  ci = i + di[c];
  cj = j + dj[c];
  ck = k + dk[c];
}

// The five tetrahedra of a cube, as corner numbers, for odd and even parity
// of i + j + k (t4mesh_block_generator.h:91-106).
__constant__ int8_t kTetOdd[5][4] = {{1, 2, 6, 3}, {3, 6, 4, 7}, {1, 4, 6, 5}, {1, 3, 4, 0}, {1, 6, 4, 3}};
__constant__ int8_t kTetEven[5][4] = {{2, 0, 5, 1}, {2, 7, 0, 3}, {2, 5, 7, 6}, {0, 7, 5, 4}, {2, 0, 7, 5}};

__device__ __forceinline__ int64_t node_id(const Body &b, int i, int j, int k) {
  return b.node0 + ((int64_t)i * (b.J + 1) + j) * (b.K + 1) + k;
}

__global__ void k_lattice_nodes(const Body *bd, int nb, int64_t N, double w, double h, double d,
                                int fixed_axis, int fixed_side, double *x0, uint8_t *interior,
                                uint8_t *fixed) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const Body b = bd[find_body<false>(bd, nb, n)];
  int64_t r = n - b.node0;
  const int k = (int)(r % (b.K + 1));
  r /= (b.K + 1);
  const int j = (int)(r % (b.J + 1));
  const int i = (int)(r / (b.J + 1));
  x0[3 * n + 0] = w * (double)i;
  x0[3 * n + 1] = h * (double)k;
  x0[3 * n + 2] = d * (double)j;
  interior[n] = (i > 0 && i < b.I && j > 0 && j < b.J && k > 0 && k < b.K) ? 1 : 0;
  fixed[n] = on_face(fixed_axis, fixed_side, b, i, j, k) ? 1 : 0;
}

__global__ void k_lattice_tets(const Body *bd, int nb, int64_t C, int32_t *tets) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  const Body b = bd[find_body<true>(bd, nb, c)];
  int64_t r = c - b.cube0;  // cubes of a body: k fastest, then j, then i
  const int k = (int)(r % b.K);
  r /= b.K;
  const int j = (int)(r % b.J);
  const int i = (int)(r / b.J);
  int32_t p[8];
  for (int q = 0; q < 8; ++q) {
    int ci, cj, ck;
    cube_corner(q, i, j, k, ci, cj, ck);
    p[q] = (int32_t)node_id(b, ci, cj, ck);
  }
  const bool odd = ((i + j + k) & 1) != 0;
  int4 *out = reinterpret_cast<int4 *>(tets) + 5 * c;
  for (int e = 0; e < 5; ++e) {
    const int8_t *q = odd ? kTetOdd[e] : kTetEven[e];
    out[e] = make_int4(p[q[0]], p[q[1]], p[q[2]], p[q[3]]);
  }
}

// Signed volume e10 . (e20 x e30) / 6 (fem_compute_volume.h:27-39), with
// each product and difference rounded on its own (no fused multiply-add), as
// a plain host evaluation rounds it.
__device__ double signed_volume(const double *x, int4 t) {
#pragma clang fp contract(off)
  const double *a = x + 3 * (int64_t)t.x, *b = x + 3 * (int64_t)t.y, *c = x + 3 * (int64_t)t.z,
               *e = x + 3 * (int64_t)t.w;
  const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  const double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  const double e3[3] = {e[0] - a[0], e[1] - a[1], e[2] - a[2]};
  const double cx = e2[1] * e3[2] - e2[2] * e3[1];
  const double cy = e2[2] * e3[0] - e2[0] * e3[2];
  const double cz = e2[0] * e3[1] - e2[1] * e3[0];
  return (e1[0] * cx + e1[1] * cy + e1[2] * cz) / 6.0;
}

// Marks the interior corners of every tetrahedron below the floor; counts
// such tetrahedra.
__global__ void k_volume_floor(const int32_t *tets, int64_t T, const double *x, double vmin,
                               const uint8_t *interior, uint8_t *mark,
                               unsigned long long *n_bad) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  const int4 q = reinterpret_cast<const int4 *>(tets)[t];
  if (!(signed_volume(x, q) < vmin)) return;
  atomicAdd(n_bad, 1ull);
  const int32_t v[4] = {q.x, q.y, q.z, q.w};
  for (int c = 0; c < 4; ++c)
    if (interior[v[c]]) mark[v[c]] = 1;
}

// x[list[m]] = x0[list[m]] + disp[m] for the listed nodes.
__global__ void k_move(const int32_t *list, int64_t n, const double *x0, const double *disp,
                       double *x) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  const int64_t v = list[m];
  for (int d = 0; d < 3; ++d) x[3 * v + d] = x0[3 * v + d] + disp[3 * m + d];
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {  // splitmix64 finaliser
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// Built-in source: displacement of node v on axis a in round r is
// (2u - 1) amp[a], u = (hash(seed, r, v, a) >> 11) 2^-53 in [0, 1).
__global__ void k_builtin_draw(const int32_t *list, int64_t n, uint64_t seed, int round,
                               double ax, double ay, double az, double *disp) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  const uint64_t v = (uint64_t)list[m];
  const double amp[3] = {ax, ay, az};
  for (int a = 0; a < 3; ++a) {
    const uint64_t key = mix64(seed + 0x9e3779b97f4a7c15ull * (1 + ((uint64_t)round << 34) + 3 * v + a));
    const double u = (double)(key >> 11) * 0x1.0p-53;
    disp[3 * m + a] = (2.0 * u - 1.0) * amp[a];
  }
}

// Nodal traction forces of the load face, gathered per node: each face node
// visits the (up to four) cubes of its body that touch it on the face, in a
// fixed order, and adds t A / 3 for every tetrahedron face of those cubes
// that lies on the plane and has the node as a corner.  Fixed nodes get 0.
__global__ void k_face_load(const Body *bd, int nb, int64_t N, const double *x, int axis, int side,
                            const uint8_t *fixed, double *f) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const Body b = bd[find_body<false>(bd, nb, n)];
  int64_t r = n - b.node0;
  const int k = (int)(r % (b.K + 1));
  r /= (b.K + 1);
  const int j = (int)(r % (b.J + 1));
  const int i = (int)(r / (b.J + 1));
  double acc[3] = {0.0, 0.0, 0.0};
  if (on_face(axis, side, b, i, j, k) && !fixed[n]) {
    const int plane = side ? axis_extent(axis, b) : 0;
    const int cube_a = side ? plane - 1 : 0;  // cube layer along the axis
    // the two in-plane lattice directions of the face
    const int u_axis = axis == 0 ? 2 : 0;       // world z (j) or x (i)
    const int v_axis = axis == 1 ? 2 : 1;       // world z (j) or y (k)
    const int u0 = axis_index(u_axis, i, j, k), v0 = axis_index(v_axis, i, j, k);
    const int ue = axis_extent(u_axis, b), ve = axis_extent(v_axis, b);
    for (int du = -1; du <= 0; ++du)
      for (int dv = -1; dv <= 0; ++dv) {
        const int cu = u0 + du, cv = v0 + dv;
        if (cu < 0 || cu >= ue || cv < 0 || cv >= ve) continue;
        int lat[3];  // cube (i, j, k) from (axis, u, v)
        lat[0] = lat[1] = lat[2] = 0;
        auto put = [&](int a, int val) { lat[a == 0 ? 0 : (a == 1 ? 2 : 1)] = val; };
        put(axis, cube_a);
        put(u_axis, cu);
        put(v_axis, cv);
        const int ci = lat[0], cj = lat[1], ck = lat[2];
        const bool odd = ((ci + cj + ck) & 1) != 0;
        for (int e = 0; e < 5; ++e) {
          const int8_t *q = odd ? kTetOdd[e] : kTetEven[e];
          int64_t on[4];
          int cnt = 0;
          bool has_n = false;
          for (int c = 0; c < 4; ++c) {
            int pi, pj, pk;
            cube_corner(q[c], ci, cj, ck, pi, pj, pk);
            if (axis_index(axis, pi, pj, pk) != plane) continue;
            const int64_t id = node_id(b, pi, pj, pk);
            on[cnt++] = id;
            has_n |= (id == n);
          }
          if (cnt != 3 || !has_n) continue;
          const double *p0 = x + 3 * on[0], *p1 = x + 3 * on[1], *p2 = x + 3 * on[2];
          const double a1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
          const double a2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
          const double cx = a1[1] * a2[2] - a1[2] * a2[1];
          const double cy = a1[2] * a2[0] - a1[0] * a2[2];
          const double cz = a1[0] * a2[1] - a1[1] * a2[0];
          const double third = 0.5 * sqrt(cx * cx + cy * cy + cz * cz) / 3.0;
          acc[0] += third * b.tx;
          acc[1] += third * b.ty;
          acc[2] += third * b.tz;
        }
      }
  }
  f[3 * n + 0] = acc[0];
  f[3 * n + 1] = acc[1];
  f[3 * n + 2] = acc[2];
}

// Owns the device buffers and the stream of one build.
struct Build {
  hipStream_t s = nullptr;
  std::vector<void *> bufs;
  hipError_t err = hipSuccess;
  template <class T>
  T *alloc(size_t n) {
    void *p = nullptr;
    if (err == hipSuccess) err = hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T));
    if (err != hipSuccess) return nullptr;
    bufs.push_back(p);
    return static_cast<T *>(p);
  }
  ~Build() {
    if (s) (void)hipStreamSynchronize(s);
    for (void *p : bufs) (void)hipFree(p);
    if (s) (void)hipStreamDestroy(s);
  }
};

inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

otfem_status alloc_failed(const Build &g) {
  return fail(OTFEM_HIP_ERROR, std::string("hipMalloc: ") + hipGetErrorString(g.err));
}

otfem_status check_desc(const otfem_blocks_desc *d, int64_t *N, int64_t *C) {
  if (!d) return fail(OTFEM_INVALID_ARG, "null description");
  if (d->num_bodies < 1 || !d->dims) return fail(OTFEM_INVALID_ARG, "need num_bodies >= 1 and dims");
  if (!(d->width > 0.0) || !(d->height > 0.0) || !(d->depth > 0.0))
    return fail(OTFEM_INVALID_ARG, "cube edges must be positive");
  int64_t n = 0, c = 0;
  for (int64_t b = 0; b < d->num_bodies; ++b) {
    const int64_t I = d->dims[3 * b], J = d->dims[3 * b + 1], K = d->dims[3 * b + 2];
    if (I < 1 || J < 1 || K < 1) return fail(OTFEM_INVALID_ARG, "every body needs I, J, K >= 1");
    n += (I + 1) * (J + 1) * (K + 1);
    c += I * J * K;
    if (n >= INT32_MAX || 5 * c >= INT32_MAX)
      return fail(OTFEM_INVALID_ARG, "too many nodes or tetrahedra for 32-bit ids");
  }
  *N = n;
  *C = c;
  return OTFEM_OK;
}

}  // namespace

extern "C" {

const char *otfem_blocks_last_error(void) { return g_err.c_str(); }

otfem_status otfem_blocks_count(const otfem_blocks_desc *d, int64_t *num_nodes, int64_t *num_tets) {
  int64_t N = 0, C = 0;
  otfem_status st = check_desc(d, &N, &C);
  if (st != OTFEM_OK) return st;
  if (num_nodes) *num_nodes = N;
  if (num_tets) *num_tets = 5 * C;
  return OTFEM_OK;
}

static otfem_status blocks_build(const otfem_blocks_desc *d, double *coords, int32_t *tets,
                                 uint8_t *fixed, double *f_ext, int64_t *body_node_offsets,
                                 int64_t *body_tet_offsets, int32_t *rounds) {
  int64_t N = 0, C = 0;
  otfem_status st = check_desc(d, &N, &C);
  if (st != OTFEM_OK) return st;
  if (!coords || !tets || !body_node_offsets || !body_tet_offsets)
    return fail(OTFEM_INVALID_ARG, "coords, tets and the body offsets are required");
  const bool has_load = d->load_axis >= 0;
  if (d->fixed_axis < -1 || d->fixed_axis > 2 || d->load_axis < -1 || d->load_axis > 2 ||
      (d->fixed_axis >= 0 && (d->fixed_side & ~1)) || (has_load && (d->load_side & ~1)))
    return fail(OTFEM_INVALID_ARG, "face axis must be -1, 0, 1 or 2 and side 0 or 1");
  if (f_ext && has_load && !d->traction) return fail(OTFEM_INVALID_ARG, "a load face needs a traction");
  if (d->jitter < 0.0 || !(d->jitter < 0.5))
    return fail(OTFEM_INVALID_ARG, "jitter must be in [0, 0.5) of the cube edge");
  const bool jitter = d->jitter > 0.0 || d->draw;
  if (jitter && d->max_rounds < 1) return fail(OTFEM_INVALID_ARG, "max_rounds must be >= 1");
  int dev = 0;
  if (hipGetDeviceCount(&dev) != hipSuccess || dev < 1) return fail(OTFEM_NO_DEVICE, "no HIP device");
  const int64_t B = d->num_bodies, T = 5 * C;

  std::vector<Body> host(B);
  body_node_offsets[0] = 0;
  body_tet_offsets[0] = 0;
  for (int64_t b = 0; b < B; ++b) {
    Body &bb = host[b];
    bb.I = d->dims[3 * b];
    bb.J = d->dims[3 * b + 1];
    bb.K = d->dims[3 * b + 2];
    bb.pad = 0;
    bb.node0 = body_node_offsets[b];
    bb.cube0 = body_tet_offsets[b] / 5;
    const double *t = d->traction ? d->traction + (d->traction_per_body ? 3 * b : 0) : nullptr;
    bb.tx = t ? t[0] : 0.0;
    bb.ty = t ? t[1] : 0.0;
    bb.tz = t ? t[2] : 0.0;
    body_node_offsets[b + 1] = bb.node0 + (int64_t)(bb.I + 1) * (bb.J + 1) * (bb.K + 1);
    body_tet_offsets[b + 1] = body_tet_offsets[b] + 5 * (int64_t)bb.I * bb.J * bb.K;
  }

  Build g;
  if (hipStreamCreate(&g.s) != hipSuccess) return fail(OTFEM_HIP_ERROR, "hipStreamCreate failed");
  Body *bd = g.alloc<Body>(B);
  double *x0 = g.alloc<double>(3 * N), *x = g.alloc<double>(3 * N);
  uint8_t *interior = g.alloc<uint8_t>(N), *fx = g.alloc<uint8_t>(N), *mark = g.alloc<uint8_t>(N);
  int32_t *tt = g.alloc<int32_t>(4 * T);
  if (g.err != hipSuccess) return alloc_failed(g);
  auto hip = [&](hipError_t e, const char *what) {
    if (e != hipSuccess && g.err == hipSuccess) g.err = e;
    if (e != hipSuccess) g_err = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipSuccess;
  };
  const int blk = 256;
  if (!hip(hipMemcpyAsync(bd, host.data(), B * sizeof(Body), hipMemcpyHostToDevice, g.s), "copy bodies"))
    return OTFEM_HIP_ERROR;
  k_lattice_nodes<<<grid_for(N, blk), blk, 0, g.s>>>(bd, (int)B, N, d->width, d->height, d->depth,
                                                    d->fixed_axis, d->fixed_side, x0, interior, fx);
  if (C > 0) k_lattice_tets<<<grid_for(C, blk), blk, 0, g.s>>>(bd, (int)B, C, tt);
  if (!hip(hipGetLastError(), "lattice kernels")) return OTFEM_HIP_ERROR;
  if (!hip(hipMemcpyAsync(x, x0, 3 * N * sizeof(double), hipMemcpyDeviceToDevice, g.s), "copy x0"))
    return OTFEM_HIP_ERROR;

  int32_t taken = 0;
  if (jitter) {
    // ordered compaction of marked nodes: list[] = ascending ids with mark set
    int32_t *list = g.alloc<int32_t>(N);
    int32_t *n_sel = g.alloc<int32_t>(1);
    unsigned long long *n_bad = g.alloc<unsigned long long>(1);
    double *disp = g.alloc<double>(3 * N);
    size_t tmp_bytes = 0;
    hipcub::CountingInputIterator<int32_t> ids(0);
    if (!hip(hipcub::DeviceSelect::Flagged(nullptr, tmp_bytes, ids, mark, list, n_sel, (int)N, g.s),
             "select size"))
      return OTFEM_HIP_ERROR;
    void *tmp = g.alloc<char>(tmp_bytes);
    if (g.err != hipSuccess) return alloc_failed(g);
    std::vector<double> hdisp;
    const double amp[3] = {d->jitter * d->width, d->jitter * d->height, d->jitter * d->depth};
    // Round 0 moves every interior node; round r >= 1 tests every
    // tetrahedron against the floor and moves the marked nodes again, until
    // a test passes or max_rounds tests have failed.  A tetrahedron below
    // the floor with no interior corner cannot be repaired: an error.
    if (!hip(hipMemcpyAsync(mark, interior, N, hipMemcpyDeviceToDevice, g.s), "copy interior"))
      return OTFEM_HIP_ERROR;
    bool done = false;
    for (int round = 0; round <= d->max_rounds; ++round) {
      if (round > 0) {
        if (!hip(hipMemsetAsync(mark, 0, N, g.s), "clear marks") ||
            !hip(hipMemsetAsync(n_bad, 0, sizeof(unsigned long long), g.s), "clear count"))
          return OTFEM_HIP_ERROR;
        if (T > 0)
          k_volume_floor<<<grid_for(T, blk), blk, 0, g.s>>>(tt, T, x, d->min_volume, interior, mark, n_bad);
        if (!hip(hipGetLastError(), "volume floor")) return OTFEM_HIP_ERROR;
      }
      if (!hip(hipcub::DeviceSelect::Flagged(tmp, tmp_bytes, ids, mark, list, n_sel, (int)N, g.s), "select"))
        return OTFEM_HIP_ERROR;
      int32_t n = 0;
      unsigned long long bad = 0;
      if (!hip(hipMemcpyAsync(&n, n_sel, sizeof(n), hipMemcpyDeviceToHost, g.s), "copy count") ||
          (round > 0 &&
           !hip(hipMemcpyAsync(&bad, n_bad, sizeof(bad), hipMemcpyDeviceToHost, g.s), "copy count")) ||
          !hip(hipStreamSynchronize(g.s), "sync"))
        return OTFEM_HIP_ERROR;
      if (round > 0) {
        if (bad == 0) {
          done = true;
          break;
        }
        if (n == 0)
          return fail(OTFEM_INVALID_ARG,
                      "jitter: a tetrahedron with no interior node is below the volume floor");
        if (round == d->max_rounds) break;
      }
      if (n == 0) continue;  // round 0 without interior nodes
      if (d->draw) {
        hdisp.resize(3 * (size_t)n);
        if (d->draw(d->draw_user, n, hdisp.data()) != 0)
          return fail(OTFEM_INVALID_ARG, "jitter: the random source failed");
        if (!hip(hipMemcpyAsync(disp, hdisp.data(), 3 * (size_t)n * sizeof(double),
                                hipMemcpyHostToDevice, g.s),
                 "copy displacements"))
          return OTFEM_HIP_ERROR;
      } else {
        k_builtin_draw<<<grid_for(n, blk), blk, 0, g.s>>>(list, n, d->seed, round, amp[0], amp[1],
                                                          amp[2], disp);
      }
      k_move<<<grid_for(n, blk), blk, 0, g.s>>>(list, n, x0, disp, x);
      if (!hip(hipGetLastError(), "move")) return OTFEM_HIP_ERROR;
      // the host buffer must stay valid until the copy has run
      if (d->draw && !hip(hipStreamSynchronize(g.s), "sync")) return OTFEM_HIP_ERROR;
      taken = round + 1;
    }
    if (!done) return fail(OTFEM_INVALID_ARG, "jitter: resampling did not converge");
  }
  if (rounds) *rounds = taken;

  double *f = nullptr;
  if (f_ext) {
    f = g.alloc<double>(3 * N);
    if (g.err != hipSuccess) return alloc_failed(g);
    k_face_load<<<grid_for(N, blk), blk, 0, g.s>>>(bd, (int)B, N, x, has_load ? d->load_axis : -1,
                                                  d->load_side, fx, f);
    if (!hip(hipGetLastError(), "face load")) return OTFEM_HIP_ERROR;
  }
  if (!hip(hipMemcpyAsync(coords, x, 3 * N * sizeof(double), hipMemcpyDeviceToHost, g.s), "copy coords") ||
      !hip(hipMemcpyAsync(tets, tt, 4 * T * sizeof(int32_t), hipMemcpyDeviceToHost, g.s), "copy tets") ||
      (fixed && !hip(hipMemcpyAsync(fixed, fx, N, hipMemcpyDeviceToHost, g.s), "copy fixed")) ||
      (f_ext &&
       !hip(hipMemcpyAsync(f_ext, f, 3 * N * sizeof(double), hipMemcpyDeviceToHost, g.s), "copy forces")) ||
      !hip(hipStreamSynchronize(g.s), "sync"))
    return OTFEM_HIP_ERROR;
  return OTFEM_OK;
}

otfem_status otfem_blocks_build(const otfem_blocks_desc *d, double *coords, int32_t *tets,
                                uint8_t *fixed, double *f_ext, int64_t *body_node_offsets,
                                int64_t *body_tet_offsets, int32_t *rounds) {
  // no C++ exception crosses the C ABI
  try {
    return blocks_build(d, coords, tets, fixed, f_ext, body_node_offsets, body_tet_offsets, rounds);
  } catch (const std::bad_alloc &) {
    try {
      g_err = "out of host memory";
    } catch (...) {
    }
    return OTFEM_OUT_OF_MEMORY;
  } catch (const std::exception &e) {
    try {
      g_err = e.what();
    } catch (...) {
    }
    return OTFEM_INVALID_ARG;
  }
}

}  // extern "C"
