// Device-side ray sources: initial ray bundles generated straight into the
// device arrays the ray programs read, with no host arrays in between.
//
// Replaces the host numpy generators of the reference:
//   * concentric_rings + multi_cumsum_inplace (utils.py:46-80, :117-175)
//     behind PointSource / ParallelBeam.generate_array (source.py:82-188);
//   * random_coords (utils.py:178-205), random=True of the same sources;
//   * fibonacci_spiral (utils.py:297-325), the notebooks' beamlet centres.
//
// Rings: the host builds the O(sqrt N) ring table (radii, angle steps, ring
// boundaries) and the positions where the reference's partitioned cumulative
// sum restarts.  Those restarts do not coincide with the ring boundaries: the
// counter of multi_cumsum_inplace lets every ring after the first absorb the
// first point of the next (restart k sits at ring boundary k plus k).  Each
// point's angle is that sum, replayed in the same order (one f64 addition
// per preceding point of its restart segment), so angles are bitwise those
// of the sequential loop.  sin/cos are the device's f64 functions.
#include "tg_common.hpp"

namespace {

constexpr int kThreads = 256;

template <typename T>
__device__ __forceinline__ void store(void* p, int64_t i, double v) {
  ((T*)p)[i] = (T)v;
}

// Index of the last entry of the sorted array a[0..n) that is <= v (a[0] <= v).
__device__ __forceinline__ int upper_index(const int64_t* __restrict__ a, int n, int64_t v) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (a[mid] <= v)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// kind 0: PointSource  (x, y) = offset, (dx, dy) = (r cos a, r sin a)
// kind 1: ParallelBeam (x, y) = (r cos a + off_x, r sin a + off_y), slopes 0
template <typename T>
__global__ __launch_bounds__(kThreads) void rings_kernel(const double* __restrict__ radius,
                                                         const double* __restrict__ step,
                                                         const int64_t* __restrict__ ring_first, int n_rings,
                                                         const int64_t* __restrict__ seg_first, int n_seg, int64_t n,
                                                         int kind, double off_x, double off_y, void* x, void* y,
                                                         void* dx, void* dy) {
  const int64_t v = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (v >= n) return;
  const int64_t s = seg_first[upper_index(seg_first, n_seg, v)];
  int ring = upper_index(ring_first, n_rings, s + 1 < n ? s + 1 : s);
  double angle = 0.0;
  for (int64_t u = s + 1; u <= v; ++u) {
    while (ring + 1 < n_rings && ring_first[ring + 1] <= u) ++ring;
    angle = __dadd_rn(angle, step[ring]);
  }
  const double r = radius[upper_index(ring_first, n_rings, v)];
  double sa, ca;
  sincos(angle, &sa, &ca);
  const double ps = __dmul_rn(r, sa), pc = __dmul_rn(r, ca);
  if (kind == 0) {
    store<T>(x, v, off_x);
    store<T>(y, v, off_y);
    store<T>(dx, v, pc);
    store<T>(dy, v, ps);
  } else {
    store<T>(x, v, __dadd_rn(pc, off_x));
    store<T>(y, v, __dadd_rn(ps, off_y));
    store<T>(dx, v, 0.0);
    store<T>(dy, v, 0.0);
  }
}

// Fibonacci spiral: r_i = radius * sqrt((i + 0.5) / denom) inside, radius on
// the boundary band (i > n - (boundary + 1)), r_0 = 0; phi_i = i * ga.
template <typename T>
__global__ __launch_bounds__(kThreads) void fibonacci_kernel(int64_t n, double radius, double band, double denom,
                                                             double ga, void* x, void* y) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const double fi = (double)i;
  double rr = fi > band ? radius : __dmul_rn(radius, sqrt(__ddiv_rn(__dadd_rn(fi, 0.5), denom)));
  if (i == 0) rr = 0.0;
  double s, c;
  sincos(__dmul_rn(fi, ga), &s, &c);
  store<T>(x, i, __dmul_rn(rr, c));
  store<T>(y, i, __dmul_rn(rr, s));
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC 2011): counter (i, stream, 0, 0), key = seed.
__device__ __forceinline__ void philox(uint32_t (&c)[4], uint32_t k0, uint32_t k1) {
  constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)M0 * c[0], p1 = (uint64_t)M1 * c[2];
    const uint32_t h0 = (uint32_t)(p0 >> 32), l0 = (uint32_t)p0, h1 = (uint32_t)(p1 >> 32), l1 = (uint32_t)p1;
    c[0] = h1 ^ c[1] ^ k0;
    c[1] = l1;
    c[2] = h0 ^ c[3] ^ k1;
    c[3] = l0;
    k0 += W0;
    k1 += W1;
  }
}

__device__ __forceinline__ double unit53(uint32_t a, uint32_t b) {
  return (double)(((uint64_t)a << 21) ^ (uint64_t)(b >> 11)) * 0x1.0p-53;  // [0, 1)
}

// Uniform points in the disc of `scale` (area-uniform: r = scale sqrt(u)),
// kind as for rings.
template <typename T>
__global__ __launch_bounds__(kThreads) void disc_kernel(int64_t n, uint64_t seed, double scale, int kind, double off_x,
                                                        double off_y, void* x, void* y, void* dx, void* dy) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  uint32_t c[4] = {(uint32_t)i, (uint32_t)((uint64_t)i >> 32), 0u, 0u};
  philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const double u = unit53(c[0], c[1]), w = unit53(c[2], c[3]);
  const double r = scale * sqrt(u);
  double s, co;
  sincos(6.283185307179586 * w, &s, &co);
  const double ps = r * s, pc = r * co;
  if (kind == 0) {
    store<T>(x, i, off_x);
    store<T>(y, i, off_y);
    store<T>(dx, i, pc);
    store<T>(dy, i, ps);
  } else {
    store<T>(x, i, pc + off_x);
    store<T>(y, i, ps + off_y);
    store<T>(dx, i, 0.0);
    store<T>(dy, i, 0.0);
  }
}

}  // namespace

extern "C" int tg_source_rings(const double* ring_radius, const double* ring_step, const int64_t* ring_first,
                               int32_t n_rings, const int64_t* seg_first, int32_t n_seg, int64_t n, int32_t kind,
                               double off_x, double off_y, int32_t dtype, void* x, void* y, void* dx, void* dy,
                               void* stream) {
  tg::clear_error();
  TG_REQUIRE(n >= 0, "tg_source_rings: n < 0");
  if (n == 0) return TG_OK;
  TG_REQUIRE(ring_radius && ring_step && ring_first && seg_first && x && y && dx && dy,
             "tg_source_rings: NULL pointer");
  TG_REQUIRE(n_rings >= 1 && n_seg >= 1, "tg_source_rings: empty ring or segment table");
  TG_REQUIRE(kind == 0 || kind == 1, "tg_source_rings: kind must be 0 (point source) or 1 (parallel beam)");
  TG_REQUIRE(dtype == TG_F32 || dtype == TG_F64, "tg_source_rings: bad dtype %d", dtype);
  const unsigned grid = (unsigned)tg::cdiv(n, kThreads);
  if (dtype == TG_F64)
    hipLaunchKernelGGL(rings_kernel<double>, dim3(grid), dim3(kThreads), 0, (hipStream_t)stream, ring_radius,
                       ring_step, ring_first, n_rings, seg_first, n_seg, n, kind, off_x, off_y, x, y, dx, dy);
  else
    hipLaunchKernelGGL(rings_kernel<float>, dim3(grid), dim3(kThreads), 0, (hipStream_t)stream, ring_radius,
                       ring_step, ring_first, n_rings, seg_first, n_seg, n, kind, off_x, off_y, x, y, dx, dy);
  TG_LAUNCH_CHECK("rings_kernel launch");
  return TG_OK;
}

extern "C" int tg_source_fibonacci(int64_t n, double radius, double band, double denom, double ga, int32_t dtype,
                                   void* x, void* y, void* stream) {
  tg::clear_error();
  TG_REQUIRE(n >= 0, "tg_source_fibonacci: n < 0");
  if (n == 0) return TG_OK;
  TG_REQUIRE(x && y, "tg_source_fibonacci: NULL pointer");
  TG_REQUIRE(dtype == TG_F32 || dtype == TG_F64, "tg_source_fibonacci: bad dtype %d", dtype);
  const unsigned grid = (unsigned)tg::cdiv(n, kThreads);
  if (dtype == TG_F64)
    hipLaunchKernelGGL(fibonacci_kernel<double>, dim3(grid), dim3(kThreads), 0, (hipStream_t)stream, n, radius, band,
                       denom, ga, x, y);
  else
    hipLaunchKernelGGL(fibonacci_kernel<float>, dim3(grid), dim3(kThreads), 0, (hipStream_t)stream, n, radius, band,
                       denom, ga, x, y);
  TG_LAUNCH_CHECK("fibonacci_kernel launch");
  return TG_OK;
}

extern "C" int tg_source_disc(int64_t n, uint64_t seed, double scale, int32_t kind, double off_x, double off_y,
                              int32_t dtype, void* x, void* y, void* dx, void* dy, void* stream) {
  tg::clear_error();
  TG_REQUIRE(n >= 0, "tg_source_disc: n < 0");
  if (n == 0) return TG_OK;
  TG_REQUIRE(x && y && dx && dy, "tg_source_disc: NULL pointer");
  TG_REQUIRE(kind == 0 || kind == 1, "tg_source_disc: kind must be 0 (point source) or 1 (parallel beam)");
  TG_REQUIRE(dtype == TG_F32 || dtype == TG_F64, "tg_source_disc: bad dtype %d", dtype);
  const unsigned grid = (unsigned)tg::cdiv(n, kThreads);
  if (dtype == TG_F64)
    hipLaunchKernelGGL(disc_kernel<double>, dim3(grid), dim3(kThreads), 0, (hipStream_t)stream, n, seed, scale, kind,
                       off_x, off_y, x, y, dx, dy);
  else
    hipLaunchKernelGGL(disc_kernel<float>, dim3(grid), dim3(kThreads), 0, (hipStream_t)stream, n, seed, scale, kind,
                       off_x, off_y, x, y, dx, dy);
  TG_LAUNCH_CHECK("disc_kernel launch");
  return TG_OK;
}
