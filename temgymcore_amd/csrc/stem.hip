// 4D-STEM shadow-image model: forward (detector coordinates, bilinear sample
// lookup, optional 4D scatter) and the descan / geometry fit (loss + adjoint)
// over every (camera length, scan position, ray) in one launch.
//
// Model (per scan position s, ray j; see include/temgym_hip.h):
//   x_s   = sp_x + (th_x + st_x) * d                       sample plane
//   x_det = x_s + Dx + (th_x + st_x + Ddx) * L_c           detector
// with Dx = A0 . v + o_x - sp_x, Ddx = A2 . v + o_dx - st_x,
// v = (sp_x, sp_y, st_x, st_y), st = pivot * sp.  Hence
//   x_det = th_x * (d + L_c) + Bx(s, c),  y_det = th_y * (d + L_c) + By(s, c)
// (reference component bodies: Scanner components.py:279-285, propagation
// propagator.py:51-72, Descanner components.py:320-372).
//
// The fit kernel streams the f32 targets exactly once (HBM-bound).  For each
// (s, c) it reduces residual moments  sum r_x, sum r_x th_x, sum r_x th_y,
// sum r_y, sum r_y th_x, sum r_y th_y, sum |r|^2  (FP32 per lane and across
// the wavefront, FP64 across wavefronts), from which every parameter gradient
// follows exactly because the detector coordinates are affine in the slopes.
// Per-scan-position gradient rows go to a workspace that a second kernel sums
// in a fixed order (deterministic FP64 result).
#include "tg_common.hpp"

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kMaxSide = 1024;
constexpr double kPi = 3.14159265358979323846;

struct ScanPos {
  double spx, spy, stx, sty;
};

// Scan position of flattened scan index s: ScanGrid.pixels_to_metres with
// square pixels `step`, rotation `rot_deg`, centre 0, no flip
// (reference coordinate_transforms.py:50-101, test_component.py:356-378).
__device__ __forceinline__ ScanPos scan_position(const tg_stem_geometry& G, const double* P, int64_t s) {
  const int64_t i = s / G.scan_w, j = s % G.scan_w;
  const double step = P[16];
  const double phi = P[17] * (kPi / 180.0);
  const double u = step * ((double)i - 0.5 * (G.scan_h - 1));
  const double v = step * ((double)j - 0.5 * (G.scan_w - 1));
  double sn, cs;
  sincos(phi, &sn, &cs);
  ScanPos p;
  p.spx = -sn * u + cs * v;
  p.spy = -(cs * u + sn * v);
  p.stx = G.pivot * p.spx;
  p.sty = G.pivot * p.spy;
  return p;
}

__device__ __forceinline__ double row_dot(const double* A, int r, const ScanPos& p) {
  return A[4 * r + 0] * p.spx + A[4 * r + 1] * p.spy + A[4 * r + 2] * p.stx + A[4 * r + 3] * p.sty;
}

// Detector-coordinate offsets (Bx, By) for camera length L.
__device__ __forceinline__ void offsets_at(const tg_stem_geometry& G, const double* A, const ScanPos& p, double L,
                                           double& bx, double& by) {
  bx = p.stx * G.defocus + (row_dot(A, 0, p) + G.offsets[0]) + (row_dot(A, 2, p) + G.offsets[2]) * L;
  by = p.sty * G.defocus + (row_dot(A, 1, p) + G.offsets[1]) + (row_dot(A, 3, p) + G.offsets[3]) * L;
}

__device__ __forceinline__ float slope_of(const tg_stem_geometry& G, int k) {
  const double pitch = 2.0 * G.semi_conv / G.n_side;
  return (float)(((double)k - 0.5 * (G.n_side - 1)) * pitch);
}

// Bilinear lookup of the sample image at (x, y) metres; neighbours outside
// the image contribute zero.
__device__ __forceinline__ float bilinear(const tg_stem_geometry& G, const float* __restrict__ img, double x,
                                          double y) {
  const double fy = G.img_m[0] * y + G.img_m[1] * x + G.img_m[2];
  const double fx = G.img_m[3] * y + G.img_m[4] * x + G.img_m[5];
  const double y0d = floor(fy), x0d = floor(fx);
  if (!(y0d >= -1.0 && y0d < (double)G.img_h && x0d >= -1.0 && x0d < (double)G.img_w)) return 0.0f;
  const int y0 = (int)y0d, x0 = (int)x0d;
  const float wy = (float)(fy - y0d), wx = (float)(fx - x0d);
  float acc = 0.0f;
#pragma unroll
  for (int a = 0; a < 2; ++a) {
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int yy = y0 + a, xx = x0 + b;
      if (yy < 0 || yy >= G.img_h || xx < 0 || xx >= G.img_w) continue;
      const float w = (a ? wy : 1.0f - wy) * (b ? wx : 1.0f - wx);
      acc = fmaf(w, img[(int64_t)yy * G.img_w + xx], acc);
    }
  }
  return acc;
}

// ---------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------
constexpr int kFwdRays = 4;  // rays per lane

__global__ __launch_bounds__(kThreads) void stem_forward_kernel(const tg_stem_geometry G, const double* __restrict__ P,
                                                                const float* __restrict__ img,
                                                                float* __restrict__ coords,
                                                                float* __restrict__ values,
                                                                float* __restrict__ image4d) {
  __shared__ float th[kMaxSide];
  __shared__ float sh_a[TG_STEM_MAX_CAM], sh_bx[TG_STEM_MAX_CAM], sh_by[TG_STEM_MAX_CAM];
  const int64_t s = blockIdx.y;
  const int64_t S = (int64_t)G.scan_h * G.scan_w;
  const int64_t J = (int64_t)G.n_side * G.n_side;
  for (int k = threadIdx.x; k < G.n_side; k += kThreads) th[k] = slope_of(G, k);
  const ScanPos p = scan_position(G, P, s);
  if (threadIdx.x < G.n_cam) {
    const double L = P[18 + threadIdx.x];
    double bx, by;
    offsets_at(G, P, p, L, bx, by);
    sh_a[threadIdx.x] = (float)(G.defocus + L);
    sh_bx[threadIdx.x] = (float)bx;
    sh_by[threadIdx.x] = (float)by;
  }
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kThreads * kFwdRays;
#pragma unroll
  for (int u = 0; u < kFwdRays; ++u) {
    const int64_t j = base + (int64_t)u * kThreads + threadIdx.x;
    if (j >= J) continue;
    const int jj = (int)j;  // J <= 2^20
    const float tx = th[jj % G.n_side], ty = th[jj / G.n_side];
    float v = 0.0f;
    if (values != nullptr || image4d != nullptr) {
      // sample plane position in f64 (scan offsets are ~1e-9 m)
      const double xs = p.spx + ((double)tx + p.stx) * G.defocus;
      const double ys = p.spy + ((double)ty + p.sty) * G.defocus;
      v = bilinear(G, img, xs, ys);
      if (values != nullptr) values[s * J + j] = v;
    }
    for (int c = 0; c < G.n_cam; ++c) {
      const float xd = fmaf(tx, sh_a[c], sh_bx[c]);
      const float yd = fmaf(ty, sh_a[c], sh_by[c]);
      if (coords != nullptr) {
        float2* o = (float2*)coords + ((int64_t)c * S + s) * J + j;
        *o = make_float2(xd, yd);
      }
      if (c == 0 && image4d != nullptr) {
        const double fy = __dadd_rn(__dadd_rn(__dmul_rn(G.det_m[0], (double)yd), __dmul_rn(G.det_m[1], (double)xd)),
                                    G.det_m[2]);
        const double fx = __dadd_rn(__dadd_rn(__dmul_rn(G.det_m[3], (double)yd), __dmul_rn(G.det_m[4], (double)xd)),
                                    G.det_m[5]);
        const int32_t iy = tg::xla_f64_to_i32(rint(fy)), ix = tg::xla_f64_to_i32(rint(fx));
        if (iy >= 0 && iy < G.det_h && ix >= 0 && ix < G.det_w)
          unsafeAtomicAdd(image4d + (s * G.det_h + iy) * G.det_w + ix, v);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// fit: forward + adjoint
// ---------------------------------------------------------------------------
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

constexpr int kMom = 7;  // sum rx, rx*tx, rx*ty, ry, ry*tx, ry*ty, |r|^2
constexpr int kPerLane = 8;
constexpr int64_t kTile = (int64_t)kThreads * kPerLane;

typedef float f4 __attribute__((ext_vector_type(4)));

// Tile `it` (camera c = it / ntile, tile it % ntile) of scan position s:
// kPerLane float4 per lane, lane-strided so each load instruction of a
// wavefront covers 1 KiB of contiguous targets.  Streamed once: non-temporal.
__device__ __forceinline__ void load_tile(const f4* __restrict__ T4, int64_t S, int64_t s, int64_t npair,
                                          int64_t ntile, int64_t it, f4 (&dst)[kPerLane]) {
  const int64_t c = it / ntile, tile = it - c * ntile;
  const int64_t row = (c * S + s) * npair;
  const int64_t q0 = tile * kTile + threadIdx.x;
#pragma unroll
  for (int k = 0; k < kPerLane; ++k) {
    const int64_t q = q0 + (int64_t)k * kThreads;
    dst[k] = q < npair ? __builtin_nontemporal_load(T4 + row + q) : f4{0.f, 0.f, 0.f, 0.f};
  }
}

// Per-scan-position gradient terms from the reduced moments M[c][k]
// (thread t < np writes row t of the workspace column s).
__device__ __forceinline__ void fit_epilogue(const tg_stem_geometry& G, const double* __restrict__ P,
                                             const ScanPos& p, const double (&mom)[TG_STEM_MAX_CAM][kWaves][kMom],
                                             int64_t s, int64_t S, double* __restrict__ partials) {
  const int np = 18 + G.n_cam + 1;
  const int t = threadIdx.x;
  if (t >= np) return;
  const double* A = P;
  const double d = G.defocus, kap = G.pivot;
  const double vv[4] = {p.spx, p.spy, p.stx, p.sty};
  double g = 0.0;
  for (int c = 0; c < G.n_cam; ++c) {
    double M[kMom];
#pragma unroll
    for (int k = 0; k < kMom; ++k) {
      double acc = 0.0;
#pragma unroll
      for (int w = 0; w < kWaves; ++w) acc += mom[c][w][k];
      M[k] = acc;
    }
    const double L = P[18 + c];
    if (t < 16) {
      const int r = t >> 2, k = t & 3;
      const double m0 = (r == 0 || r == 2) ? M[0] : M[3];
      const double scale = (r >= 2) ? L : 1.0;
      g += 2.0 * vv[k] * scale * m0;
    } else if (t == 16 || t == 17) {
      // d(Bx, By) / d(spx, spy), with st = pivot * sp
      const double dbx_dx = kap * d + (A[0] + A[2] * kap) + (A[8] + A[10] * kap) * L;
      const double dbx_dy = (A[1] + A[3] * kap) + (A[9] + A[11] * kap) * L;
      const double dby_dx = (A[4] + A[6] * kap) + (A[12] + A[14] * kap) * L;
      const double dby_dy = kap * d + (A[5] + A[7] * kap) + (A[13] + A[15] * kap) * L;
      double dsx, dsy;
      if (t == 16) {  // step: sp is proportional to it
        dsx = p.spx / P[16];
        dsy = p.spy / P[16];
      } else {        // rotation in degrees: d sp_x = sp_y dphi, d sp_y = -sp_x dphi
        dsx = p.spy * (kPi / 180.0);
        dsy = -p.spx * (kPi / 180.0);
      }
      g += 2.0 * ((dbx_dx * dsx + dbx_dy * dsy) * M[0] + (dby_dx * dsx + dby_dy * dsy) * M[3]);
    } else if (t == 18 + c) {
      const double dbx_dl = row_dot(A, 2, p) + G.offsets[2];
      const double dby_dl = row_dot(A, 3, p) + G.offsets[3];
      g += 2.0 * (M[1] + dbx_dl * M[0] + M[5] + dby_dl * M[3]);
    } else if (t == np - 1) {
      g += M[6];
    }
  }
  partials[(int64_t)t * S + s] = g;
}

// One workgroup per scan position.  The targets are streamed in tiles of
// kPerLane float4 (ray pairs) per lane, double-buffered across tiles and
// camera lengths, so the per-camera wavefront reductions never leave the load
// pipeline empty.  (A persistent variant that also prefetched across scan
// positions needed 140 VGPRs, fell to 3 waves per SIMD and was slower.)
__global__ __launch_bounds__(kThreads) void stem_fit_kernel(const tg_stem_geometry G, const double* __restrict__ P,
                                                            const float* __restrict__ T,
                                                            double* __restrict__ partials) {
  __shared__ float th[kMaxSide];
  __shared__ double mom[TG_STEM_MAX_CAM][kWaves][kMom];
  const int64_t S = (int64_t)G.scan_h * G.scan_w;
  const int64_t J = (int64_t)G.n_side * G.n_side;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int k = threadIdx.x; k < G.n_side; k += kThreads) th[k] = slope_of(G, k);
  __syncthreads();
  const int64_t npair = J >> 1;  // two rays per float4 of targets (n_side even)
  const int half = G.n_side >> 1, drow = kThreads / half, dcol = kThreads % half;
  const int64_t ntile = (npair + kTile - 1) / kTile;
  const int64_t nit = (int64_t)G.n_cam * ntile;
  const f4* __restrict__ T4 = (const f4*)T;
  f4 cur[kPerLane], nxt[kPerLane];
  const int64_t s = blockIdx.x;
  load_tile(T4, S, s, npair, ntile, 0, cur);
  {
    const ScanPos p = scan_position(G, P, s);
    float a_hi = 0.f, a_lo = 0.f, bx_hi = 0.f, bx_lo = 0.f, by_hi = 0.f, by_lo = 0.f;
    float m0 = 0.f, m1 = 0.f, m2 = 0.f, m3 = 0.f, m4 = 0.f, m5 = 0.f, m6 = 0.f;
    for (int64_t it = 0; it < nit; ++it) {
      if (it + 1 < nit) load_tile(T4, S, s, npair, ntile, it + 1, nxt);
      const int c = (int)(it / ntile);
      const int64_t tile = it - (int64_t)c * ntile;
      if (tile == 0) {
        const double L = P[18 + c];
        double bx, by;
        offsets_at(G, P, p, L, bx, by);
        // double-float split of the slope scale and offsets: the residual
        //   r = th*a + b - t = (fma(th, a_hi, -t) + b_hi) + fma(th, a_lo, b_lo)
        // is exact to about one ulp of r itself (fma(th, a_hi, -t) rounds once),
        // not to one ulp of the ~1e-4 m detector coordinate.
        const double ad = G.defocus + L;
        a_hi = (float)ad;
        a_lo = (float)(ad - (double)a_hi);
        bx_hi = (float)bx;
        bx_lo = (float)(bx - (double)bx_hi);
        by_hi = (float)by;
        by_lo = (float)(by - (double)by_hi);
        m0 = m1 = m2 = m3 = m4 = m5 = m6 = 0.f;
      }
      // ray pair q = (row iy, columns 2*hc, 2*hc+1); stepped without division
      const int64_t q0 = tile * kTile + threadIdx.x;
      int iy = (int)(q0 / half), hc = (int)(q0 % half);
#pragma unroll
      for (int k = 0; k < kPerLane; ++k) {
        const bool valid = q0 + (int64_t)k * kThreads < npair;
        if (valid) {
          const f4 t = cur[k];
          const float tx0 = th[2 * hc], tx1 = th[2 * hc + 1], ty = th[iy];
          const float rx0 = (fmaf(tx0, a_hi, -t.x) + bx_hi) + fmaf(tx0, a_lo, bx_lo);
          const float ry0 = (fmaf(ty, a_hi, -t.y) + by_hi) + fmaf(ty, a_lo, by_lo);
          const float rx1 = (fmaf(tx1, a_hi, -t.z) + bx_hi) + fmaf(tx1, a_lo, bx_lo);
          const float ry1 = (fmaf(ty, a_hi, -t.w) + by_hi) + fmaf(ty, a_lo, by_lo);
          m0 += rx0 + rx1;
          m1 = fmaf(rx0, tx0, fmaf(rx1, tx1, m1));
          m2 = fmaf(rx0 + rx1, ty, m2);
          m3 += ry0 + ry1;
          m4 = fmaf(ry0, tx0, fmaf(ry1, tx1, m4));
          m5 = fmaf(ry0 + ry1, ty, m5);
          m6 = fmaf(rx0, rx0, fmaf(ry0, ry0, fmaf(rx1, rx1, fmaf(ry1, ry1, m6))));
        }
        hc += dcol;
        iy += drow;
        if (hc >= half) {
          hc -= half;
          ++iy;
        }
      }
      if (tile == ntile - 1) {
        const float w0 = wave_sum(m0), w1 = wave_sum(m1), w2 = wave_sum(m2), w3 = wave_sum(m3);
        const float w4 = wave_sum(m4), w5 = wave_sum(m5), w6 = wave_sum(m6);
        if (lane == 0) {
          mom[c][wave][0] = w0;
          mom[c][wave][1] = w1;
          mom[c][wave][2] = w2;
          mom[c][wave][3] = w3;
          mom[c][wave][4] = w4;
          mom[c][wave][5] = w5;
          mom[c][wave][6] = w6;
        }
      }
#pragma unroll
      for (int k = 0; k < kPerLane; ++k) cur[k] = nxt[k];
    }
    __syncthreads();
    fit_epilogue(G, P, p, mom, s, S, partials);
  }
}

// Fixed-order FP64 row sums of the [np, rows] workspace (one block per row,
// coalesced reads, wavefront shuffles then one LDS step).
constexpr int kSumThreads = 1024;

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ __launch_bounds__(kSumThreads) void row_sum_kernel(const double* __restrict__ partials, int64_t rows,
                                                              double* __restrict__ out) {
  __shared__ double red[kSumThreads / 64];
  const double* __restrict__ src = partials + (int64_t)blockIdx.x * rows;
  double acc = 0.0;
  for (int64_t r = threadIdx.x; r < rows; r += kSumThreads) acc += src[r];
  acc = wave_sum_f64(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x < 64) {
    double v = threadIdx.x < kSumThreads / 64 ? red[threadIdx.x] : 0.0;
    v = wave_sum_f64(v);
    if (threadIdx.x == 0) out[blockIdx.x] = v;
  }
}

int validate(const tg_stem_geometry* g, const double* params) {
  TG_REQUIRE(g != nullptr && params != nullptr, "stem: NULL geometry or params");
  TG_REQUIRE(g->scan_h > 0 && g->scan_w > 0 && (int64_t)g->scan_h * g->scan_w <= 65535,
             "stem: scan grid %dx%d must have 1..65535 positions", g->scan_h, g->scan_w);
  TG_REQUIRE(g->n_side >= 2 && g->n_side <= kMaxSide && (g->n_side % 2) == 0,
             "stem: n_side %d must be even and in [2, %d]", g->n_side, kMaxSide);
  TG_REQUIRE(g->n_cam >= 1 && g->n_cam <= TG_STEM_MAX_CAM, "stem: n_cam %d out of range", g->n_cam);
  return TG_OK;
}

}  // namespace

extern "C" int tg_stem_forward(const tg_stem_geometry* g, const double* params, const float* image, float* coords,
                               float* values, float* image4d, void* stream) {
  tg::clear_error();
  int st = validate(g, params);
  if (st != TG_OK) return st;
  TG_REQUIRE(image != nullptr || (values == nullptr && image4d == nullptr),
             "tg_stem_forward: sample image required for values / image4d");
  TG_REQUIRE(image == nullptr || (g->img_h > 0 && g->img_w > 0), "tg_stem_forward: bad sample image shape");
  TG_REQUIRE(image4d == nullptr || (g->det_h > 0 && g->det_w > 0), "tg_stem_forward: bad detector shape");
  if (coords == nullptr && values == nullptr && image4d == nullptr) return TG_OK;
  const int64_t J = (int64_t)g->n_side * g->n_side;
  dim3 grid((unsigned)tg::cdiv(J, (int64_t)kThreads * kFwdRays), (unsigned)((int64_t)g->scan_h * g->scan_w));
  hipLaunchKernelGGL(stem_forward_kernel, grid, dim3(kThreads), 0, (hipStream_t)stream, *g, params, image, coords,
                     values, image4d);
  TG_LAUNCH_CHECK("stem_forward_kernel launch");
  return TG_OK;
}

extern "C" int tg_stem_fit(const tg_stem_geometry* g, const double* params, const float* targets, double* partials,
                           double* out, void* stream) {
  tg::clear_error();
  int st = validate(g, params);
  if (st != TG_OK) return st;
  TG_REQUIRE(targets != nullptr && partials != nullptr && out != nullptr, "tg_stem_fit: NULL buffer");
  TG_REQUIRE(((uintptr_t)targets & 15) == 0, "tg_stem_fit: targets must be 16-byte aligned");
  const int64_t S = (int64_t)g->scan_h * g->scan_w;
  const int np = 18 + g->n_cam + 1;
  hipLaunchKernelGGL(stem_fit_kernel, dim3((unsigned)S), dim3(kThreads), 0, (hipStream_t)stream, *g, params, targets,
                     partials);
  TG_LAUNCH_CHECK("stem_fit_kernel launch");
  hipLaunchKernelGGL(row_sum_kernel, dim3((unsigned)np), dim3(kSumThreads), 0, (hipStream_t)stream, partials, S, out);
  TG_LAUNCH_CHECK("row_sum_kernel launch");
  return TG_OK;
}
