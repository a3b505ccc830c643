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
#include <mutex>

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
// Sample-image pixel coordinates of the sample-plane position of slope
// (tx, ty) at one scan position.  The position is affine in the slope,
//   x = spx + (tx + stx) d,  y = spy + (ty + sty) d  (d = defocus),
// so (fy, fx) = origin + slope terms, with the origin and the four slope
// coefficients folded per scan position (FP64).
struct SampleMap {
  double fy0, fx0, ky_x, ky_y, kx_x, kx_y;
};

__device__ __forceinline__ SampleMap sample_map(const tg_stem_geometry& G, const ScanPos& p) {
  const double xc = p.spx + p.stx * G.defocus, yc = p.spy + p.sty * G.defocus;
  SampleMap m;
  m.fy0 = G.img_m[0] * yc + G.img_m[1] * xc + G.img_m[2];
  m.fx0 = G.img_m[3] * yc + G.img_m[4] * xc + G.img_m[5];
  m.ky_y = G.img_m[0] * G.defocus;
  m.ky_x = G.img_m[1] * G.defocus;
  m.kx_y = G.img_m[3] * G.defocus;
  m.kx_x = G.img_m[4] * G.defocus;
  return m;
}

__device__ __forceinline__ void sample_pixel(const SampleMap& m, float tx, float ty, double& fy, double& fx) {
  fy = fma(m.ky_y, (double)ty, fma(m.ky_x, (double)tx, m.fy0));
  fx = fma(m.kx_y, (double)ty, fma(m.kx_x, (double)tx, m.fx0));
}

// The sample-image region one scan position's rays read: kImgTile^2 pixels
// around the footprint centre, staged in LDS (pixels outside the image are
// never read from it: the lookup keeps the image-bounds test).
constexpr int kImgTile = 32;

// Bilinear lookup at pixel coordinates (fy, fx); neighbours outside the
// image contribute zero (the lookup is 0 when (floor fy, floor fx) is not in
// [-1, h) x [-1, w)).  The four neighbours come from the LDS tile when they
// all lie in it, else from global memory (same weights, same order).
__device__ __forceinline__ float bilinear_tiled(const tg_stem_geometry& G, const float* __restrict__ img,
                                                const float* tile, int oy, int ox, double fy, double fx) {
  const double y0d = floor(fy), x0d = floor(fx);
  if (!(y0d >= -1.0 && y0d < (double)G.img_h && x0d >= -1.0 && x0d < (double)G.img_w)) return 0.0f;
  const int y0 = (int)y0d, x0 = (int)x0d;
  const bool in_tile = y0 >= oy && y0 + 1 < oy + kImgTile && x0 >= ox && x0 + 1 < ox + kImgTile;
  const float wy = (float)(fy - y0d), wx = (float)(fx - x0d);
  float acc = 0.0f;
#pragma unroll
  for (int a = 0; a < 2; ++a) {
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int yy = y0 + a, xx = x0 + b;
      if (yy < 0 || yy >= G.img_h || xx < 0 || xx >= G.img_w) continue;
      const float w = (a ? wy : 1.0f - wy) * (b ? wx : 1.0f - wx);
      const float v = in_tile ? tile[(yy - oy) * kImgTile + (xx - ox)] : img[(int64_t)yy * G.img_w + xx];
      acc = fmaf(w, v, acc);
    }
  }
  return acc;
}

// ---------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------
// rays per lane: a workgroup covers 4096 rays (one scan position of the
// 64x64 workload), so its set-up (slopes, scan position, camera offsets)
// is paid once per 4096 rays rather than per 1024
constexpr int kFwdRays = 16;

__global__ __launch_bounds__(kThreads) void stem_forward_kernel(const tg_stem_geometry G, const double* __restrict__ P,
                                                                const float* __restrict__ img,
                                                                float* __restrict__ coords,
                                                                float* __restrict__ values,
                                                                float* __restrict__ image4d) {
  __shared__ float th[kMaxSide];
  __shared__ float sh_a[TG_STEM_MAX_CAM], sh_bx[TG_STEM_MAX_CAM], sh_by[TG_STEM_MAX_CAM];
  __shared__ float tile[kImgTile * kImgTile];
  const int64_t s = blockIdx.y;
  const int64_t S = (int64_t)G.scan_h * G.scan_w;
  const int64_t J = (int64_t)G.n_side * G.n_side;
  for (int k = threadIdx.x; k < G.n_side; k += kThreads) th[k] = slope_of(G, k);
  const ScanPos p = scan_position(G, P, s);
  const SampleMap smap = sample_map(G, p);
  // the image tile around the footprint centre (origins far outside the
  // image only send every lookup to global memory)
  const int oy = (int)fmin(fmax(floor(smap.fy0) - kImgTile / 2, -4.0 * kImgTile), (double)G.img_h + 4.0 * kImgTile);
  const int ox = (int)fmin(fmax(floor(smap.fx0) - kImgTile / 2, -4.0 * kImgTile), (double)G.img_w + 4.0 * kImgTile);
  // camera length of this thread's camera constants, loaded before the tile
  // so that both arrive in one round trip
  const double Lc = P[18 + min((int)threadIdx.x, G.n_cam - 1)];
  if (img != nullptr) {
    // all of a thread's tile loads are issued before the first is used
    // (clamped indices, unpredicated; pixels outside the image become 0):
    // predicated loads in a loop made each one wait for the previous
    constexpr int kPerThread = kImgTile * kImgTile / kThreads;
    static_assert(kPerThread * kThreads == kImgTile * kImgTile, "tile loads per thread");
    float tv[kPerThread];
#pragma unroll
    for (int i = 0; i < kPerThread; ++i) {
      const int k = (int)threadIdx.x + i * kThreads;
      const int yy = min(max(oy + k / kImgTile, 0), G.img_h - 1), xx = min(max(ox + k % kImgTile, 0), G.img_w - 1);
      tv[i] = img[(int64_t)yy * G.img_w + xx];
    }
#pragma unroll
    for (int i = 0; i < kPerThread; ++i) {
      const int k = (int)threadIdx.x + i * kThreads;
      const int yy = oy + k / kImgTile, xx = ox + k % kImgTile;
      tile[k] = (yy >= 0 && yy < G.img_h && xx >= 0 && xx < G.img_w) ? tv[i] : 0.0f;
    }
  }
  if (threadIdx.x < G.n_cam) {
    double bx, by;
    offsets_at(G, P, p, Lc, bx, by);
    sh_a[threadIdx.x] = (float)(G.defocus + Lc);
    sh_bx[threadIdx.x] = (float)bx;
    sh_by[threadIdx.x] = (float)by;
  }
  __syncthreads();
  // the tile lies inside the image: lookups whose neighbours are all in the
  // tile need no image-bounds tests
  const bool tile_inside = oy >= 0 && ox >= 0 && oy + kImgTile <= G.img_h && ox + kImgTile <= G.img_w;
  const int jb = blockIdx.x * kThreads * kFwdRays;  // J <= 2^20
  // per-workgroup output bases; rays index them with 32-bit offsets
  float* __restrict__ vrow = values != nullptr ? values + s * J : nullptr;
  float2* __restrict__ crow = coords != nullptr ? (float2*)coords + s * J : nullptr;
  const int64_t cstride = S * J;  // between camera lengths, in float2
  // ray j = (row iy, column ix), stepped by kThreads without dividing
  const int n = G.n_side, drow = kThreads / n, dcol = kThreads % n;
  const int j0 = jb + (int)threadIdx.x;
  int iy = j0 / n, ix = j0 - (j0 / n) * n;
  const bool need_v = values != nullptr || image4d != nullptr;
#pragma unroll 4
  for (int u = 0; u < kFwdRays; ++u, ix += dcol, iy += drow) {
    if (ix >= n) {
      ix -= n;
      ++iy;
    }
    const int j = j0 + u * kThreads;
    if (j >= (int)J) continue;
    const float tx = th[ix], ty = th[iy];
    float v = 0.0f;
    if (need_v) {
      double fy, fx;
      sample_pixel(smap, tx, ty, fy, fx);
      const double y0d = floor(fy), x0d = floor(fx);
      // tile coordinates (clamped first: positions far off the image or NaN
      // only ever take the general path)
      const int y0 = (int)fmin(fmax(y0d, -1e8), 1e8) - oy, x0 = (int)fmin(fmax(x0d, -1e8), 1e8) - ox;
      if (tile_inside && y0 >= 0 && y0 < kImgTile - 1 && x0 >= 0 && x0 < kImgTile - 1) {
        // all four neighbours in the tile (hence in the image): bilinear_tiled's
        // accumulation without its bounds tests
        const float wy = (float)(fy - y0d), wx = (float)(fx - x0d);
        const int t = y0 * kImgTile + x0;
        v = fmaf((1.0f - wy) * (1.0f - wx), tile[t], 0.0f);
        v = fmaf((1.0f - wy) * wx, tile[t + 1], v);
        v = fmaf(wy * (1.0f - wx), tile[t + kImgTile], v);
        v = fmaf(wy * wx, tile[t + kImgTile + 1], v);
      } else {
        v = bilinear_tiled(G, img, tile, oy, ox, fy, fx);
      }
      if (vrow != nullptr) vrow[j] = v;
    }
    float2* __restrict__ co = crow;
    for (int c = 0; c < G.n_cam; ++c, co += cstride) {
      const float xd = fmaf(tx, sh_a[c], sh_bx[c]);
      const float yd = fmaf(ty, sh_a[c], sh_by[c]);
      if (co != nullptr) co[j] = make_float2(xd, yd);
      if (c == 0 && image4d != nullptr) {
        const double fy = __dadd_rn(__dadd_rn(__dmul_rn(G.det_m[0], (double)yd), __dmul_rn(G.det_m[1], (double)xd)),
                                    G.det_m[2]);
        const double fx = __dadd_rn(__dadd_rn(__dmul_rn(G.det_m[3], (double)yd), __dmul_rn(G.det_m[4], (double)xd)),
                                    G.det_m[5]);
        const int32_t py = tg::xla_f64_to_i32(rint(fy)), px = tg::xla_f64_to_i32(rint(fx));
        if (py >= 0 && py < G.det_h && px >= 0 && px < G.det_w)
          unsafeAtomicAdd(image4d + (s * G.det_h + py) * G.det_w + px, v);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// shadow-image back-projection: the adjoint of the 4D deposit above
// ---------------------------------------------------------------------------
// image[q] += sum over rays (s, j) whose detector pixel (camera `cam`) lies on
// the detector of data4d[s, pixel] * w_q(s, j), with w_q the bilinear weight
// of sample pixel q at the ray's sample-plane position (the same weights,
// bounds and detector binning as stem_forward_kernel); weight[q] += w_q.
// One scan position's sample footprint is a few pixels across
// (defocus x semi-convergence), so each workgroup accumulates into a
// kBpTile^2 LDS tile centred on it and flushes that once; contributions
// outside the tile go to global memory directly.  FP64 accumulation.
//
// Each thread takes a kBpBlock x kBpBlock block of adjacent rays, whose
// sample positions span well under a pixel on the usual geometries: the
// block's bilinear deposits then fall in a 3 x 3 pixel window anchored at
// the floor of the block's smallest coordinates, accumulated in registers and
// added to the tile once (9 LDS atomics per 16 rays instead of 64; same-
// address LDS atomics from the lanes of a wavefront serialise, and bound the
// one-ray-at-a-time form).  A ray whose cell falls outside the window (wide
// footprints) deposits its four neighbours directly to the LDS tile (or to
// global memory outside it).
constexpr int kBpTile = 32;
constexpr int kBpBlock = 4;                        // rays per side of one thread's block
constexpr int kBpRayTile = 16 * kBpBlock;          // rays per side of one workgroup (16 x 16 blocks)
static_assert(kThreads == 256, "one workgroup = 16 x 16 ray blocks");

// Destination of a deposit to sample pixel (yy, xx): its index in the LDS
// tile, kBpGlobal for global memory, kBpSkip outside the image.  (The atomics
// stay at the call sites on the __shared__ arrays themselves, so that they
// compile to LDS atomics rather than flat ones.)
constexpr int kBpGlobal = -1, kBpSkip = -2;

__device__ __forceinline__ int bp_slot(const tg_stem_geometry& G, int oy, int ox, int yy, int xx) {
  if (yy < 0 || yy >= G.img_h || xx < 0 || xx >= G.img_w) return kBpSkip;
  const int ty_ = yy - oy, tx_ = xx - ox;
  return ty_ >= 0 && ty_ < kBpTile && tx_ >= 0 && tx_ < kBpTile ? ty_ * kBpTile + tx_ : kBpGlobal;
}

// WEIGHT: also accumulate weight[q] (the sum of w_q), else weight is unused.
template <bool WEIGHT>
__global__ __launch_bounds__(kThreads) void stem_backproject_kernel(const tg_stem_geometry G,
                                                                    const double* __restrict__ P,
                                                                    const float* __restrict__ data, int cam,
                                                                    double* __restrict__ image,
                                                                    double* __restrict__ weight) {
  __shared__ float th[kMaxSide];
  __shared__ double t_img[kBpTile * kBpTile];
  __shared__ double t_w[WEIGHT ? kBpTile * kBpTile : 1];
  __shared__ float sh_a, sh_bx, sh_by;
  const int64_t s = blockIdx.y;
  const int n = G.n_side;
  for (int k = threadIdx.x; k < n; k += kThreads) th[k] = slope_of(G, k);
  for (int k = threadIdx.x; k < kBpTile * kBpTile; k += kThreads) {
    t_img[k] = 0.0;
    if (WEIGHT) t_w[k] = 0.0;
  }
  const ScanPos p = scan_position(G, P, s);
  if (threadIdx.x == 0) {
    const double L = P[18 + cam];
    double bx, by;
    offsets_at(G, P, p, L, bx, by);
    sh_a = (float)(G.defocus + L);
    sh_bx = (float)bx;
    sh_by = (float)by;
  }
  // tile origin: the footprint centre (zero slope) minus half a tile
  const SampleMap smap = sample_map(G, p);
  const double oyd = floor(smap.fy0) - kBpTile / 2, oxd = floor(smap.fx0) - kBpTile / 2;
  // origins far outside the image only route everything to global memory
  const int oy = (int)fmin(fmax(oyd, -4.0 * kBpTile), (double)G.img_h + 4.0 * kBpTile);
  const int ox = (int)fmin(fmax(oxd, -4.0 * kBpTile), (double)G.img_w + 4.0 * kBpTile);
  __syncthreads();
  // this thread's block of rays: rows iy0.., columns ix0.. (clipped to n)
  const int ntx = (n + kBpRayTile - 1) / kBpRayTile;
  const int iy0 = (int)(blockIdx.x / ntx) * kBpRayTile + (int)(threadIdx.x / 16) * kBpBlock;
  const int ix0 = (int)(blockIdx.x % ntx) * kBpRayTile + (int)(threadIdx.x % 16) * kBpBlock;
  if (iy0 < n && ix0 < n) {
    // The sample position fma(ky_y, ty, fma(ky_x, tx, fy0)) (likewise fx) and
    // the detector coordinate ((m0 yd) + (m1 xd)) + m2 (likewise fx) take
    // their inner terms from the ray's column (tx, xd) or row (ty, yd) alone:
    // computed once per row and column of the block, with the same roundings
    // as the per-ray expressions of stem_forward_kernel.  Rows and columns
    // past n repeat the last one (their rays are masked out below).
    double cy[kBpBlock], cx[kBpBlock], tyd[kBpBlock];
    double ry0[kBpBlock], ry3[kBpBlock], cx1[kBpBlock], cx4[kBpBlock];
#pragma unroll
    for (int k = 0; k < kBpBlock; ++k) {
      const float tx = th[min(ix0 + k, n - 1)], ty = th[min(iy0 + k, n - 1)];
      cy[k] = fma(smap.ky_x, (double)tx, smap.fy0);
      cx[k] = fma(smap.kx_x, (double)tx, smap.fx0);
      tyd[k] = (double)ty;
      const float xd = fmaf(tx, sh_a, sh_bx), yd = fmaf(ty, sh_a, sh_by);
      ry0[k] = __dmul_rn(G.det_m[0], (double)yd);
      ry3[k] = __dmul_rn(G.det_m[3], (double)yd);
      cx1[k] = __dmul_rn(G.det_m[1], (double)xd);
      cx4[k] = __dmul_rn(G.det_m[4], (double)xd);
    }
    // window anchor: floor of the smallest sample coordinates over the
    // block's corners (the position is affine in the slopes)
    constexpr int L = kBpBlock - 1;
    const double wy0 = floor(fmin(fmin(fma(smap.ky_y, tyd[0], cy[0]), fma(smap.ky_y, tyd[0], cy[L])),
                                  fmin(fma(smap.ky_y, tyd[L], cy[0]), fma(smap.ky_y, tyd[L], cy[L]))));
    const double wx0 = floor(fmin(fmin(fma(smap.kx_y, tyd[0], cx[0]), fma(smap.kx_y, tyd[0], cx[L])),
                                  fmin(fma(smap.kx_y, tyd[L], cx[0]), fma(smap.kx_y, tyd[L], cx[L]))));
    // anchors far outside the image: every ray deposits directly (and skips)
    const bool anchored = wy0 > -4.0 && wy0 < (double)G.img_h + 4.0 && wx0 > -4.0 && wx0 < (double)G.img_w + 4.0;
    const int Y = anchored ? (int)wy0 : 0, X = anchored ? (int)wx0 : 0;
    double acc[9], accw[WEIGHT ? 9 : 1];
#pragma unroll
    for (int k = 0; k < 9; ++k) acc[k] = 0.0;
#pragma unroll
    for (int k = 0; k < (WEIGHT ? 9 : 1); ++k) accw[k] = 0.0;
    // the block's 16 detector values, all loads issued before the first is
    // used (the position's first pixel stands in for a ray off the detector
    // or past n; `hit` says which rays are real): a load per ray behind its
    // bounds test waited for each in turn.  One scan position's detector
    // image is addressed by a 32-bit byte offset from a uniform base (the
    // host bounds det_h * det_w by 2^30).
    const char* __restrict__ dimg = (const char*)(data + s * G.det_h * G.det_w);
    float dv[kBpBlock * kBpBlock];
    uint32_t hit = 0;
#pragma unroll
    for (int r = 0; r < kBpBlock; ++r) {
#pragma unroll
      for (int c = 0; c < kBpBlock; ++c) {
        const double dfy = __dadd_rn(__dadd_rn(ry0[r], cx1[c]), G.det_m[2]);
        const double dfx = __dadd_rn(__dadd_rn(ry3[r], cx4[c]), G.det_m[5]);
        const int32_t py = tg::xla_f64_to_i32(rint(dfy)), px = tg::xla_f64_to_i32(rint(dfx));
        // (& rather than &&: one straight-line test, no exec-mask branches)
        const bool ok = (iy0 + r < n) & (ix0 + c < n) & ((uint32_t)py < (uint32_t)G.det_h) &
                        ((uint32_t)px < (uint32_t)G.det_w);
        dv[r * kBpBlock + c] = *(const float*)(dimg + (ok ? (uint32_t)(py * G.det_w + px) << 2 : 0u));
        hit |= (uint32_t)ok << (r * kBpBlock + c);
      }
    }
    __builtin_amdgcn_sched_barrier(0);  // all 16 loads in flight before the deposits
    // Per ray, straight-line: the window deposit runs for every ray with its
    // weights and value selected to zero unless the ray is real, on the
    // image and inside the window.  Rays on the image outside the window
    // (wide footprints; rare) are marked in `direct` and deposited by the
    // loop below.  (Conditional blocks in this unrolled loop would let the
    // compiler sink the gathered loads into them and keep the rays' values
    // live across them.)
    uint32_t direct = 0;
#pragma unroll
    for (int r = 0; r < kBpBlock; ++r) {
#pragma unroll
      for (int c = 0; c < kBpBlock; ++c) {
        const bool real = (hit >> (r * kBpBlock + c)) & 1u;
        // bilinear weights at the sample-plane position, as in bilinear_tiled()
        const double fy = fma(smap.ky_y, tyd[r], cy[c]), fx = fma(smap.kx_y, tyd[r], cx[c]);
        const double y0d = floor(fy), x0d = floor(fx);
        const bool on_img = (y0d >= -1.0) & (y0d < (double)G.img_h) & (x0d >= -1.0) & (x0d < (double)G.img_w);
        // (positions off the image, NaN included, convert a stand-in 0)
        const int y0 = (int)(on_img ? y0d : 0.0), x0 = (int)(on_img ? x0d : 0.0);
        const float wy = (float)(fy - y0d), wx = (float)(fx - x0d);
        const int dy = y0 - Y, dx = x0 - X;
        const bool win = real & on_img & anchored & ((unsigned)dy <= 1u) & ((unsigned)dx <= 1u);
        direct |= (uint32_t)(real & on_img & !win) << (r * kBpBlock + c);
        // separable weights placed in the 3 x 3 window: row i of the window
        // carries (1 - wy) at dy and wy at dy + 1, likewise columns.  rx * d
        // is exact in FP64 (two 24-bit significands), so each deposit
        // ry * rx * d is rounded once, by the FMA into the sum.
        float ry[3] = {win & (dy == 0) ? 1.0f - wy : 0.0f, win ? (dy == 0 ? wy : 1.0f - wy) : 0.0f,
                       win & (dy != 0) ? wy : 0.0f};
        float rx[3] = {win & (dx == 0) ? 1.0f - wx : 0.0f, win ? (dx == 0 ? wx : 1.0f - wx) : 0.0f,
                       win & (dx != 0) ? wx : 0.0f};
        float df = win ? dv[r * kBpBlock + c] : 0.0f;
        // (the selects stay in f32: without these empty statements the
        // compiler converts first and selects both halves of each double)
#pragma unroll
        for (int k = 0; k < 3; ++k) asm volatile("" : "+v"(ry[k]), "+v"(rx[k]));
        asm volatile("" : "+v"(df));
        const double d = (double)df;
        double rxd[3], t[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          rxd[k] = (double)rx[k];
          t[k] = rxd[k] * d;
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const double ryd = (double)ry[i];
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            acc[3 * i + k] = fma(ryd, t[k], acc[3 * i + k]);
            if (WEIGHT) accw[3 * i + k] = fma(ryd, rxd[k], accw[3 * i + k]);
          }
        }
        // one ray at a time: instruction selection would otherwise defer
        // every ray's sums past the last ray's set-up and hold all their
        // weights live at once.  (The empty statements emit nothing; they
        // make this ray's sums complete before the next ray's begin.)
#pragma unroll
        for (int k = 0; k < 9; ++k) {
          asm volatile("" : "+v"(acc[k]));
          if (WEIGHT) asm volatile("" : "+v"(accw[k]));
        }
      }
    }
    // direct deposits: the ray's position and detector value again, by the
    // same operations, then its four bilinear neighbours straight to the LDS
    // tile (or global memory outside it)
    while (direct != 0u) {
      const int k = __builtin_ctz(direct);
      direct &= direct - 1u;
      const int r = k / kBpBlock, c = k % kBpBlock;
      const float tx = th[ix0 + c], ty = th[iy0 + r];  // real rays lie inside n
      double fy, fx;
      sample_pixel(smap, tx, ty, fy, fx);
      const float xd = fmaf(tx, sh_a, sh_bx), yd = fmaf(ty, sh_a, sh_by);
      const double dfy = __dadd_rn(__dadd_rn(__dmul_rn(G.det_m[0], (double)yd), __dmul_rn(G.det_m[1], (double)xd)),
                                   G.det_m[2]);
      const double dfx = __dadd_rn(__dadd_rn(__dmul_rn(G.det_m[3], (double)yd), __dmul_rn(G.det_m[4], (double)xd)),
                                   G.det_m[5]);
      const int32_t py = tg::xla_f64_to_i32(rint(dfy)), px = tg::xla_f64_to_i32(rint(dfx));
      const double dd = (double)*(const float*)(dimg + ((uint32_t)(py * G.det_w + px) << 2));
      const double y0d = floor(fy), x0d = floor(fx);
      const int y0 = (int)y0d, x0 = (int)x0d;  // on the image: marked rays only
      const float wy = (float)(fy - y0d), wx = (float)(fx - x0d);
#pragma unroll
      for (int a = 0; a < 2; ++a) {
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          const double w = (double)(a ? wy : 1.0f - wy) * (double)(b ? wx : 1.0f - wx);  // exact
          const int q = bp_slot(G, oy, ox, y0 + a, x0 + b);
          if (q >= 0) {
            atomicAdd(&t_img[q], w * dd);
            if (WEIGHT) atomicAdd(&t_w[q], w);
          } else if (q == kBpGlobal) {
            const int64_t gq = (int64_t)(y0 + a) * G.img_w + (x0 + b);
            unsafeAtomicAdd(image + gq, w * dd);
            if (WEIGHT) unsafeAtomicAdd(weight + gq, w);
          }
        }
      }
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      // untouched window pixels hold exact zeros (as do zero deposits, which
      // change no sum)
      if (WEIGHT ? accw[k] == 0.0 : acc[k] == 0.0) continue;
      const int yy = Y + k / 3, xx = X + k % 3;
      const int q = bp_slot(G, oy, ox, yy, xx);
      if (q >= 0) {
        atomicAdd(&t_img[q], acc[k]);
        if (WEIGHT) atomicAdd(&t_w[q], accw[k]);
      } else if (q == kBpGlobal) {
        const int64_t gq = (int64_t)yy * G.img_w + xx;
        unsafeAtomicAdd(image + gq, acc[k]);
        if (WEIGHT) unsafeAtomicAdd(weight + gq, accw[k]);
      }
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < kBpTile * kBpTile; k += kThreads) {
    const int yy = oy + k / kBpTile, xx = ox + k % kBpTile;
    if (yy < 0 || yy >= G.img_h || xx < 0 || xx >= G.img_w) continue;  // never written
    const int64_t q = (int64_t)yy * G.img_w + xx;
    if (t_img[k] != 0.0) unsafeAtomicAdd(image + q, t_img[k]);
    if (WEIGHT && t_w[k] != 0.0) unsafeAtomicAdd(weight + q, t_w[k]);
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
// Indices past the row's end are clamped to its last float4 (the tile sums
// skip those pairs), so every load is issued unpredicated: predicated loads
// made the compiler wait for all outstanding loads before the first use.
__device__ __forceinline__ void load_tile(const f4* __restrict__ T4, int64_t S, int64_t s, int64_t npair,
                                          int64_t ntile, int64_t it, f4 (&dst)[kPerLane]) {
  const int64_t c = it / ntile, tile = it - c * ntile;
  const int64_t row = (c * S + s) * npair;
  const int64_t q0 = tile * kTile + threadIdx.x;
#pragma unroll
  for (int k = 0; k < kPerLane; ++k) {
    const int64_t q = q0 + (int64_t)k * kThreads;
    dst[k] = __builtin_nontemporal_load(T4 + row + (q < npair ? q : npair - 1));
  }
}

// Per-scan-position gradient terms from the per-camera moments M[c][k]
// (FP64 sums of the wavefront partials, waves in order); thread t < np
// writes entry t of the workspace row s (the np terms of one scan position
// are contiguous: one short coalesced store per workgroup, where a column
// layout cost 27 scattered 8-byte stores and ~9 us per pass).  Every term
// is linear in the moments with coefficients affine in the camera length
// L_c, so entries 0..17 need only the camera sums of the x and y residual
// totals, plain and L-weighted; entry 18 + c takes camera c alone and the
// last entry sums |r|^2.
__device__ __forceinline__ void fit_epilogue(const tg_stem_geometry& G, const double* __restrict__ P,
                                             const ScanPos& p, const double (&M)[TG_STEM_MAX_CAM][kMom], int64_t s,
                                             int64_t S, double* __restrict__ partials) {
  const int np = 18 + G.n_cam + 1;
  const int t = threadIdx.x;
  if (t >= np) return;
  const double* A = P;
  const double* L = P + 18;
  double g = 0.0;
  if (t == np - 1) {
    for (int c = 0; c < G.n_cam; ++c) g += M[c][6];
  } else if (t >= 18) {
    const int c = t - 18;
    const double dbx_dl = row_dot(A, 2, p) + G.offsets[2];
    const double dby_dl = row_dot(A, 3, p) + G.offsets[3];
    g = 2.0 * (M[c][1] + dbx_dl * M[c][0] + M[c][5] + dby_dl * M[c][3]);
  } else {
    double sx = 0.0, sxl = 0.0, sy = 0.0, syl = 0.0;
    for (int c = 0; c < G.n_cam; ++c) {
      sx += M[c][0];
      sxl += L[c] * M[c][0];
      sy += M[c][3];
      syl += L[c] * M[c][3];
    }
    if (t < 16) {
      // d/dA[r][k] of sum_c |r_c|^2: rows 0, 1 act on the offsets (x, y),
      // rows 2, 3 on the slopes (scaled by L_c); column k picks v[k]
      const int r = t >> 2, k = t & 3;
      const double vk = k == 0 ? p.spx : k == 1 ? p.spy : k == 2 ? p.stx : p.sty;
      const double m = r == 0 ? sx : r == 1 ? sy : r == 2 ? sxl : syl;
      g = 2.0 * vk * m;
    } else {
      // d(Bx, By) / d(spx, spy) = a + b L_c, with st = pivot * sp
      const double d = G.defocus, kap = G.pivot;
      const double ax_x = kap * d + (A[0] + A[2] * kap), bx_x = A[8] + A[10] * kap;
      const double ax_y = A[1] + A[3] * kap, bx_y = A[9] + A[11] * kap;
      const double ay_x = A[4] + A[6] * kap, by_x = A[12] + A[14] * kap;
      const double ay_y = kap * d + (A[5] + A[7] * kap), by_y = A[13] + A[15] * kap;
      double dsx, dsy;
      if (t == 16) {
        // step: sp = step * (rotated integer pixel offsets), so d sp / d step
        // is the offsets themselves (finite at step = 0, where sp / step is 0/0)
        const int64_t i = s / G.scan_w, j = s % G.scan_w;
        const double u1 = (double)i - 0.5 * (G.scan_h - 1);
        const double v1 = (double)j - 0.5 * (G.scan_w - 1);
        double sn, cs;
        sincos(P[17] * (kPi / 180.0), &sn, &cs);
        dsx = -sn * u1 + cs * v1;
        dsy = -(cs * u1 + sn * v1);
      } else {        // rotation in degrees: d sp_x = sp_y dphi, d sp_y = -sp_x dphi
        dsx = p.spy * (kPi / 180.0);
        dsy = -p.spx * (kPi / 180.0);
      }
      g = 2.0 * ((ax_x * dsx + ax_y * dsy) * sx + (bx_x * dsx + bx_y * dsy) * sxl + (ay_x * dsx + ay_y * dsy) * sy +
                 (by_x * dsx + by_y * dsy) * syl);
    }
  }
  partials[s * np + t] = g;
}

// Camera-length constants of the residual r = th a + b - t as double-float
// splits: r = (fma(th, a_hi, -t) + b_hi) + fma(th, a_lo, b_lo) is exact to
// about one ulp of r itself (fma(th, a_hi, -t) rounds once), not to one ulp
// of the ~1e-4 m detector coordinate.
struct CamConst {
  float a_hi, a_lo, bx_hi, bx_lo, by_hi, by_lo;
};

__device__ __forceinline__ CamConst cam_const(const tg_stem_geometry& G, const double* __restrict__ P,
                                              const ScanPos& p, int c) {
  const double L = P[18 + c];
  double bx, by;
  offsets_at(G, P, p, L, bx, by);
  const double ad = G.defocus + L;
  CamConst k;
  k.a_hi = (float)ad;
  k.a_lo = (float)(ad - (double)k.a_hi);
  k.bx_hi = (float)bx;
  k.bx_lo = (float)(bx - (double)k.bx_hi);
  k.by_hi = (float)by;
  k.by_lo = (float)(by - (double)k.by_hi);
  return k;
}

// Adds one tile's residual moments to m[7] (per lane, FP32).
__device__ __forceinline__ void accumulate_tile(const tg_stem_geometry& G, const float* __restrict__ th,
                                                const CamConst& K, const f4 (&cur)[kPerLane], int64_t tile,
                                                int64_t npair, float (&m)[kMom]) {
  const int half = G.n_side >> 1, drow = kThreads / half, dcol = kThreads % half;
  // ray pair q = (row iy, columns 2*hc, 2*hc+1); stepped without division
  const int64_t q0 = tile * kTile + threadIdx.x;
  int iy = (int)(q0 / half), hc = (int)(q0 % half);
  if (dcol == 0 && (tile + 1) * kTile <= npair) {
    // full tile and a column count that divides the workgroup: this lane's
    // two columns are fixed, rows advance by drow per pair
    const float tx0 = th[2 * hc], tx1 = th[2 * hc + 1];
#pragma unroll
    for (int k = 0; k < kPerLane; ++k) {
      const f4 t = cur[k];
      const float ty = th[iy + k * drow];
      const float rx0 = (fmaf(tx0, K.a_hi, -t.x) + K.bx_hi) + fmaf(tx0, K.a_lo, K.bx_lo);
      const float ry0 = (fmaf(ty, K.a_hi, -t.y) + K.by_hi) + fmaf(ty, K.a_lo, K.by_lo);
      const float rx1 = (fmaf(tx1, K.a_hi, -t.z) + K.bx_hi) + fmaf(tx1, K.a_lo, K.bx_lo);
      const float ry1 = (fmaf(ty, K.a_hi, -t.w) + K.by_hi) + fmaf(ty, K.a_lo, K.by_lo);
      m[0] += rx0 + rx1;
      m[1] = fmaf(rx0, tx0, fmaf(rx1, tx1, m[1]));
      m[2] = fmaf(rx0 + rx1, ty, m[2]);
      m[3] += ry0 + ry1;
      m[4] = fmaf(ry0, tx0, fmaf(ry1, tx1, m[4]));
      m[5] = fmaf(ry0 + ry1, ty, m[5]);
      m[6] = fmaf(rx0, rx0, fmaf(ry0, ry0, fmaf(rx1, rx1, fmaf(ry1, ry1, m[6]))));
    }
    return;
  }
#pragma unroll
  for (int k = 0; k < kPerLane; ++k) {
    if (q0 + (int64_t)k * kThreads < npair) {
      const f4 t = cur[k];
      const float tx0 = th[2 * hc], tx1 = th[2 * hc + 1], ty = th[iy];
      const float rx0 = (fmaf(tx0, K.a_hi, -t.x) + K.bx_hi) + fmaf(tx0, K.a_lo, K.bx_lo);
      const float ry0 = (fmaf(ty, K.a_hi, -t.y) + K.by_hi) + fmaf(ty, K.a_lo, K.by_lo);
      const float rx1 = (fmaf(tx1, K.a_hi, -t.z) + K.bx_hi) + fmaf(tx1, K.a_lo, K.bx_lo);
      const float ry1 = (fmaf(ty, K.a_hi, -t.w) + K.by_hi) + fmaf(ty, K.a_lo, K.by_lo);
      m[0] += rx0 + rx1;
      m[1] = fmaf(rx0, tx0, fmaf(rx1, tx1, m[1]));
      m[2] = fmaf(rx0 + rx1, ty, m[2]);
      m[3] += ry0 + ry1;
      m[4] = fmaf(ry0, tx0, fmaf(ry1, tx1, m[4]));
      m[5] = fmaf(ry0 + ry1, ty, m[5]);
      m[6] = fmaf(rx0, rx0, fmaf(ry0, ry0, fmaf(rx1, rx1, fmaf(ry1, ry1, m[6]))));
    }
    hc += dcol;
    iy += drow;
    if (hc >= half) {
      hc -= half;
      ++iy;
    }
  }
}

// Wavefront sums of m[7] into wm[wave][k] (lane 0 writes).
__device__ __forceinline__ void wave_moments(const float (&m)[kMom], double (&wm)[kWaves][kMom]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < kMom; ++k) {
    const float v = wave_sum(m[k]);
    if (lane == 0) wm[wave][k] = v;
  }
}

__device__ __forceinline__ void fill_slopes(const tg_stem_geometry& G, float* th) {
  for (int k = threadIdx.x; k < G.n_side; k += kThreads) th[k] = slope_of(G, k);
}

// Variant "scan": one workgroup per scan position streams the tiles of every
// camera length, double-buffered across tiles and cameras.  (A persistent
// variant that also prefetched across scan positions needed 140 VGPRs, fell
// to 3 waves per SIMD and was slower.)
__global__ __launch_bounds__(kThreads) void stem_fit_kernel(const tg_stem_geometry G, const double* __restrict__ P,
                                                            const float* __restrict__ T,
                                                            double* __restrict__ partials) {
  __shared__ float th[kMaxSide];
  __shared__ double wm[TG_STEM_MAX_CAM][kWaves][kMom];
  __shared__ double M[TG_STEM_MAX_CAM][kMom];
  __shared__ double sp[18 + TG_STEM_MAX_CAM];  // the parameters, read by the set-up and the epilogue
  const int64_t S = (int64_t)G.scan_h * G.scan_w;
  const int64_t J = (int64_t)G.n_side * G.n_side;
  fill_slopes(G, th);
  if (threadIdx.x < 18 + G.n_cam) sp[threadIdx.x] = P[threadIdx.x];
  __syncthreads();
  const int64_t npair = J >> 1;  // two rays per float4 of targets (n_side even)
  const int64_t ntile = (npair + kTile - 1) / kTile;
  const int64_t nit = (int64_t)G.n_cam * ntile;
  const f4* __restrict__ T4 = (const f4*)T;
  f4 cur[kPerLane], nxt[kPerLane];
  const int64_t s = blockIdx.x;
  load_tile(T4, S, s, npair, ntile, 0, cur);
  const ScanPos p = scan_position(G, sp, s);
  CamConst K = {};
  float m[kMom];
  for (int64_t it = 0; it < nit; ++it) {
    if (it + 1 < nit) load_tile(T4, S, s, npair, ntile, it + 1, nxt);
    const int c = (int)(it / ntile);
    const int64_t tile = it - (int64_t)c * ntile;
    if (tile == 0) {
      K = cam_const(G, sp, p, c);
#pragma unroll
      for (int k = 0; k < kMom; ++k) m[k] = 0.f;
    }
    accumulate_tile(G, th, K, cur, tile, npair, m);
    if (tile == ntile - 1) wave_moments(m, wm[c]);
#pragma unroll
    for (int k = 0; k < kPerLane; ++k) cur[k] = nxt[k];
  }
  __syncthreads();
  if (threadIdx.x < G.n_cam * kMom) {
    const int c = threadIdx.x / kMom, k = threadIdx.x % kMom;
    double acc = 0.0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) acc += wm[c][w][k];
    M[c][k] = acc;
  }
  __syncthreads();
  fit_epilogue(G, sp, p, M, s, S, partials);
}

// Variant "camera": one workgroup per (scan position, camera length) --
// n_cam times more, shorter work items, for scans too small to fill the
// device one workgroup per position.  Each pays the per-workgroup set-up
// (slopes, scan position, camera constants) without a second tile to
// overlap it with, so on large scans it is slower than "scan".  Writes the camera's FP64 moments; stem_fit_epilogue
// finishes each scan position with the same arithmetic as the scan variant.
__global__ __launch_bounds__(kThreads) void stem_fit_cam_kernel(const tg_stem_geometry G,
                                                                const double* __restrict__ P,
                                                                const float* __restrict__ T,
                                                                double* __restrict__ cam_mom) {
  __shared__ float th[kMaxSide];
  __shared__ double wm[kWaves][kMom];
  const int64_t S = (int64_t)G.scan_h * G.scan_w;
  const int64_t J = (int64_t)G.n_side * G.n_side;
  fill_slopes(G, th);
  __syncthreads();
  const int64_t npair = J >> 1;
  const int64_t ntile = (npair + kTile - 1) / kTile;
  const f4* __restrict__ T4 = (const f4*)T;
  const int64_t s = blockIdx.x;
  const int c = blockIdx.y;
  f4 cur[kPerLane], nxt[kPerLane];
  load_tile(T4, S, s, npair, ntile, (int64_t)c * ntile, cur);
  const ScanPos p = scan_position(G, P, s);
  const CamConst K = cam_const(G, P, p, c);
  float m[kMom];
#pragma unroll
  for (int k = 0; k < kMom; ++k) m[k] = 0.f;
  for (int64_t tile = 0; tile < ntile; ++tile) {
    if (tile + 1 < ntile) load_tile(T4, S, s, npair, ntile, (int64_t)c * ntile + tile + 1, nxt);
    accumulate_tile(G, th, K, cur, tile, npair, m);
#pragma unroll
    for (int k = 0; k < kPerLane; ++k) cur[k] = nxt[k];
  }
  wave_moments(m, wm);
  __syncthreads();
  if (threadIdx.x < kMom) {
    double acc = 0.0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) acc += wm[w][threadIdx.x];
    cam_mom[(s * G.n_cam + c) * kMom + threadIdx.x] = acc;
  }
}

__global__ __launch_bounds__(64) void stem_fit_epilogue(const tg_stem_geometry G, const double* __restrict__ P,
                                                        const double* __restrict__ cam_mom,
                                                        double* __restrict__ partials) {
  __shared__ double M[TG_STEM_MAX_CAM][kMom];
  const int64_t S = (int64_t)G.scan_h * G.scan_w;
  const int64_t s = blockIdx.x;
  for (int i = threadIdx.x; i < G.n_cam * kMom; i += 64) M[i / kMom][i % kMom] = cam_mom[s * G.n_cam * kMom + i];
  __syncthreads();
  const ScanPos p = scan_position(G, P, s);
  fit_epilogue(G, P, p, M, s, S, partials);
}

// Fixed-order FP64 column sums of the [rows, cols] workspace (one block per
// column, strided reads of the L2-resident terms, wavefront shuffles then
// one LDS step).
constexpr int kSumThreads = 1024;

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ __launch_bounds__(kSumThreads) void row_sum_kernel(const double* __restrict__ partials, int64_t rows,
                                                              int64_t cols, double* __restrict__ out) {
  __shared__ double red[kSumThreads / 64];
  const double* __restrict__ src = partials + blockIdx.x;  // column blockIdx.x of [rows, cols]
  double acc = 0.0;
  int64_t r = threadIdx.x;
  // four loads in flight per step (a plain loop waited for each), summed in
  // the same order as one at a time
  for (; r + 3 * kSumThreads < rows; r += 4 * kSumThreads) {
    const double a = src[r * cols], b = src[(r + kSumThreads) * cols];
    const double c = src[(r + 2 * kSumThreads) * cols], d = src[(r + 3 * kSumThreads) * cols];
    acc += a;
    acc += b;
    acc += c;
    acc += d;
  }
  for (; r < rows; r += kSumThreads) acc += src[r * cols];
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

extern "C" int tg_stem_backproject(const tg_stem_geometry* g, const double* params, const float* data4d,
                                   int32_t cam, double* image, double* weight, void* stream) {
  tg::clear_error();
  int st = validate(g, params);
  if (st != TG_OK) return st;
  TG_REQUIRE(data4d != nullptr && image != nullptr, "tg_stem_backproject: NULL data or image");
  TG_REQUIRE(cam >= 0 && cam < g->n_cam, "tg_stem_backproject: camera %d out of range", cam);
  TG_REQUIRE(g->img_h > 0 && g->img_w > 0 && g->det_h > 0 && g->det_w > 0,
             "tg_stem_backproject: bad sample or detector shape");
  TG_REQUIRE((int64_t)g->det_h * g->det_w <= (int64_t)1 << 30,
             "tg_stem_backproject: detector of %lld pixels exceeds 2^30", (long long)g->det_h * g->det_w);
  const int64_t ntile = tg::cdiv((int64_t)g->n_side, (int64_t)kBpRayTile);
  dim3 grid((unsigned)(ntile * ntile), (unsigned)((int64_t)g->scan_h * g->scan_w));
  if (weight != nullptr)
    hipLaunchKernelGGL(stem_backproject_kernel<true>, grid, dim3(kThreads), 0, (hipStream_t)stream, *g, params, data4d,
                       cam, image, weight);
  else
    hipLaunchKernelGGL(stem_backproject_kernel<false>, grid, dim3(kThreads), 0, (hipStream_t)stream, *g, params,
                       data4d, cam, image, weight);
  TG_LAUNCH_CHECK("stem_backproject_kernel launch");
  return TG_OK;
}

extern "C" int64_t tg_stem_fit_workspace(const tg_stem_geometry* g) {
  if (g == nullptr || g->scan_h <= 0 || g->scan_w <= 0 || g->n_cam < 1 || g->n_cam > TG_STEM_MAX_CAM) return 0;
  const int64_t S = (int64_t)g->scan_h * g->scan_w;
  return (int64_t)sizeof(double) * S * ((18 + g->n_cam + 1) + (int64_t)g->n_cam * kMom);
}

namespace {

// The per-scan-position terms [S, np] at the head of the workspace (the
// launches of tg_stem_fit before its final reduction).
int fit_terms(const tg_stem_geometry* g, const double* params, const float* targets, double* workspace,
              hipStream_t stream) {
  TG_REQUIRE(targets != nullptr && workspace != nullptr, "tg_stem_fit: NULL buffer");
  TG_REQUIRE(((uintptr_t)targets & 15) == 0, "tg_stem_fit: targets must be 16-byte aligned");
  const int64_t S = (int64_t)g->scan_h * g->scan_w;
  const int np = 18 + g->n_cam + 1;
  double* partials = workspace;                   // [S, np]
  double* cam_mom = workspace + (int64_t)np * S;  // [S, n_cam, 7]
  // launch variant: the autotuner's choice, else one workgroup per scan
  // position when that gives every CU several workgroups, and one per (scan
  // position, camera length) for small scans
  const int64_t key[4] = {S, (int64_t)g->n_side * g->n_side, g->n_cam, 0};
  int32_t variant = S >= 4LL * tg::cu_count() ? TG_STEM_FIT_SCAN : TG_STEM_FIT_CAMERA;
  int32_t tuned = 0;
  if (tg::tune_lookup(TG_TUNE_STEM_FIT, key, tuned)) variant = tuned;
  if (variant == TG_STEM_FIT_SCAN) {
    hipLaunchKernelGGL(stem_fit_kernel, dim3((unsigned)S), dim3(kThreads), 0, (hipStream_t)stream, *g, params,
                       targets, partials);
    TG_LAUNCH_CHECK("stem_fit_kernel launch");
  } else {
    hipLaunchKernelGGL(stem_fit_cam_kernel, dim3((unsigned)S, (unsigned)g->n_cam), dim3(kThreads), 0,
                       (hipStream_t)stream, *g, params, targets, cam_mom);
    TG_LAUNCH_CHECK("stem_fit_cam_kernel launch");
    hipLaunchKernelGGL(stem_fit_epilogue, dim3((unsigned)S), dim3(64), 0, (hipStream_t)stream, *g, params, cam_mom,
                       partials);
    TG_LAUNCH_CHECK("stem_fit_epilogue launch");
  }
  return TG_OK;
}

}  // namespace

extern "C" int tg_stem_fit_terms(const tg_stem_geometry* g, const double* params, const float* targets,
                                 double* workspace, void* stream) {
  tg::clear_error();
  int st = validate(g, params);
  if (st != TG_OK) return st;
  return fit_terms(g, params, targets, workspace, (hipStream_t)stream);
}

extern "C" int tg_stem_fit(const tg_stem_geometry* g, const double* params, const float* targets, double* workspace,
                           double* out, void* stream) {
  tg::clear_error();
  int st = validate(g, params);
  if (st != TG_OK) return st;
  TG_REQUIRE(out != nullptr, "tg_stem_fit: NULL buffer");
  st = fit_terms(g, params, targets, workspace, (hipStream_t)stream);
  if (st != TG_OK) return st;
  const int64_t S = (int64_t)g->scan_h * g->scan_w;
  const int np = 18 + g->n_cam + 1;
  hipLaunchKernelGGL(row_sum_kernel, dim3((unsigned)np), dim3(kSumThreads), 0, (hipStream_t)stream, workspace, S,
                     (int64_t)np, out);
  TG_LAUNCH_CHECK("row_sum_kernel launch");
  return TG_OK;
}

// ---------------------------------------------------------------------------
// tg_stem_fit_noalloc: tg_stem_fit with its workspace in a device array of
// this code object instead of a caller buffer, for callers that cannot
// allocate (HIP graph capture, allocation-free serving loops).  Each GPU holds
// its own copy of the array until the process ends.  Host code never passes
// the array's name to a kernel (on the host it is only a stand-in for the
// symbol): the device address comes from hipGetSymbolAddress, once per
// device, with that device current.
// ---------------------------------------------------------------------------

namespace {

constexpr int64_t kNoAllocScratchBytes = 8LL << 20;  // 8 MiB per GPU
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kScratchDevices] = {};

// Device address of this device's scratch (the current device, which must be
// the stream's device).
int scratch_address(hipStream_t stream, double** out) {
  int dev = -1;
  TG_HIP_CHECK(hipGetDevice(&dev));
  TG_REQUIRE(dev >= 0 && dev < kScratchDevices, "tg_stem_fit_noalloc: device %d out of range", dev);
  if (stream != nullptr) {
    int sdev = -1;
    TG_HIP_CHECK(hipStreamGetDevice(stream, &sdev));
    TG_REQUIRE(sdev == dev, "tg_stem_fit_noalloc: the stream belongs to device %d, the current device is %d", sdev,
               dev);
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    TG_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    TG_REQUIRE(p != nullptr, "tg_stem_fit_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return TG_OK;
}

}  // namespace

extern "C" int64_t tg_stem_fit_noalloc_max_bytes(void) { return kNoAllocScratchBytes; }

extern "C" int tg_stem_fit_noalloc(const tg_stem_geometry* g, const double* params, const float* targets,
                                   double* out, void* stream) {
  tg::clear_error();
  int st = validate(g, params);
  if (st != TG_OK) return st;
  TG_REQUIRE(out != nullptr, "tg_stem_fit_noalloc: NULL buffer");
  const int64_t need = tg_stem_fit_workspace(g);
  TG_REQUIRE(need > 0 && need <= kNoAllocScratchBytes,
             "tg_stem_fit_noalloc: this geometry needs %lld bytes of scratch, more than the %lld available "
             "(use tg_stem_fit with a caller workspace)",
             (long long)need, (long long)kNoAllocScratchBytes);
  double* ws = nullptr;
  st = scratch_address((hipStream_t)stream, &ws);
  if (st != TG_OK) return st;
  // every word of the workspace a launch reads is written earlier in the same
  // call (per-position rows, per-camera moments): nothing needs clearing
  st = fit_terms(g, params, targets, ws, (hipStream_t)stream);
  if (st != TG_OK) return st;
  const int64_t S = (int64_t)g->scan_h * g->scan_w;
  const int np = 18 + g->n_cam + 1;
  hipLaunchKernelGGL(row_sum_kernel, dim3((unsigned)np), dim3(kSumThreads), 0, (hipStream_t)stream, ws, S,
                     (int64_t)np, out);
  TG_LAUNCH_CHECK("row_sum_kernel launch");
  return TG_OK;
}
