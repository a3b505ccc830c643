// Gaussian-beamlet field sums (reference gaussian.py:225-427).
//
// Stage 1 (beamlet_coef_kernel, one lane per beamlet, f64): the reference's
// per-beamlet 2x2 complex algebra -- Q1^-1 from waist/radius/theta
// (gaussian.py:138-177), Q2^-1 = solve(A + B Q1^-1, C + D Q1^-1)
// (gaussian.py:92-96), the misalignment phases phi1, phi2, the offset term
// 2 r.f and the prefactor amp / sqrt(det(A + B Q1^-1)) (gaussian.py:276-316),
// or the input-plane Gaussian (gaussian.py:402-407) -- is folded together
// with the k*pathlength phase and the Grid's pixel->metre map into
//   Phi(px, py) = a px^2 + b px py + c py^2 + d px + e py + f   (complex),
// so that a beamlet's field at pixel (py, px) is exp(Phi).  The envelope
// exp(Re Phi) gives the culling ellipse.
//
// Stage 2 (beamlet_sum_kernel): a workgroup owns a 64x64 detector tile and a
// slice of the beamlets.  Beamlets whose cutoff ellipse overlaps the tile are
// compacted (deterministically, by wavefront ballots) into an LDS list.  Each
// lane owns a run of 16 pixels of one row; per beamlet it evaluates Phi in
// f64 at the first pixel of the run that lies inside the ellipse (phase
// reduced to revolutions in double-double), then walks the run with the
// complex multiplicative recurrence F <- F R, R <- R K (K = exp(2a)): two
// complex f32 products per pixel, no transcendental per pixel.  Partial tiles
// of the beamlet slices are summed in a fixed order by a final kernel.
#include "beamlet_common.hpp"

namespace {

using namespace tgb;

__global__ __launch_bounds__(256) void beamlet_coef_kernel(const double* __restrict__ params, int64_t n,
                                                           const double* __restrict__ abcd, int64_t abcd_stride,
                                                           int32_t mode, double m0, double m1, double m2, double m3,
                                                           double m4, double m5, float cut,
                                                           tg_beamlet_coef* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* p = params + i * TG_BEAMLET_NPARAM;
  const double m[6] = {m0, m1, m2, m3, m4, m5};
  const double* T = mode == TG_BEAMLET_PROPAGATED ? abcd + i * abcd_stride : nullptr;
  out[i] = make_coef(p, T, mode, m, mk(p[5], p[6]), cut);
}

__global__ __launch_bounds__(kThreads) void beamlet_sum_kernel(const tg_beamlet_coef* __restrict__ coef, int64_t n,
                                                               int64_t H, int64_t W, int32_t ntx, int32_t ntiles,
                                                               int32_t nsplit, float2* __restrict__ ws) {
  __shared__ TileLists sh;
  const int tile = blockIdx.x % ntiles, split = blockIdx.x / ntiles;
  const int32_t tx0 = (tile % ntx) * kTileW, ty0 = (tile / ntx) * kTileH;
  const int64_t b_begin = n * split / nsplit, b_end = n * (split + 1) / nsplit;
  const int t = threadIdx.x;
  const int32_t row = ty0 + t / kLanesPerRow;
  const int32_t x0 = tx0 + (t % kLanesPerRow) * kRun;
  float2 acc[kRun];
#pragma unroll
  for (int j = 0; j < kRun; ++j) acc[j] = make_float2(0.f, 0.f);
  tile_sum(coef, b_begin, b_end, tx0, ty0, (int32_t)H, (int32_t)W, sh, acc);
  if (row < H) {
    float2* dst = ws + ((int64_t)split * H + row) * W;
#pragma unroll
    for (int j = 0; j < kRun; ++j)
      if (x0 + j < W) dst[x0 + j] = acc[j];
  }
}

__global__ __launch_bounds__(256) void split_sum_kernel(const float2* __restrict__ ws, int64_t npix, int32_t nsplit,
                                                        float2* __restrict__ field) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npix) return;
  double re = 0.0, im = 0.0;
  for (int s = 0; s < nsplit; ++s) {
    const float2 v = ws[(int64_t)s * npix + i];
    re += v.x;
    im += v.y;
  }
  field[i] = make_float2((float)re, (float)im);
}

int splits_for(int64_t n, int64_t H, int64_t W) {
  const int64_t ntiles = tg::cdiv(H, kTileH) * tg::cdiv(W, kTileW);
  // Many small work items (tile x beamlet slice) let the hardware scheduler
  // balance tiles near the centre of the pattern, which overlap many more
  // beamlets than the edge tiles.
  const int64_t want = 64LL * tg::cu_count();
  int64_t s = tg::cdiv(want, ntiles);
  const int64_t by_n = tg::cdiv(n, 256);  // at least ~256 beamlets per slice
  if (s > by_n) s = by_n;
  if (s > 128) s = 128;
  if (s < 1) s = 1;
  return (int)s;
}

}  // namespace

extern "C" int tg_beamlet_coefficients(const double* params, int64_t n, const double* abcd, int64_t abcd_stride,
                                       int32_t mode, const double* pix2m, double nsigma, tg_beamlet_coef* out,
                                       void* stream) {
  tg::clear_error();
  TG_REQUIRE(n >= 0, "tg_beamlet_coefficients: negative count");
  TG_REQUIRE(mode == TG_BEAMLET_PROPAGATED || mode == TG_BEAMLET_INPUT, "tg_beamlet_coefficients: bad mode");
  TG_REQUIRE(pix2m != nullptr, "tg_beamlet_coefficients: NULL pixel transform");
  if (n == 0) return TG_OK;
  TG_REQUIRE(params != nullptr && out != nullptr, "tg_beamlet_coefficients: NULL buffer");
  TG_REQUIRE(mode == TG_BEAMLET_INPUT || (abcd != nullptr && (abcd_stride == 0 || abcd_stride == 25)),
             "tg_beamlet_coefficients: propagated mode needs abcd with stride 0 or 25");
  const float cut = nsigma > 0.0 ? (float)(nsigma * nsigma) : 80.0f;
  hipLaunchKernelGGL(beamlet_coef_kernel, dim3((unsigned)tg::cdiv(n, 256)), dim3(256), 0, (hipStream_t)stream, params,
                     n, abcd, abcd_stride, mode, pix2m[0], pix2m[1], pix2m[2], pix2m[3], pix2m[4], pix2m[5], cut, out);
  TG_LAUNCH_CHECK("beamlet_coef_kernel launch");
  return TG_OK;
}

extern "C" int64_t tg_beamlet_sum_workspace(int64_t n, int64_t height, int64_t width) {
  if (n < 0 || height <= 0 || width <= 0) return 0;
  return (int64_t)splits_for(n, height, width) * height * width * (int64_t)sizeof(float2);
}

extern "C" int tg_beamlet_sum(const tg_beamlet_coef* coef, int64_t n, int64_t height, int64_t width, float* field,
                              void* workspace, int64_t workspace_bytes, void* stream) {
  tg::clear_error();
  TG_REQUIRE(n >= 0 && height > 0 && width > 0, "tg_beamlet_sum: bad sizes");
  TG_REQUIRE(field != nullptr && workspace != nullptr, "tg_beamlet_sum: NULL buffer");
  TG_REQUIRE(n == 0 || coef != nullptr, "tg_beamlet_sum: NULL coefficients");
  TG_REQUIRE(n <= (int64_t)1 << 31, "tg_beamlet_sum: too many beamlets");
  const int nsplit = splits_for(n, height, width);
  const int64_t need = (int64_t)nsplit * height * width * (int64_t)sizeof(float2);
  TG_REQUIRE(workspace_bytes >= need, "tg_beamlet_sum: workspace %lld bytes < %lld", (long long)workspace_bytes,
             (long long)need);
  const int64_t ntx = tg::cdiv(width, kTileW), nty = tg::cdiv(height, kTileH);
  const int64_t ntiles = ntx * nty;
  TG_REQUIRE(ntiles * nsplit < (int64_t)1 << 31, "tg_beamlet_sum: detector too large");
  hipLaunchKernelGGL(beamlet_sum_kernel, dim3((unsigned)(ntiles * nsplit)), dim3(kThreads), 0, (hipStream_t)stream,
                     coef, n, height, width, (int32_t)ntx, (int32_t)ntiles, (int32_t)nsplit, (float2*)workspace);
  TG_LAUNCH_CHECK("beamlet_sum_kernel launch");
  const int64_t npix = height * width;
  hipLaunchKernelGGL(split_sum_kernel, dim3((unsigned)tg::cdiv(npix, 256)), dim3(256), 0, (hipStream_t)stream,
                     (const float2*)workspace, npix, (int32_t)nsplit, (float2*)field);
  TG_LAUNCH_CHECK("split_sum_kernel launch");
  return TG_OK;
}
