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
#include "tg_common.hpp"

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kInv2PiHi = 0.15915494309189535;     // f64(1 / (2 pi))
constexpr double kInv2PiLo = -9.4243388001133859e-18;  // 1/(2 pi) - kInv2PiHi
constexpr float kUnbounded = 3.0e38f;

// ---------------------------------------------------------------------------
// f64 complex 2x2 algebra
// ---------------------------------------------------------------------------
struct cd {
  double re, im;
};
__device__ __forceinline__ cd mk(double r, double i = 0.0) { return cd{r, i}; }
__device__ __forceinline__ cd add(cd a, cd b) { return cd{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cd sub(cd a, cd b) { return cd{a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cd mul(cd a, cd b) { return cd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cd scl(cd a, double s) { return cd{a.re * s, a.im * s}; }
__device__ __forceinline__ cd cdiv(cd a, cd b) {
  const double den = b.re * b.re + b.im * b.im;
  return cd{(a.re * b.re + a.im * b.im) / den, (a.im * b.re - a.re * b.im) / den};
}
__device__ __forceinline__ cd times_i(cd a) { return cd{-a.im, a.re}; }
__device__ __forceinline__ cd clog(cd a) { return cd{0.5 * log(a.re * a.re + a.im * a.im), atan2(a.im, a.re)}; }
__device__ __forceinline__ bool finite(cd a) { return isfinite(a.re) && isfinite(a.im); }

struct cm {
  cd m00, m01, m10, m11;
};
__device__ __forceinline__ cm cmul(const cm& a, const cm& b) {
  return cm{add(mul(a.m00, b.m00), mul(a.m01, b.m10)), add(mul(a.m00, b.m01), mul(a.m01, b.m11)),
            add(mul(a.m10, b.m00), mul(a.m11, b.m10)), add(mul(a.m10, b.m01), mul(a.m11, b.m11))};
}
__device__ __forceinline__ cm cadd(const cm& a, const cm& b) {
  return cm{add(a.m00, b.m00), add(a.m01, b.m01), add(a.m10, b.m10), add(a.m11, b.m11)};
}
__device__ __forceinline__ cd cdet(const cm& a) { return sub(mul(a.m00, a.m11), mul(a.m01, a.m10)); }
// Inverse; `ok` false when singular or non-finite (the reference's
// nan_to_num(solve(M, I)) then yields the zero matrix).
__device__ __forceinline__ cm cinv(const cm& a, bool& ok) {
  const cd d = cdet(a);
  ok = (d.re != 0.0 || d.im != 0.0);
  const cm r{cdiv(a.m11, d), cdiv(scl(a.m01, -1.0), d), cdiv(scl(a.m10, -1.0), d), cdiv(a.m00, d)};
  ok = ok && finite(r.m00) && finite(r.m01) && finite(r.m10) && finite(r.m11);
  return r;
}
__device__ __forceinline__ cm real_cm(double a, double b, double c, double d) {
  return cm{mk(a), mk(b), mk(c), mk(d)};
}
__device__ __forceinline__ cm zero_cm() { return real_cm(0, 0, 0, 0); }
__device__ __forceinline__ void cmv(const cm& a, const cd (&v)[2], cd (&o)[2]) {
  o[0] = add(mul(a.m00, v[0]), mul(a.m01, v[1]));
  o[1] = add(mul(a.m10, v[0]), mul(a.m11, v[1]));
}
__device__ __forceinline__ cd dot(const cd (&u)[2], const cd (&v)[2]) { return add(mul(u[0], v[0]), mul(u[1], v[1])); }
__device__ __forceinline__ cm transpose(const cm& a) { return cm{a.m00, a.m10, a.m01, a.m11}; }
__device__ __forceinline__ cd sqrt_principal(cd a) {
  const double r = sqrt(hypot(a.re, a.im));
  const double t = 0.5 * atan2(a.im, a.re);
  double s, c;
  sincos(t, &s, &c);
  return cd{r * c, r * s};
}

// ---------------------------------------------------------------------------
// stage 1: coefficients
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void beamlet_coef_kernel(const double* __restrict__ params, int64_t n,
                                                           const double* __restrict__ abcd, int64_t abcd_stride,
                                                           int32_t mode, double m0, double m1, double m2, double m3,
                                                           double m4, double m5, float cut,
                                                           tg_beamlet_coef* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* p = params + i * TG_BEAMLET_NPARAM;
  const double x = p[0], y = p[1], dx = p[2], dy = p[3], pl = p[4];
  const cd amp = mk(p[5], p[6]);
  const double wx = p[7], wy = p[8], rx = p[9], ry = p[10], lam = p[11], th = p[12];
  const double k = 2.0 * kPi / lam;
  // 1/q per principal axis (gaussian.py:138-154)
  const double imx = lam / (kPi * wx * wx), imy = lam / (kPi * wy * wy);
  const cd qx = mk(isinf(rx) ? 0.0 : -1.0 / rx, imx);
  const cd qy = mk(isinf(ry) ? 0.0 : -1.0 / ry, imy);
  double sn, cs;
  sincos(th, &sn, &cs);
  // R(theta) diag(qx, qy) R(theta)^T, R = [[c, -s], [s, c]] (gaussian.py:157-177)
  const cm Q1i{add(scl(qx, cs * cs), scl(qy, sn * sn)), scl(sub(qx, qy), cs * sn), scl(sub(qx, qy), cs * sn),
               add(scl(qx, sn * sn), scl(qy, cs * cs))};
  cm Qs;            // symmetric quadratic form in r
  cd L[2], c0, pref;
  double ex = 0.0, ey = 0.0;
  const cd r1m[2] = {mk(x), mk(y)};
  const cd t1m[2] = {mk(dx), mk(dy)};
  if (mode == TG_BEAMLET_PROPAGATED) {
    const double* T = abcd + i * abcd_stride;
    const cm A = real_cm(T[0], T[1], T[5], T[6]);
    const cm B = real_cm(T[2], T[3], T[7], T[8]);
    const cm C = real_cm(T[10], T[11], T[15], T[16]);
    const cm D = real_cm(T[12], T[13], T[17], T[18]);
    ex = T[4];
    ey = T[9];
    const cd f[2] = {mk(T[14]), mk(T[19])};
    const cm M1 = cadd(A, cmul(B, Q1i));
    const cm M2 = cadd(C, cmul(D, Q1i));
    bool ok;
    const cm Q2 = cmul(cinv(M1, ok), M2);  // solve(A + B Q1i, C + D Q1i)
    bool okb, okq;
    cm Binv = cinv(B, okb);
    if (!okb) Binv = zero_cm();
    cm Q1 = cinv(Q1i, okq);
    if (!okq) Q1 = zero_cm();
    cd r2m[2], tmp[2];
    cmv(A, r1m, r2m);
    cmv(B, t1m, tmp);
    r2m[0] = add(r2m[0], tmp[0]);
    r2m[1] = add(r2m[1], tmp[1]);
    pref = cdiv(amp, sqrt_principal(cdet(M1)));
    const cm ABinv = cmul(A, Binv);
    bool okn;
    const cm Mm = cmul(Q1, cinv(cmul(B, cadd(cmul(A, Q1), B)), okn));
    // linear: -2 Binv^T r1m + 2 Mm^T r2m + 2 f ; constant: r1m^T ABinv r1m - r2m^T Mm r2m
    cd a1[2], a2[2], a3[2], a4[2];
    cmv(transpose(Binv), r1m, a1);
    cmv(transpose(Mm), r2m, a2);
    L[0] = add(add(scl(a1[0], -2.0), scl(a2[0], 2.0)), scl(f[0], 2.0));
    L[1] = add(add(scl(a1[1], -2.0), scl(a2[1], 2.0)), scl(f[1], 2.0));
    cmv(ABinv, r1m, a3);
    cmv(Mm, r2m, a4);
    c0 = sub(dot(r1m, a3), dot(r2m, a4));
    Qs = cm{Q2.m00, scl(add(Q2.m01, Q2.m10), 0.5), scl(add(Q2.m01, Q2.m10), 0.5), Q2.m11};
  } else {
    // (r - m)^T Q (r - m) + 2 (r - m).t  (gaussian.py:402-407)
    Qs = cm{Q1i.m00, scl(add(Q1i.m01, Q1i.m10), 0.5), scl(add(Q1i.m01, Q1i.m10), 0.5), Q1i.m11};
    cd qm[2];
    cmv(Qs, r1m, qm);
    L[0] = add(scl(qm[0], -2.0), scl(t1m[0], 2.0));
    L[1] = add(scl(qm[1], -2.0), scl(t1m[1], 2.0));
    c0 = sub(dot(r1m, qm), scl(dot(t1m, r1m), 2.0));
    pref = amp;
  }
  // r = G p + r0, p = (px, py): x_m = m3 py + m4 px + m5, y_m = m0 py + m1 px + m2
  const double G00 = m4, G01 = m3, G10 = m1, G11 = m0;
  const cd r0[2] = {mk(m5 - ex), mk(m2 - ey)};
  // Qp = G^T Qs G, Lp = 2 G^T Qs r0 + G^T L, cp = r0^T Qs r0 + L.r0 + c0
  const cm G = real_cm(G00, G01, G10, G11);
  const cm Qp = cmul(transpose(G), cmul(Qs, G));
  cd qr0[2], t2[2], t3[2];
  cmv(Qs, r0, qr0);
  cmv(transpose(G), qr0, t2);
  cmv(transpose(G), L, t3);
  const cd Lp0 = add(scl(t2[0], 2.0), t3[0]);
  const cd Lp1 = add(scl(t2[1], 2.0), t3[1]);
  const cd cp = add(add(dot(r0, qr0), dot(L, r0)), c0);
  const double h = 0.5 * k;
  const cd ca = times_i(scl(Qp.m00, h));
  const cd cb = times_i(scl(add(Qp.m01, Qp.m10), h));
  const cd cc = times_i(scl(Qp.m11, h));
  const cd cdd = times_i(scl(Lp0, h));
  const cd ce = times_i(scl(Lp1, h));
  const cd psi = mk(0.0, k * pl);  // k * pathlength of the input ray (gaussian.py:254-255)
  const cd cf = add(add(times_i(scl(cp, h)), psi), clog(pref));
  tg_beamlet_coef o;
  const cd cs6[6] = {ca, cb, cc, cdd, ce, cf};
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    o.phi[2 * j] = cs6[j].re;
    o.phi[2 * j + 1] = cs6[j].im;
  }
  const double ke = exp(2.0 * ca.re);
  double ks, kc;
  sincos(2.0 * ca.im, &ks, &kc);
  o.k_re = (float)(ke * kc);
  o.k_im = (float)(ke * ks);
  // envelope Re Phi = -p^T S p + g.p + const, S = -Re [[a, b/2], [b/2, c]]
  const double s00 = -ca.re, s01 = -0.5 * cb.re, s11 = -cc.re;
  const double det = s00 * s11 - s01 * s01;
  if (s00 > 0.0 && det > 0.0 && isfinite(det)) {
    const double g0 = cdd.re, g1 = ce.re;
    o.cx = (float)((s11 * g0 - s01 * g1) / (2.0 * det));
    o.cy = (float)((s00 * g1 - s01 * g0) / (2.0 * det));
    // s00 u^2 + 2 s01 u v + s11 v^2 <= cut  <=>  u = -alpha v +- sqrt(beta - gamma v^2)
    o.alpha = (float)(s01 / s00);
    o.beta = (float)((double)cut / s00);
    o.gamma = (float)(det / (s00 * s00));
    o.cut = cut;
    o.box_w = (float)sqrt((double)cut * s11 / det);
    o.box_h = (float)sqrt((double)cut * s00 / det);
  } else {  // no decaying envelope: every pixel
    o.cx = o.cy = 0.0f;
    o.alpha = o.beta = o.gamma = 0.0f;
    o.cut = cut;
    o.box_w = o.box_h = kUnbounded;
  }
  out[i] = o;
}

// ---------------------------------------------------------------------------
// stage 2: tile sums
// ---------------------------------------------------------------------------
constexpr int kRun = 16;                  // pixels per lane (one row segment)
constexpr int kLanesPerRow = 4;
constexpr int kTileW = kRun * kLanesPerRow;  // tile is kTileW x kTileH pixels
constexpr int kTileH = 64;
constexpr int kThreads = kLanesPerRow * kTileH;  // 256
constexpr int kWaves = kThreads / 64;

// (x - rint(x)) of x * (1 / 2 pi) in double-double: the phase in revolutions,
// in [-0.5, 0.5], accurate for |x| up to ~1e12 rad.
__device__ __forceinline__ float revolutions(double x) {
  const double hi = x * kInv2PiHi;
  const double err = fma(x, kInv2PiHi, -hi) + x * kInv2PiLo;
  const double fr = hi - rint(hi);
  return (float)(fr + err);
}

// exp(re + i im) for f64 (re, im), im reduced in revolutions
__device__ __forceinline__ float2 cexp_f(double re, double im) {
  const float m = __expf((float)re);
  const float rv = revolutions(im);
  return make_float2(m * __builtin_amdgcn_cosf(rv), m * __builtin_amdgcn_sinf(rv));
}

__device__ __forceinline__ float2 cmulf(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}

__global__ __launch_bounds__(kThreads) void beamlet_sum_kernel(const tg_beamlet_coef* __restrict__ coef, int64_t n,
                                                               int64_t H, int64_t W, int32_t ntx, int32_t ntiles,
                                                               int32_t nsplit, float2* __restrict__ ws) {
  __shared__ int32_t list[kThreads];
  __shared__ int32_t wave_cnt[kWaves];
  const int tile = blockIdx.x % ntiles, split = blockIdx.x / ntiles;
  const int64_t tx0 = (int64_t)(tile % ntx) * kTileW, ty0 = (int64_t)(tile / ntx) * kTileH;
  const int64_t b_begin = n * split / nsplit, b_end = n * (split + 1) / nsplit;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int64_t row = ty0 + t / kLanesPerRow;
  const int64_t x0 = tx0 + (int64_t)(t % kLanesPerRow) * kRun;
  float2 acc[kRun];
#pragma unroll
  for (int j = 0; j < kRun; ++j) acc[j] = make_float2(0.f, 0.f);
  const float ftx0 = (float)tx0, fty0 = (float)ty0, ftx1 = (float)(tx0 + kTileW - 1), fty1 = (float)(ty0 + kTileH - 1);
  for (int64_t b0 = b_begin; b0 < b_end; b0 += kThreads) {
    // compact the beamlets of this chunk whose cutoff ellipse meets the tile
    const int64_t b = b0 + t;
    bool hit = false;
    if (b < b_end) {
      const tg_beamlet_coef& c = coef[b];
      hit = c.cx + c.box_w >= ftx0 && c.cx - c.box_w <= ftx1 && c.cy + c.box_h >= fty0 && c.cy - c.box_h <= fty1;
    }
    const uint64_t mask = __ballot(hit);
    if (lane == 0) wave_cnt[wave] = __popcll(mask);
    __syncthreads();
    int off = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      off += (w < wave) ? wave_cnt[w] : 0;
      total += wave_cnt[w];
    }
    if (hit) list[off + __popcll(mask & ((1ull << lane) - 1ull))] = (int32_t)(b - b0);
    __syncthreads();
    for (int q = 0; q < total; ++q) {
      const int64_t bi = b0 + __builtin_amdgcn_readfirstlane(list[q]);
      const tg_beamlet_coef& c = coef[bi];
      if (row >= H) continue;
      // this lane's pixels inside the cutoff ellipse: [xs, xe]
      int64_t xs = x0, xe = x0 + kRun - 1;
      if (xe >= W) xe = W - 1;
      if (c.gamma > 0.0f) {
        const float v = (float)row - c.cy;
        const float t2 = fmaf(-c.gamma * v, v, c.beta);
        if (t2 < 0.0f) continue;
        const float sq = sqrtf(t2);
        const float mid = fmaf(-c.alpha, v, c.cx);
        const float lo = ceilf(mid - sq), hi = floorf(mid + sq);
        if (lo > (float)xs) xs = (lo > (float)(x0 + kRun)) ? x0 + kRun : (int64_t)lo;
        if (hi < (float)xe) xe = (hi < (float)(x0 - 1)) ? x0 - 1 : (int64_t)hi;
      }
      if (xs > xe) continue;
      // Phi and the first step ratio at (xs, row), in f64
      const double px = (double)xs, py = (double)row;
      const double* a = c.phi;
      const double pre = ((a[0] * px + a[2] * py + a[6]) * px + (a[4] * py + a[8]) * py) + a[10];
      const double pim = ((a[1] * px + a[3] * py + a[7]) * px + (a[5] * py + a[9]) * py) + a[11];
      const double rre = a[0] * (2.0 * px + 1.0) + a[2] * py + a[6];
      const double rim = a[1] * (2.0 * px + 1.0) + a[3] * py + a[7];
      float2 F = cexp_f(pre, pim);
      float2 R = cexp_f(rre, rim);
      const float2 K = make_float2(c.k_re, c.k_im);
#pragma unroll
      for (int j = 0; j < kRun; ++j) {
        const int64_t pxj = x0 + j;
        if (pxj >= xs && pxj <= xe) {
          acc[j].x += F.x;
          acc[j].y += F.y;
          F = cmulf(F, R);
          R = cmulf(R, K);
        }
      }
    }
    __syncthreads();  // list is rewritten by the next chunk
  }
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
