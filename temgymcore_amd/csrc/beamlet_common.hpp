// Device code shared by the Gaussian-beamlet kernels (beamlet.hip) and the
// batched Gaussian fits (gauss_fit.hip): f64 complex 2x2 algebra, the folding
// of one beamlet into the complex quadratic Phi(px, py) of tg_beamlet_coef,
// and the f32 recurrence that sums exp(Phi) over a 64x64 detector tile.
#pragma once

#include "tg_common.hpp"

namespace tgb {

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
// one beamlet -> Phi(px, py)
// ---------------------------------------------------------------------------
// p: parameter row (TG_BEAMLET_NPARAM f64); T: its 5x5 ray-transfer matrix
// (propagated mode only); m: pixel->metre map, y_m = m0 py + m1 px + m2,
// x_m = m3 py + m4 px + m5; amp: the complex amplitude folded into Phi's
// constant term; cut: culling level (Re Phi >= peak - cut).
__device__ inline tg_beamlet_coef make_coef(const double* __restrict__ p, const double* __restrict__ T, int32_t mode,
                                            const double (&m)[6], cd amp, float cut) {
  const double x = p[0], y = p[1], dx = p[2], dy = p[3], pl = p[4];
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
  const double G00 = m[4], G01 = m[3], G10 = m[1], G11 = m[0];
  const cd r0[2] = {mk(m[5] - ex), mk(m[2] - ey)};
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
  return o;
}

// ---------------------------------------------------------------------------
// tile sums
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
  const double err = fma(x, kInv2PiLo, fma(x, kInv2PiHi, -hi));
  const double fr = hi - rint(hi);
  return (float)(fr + err);
}

// exp(re + i im) for f64 (re, im), im reduced in revolutions
__device__ __forceinline__ float2 cexp_f(double re, double im) {
  const float mag = __expf((float)re);
  const float rv = revolutions(im);
  return make_float2(mag * __builtin_amdgcn_cosf(rv), mag * __builtin_amdgcn_sinf(rv));
}

__device__ __forceinline__ float2 cmulf(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}

// Pixels [xs, xe] of row `row` within the run [x0, x0 + kRun - 1] (and < W)
// that lie inside the beamlet's cutoff ellipse; false when there are none.
// Pixel indices are 32-bit (images are far below 2^31 pixels per side).
__device__ __forceinline__ bool row_extent(const tg_beamlet_coef& c, int32_t row, int32_t x0, int32_t W, int32_t& xs,
                                           int32_t& xe) {
  xs = x0;
  xe = min(x0 + kRun - 1, W - 1);
  if (c.gamma > 0.0f) {
    const float v = (float)row - c.cy;
    const float t2 = fmaf(-c.gamma * v, v, c.beta);
    if (t2 < 0.0f) return false;
    const float sq = __builtin_amdgcn_sqrtf(t2);
    const float mid = fmaf(-c.alpha, v, c.cx);
    // clamp in float first so the conversions stay in int32 range
    const float lo = fminf(fmaxf(ceilf(mid - sq), (float)x0), (float)(x0 + kRun));
    const float hi = fmaxf(fminf(floorf(mid + sq), (float)xe), (float)(x0 - 1));
    xs = (int32_t)lo;
    xe = (int32_t)hi;
  }
  return xs <= xe;
}

// exp(Phi) at (xs, row) and the first x-step ratio exp(Phi(xs + 1) - Phi(xs)),
// with Phi evaluated in f64.
__device__ __forceinline__ void run_start(const tg_beamlet_coef& c, int32_t xs, int32_t row, float2& F, float2& R) {
  const double px = (double)xs, py = (double)row, px2 = 2.0 * px + 1.0;
  const double* a = c.phi;
  // Phi = (a px + b py + d) px + (c py + e) py + f, and Phi(px + 1) - Phi(px)
  const double lre = fma(a[2], py, a[6]), lim = fma(a[3], py, a[7]);
  const double pre = fma(fma(a[0], px, lre), px, fma(fma(a[4], py, a[8]), py, a[10]));
  const double pim = fma(fma(a[1], px, lim), px, fma(fma(a[5], py, a[9]), py, a[11]));
  F = cexp_f(pre, pim);
  R = cexp_f(fma(a[0], px2, lre), fma(a[1], px2, lim));
}

// Shared memory of tile_sum.
struct TileLists {
  int32_t list[kThreads];
  int32_t wave_cnt[kWaves];
};

// Adds the beamlets coef[b_begin, b_end) to the tile whose top-left pixel is
// (ty0, tx0): lane t owns row ty0 + t / kLanesPerRow and the kRun pixels from
// x0 = tx0 + (t % kLanesPerRow) kRun, accumulated into acc.  Beamlets whose
// cutoff box misses the tile are skipped after a deterministic ballot
// compaction; every thread of the block must call this.
__device__ __forceinline__ void tile_sum(const tg_beamlet_coef* __restrict__ coef, int64_t b_begin, int64_t b_end,
                                         int32_t tx0, int32_t ty0, int32_t H, int32_t W, TileLists& sh,
                                         float2 (&acc)[kRun]) {
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int32_t row = ty0 + t / kLanesPerRow;
  const int32_t x0 = tx0 + (t % kLanesPerRow) * kRun;
  const float ftx0 = (float)tx0, fty0 = (float)ty0, ftx1 = (float)(tx0 + kTileW - 1), fty1 = (float)(ty0 + kTileH - 1);
  for (int64_t b0 = b_begin; b0 < b_end; b0 += kThreads) {
    const int64_t b = b0 + t;
    bool hit = false;
    if (b < b_end) {
      const tg_beamlet_coef& c = coef[b];
      hit = c.cx + c.box_w >= ftx0 && c.cx - c.box_w <= ftx1 && c.cy + c.box_h >= fty0 && c.cy - c.box_h <= fty1;
    }
    const uint64_t mask = __ballot(hit);
    if (lane == 0) sh.wave_cnt[wave] = __popcll(mask);
    __syncthreads();
    int off = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      off += (w < wave) ? sh.wave_cnt[w] : 0;
      total += sh.wave_cnt[w];
    }
    if (hit) sh.list[off + __popcll(mask & ((1ull << lane) - 1ull))] = (int32_t)(b - b0);
    __syncthreads();
    for (int q = 0; q < total; ++q) {
      const int64_t bi = b0 + __builtin_amdgcn_readfirstlane(sh.list[q]);
      const tg_beamlet_coef& c = coef[bi];
      if (row >= H) continue;
      int32_t xs, xe;
      if (!row_extent(c, row, x0, W, xs, xe)) continue;
      float2 F, R;
      run_start(c, xs, row, F, R);
      const float2 K = make_float2(c.k_re, c.k_im);
#pragma unroll
      for (int j = 0; j < kRun; ++j) {
        const int32_t pxj = x0 + j;
        if (pxj >= xs && pxj <= xe) {
          acc[j].x += F.x;
          acc[j].y += F.y;
          F = cmulf(F, R);
          R = cmulf(R, K);
        }
      }
    }
    __syncthreads();  // the list is rewritten by the next chunk
  }
}

}  // namespace tgb
