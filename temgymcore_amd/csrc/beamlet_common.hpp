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
// forward-mode dual numbers (value, tangent) for derivatives of the f64
// coefficient algebra below
// ---------------------------------------------------------------------------
// the real functions stay visible next to the dual overloads below
using ::atan2;
using ::hypot;
using ::log;
using ::sincos;
using ::sqrt;

struct dual {
  double v, t;
  __device__ constexpr dual(double v_ = 0.0, double t_ = 0.0) : v(v_), t(t_) {}
};
__device__ __forceinline__ dual operator+(dual a, dual b) { return dual(a.v + b.v, a.t + b.t); }
__device__ __forceinline__ dual operator-(dual a, dual b) { return dual(a.v - b.v, a.t - b.t); }
__device__ __forceinline__ dual operator-(dual a) { return dual(-a.v, -a.t); }
__device__ __forceinline__ dual operator*(dual a, dual b) { return dual(a.v * b.v, a.t * b.v + a.v * b.t); }
__device__ __forceinline__ dual operator/(dual a, dual b) {
  const double q = a.v / b.v;
  return dual(q, (a.t - q * b.t) / b.v);
}
__device__ __forceinline__ dual operator+(dual a, double b) { return dual(a.v + b, a.t); }
__device__ __forceinline__ dual operator+(double a, dual b) { return dual(a + b.v, b.t); }
__device__ __forceinline__ dual operator-(dual a, double b) { return dual(a.v - b, a.t); }
__device__ __forceinline__ dual operator-(double a, dual b) { return dual(a - b.v, -b.t); }
__device__ __forceinline__ dual operator*(dual a, double b) { return dual(a.v * b, a.t * b); }
__device__ __forceinline__ dual operator*(double a, dual b) { return dual(a * b.v, a * b.t); }
__device__ __forceinline__ dual operator/(dual a, double b) { return dual(a.v / b, a.t / b); }
__device__ __forceinline__ dual operator/(double a, dual b) {
  const double q = a / b.v;
  return dual(q, -q * b.t / b.v);
}
__device__ __forceinline__ dual sqrt(dual a) {
  const double r = ::sqrt(a.v);
  return dual(r, a.t / (2.0 * r));
}
__device__ __forceinline__ dual log(dual a) { return dual(::log(a.v), a.t / a.v); }
__device__ __forceinline__ dual atan2(dual y, dual x) {
  return dual(::atan2(y.v, x.v), (x.v * y.t - y.v * x.t) / (x.v * x.v + y.v * y.v));
}
__device__ __forceinline__ dual hypot(dual a, dual b) {
  const double h = ::hypot(a.v, b.v);
  return dual(h, (a.v * a.t + b.v * b.t) / h);
}
__device__ __forceinline__ void sincos(dual a, dual* s, dual* c) {
  double sv, cv;
  ::sincos(a.v, &sv, &cv);
  *s = dual(sv, cv * a.t);
  *c = dual(cv, -sv * a.t);
}
__device__ __forceinline__ double val(double a) { return a; }
__device__ __forceinline__ double val(dual a) { return a.v; }

// ---------------------------------------------------------------------------
// complex 2x2 algebra over a real scalar type S (double, or dual)
// ---------------------------------------------------------------------------
template <class S>
struct cx {
  S re, im;
};
using cd = cx<double>;
template <class S>
__device__ __forceinline__ cx<S> mk(S r) {
  return cx<S>{r, S(0.0)};
}
template <class S>
__device__ __forceinline__ cx<S> mk(S r, S i) {
  return cx<S>{r, i};
}
template <class S>
__device__ __forceinline__ cx<S> add(cx<S> a, cx<S> b) {
  return cx<S>{a.re + b.re, a.im + b.im};
}
template <class S>
__device__ __forceinline__ cx<S> sub(cx<S> a, cx<S> b) {
  return cx<S>{a.re - b.re, a.im - b.im};
}
template <class S>
__device__ __forceinline__ cx<S> mul(cx<S> a, cx<S> b) {
  return cx<S>{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class S, class F>
__device__ __forceinline__ cx<S> scl(cx<S> a, F s) {
  return cx<S>{a.re * s, a.im * s};
}
template <class S>
__device__ __forceinline__ cx<S> cdiv(cx<S> a, cx<S> b) {
  const S den = b.re * b.re + b.im * b.im;
  return cx<S>{(a.re * b.re + a.im * b.im) / den, (a.im * b.re - a.re * b.im) / den};
}
template <class S>
__device__ __forceinline__ cx<S> times_i(cx<S> a) {
  return cx<S>{-a.im, a.re};
}
template <class S>
__device__ __forceinline__ cx<S> clog(cx<S> a) {
  return cx<S>{0.5 * log(a.re * a.re + a.im * a.im), atan2(a.im, a.re)};
}
template <class S>
__device__ __forceinline__ bool finite(cx<S> a) {
  return isfinite(val(a.re)) && isfinite(val(a.im));
}

template <class S>
struct cmat {
  cx<S> m00, m01, m10, m11;
};
using cm = cmat<double>;
template <class S>
__device__ __forceinline__ cmat<S> cmul(const cmat<S>& a, const cmat<S>& b) {
  return cmat<S>{add(mul(a.m00, b.m00), mul(a.m01, b.m10)), add(mul(a.m00, b.m01), mul(a.m01, b.m11)),
                 add(mul(a.m10, b.m00), mul(a.m11, b.m10)), add(mul(a.m10, b.m01), mul(a.m11, b.m11))};
}
template <class S>
__device__ __forceinline__ cmat<S> cadd(const cmat<S>& a, const cmat<S>& b) {
  return cmat<S>{add(a.m00, b.m00), add(a.m01, b.m01), add(a.m10, b.m10), add(a.m11, b.m11)};
}
template <class S>
__device__ __forceinline__ cx<S> cdet(const cmat<S>& a) {
  return sub(mul(a.m00, a.m11), mul(a.m01, a.m10));
}
// Inverse; `ok` false when singular or non-finite (the reference's
// nan_to_num(solve(M, I)) then yields the zero matrix).
template <class S>
__device__ __forceinline__ cmat<S> cinv(const cmat<S>& a, bool& ok) {
  const cx<S> d = cdet(a);
  ok = (val(d.re) != 0.0 || val(d.im) != 0.0);
  const cmat<S> r{cdiv(a.m11, d), cdiv(scl(a.m01, -1.0), d), cdiv(scl(a.m10, -1.0), d), cdiv(a.m00, d)};
  ok = ok && finite(r.m00) && finite(r.m01) && finite(r.m10) && finite(r.m11);
  return r;
}
template <class S>
__device__ __forceinline__ cmat<S> real_cm(S a, S b, S c, S d) {
  return cmat<S>{mk(a), mk(b), mk(c), mk(d)};
}
template <class S>
__device__ __forceinline__ cmat<S> zero_cm() {
  return real_cm(S(0.0), S(0.0), S(0.0), S(0.0));
}
template <class S>
__device__ __forceinline__ void cmv(const cmat<S>& a, const cx<S> (&v)[2], cx<S> (&o)[2]) {
  o[0] = add(mul(a.m00, v[0]), mul(a.m01, v[1]));
  o[1] = add(mul(a.m10, v[0]), mul(a.m11, v[1]));
}
template <class S>
__device__ __forceinline__ cx<S> dot(const cx<S> (&u)[2], const cx<S> (&v)[2]) {
  return add(mul(u[0], v[0]), mul(u[1], v[1]));
}
template <class S>
__device__ __forceinline__ cmat<S> transpose(const cmat<S>& a) {
  return cmat<S>{a.m00, a.m10, a.m01, a.m11};
}
template <class S>
__device__ __forceinline__ cx<S> sqrt_principal(cx<S> a) {
  const S r = sqrt(hypot(a.re, a.im));
  const S t = 0.5 * atan2(a.im, a.re);
  S s, c;
  sincos(t, &s, &c);
  return cx<S>{r * c, r * s};
}

// ---------------------------------------------------------------------------
// one beamlet -> Phi(px, py)
// ---------------------------------------------------------------------------
// Phi's six complex coefficients (a, b, c, d, e, f) of
//   Phi(px, py) = a px^2 + b px py + c py^2 + d px + e py + f,
// so that the beamlet's field at pixel (py, px) is exp(Phi).
// p: the first 13 columns of the parameter row; T: its 5x5 ray-transfer
// matrix, row-major (propagated mode only); m: pixel->metre map,
// y_m = m0 py + m1 px + m2, x_m = m3 py + m4 px + m5; amp: the complex
// amplitude folded into f.  S = double, or dual for derivatives.
template <class S>
__device__ inline void phi_coefficients(const S (&p)[13], const S* T, int32_t mode, const double (&m)[6],
                                        cx<S> amp, cx<S> (&out)[6]) {
  const S x = p[0], y = p[1], dx = p[2], dy = p[3], pl = p[4];
  const S wx = p[7], wy = p[8], rx = p[9], ry = p[10], lam = p[11], th = p[12];
  const S k = 2.0 * kPi / lam;
  // 1/q per principal axis (gaussian.py:138-154)
  const S imx = lam / (kPi * wx * wx), imy = lam / (kPi * wy * wy);
  const cx<S> qx = mk(isinf(val(rx)) ? S(0.0) : -1.0 / rx, imx);
  const cx<S> qy = mk(isinf(val(ry)) ? S(0.0) : -1.0 / ry, imy);
  S sn, cs;
  sincos(th, &sn, &cs);
  // R(theta) diag(qx, qy) R(theta)^T, R = [[c, -s], [s, c]] (gaussian.py:157-177)
  const cmat<S> Q1i{add(scl(qx, cs * cs), scl(qy, sn * sn)), scl(sub(qx, qy), cs * sn), scl(sub(qx, qy), cs * sn),
                    add(scl(qx, sn * sn), scl(qy, cs * cs))};
  cmat<S> Qs;  // symmetric quadratic form in r
  cx<S> L[2], c0, pref;
  S ex = S(0.0), ey = S(0.0);
  const cx<S> r1m[2] = {mk(x), mk(y)};
  const cx<S> t1m[2] = {mk(dx), mk(dy)};
  if (mode == TG_BEAMLET_PROPAGATED) {
    const cmat<S> A = real_cm(T[0], T[1], T[5], T[6]);
    const cmat<S> B = real_cm(T[2], T[3], T[7], T[8]);
    const cmat<S> C = real_cm(T[10], T[11], T[15], T[16]);
    const cmat<S> D = real_cm(T[12], T[13], T[17], T[18]);
    ex = T[4];
    ey = T[9];
    const cx<S> f[2] = {mk(T[14]), mk(T[19])};
    const cmat<S> M1 = cadd(A, cmul(B, Q1i));
    const cmat<S> M2 = cadd(C, cmul(D, Q1i));
    bool ok;
    const cmat<S> Q2 = cmul(cinv(M1, ok), M2);  // solve(A + B Q1i, C + D Q1i)
    bool okb, okq;
    cmat<S> Binv = cinv(B, okb);
    if (!okb) Binv = zero_cm<S>();
    cmat<S> Q1 = cinv(Q1i, okq);
    if (!okq) Q1 = zero_cm<S>();
    cx<S> r2m[2], tmp[2];
    cmv(A, r1m, r2m);
    cmv(B, t1m, tmp);
    r2m[0] = add(r2m[0], tmp[0]);
    r2m[1] = add(r2m[1], tmp[1]);
    pref = cdiv(amp, sqrt_principal(cdet(M1)));
    const cmat<S> ABinv = cmul(A, Binv);
    bool okn;
    const cmat<S> Mm = cmul(Q1, cinv(cmul(B, cadd(cmul(A, Q1), B)), okn));
    // linear: -2 Binv^T r1m + 2 Mm^T r2m + 2 f ; constant: r1m^T ABinv r1m - r2m^T Mm r2m
    cx<S> a1[2], a2[2], a3[2], a4[2];
    cmv(transpose(Binv), r1m, a1);
    cmv(transpose(Mm), r2m, a2);
    L[0] = add(add(scl(a1[0], -2.0), scl(a2[0], 2.0)), scl(f[0], 2.0));
    L[1] = add(add(scl(a1[1], -2.0), scl(a2[1], 2.0)), scl(f[1], 2.0));
    cmv(ABinv, r1m, a3);
    cmv(Mm, r2m, a4);
    c0 = sub(dot(r1m, a3), dot(r2m, a4));
    Qs = cmat<S>{Q2.m00, scl(add(Q2.m01, Q2.m10), 0.5), scl(add(Q2.m01, Q2.m10), 0.5), Q2.m11};
  } else {
    // (r - m)^T Q (r - m) + 2 (r - m).t  (gaussian.py:402-407)
    Qs = cmat<S>{Q1i.m00, scl(add(Q1i.m01, Q1i.m10), 0.5), scl(add(Q1i.m01, Q1i.m10), 0.5), Q1i.m11};
    cx<S> qm[2];
    cmv(Qs, r1m, qm);
    L[0] = add(scl(qm[0], -2.0), scl(t1m[0], 2.0));
    L[1] = add(scl(qm[1], -2.0), scl(t1m[1], 2.0));
    c0 = sub(dot(r1m, qm), scl(dot(t1m, r1m), 2.0));
    pref = amp;
  }
  // r = G p + r0, p = (px, py): x_m = m3 py + m4 px + m5, y_m = m0 py + m1 px + m2
  const cmat<S> G = real_cm(S(m[4]), S(m[3]), S(m[1]), S(m[0]));
  const cx<S> r0[2] = {mk(m[5] - ex), mk(m[2] - ey)};
  // Qp = G^T Qs G, Lp = 2 G^T Qs r0 + G^T L, cp = r0^T Qs r0 + L.r0 + c0
  const cmat<S> Qp = cmul(transpose(G), cmul(Qs, G));
  cx<S> qr0[2], t2[2], t3[2];
  cmv(Qs, r0, qr0);
  cmv(transpose(G), qr0, t2);
  cmv(transpose(G), L, t3);
  const cx<S> Lp0 = add(scl(t2[0], 2.0), t3[0]);
  const cx<S> Lp1 = add(scl(t2[1], 2.0), t3[1]);
  const cx<S> cp = add(add(dot(r0, qr0), dot(L, r0)), c0);
  const S h = 0.5 * k;
  out[0] = times_i(scl(Qp.m00, h));
  out[1] = times_i(scl(add(Qp.m01, Qp.m10), h));
  out[2] = times_i(scl(Qp.m11, h));
  out[3] = times_i(scl(Lp0, h));
  out[4] = times_i(scl(Lp1, h));
  const cx<S> psi = mk(S(0.0), k * pl);  // k * pathlength of the input ray (gaussian.py:254-255)
  out[5] = add(add(times_i(scl(cp, h)), psi), clog(pref));
}

// The coefficient record of one beamlet: Phi plus the f32 data of its
// recurrence (K = exp(2a)) and of its cutoff ellipse.  p: parameter row
// (TG_BEAMLET_NPARAM f64); cut: culling level (Re Phi >= peak - cut).
__device__ inline tg_beamlet_coef make_coef(const double* __restrict__ p, const double* __restrict__ T, int32_t mode,
                                            const double (&m)[6], cd amp, float cut) {
  double pp[13];
#pragma unroll
  for (int j = 0; j < 13; ++j) pp[j] = p[j];
  cd c6[6];
  phi_coefficients<double>(pp, T, mode, m, amp, c6);
  const cd ca = c6[0], cb = c6[1], cc = c6[2], cdd = c6[3], ce = c6[4];
  tg_beamlet_coef o;
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    o.phi[2 * j] = c6[j].re;
    o.phi[2 * j + 1] = c6[j].im;
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

// Pixels [xs, xe] of row `row` within the run [x0, x0 + RUN - 1] (and < W)
// that lie inside the beamlet's cutoff ellipse; false when there are none.
// Pixel indices are 32-bit (images are far below 2^31 pixels per side).
template <int RUN = kRun>
__device__ __forceinline__ bool row_extent(const tg_beamlet_coef& c, int32_t row, int32_t x0, int32_t W, int32_t& xs,
                                           int32_t& xe) {
  xs = x0;
  xe = min(x0 + RUN - 1, W - 1);
  if (c.gamma > 0.0f) {
    const float v = (float)row - c.cy;
    const float t2 = fmaf(-c.gamma * v, v, c.beta);
    if (t2 < 0.0f) return false;
    const float sq = __builtin_amdgcn_sqrtf(t2);
    const float mid = fmaf(-c.alpha, v, c.cx);
    // clamp in float first so the conversions stay in int32 range
    const float lo = fminf(fmaxf(ceilf(mid - sq), (float)x0), (float)(x0 + RUN));
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

// Tile shapes of the recurrence tile sum (tg_tile_shape): LPR lanes per row
// of the tile, each walking RUN pixels, kThreads / LPR rows.
struct TileDims {
  int32_t w, h;
};
__host__ __device__ inline TileDims tile_dims(int32_t shape) {
  switch (shape) {
    case TG_TILE_128x32_RUN16: return {128, 32};
    case TG_TILE_32x128_RUN16: return {32, 128};
    case TG_TILE_64x32_RUN8: return {64, 32};
    default: return {kTileW, kTileH};
  }
}

// Adds the beamlets coef[b_begin, b_end) to the tile whose top-left pixel is
// (ty0, tx0): lane t owns row ty0 + t / LPR and the RUN pixels from
// x0 = tx0 + (t % LPR) RUN, accumulated into acc.  Beamlets whose cutoff box
// misses the tile are skipped after a deterministic ballot compaction; every
// thread of the block must call this.
template <int LPR, int RUN>
__device__ __forceinline__ void tile_sum(const tg_beamlet_coef* __restrict__ coef, int64_t b_begin, int64_t b_end,
                                         int32_t tx0, int32_t ty0, int32_t H, int32_t W, TileLists& sh,
                                         float2 (&acc)[RUN]) {
  constexpr int TW = LPR * RUN, TH = kThreads / LPR;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int32_t row = ty0 + t / LPR;
  const int32_t x0 = tx0 + (t % LPR) * RUN;
  const float ftx0 = (float)tx0, fty0 = (float)ty0, ftx1 = (float)(tx0 + TW - 1), fty1 = (float)(ty0 + TH - 1);
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
      if (!row_extent<RUN>(c, row, x0, W, xs, xe)) continue;
      float2 F, R;
      run_start(c, xs, row, F, R);
      const float2 K = make_float2(c.k_re, c.k_im);
#pragma unroll
      for (int j = 0; j < RUN; ++j) {
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

// One tile of a recurrence tile sum, stored into the rows of `dst_rows`
// (row r of the image at dst_rows + r * W): the body of the tile-sum kernels.
template <int LPR, int RUN>
__device__ __forceinline__ void tile_sum_store(const tg_beamlet_coef* __restrict__ coef, int64_t b_begin,
                                               int64_t b_end, int32_t tx0, int32_t ty0, int32_t H, int32_t W,
                                               TileLists& sh, float2* __restrict__ dst_rows) {
  const int t = threadIdx.x;
  const int32_t row = ty0 + t / LPR;
  const int32_t x0 = tx0 + (t % LPR) * RUN;
  float2 acc[RUN];
#pragma unroll
  for (int j = 0; j < RUN; ++j) acc[j] = make_float2(0.f, 0.f);
  tile_sum<LPR, RUN>(coef, b_begin, b_end, tx0, ty0, H, W, sh, acc);
  if (row < H) {
    float2* dst = dst_rows + (int64_t)row * W;
#pragma unroll
    for (int j = 0; j < RUN; ++j)
      if (x0 + j < W) dst[x0 + j] = acc[j];
  }
}

// Centre pixel of a beamlet's moments: its envelope centre rounded and
// clamped into the image (NaN -> 0).  The moment and gradient kernels must
// agree on it.
__device__ __forceinline__ void centre_pixel(const tg_beamlet_coef& c, int32_t H, int32_t W, float& pcx, float& pcy) {
  pcx = fminf(fmaxf(rintf(c.cx), 0.0f), (float)(W - 1));
  pcy = fminf(fmaxf(rintf(c.cy), 0.0f), (float)(H - 1));
}

constexpr int kMoments = 12;  // 6 complex moments

// One wavefront's partial moments of beamlet c (wave-uniform record) over
// the pixels of its cutoff ellipse in an H x W image of adjoint weights:
//   M = sum over pixels of W E {1, u, v, u^2, uv, v^2}   (re, im interleaved)
// with E = exp(Phi_c), (u, v) = pixel offsets from centre_pixel(c).  Lane
// `lane` of 64 takes every 64th 16-pixel run; f32 within a run, f64 across
// runs.  `wbase` needs kRun readable elements past the image (tail padding).
__device__ inline void beamlet_moments(const tg_beamlet_coef& c, int32_t H, int32_t W,
                                       const float2* __restrict__ wbase, int lane, double (&M)[kMoments]) {
  // pixel box of the cutoff ellipse, clipped to the image (NaN -> whole image)
  const float fc0 = fmaxf(ceilf(c.cx - c.box_w), 0.0f), fc1 = fminf(floorf(c.cx + c.box_w), (float)(W - 1));
  const float fr0 = fmaxf(ceilf(c.cy - c.box_h), 0.0f), fr1 = fminf(floorf(c.cy + c.box_h), (float)(H - 1));
  float pcx, pcy;
  centre_pixel(c, H, W, pcx, pcy);
#pragma unroll
  for (int k = 0; k < kMoments; ++k) M[k] = 0.0;
  if (fc0 <= fc1 && fr0 <= fr1) {
    const int32_t c0 = (int32_t)fc0, r0 = (int32_t)fr0;
    const int32_t nrows = (int32_t)fr1 - r0 + 1;
    const int32_t rpr = ((int32_t)fc1 - c0 + kRun) / kRun;  // 16-pixel runs per row
    // lane -> (row offset, run) of task `lane`, then advance by 64 tasks per
    // step without dividing: 64 = q rpr + rem
    int32_t rr = lane / rpr, run = lane - (lane / rpr) * rpr;
    const int32_t q = 64 / rpr, rem = 64 - q * rpr;
    const float2 K = make_float2(c.k_re, c.k_im);
    while (rr < nrows) {
      const int32_t row = r0 + rr;
      const int32_t x0 = c0 + run * kRun;
      int32_t xs, xe;
      if (row_extent(c, row, x0, W, xs, xe)) {
        // the pixels xs .. xe (at most kRun) of this run, as xs + k: their 16
        // weights are loaded at once, before the f64 run-start math so that
        // their latency overlaps it; reads past xe land in the row, the next
        // row or the buffer's kRun-element tail padding (xs <= W - 1) and are
        // masked out below
        const float2* wrun = wbase + (int64_t)row * W + xs;
        float2 w[kRun];
#pragma unroll
        for (int k = 0; k < kRun; ++k) w[k] = wrun[k];
        // keep all 16 loads ahead of the math: one memory round trip per run
        __builtin_amdgcn_sched_barrier(0);
        const int32_t cnt = xe - xs;  // last active k
        float2 F, R;
        run_start(c, xs, row, F, R);
        float2 s0 = make_float2(0.f, 0.f), s1 = s0, s2 = s0;
        // branch-free: steps past cnt are computed and discarded (select, so
        // an overflowing tail cannot leak into the sums)
#pragma unroll
        for (int k = 0; k < kRun; ++k) {
          float2 P = cmulf(w[k], F);
          if (k > 0) {
            const bool on = k <= cnt;
            P.x = on ? P.x : 0.0f;
            P.y = on ? P.y : 0.0f;
          }
          s0.x += P.x;
          s0.y += P.y;
          s1.x = fmaf((float)k, P.x, s1.x);
          s1.y = fmaf((float)k, P.y, s1.y);
          s2.x = fmaf((float)(k * k), P.x, s2.x);
          s2.y = fmaf((float)(k * k), P.y, s2.y);
          if (k + 1 < kRun) {
            F = cmulf(F, R);
            R = cmulf(R, K);
          }
        }
        // run sums over k = x - xs -> moments in u = x - pcx = ub + k, v = row - pcy
        const double ub = (double)xs - (double)pcx, v = (double)row - (double)pcy, vv = v * v;
        const double a0r = s0.x, a0i = s0.y, a1r = s1.x, a1i = s1.y, a2r = s2.x, a2i = s2.y;
        const double u1r = fma(ub, a0r, a1r), u1i = fma(ub, a0i, a1i);  // sum u P
        M[0] += a0r;
        M[1] += a0i;
        M[2] += u1r;
        M[3] += u1i;
        M[4] = fma(v, a0r, M[4]);
        M[5] = fma(v, a0i, M[5]);
        M[6] += fma(ub, u1r + a1r, a2r);  // sum u^2 P = ub (ub a0 + 2 a1) + a2
        M[7] += fma(ub, u1i + a1i, a2i);
        M[8] = fma(v, u1r, M[8]);
        M[9] = fma(v, u1i, M[9]);
        M[10] = fma(vv, a0r, M[10]);
        M[11] = fma(vv, a0i, M[11]);
      }
      run += rem;
      rr += q;
      if (run >= rpr) {
        run -= rpr;
        ++rr;
      }
    }
  }
}

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

}  // namespace tgb
