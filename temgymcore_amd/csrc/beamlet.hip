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
//
// Backward (tg_beamlet_backward): for a real loss L of the field, with the
// cotangent g = dL/dRe(field) + i dL/dIm(field) and W = conj(g),
//   dL/dtheta = Re sum_pixels W du_b/dtheta,  u_b = amp exp(Phi'_b(px, py)),
// Phi' being Phi with unit amplitude.  One wavefront per beamlet reduces the
// six complex moments sum W exp(Phi') {1, u, v, u^2, uv, v^2} over its cutoff
// ellipse (the kernel shared with the batched fits); then, per beamlet in
// f64, the moments give G_k = sum W exp(Phi') m_k for the six monomials
// m_k = px^2, px py, py^2, px, py, 1 of Phi, and
//   dL/dtheta = Re[amp sum_k G_k dc_k/dtheta]
// with dc_k/dtheta from forward-mode dual numbers through the same
// coefficient algebra (one pass per parameter and ray-transfer entry).
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

// ---------------------------------------------------------------------------
// Path selection: kFlagBlocks workgroups each write whether their share of
// the beamlets holds one that is not separable (every entry rewritten per
// launch, no atomics); every tile-sum workgroup reads all of them.
// ---------------------------------------------------------------------------
constexpr int kFlagBlocks = 256;

// exp(Phi) = Ex(px) Ey(py) with bounded factors: no px*py term, an envelope
// that decays along both axes, finite coefficients.
__device__ __forceinline__ bool separable(const tg_beamlet_coef& c) {
  bool fin = true;
#pragma unroll
  for (int j = 0; j < 12; ++j) fin = fin && isfinite(c.phi[j]);
  return fin && c.phi[2] == 0.0 && c.phi[3] == 0.0 && c.phi[0] < 0.0 && c.phi[4] < 0.0;
}

// The factors of a separable beamlet (b = 0):
//   Ex = exp(ar (px - cx)^2 + i (ai px^2 + di px)),
//   Ey = exp(cr (py - cy)^2 + g0 + i (ci py^2 + ei py + fi)),
// with a = ar + i ai etc., cx = -dr / (2 ar), cy = -er / (2 cr) and
// g0 = fr - ar cx^2 - cr cy^2 (the log of the peak amplitude): |Ex| <= 1 and
// |Ey| <= the peak.  One axis is
//   exp(q2 (p - cen)^2 + g + i (q1 p^2 + l1 p + c0)).
struct SepAxis {
  double q2, q1, l1, c0, cen, g;
};
struct SepParams {
  SepAxis x, y;
};
static_assert(sizeof(SepParams) == 96, "SepParams layout");

__device__ __forceinline__ SepParams sep_params(const tg_beamlet_coef& c) {
  SepParams q;
  const double ar = c.phi[0], cr = c.phi[4];
  const double cx = -c.phi[6] / (2.0 * ar), cy = -c.phi[8] / (2.0 * cr);
  q.x = {ar, c.phi[1], c.phi[7], 0.0, cx, 0.0};
  q.y = {cr, c.phi[5], c.phi[9], c.phi[11], cy, c.phi[10] - ar * cx * cx - cr * cy * cy};
  return q;
}

// flags[blockIdx.x]: some beamlet of this workgroup's share is not separable
// (or force_valu); par[i]: the factors of beamlet i when it is separable.
__global__ __launch_bounds__(256) void sep_flag_kernel(const tg_beamlet_coef* __restrict__ coef, int64_t n,
                                                       int32_t force_valu, int32_t* __restrict__ flags,
                                                       SepParams* __restrict__ par) {
  int bad = force_valu;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n && !force_valu; i += (int64_t)gridDim.x * 256) {
    const tg_beamlet_coef& c = coef[i];
    if (separable(c))
      par[i] = sep_params(c);
    else
      bad = 1;
  }
  bad = __syncthreads_or(bad);
  if (threadIdx.x == 0) flags[blockIdx.x] = bad;
}

// True (in every thread of the workgroup) when the launch takes the
// matrix-core path.
__device__ __forceinline__ bool matrix_path(const int32_t* __restrict__ flags) {
  const int bad = threadIdx.x < kFlagBlocks ? flags[threadIdx.x] : 0;
  return !__syncthreads_or(bad);
}

template <int LPR, int RUN>
__global__ __launch_bounds__(kThreads) void beamlet_sum_kernel(const tg_beamlet_coef* __restrict__ coef, int64_t n,
                                                               int64_t H, int64_t W, int32_t ntx, int32_t ntiles,
                                                               int32_t nsplit, const int32_t* __restrict__ flags,
                                                               float2* __restrict__ ws) {
  __shared__ TileLists sh;
  if (matrix_path(flags)) return;  // beamlet_sum_sep_kernel covers this launch
  const int split = blockIdx.x % nsplit, tile = blockIdx.x / nsplit;  // slices of a tile adjacent
  const int32_t tx0 = (tile % ntx) * (LPR * RUN), ty0 = (tile / ntx) * (kThreads / LPR);
  const int64_t b_begin = n * split / nsplit, b_end = n * (split + 1) / nsplit;
  tile_sum_store<LPR, RUN>(coef, b_begin, b_end, tx0, ty0, (int32_t)H, (int32_t)W, sh, ws + (int64_t)split * H * W);
}

// ---------------------------------------------------------------------------
// Separable beamlets on the matrix cores.  A tile is the complex matrix
// product  T[m][n] = sum_k Ey_k(ty0 + m) Ex_k(tx0 + n)  over the beamlets whose
// cutoff box meets it: three real f32 MFMA products per k (sep_mfma).
// Wave w owns the 32x32 block at rows 32 (w / 2), columns
// 32 (w % 2) of the 64x64 tile.  The factors of kSepK beamlets at a time are
// staged in LDS, double-buffered: a wave evaluates chunk c + 1 into one
// buffer while issuing chunk c's products from the other, one barrier per
// chunk.  (On gfx950 the vector instructions and the f32 MFMAs of a SIMD do
// not execute concurrently -- measured: their times add, also across waves
// -- so the evaluation's instruction count is what it costs.)  Each thread
// evaluates one 16-pixel segment of one factor directly, without branches:
//   exp2(q2' (p - cen)^2 + g') (cos, sin)(2 pi theta(p)),
//   theta(p0 + j) = theta0 + j A + j^2 B   (revolutions),
// theta0 being the segment start's phase reduced in double-double.  theta0,
// A and B are reduced modulo one revolution as 32-bit fixed-point fractions
// (exact whatever the size of the linear and quadratic phase terms; an f32
// sum theta0 + jA + j^2 B lost 1e-5 revolutions on strongly curved
// input-plane beamlets), and seed a unit-phasor recurrence (two transcendental
// pairs per segment instead of a sine and cosine per value).  Parameters are
// loaded one chunk ahead of their use.
// ---------------------------------------------------------------------------
typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int kSepK = 32;          // beamlets per LDS chunk (the product's K)
constexpr int kSepSeg = 16;        // pixels per generated segment
constexpr int kSepRow = kTileW + 1;  // LDS row pitch (float2): a bank offset per beamlet row
constexpr int kSepBlock = 2048;    // beamlets per list build
constexpr double kLog2e = 1.4426950408889634;
static_assert(kTileW == 64 && kTileH == 64 && kThreads == 256 && kThreads >= kFlagBlocks &&
                  kSepK * 2 * (kTileW / kSepSeg) == kThreads && kSepSeg == kSepK / 2 &&
                  kSepBlock % kThreads == 0,
              "the matrix-core tile sum assumes 64x64 tiles, four waves, one segment per thread and one "
              "segment value per k pair");

// The axis record (q2, q1), (l1, c0), (cen, g) of beamlet i.
__device__ __forceinline__ void load_axis(const SepParams* __restrict__ par, int64_t i, int axis, double2& v0,
                                          double2& v1, double2& v2) {
  const double2* src = reinterpret_cast<const double2*>(par + i) + 3 * axis;
  v0 = src[0];
  v1 = src[1];
  v2 = src[2];
}

// Evaluation state of one 16-pixel segment starting at pixel p0: the
// log2-scaled envelope, evaluated directly, and the unit phasor
// z_j = exp(2 pi i theta(p0 + j)) stepped by the unit ratios
// r_j = z_{j+1} / z_j, r_{j+1} = r_j k (|z| = 1: nothing to underflow).  The
// seeds come from theta0, A and B reduced modulo one revolution as 32-bit
// fixed-point fractions (exact wrapping integer arithmetic), so the phase
// error is that of 16 f32 rotations (CPU emulation of this scheme: worst
// 1.4e-5 rad, |z| within 6e-6 of 1).  An invalid segment evaluates to zero.
struct SegGen {
  float2 z, r, k;
  float t0, q2l, gl;
};

// r (revolutions, |r| < 2^52) modulo one, in 2^-32 units
__device__ __forceinline__ uint32_t fixed_revolutions(double r) {
  return (uint32_t)(int64_t)rint((r - rint(r)) * 4294967296.0);
}

// exp(2 pi i u 2^-32)
__device__ __forceinline__ float2 phasor(uint32_t u) {
  const float t = (float)(int32_t)u * 0x1p-32f;  // [-0.5, 0.5) revolutions
  return make_float2(__builtin_amdgcn_cosf(t), __builtin_amdgcn_sinf(t));
}

__device__ __forceinline__ SegGen seg_setup(double2 v0, double2 v1, double2 v2, double p0, bool valid) {
  const double q2 = v0.x, q1 = v0.y, l1 = v1.x, c0 = v1.y, cen = v2.x, g = v2.y;
  // theta(p0 + j) = theta0 + j A + j^2 B  (revolutions)
  const double x = fma(fma(q1, p0, l1), p0, c0);
  const double hi = x * kInv2PiHi;
  const double theta0 = (hi - rint(hi)) + fma(x, kInv2PiLo, fma(x, kInv2PiHi, -hi));
  const uint32_t a = fixed_revolutions(fma(2.0 * q1, p0, l1) * kInv2PiHi);
  const uint32_t b = fixed_revolutions(q1 * kInv2PiHi);
  SegGen s;
  s.z = phasor(fixed_revolutions(theta0));
  s.r = phasor(a + b);  // theta(1) - theta(0) = A + B
  s.k = phasor(2u * b);  // second difference 2 B
  s.t0 = (float)(p0 - cen);
  s.q2l = (float)(q2 * kLog2e);
  s.gl = valid ? (float)(g * kLog2e) : -INFINITY;
  return s;
}

// The segment's value at pixel p0 + j (j = 0, 1, ... in turn).
__device__ __forceinline__ float2 seg_next(SegGen& s, int j) {
  const float t = s.t0 + (float)j;
  const float mag = __builtin_amdgcn_exp2f(fmaf(s.q2l * t, t, s.gl));
  const float2 v = make_float2(mag * s.z.x, mag * s.z.y);
  s.z = cmulf(s.z, s.r);
  s.r = cmulf(s.r, s.k);
  return v;
}

// One k pair of the complex product a b (a = Ey_k(m), b = Ex_k(n): this
// lane's operands) in three real products (Gauss): p1 += a.re b.re,
// p2 += a.im b.im, p3 += (a.re + a.im)(b.re + b.im); then Re = p1 - p2 and
// Im = p3 - (p1 + p2).  Three MFMAs per k pair instead of four (no vector
// instruction and MFMA of one SIMD run together, so the two adds cost far
// less than the fourth product).
__device__ __forceinline__ void sep_mfma(float2 a, float2 b, floatx16& p1, floatx16& p2, floatx16& p3) {
  p1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a.x, b.x, p1, 0, 0, 0);
  p2 = __builtin_amdgcn_mfma_f32_32x32x2f32(a.y, b.y, p2, 0, 0, 0);
  p3 = __builtin_amdgcn_mfma_f32_32x32x2f32(a.x + a.y, b.x + b.y, p3, 0, 0, 0);
}

__global__ __launch_bounds__(kThreads) void beamlet_sum_sep_kernel(const tg_beamlet_coef* __restrict__ coef,
                                                                   const SepParams* __restrict__ par, int64_t n,
                                                                   int64_t H, int64_t W, int32_t ntx, int32_t ntiles,
                                                                   int32_t nsplit, const int32_t* __restrict__ flags,
                                                                   float2* __restrict__ ws) {
  __shared__ int32_t list[kSepBlock];
  __shared__ int32_t wave_cnt[kWaves];
  __shared__ float2 fac[2][2][kSepK][kSepRow];  // [buffer][x: Ex_k(tx0 + n), y: Ey_k(ty0 + m)][k][pixel]
  if (!matrix_path(flags)) return;  // beamlet_sum_kernel covers this launch
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  // the slices of one tile are consecutive workgroups: the heavy tiles near
  // the pattern's centre are spread through the dispatch order instead of
  // all ending it
  const int split = blockIdx.x % nsplit, tile = blockIdx.x / nsplit;
  const int32_t tx0 = (tile % ntx) * kTileW, ty0 = (tile / ntx) * kTileH;
  const int64_t b_begin = n * split / nsplit, b_end = n * (split + 1) / nsplit;
  const float ftx0 = (float)tx0, fty0 = (float)ty0, ftx1 = (float)(tx0 + kTileW - 1), fty1 = (float)(ty0 + kTileH - 1);
  const int mb = (wave >> 1) * 32, nb = (wave & 1) * 32;
  // this thread's segment: beamlet gk of a chunk, axis (waves 0-1: x, 2-3: y), pixels p0 .. p0 + 15
  const int gk = t & (kSepK - 1), gaxis = t >> 7, s0 = ((t >> 5) & 3) * kSepSeg;
  const double p0 = (double)((gaxis == 0 ? tx0 : ty0) + s0);
  floatx16 p1, p2, p3;
#pragma unroll
  for (int j = 0; j < 16; ++j) p1[j] = p2[j] = p3[j] = 0.0f;
  for (int64_t blk = b_begin; blk < b_end; blk += kSepBlock) {
    // the beamlets of [blk, blk + kSepBlock) whose cutoff ellipse meets the
    // tile, in order.  A separable beamlet's ellipse is axis-aligned
    // (alpha = 0): u^2 + gamma v^2 <= beta, so the tile's pixel centre nearest
    // (cx, cy) decides, exactly -- tiles in the corners of the ellipse's
    // bounding box are skipped.
    int cnt = 0;
    for (int64_t b0 = blk; b0 < b_end && b0 < blk + kSepBlock; b0 += kThreads) {
      const int64_t b = b0 + t;
      bool hit = false;
      if (b < b_end) {
        const tg_beamlet_coef& c = coef[b];
        const float du = fmaxf(fmaxf(ftx0 - c.cx, c.cx - ftx1), 0.0f);
        const float dv = fmaxf(fmaxf(fty0 - c.cy, c.cy - fty1), 0.0f);
        hit = fmaf(c.gamma * dv, dv, du * du) <= c.beta;
      }
      const uint64_t mask = __ballot(hit);
      if (lane == 0) wave_cnt[wave] = __popcll(mask);
      __syncthreads();
      int off = cnt, total = 0;
#pragma unroll
      for (int w = 0; w < kWaves; ++w) {
        off += (w < wave) ? wave_cnt[w] : 0;
        total += wave_cnt[w];
      }
      if (hit) list[off + __popcll(mask & ((1ull << lane) - 1ull))] = (int32_t)(b - blk);
      cnt += total;
      __syncthreads();  // wave_cnt is rewritten next; the list is complete after the last pass
    }
    if (cnt == 0) continue;
    const int nch = (cnt + kSepK - 1) / kSepK;
    // chunk q's beamlet for this thread (past cnt: the last one again, evaluated as zero)
    auto beamlet = [&](int q) -> int64_t { return blk + list[min(q * kSepK + gk, cnt - 1)]; };
    double2 v0, v1, v2;  // the parameters of the chunk evaluated next
    load_axis(par, beamlet(0), gaxis, v0, v1, v2);
    {
      SegGen sg = seg_setup(v0, v1, v2, p0, gk < cnt);
      float2* dst = &fac[0][gaxis][gk][s0];
#pragma unroll
      for (int j = 0; j < kSepSeg; ++j) dst[j] = seg_next(sg, j);
    }
    load_axis(par, beamlet(1), gaxis, v0, v1, v2);
    __syncthreads();
    for (int c = 0; c < nch; ++c) {
      const int cur = c & 1;
      // A[m][k] = Ey_k(ty0 + mb + m) from lane m + 32 k, B[k][n] = Ex_k(tx0 + nb + n)
      // from lane n + 32 k (v_mfma_f32_32x32x2f32 operand layout); every
      // chunk runs all kSepK (zero past cnt)
      // (operands one k pair ahead, so LDS latency hides behind the products)
      const float2* ea = &fac[cur][1][lane >> 5][mb + (lane & 31)];
      const float2* eb = &fac[cur][0][lane >> 5][nb + (lane & 31)];
      float2 a = ea[0], b = eb[0];
      if (c + 1 < nch) {
        // chunk c + 1 into the other buffer, one segment value per k pair,
        // and chunk c + 2's parameters in flight
        SegGen sg = seg_setup(v0, v1, v2, p0, (c + 1) * kSepK + gk < cnt);
        load_axis(par, beamlet(c + 2), gaxis, v0, v1, v2);
        // keep these loads here: sunk below the products (as the scheduler
        // does by itself), the next chunk waited a whole memory round trip
        __builtin_amdgcn_sched_barrier(0);
        // (the values are stored after the products: a store into the
        // other buffer between them would order every later operand read
        // behind it)
        float2 val[kSepSeg];
#pragma unroll
        for (int kk = 0; kk < kSepK; kk += 2) {
          float2 an = a, bn = b;
          if (kk + 2 < kSepK) {
            an = ea[(kk + 2) * kSepRow];
            bn = eb[(kk + 2) * kSepRow];
          }
          sep_mfma(a, b, p1, p2, p3);
          val[kk >> 1] = seg_next(sg, kk >> 1);
          a = an;
          b = bn;
          // issue order: products and the value's dependent vector
          // instructions alternate (each covers the other's latency)
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // one MFMA
            __builtin_amdgcn_sched_group_barrier(0x002, 4, 0);  // four VALU
          }
        }
        float2* dst = &fac[cur ^ 1][gaxis][gk][s0];
#pragma unroll
        for (int j = 0; j < kSepSeg; ++j) dst[j] = val[j];
      } else {
#pragma unroll
        for (int kk = 0; kk < kSepK; kk += 2) {
          float2 an = a, bn = b;
          if (kk + 2 < kSepK) {
            an = ea[(kk + 2) * kSepRow];
            bn = eb[(kk + 2) * kSepRow];
          }
          sep_mfma(a, b, p1, p2, p3);
          a = an;
          b = bn;
        }
      }
      __syncthreads();  // chunk c + 1 is complete; the next iteration rewrites buffer `cur`
    }
  }
  // accumulator j of lane l holds row 8 (j / 4) + 4 (l / 32) + j % 4, column l % 32
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int64_t py = ty0 + mb + 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
    const int64_t px = tx0 + nb + (lane & 31);
    if (py < H && px < W) ws[((int64_t)split * H + py) * W + px] = make_float2(p1[j] - p2[j], p3[j] - (p1[j] + p2[j]));
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

// ---------------------------------------------------------------------------
// backward
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void beamlet_unit_kernel(const double* __restrict__ params, int64_t n,
                                                           const double* __restrict__ abcd, int64_t abcd_stride,
                                                           int32_t mode, double m0, double m1, double m2, double m3,
                                                           double m4, double m5, float cut,
                                                           tg_beamlet_coef* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* p = params + i * TG_BEAMLET_NPARAM;
  const double m[6] = {m0, m1, m2, m3, m4, m5};
  const double* T = mode == TG_BEAMLET_PROPAGATED ? abcd + i * abcd_stride : nullptr;
  out[i] = make_coef(p, T, mode, m, mk(1.0, 0.0), cut);
}

// w = conj(g) over the image, then kRun zeros of tail padding
__global__ __launch_bounds__(256) void conj_pad_kernel(const float2* __restrict__ g, int64_t npix,
                                                       float2* __restrict__ w) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < npix) {
    const float2 v = g[i];
    w[i] = make_float2(v.x, -v.y);
  } else if (i < npix + kRun) {
    w[i] = make_float2(0.0f, 0.0f);
  }
}

constexpr int kBeamletsPerMomentBlock = 4;  // one wavefront each

__global__ __launch_bounds__(256) void beamlet_moment_kernel(const tg_beamlet_coef* __restrict__ unit, int64_t n,
                                                             int32_t H, int32_t W, const float2* __restrict__ w,
                                                             double* __restrict__ moments) {
  const int lane = threadIdx.x & 63;
  // one beamlet per wavefront: its index and record are wave-uniform
  const int64_t b = (int64_t)blockIdx.x * kBeamletsPerMomentBlock + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (b >= n) return;  // whole wavefronts leave; no block barrier below
  double M[kMoments];
  beamlet_moments(unit[b], H, W, w, lane, M);
#pragma unroll
  for (int k = 0; k < kMoments; ++k) M[k] = wave_sum_f64(M[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < kMoments; ++k) moments[b * kMoments + k] = M[k];
  }
}

__global__ __launch_bounds__(64) void beamlet_grad_kernel(const double* __restrict__ params, int64_t n,
                                                          const double* __restrict__ abcd, int64_t abcd_stride,
                                                          int32_t mode, double m0, double m1, double m2, double m3,
                                                          double m4, double m5, int32_t H, int32_t W,
                                                          const tg_beamlet_coef* __restrict__ unit,
                                                          const double* __restrict__ moments,
                                                          double* __restrict__ grad_params,
                                                          double* __restrict__ grad_abcd) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* p = params + i * TG_BEAMLET_NPARAM;
  const bool prop = mode == TG_BEAMLET_PROPAGATED;
  const double* T = prop ? abcd + i * abcd_stride : nullptr;
  const double m[6] = {m0, m1, m2, m3, m4, m5};
  float fcx, fcy;
  centre_pixel(unit[i], H, W, fcx, fcy);
  const double pcx = fcx, pcy = fcy;
  const double* Mo = moments + i * kMoments;
  const cd S0 = mk(Mo[0], Mo[1]), Su = mk(Mo[2], Mo[3]), Sv = mk(Mo[4], Mo[5]);
  const cd Suu = mk(Mo[6], Mo[7]), Suv = mk(Mo[8], Mo[9]), Svv = mk(Mo[10], Mo[11]);
  // G_k = sum W exp(Phi') m_k, monomials of (px, py) = (pcx + u, pcy + v)
  cd G[6];
  G[0] = add(add(Suu, scl(Su, 2.0 * pcx)), scl(S0, pcx * pcx));
  G[1] = add(add(Suv, scl(Su, pcy)), add(scl(Sv, pcx), scl(S0, pcx * pcy)));
  G[2] = add(add(Svv, scl(Sv, 2.0 * pcy)), scl(S0, pcy * pcy));
  G[3] = add(Su, scl(S0, pcx));
  G[4] = add(Sv, scl(S0, pcy));
  G[5] = S0;
  const double ar = p[5], ai = p[6];
  double* gp = grad_params + i * TG_BEAMLET_NPARAM;
  gp[5] = S0.re;   // u = amp exp(Phi'): dL/d amp_re = Re sum W exp(Phi')
  gp[6] = -S0.im;  //                    dL/d amp_im = Re(i sum W exp(Phi'))
  gp[13] = gp[14] = gp[15] = 0.0;
  // directions: parameter columns 0-4 and 7-12, then ray-transfer entries 0-19
  const int n_dir = prop ? 11 + 20 : 11;
  for (int q = 0; q < n_dir; ++q) {
    const int col = q < 5 ? q : (q < 11 ? q + 2 : -1);
    const int ent = q < 11 ? -1 : q - 11;
    dual pd[13], Td[25];
#pragma unroll
    for (int k = 0; k < 13; ++k) pd[k] = dual(p[k], k == col ? 1.0 : 0.0);
    if (prop) {
#pragma unroll
      for (int e = 0; e < 25; ++e) Td[e] = dual(T[e], e == ent ? 1.0 : 0.0);
    }
    cx<dual> d6[6];
    phi_coefficients<dual>(pd, Td, mode, m, mk(dual(1.0), dual(0.0)), d6);
    double sr = 0.0, si = 0.0;  // sum_k G_k dc_k
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      sr += G[k].re * d6[k].re.t - G[k].im * d6[k].im.t;
      si += G[k].re * d6[k].im.t + G[k].im * d6[k].re.t;
    }
    const double g = ar * sr - ai * si;  // Re[amp sum_k G_k dc_k]
    if (col >= 0)
      gp[col] = g;
    else if (grad_abcd)
      grad_abcd[i * 25 + ent] = g;
  }
  if (grad_abcd) {
    for (int e = prop ? 20 : 0; e < 25; ++e) grad_abcd[i * 25 + e] = 0.0;
  }
}

int splits_for(int64_t n, int64_t H, int64_t W) {
  const int64_t key[4] = {n, H, W, 0};
  int32_t tuned = 0;
  if (tg::tune_lookup(TG_TUNE_BEAMLET_SPLITS, key, tuned)) {
    int64_t s = tuned;
    if (s > 128) s = 128;
    if (s > n && n > 0) s = n;
    if (s < 1) s = 1;
    return (int)s;
  }
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

// The recurrence tile sum (beamlet_sum_kernel) for a tile shape
// (tg_tile_shape): one workgroup per (tile, beamlet slice).
int launch_recurrence_sum(int32_t shape, const tg_beamlet_coef* coef, int64_t n, int64_t H, int64_t W, int nsplit,
                          const int32_t* flags, float2* ws, hipStream_t s) {
  const TileDims d = tile_dims(shape);
  const int64_t ntx = tg::cdiv(W, (int64_t)d.w), ntiles = ntx * tg::cdiv(H, (int64_t)d.h);
  TG_REQUIRE(ntiles * nsplit < (int64_t)1 << 31, "tg_beamlet_sum: detector too large");
  const dim3 grid((unsigned)(ntiles * nsplit));
  switch (shape) {
    case TG_TILE_128x32_RUN16:
      hipLaunchKernelGGL((beamlet_sum_kernel<8, 16>), grid, dim3(kThreads), 0, s, coef, n, H, W, (int32_t)ntx,
                         (int32_t)ntiles, (int32_t)nsplit, flags, ws);
      break;
    case TG_TILE_32x128_RUN16:
      hipLaunchKernelGGL((beamlet_sum_kernel<2, 16>), grid, dim3(kThreads), 0, s, coef, n, H, W, (int32_t)ntx,
                         (int32_t)ntiles, (int32_t)nsplit, flags, ws);
      break;
    case TG_TILE_64x32_RUN8:
      hipLaunchKernelGGL((beamlet_sum_kernel<8, 8>), grid, dim3(kThreads), 0, s, coef, n, H, W, (int32_t)ntx,
                         (int32_t)ntiles, (int32_t)nsplit, flags, ws);
      break;
    default:
      hipLaunchKernelGGL((beamlet_sum_kernel<kLanesPerRow, kRun>), grid, dim3(kThreads), 0, s, coef, n, H, W,
                         (int32_t)ntx, (int32_t)ntiles, (int32_t)nsplit, flags, ws);
  }
  TG_LAUNCH_CHECK("beamlet_sum_kernel launch");
  return TG_OK;
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

// Workspace of tg_beamlet_sum: [nsplit, H, W] complex64 slice tiles (padded to
// 256 bytes), the kFlagBlocks path flags, then n SepParams records.
int64_t slices_bytes(int nsplit, int64_t H, int64_t W) {
  return ((int64_t)nsplit * H * W * (int64_t)sizeof(float2) + 255) / 256 * 256;
}

int64_t sum_workspace_bytes(int nsplit, int64_t n, int64_t H, int64_t W) {
  return slices_bytes(nsplit, H, W) + kFlagBlocks * (int64_t)sizeof(int32_t) + n * (int64_t)sizeof(SepParams);
}

extern "C" int64_t tg_beamlet_sum_workspace(int64_t n, int64_t height, int64_t width) {
  if (n < 0 || height <= 0 || width <= 0) return 0;
  return sum_workspace_bytes(splits_for(n, height, width), n, height, width);
}

extern "C" int tg_beamlet_sum(const tg_beamlet_coef* coef, int64_t n, int64_t height, int64_t width, float* field,
                              void* workspace, int64_t workspace_bytes, void* stream) {
  tg::clear_error();
  TG_REQUIRE(n >= 0 && height > 0 && width > 0, "tg_beamlet_sum: bad sizes");
  TG_REQUIRE(field != nullptr && workspace != nullptr, "tg_beamlet_sum: NULL buffer");
  TG_REQUIRE(n == 0 || coef != nullptr, "tg_beamlet_sum: NULL coefficients");
  TG_REQUIRE(n <= (int64_t)1 << 31, "tg_beamlet_sum: too many beamlets");
  const int nsplit = splits_for(n, height, width);
  const int64_t slices = slices_bytes(nsplit, height, width);
  const int64_t need = sum_workspace_bytes(nsplit, n, height, width);
  TG_REQUIRE(workspace_bytes >= need, "tg_beamlet_sum: workspace %lld bytes < %lld", (long long)workspace_bytes,
             (long long)need);
  int st = TG_OK;
  const int64_t ntx = tg::cdiv(width, kTileW), nty = tg::cdiv(height, kTileH);
  const int64_t ntiles = ntx * nty;
  TG_REQUIRE(ntiles * nsplit < (int64_t)1 << 31, "tg_beamlet_sum: detector too large");
  const int64_t key[4] = {n, height, width, 0};
  int32_t path = TG_BEAMLET_PATH_AUTO;
  tg::tune_lookup(TG_TUNE_BEAMLET_PATH, key, path);
  int32_t* flags = (int32_t*)((char*)workspace + slices);
  SepParams* par = (SepParams*)(flags + kFlagBlocks);
  // which tile sum runs is decided on the device (no host round trip, so the
  // call stays graph-capturable); the other kernel's workgroups exit at once
  hipLaunchKernelGGL(sep_flag_kernel, dim3(kFlagBlocks), dim3(256), 0, (hipStream_t)stream, coef, n,
                     (int32_t)(path == TG_BEAMLET_PATH_VALU), flags, par);
  TG_LAUNCH_CHECK("sep_flag_kernel launch");
  hipLaunchKernelGGL(beamlet_sum_sep_kernel, dim3((unsigned)(ntiles * nsplit)), dim3(kThreads), 0,
                     (hipStream_t)stream, coef, (const SepParams*)par, n, height, width, (int32_t)ntx,
                     (int32_t)ntiles, (int32_t)nsplit, (const int32_t*)flags, (float2*)workspace);
  TG_LAUNCH_CHECK("beamlet_sum_sep_kernel launch");
  int32_t shape = TG_TILE_64x64_RUN16;
  tg::tune_lookup(TG_TUNE_TILE_SHAPE, key, shape);
  st = launch_recurrence_sum(shape, coef, n, height, width, nsplit, flags, (float2*)workspace, (hipStream_t)stream);
  if (st != TG_OK) return st;
  const int64_t npix = height * width;
  hipLaunchKernelGGL(split_sum_kernel, dim3((unsigned)tg::cdiv(npix, 256)), dim3(256), 0, (hipStream_t)stream,
                     (const float2*)workspace, npix, (int32_t)nsplit, (float2*)field);
  TG_LAUNCH_CHECK("split_sum_kernel launch");
  return TG_OK;
}

extern "C" int64_t tg_beamlet_backward_workspace(int64_t n, int64_t height, int64_t width) {
  if (n < 0 || height <= 0 || width <= 0) return 0;
  auto al = [](int64_t b) { return (b + 255) & ~(int64_t)255; };
  return al(n * (int64_t)sizeof(tg_beamlet_coef)) + al((height * width + kRun) * (int64_t)sizeof(float2)) +
         al(n * kMoments * (int64_t)sizeof(double));
}

extern "C" int tg_beamlet_backward(const double* params, int64_t n, const double* abcd, int64_t abcd_stride,
                                   int32_t mode, const double* pix2m, double nsigma, int64_t height, int64_t width,
                                   const float* grad_field, double* grad_params, double* grad_abcd, void* workspace,
                                   int64_t workspace_bytes, void* stream) {
  tg::clear_error();
  TG_REQUIRE(n >= 0 && height > 0 && width > 0, "tg_beamlet_backward: bad sizes");
  TG_REQUIRE(height < ((int64_t)1 << 31) && width < ((int64_t)1 << 31) && height * width < ((int64_t)1 << 40),
             "tg_beamlet_backward: image too large");
  TG_REQUIRE(mode == TG_BEAMLET_PROPAGATED || mode == TG_BEAMLET_INPUT, "tg_beamlet_backward: bad mode");
  TG_REQUIRE(pix2m != nullptr, "tg_beamlet_backward: NULL pixel transform");
  TG_REQUIRE(grad_abcd == nullptr || mode == TG_BEAMLET_PROPAGATED,
             "tg_beamlet_backward: ray-transfer gradients need propagated mode");
  if (n == 0) return TG_OK;
  TG_REQUIRE(params != nullptr && grad_field != nullptr && grad_params != nullptr && workspace != nullptr,
             "tg_beamlet_backward: NULL buffer");
  TG_REQUIRE(mode == TG_BEAMLET_INPUT || (abcd != nullptr && (abcd_stride == 0 || abcd_stride == 25)),
             "tg_beamlet_backward: propagated mode needs abcd with stride 0 or 25");
  TG_REQUIRE(reinterpret_cast<uintptr_t>(workspace) % 256 == 0, "tg_beamlet_backward: workspace not 256-byte aligned");
  const int64_t need = tg_beamlet_backward_workspace(n, height, width);
  TG_REQUIRE(workspace_bytes >= need, "tg_beamlet_backward: workspace %lld bytes < %lld", (long long)workspace_bytes,
             (long long)need);
  auto al = [](int64_t b) { return (b + 255) & ~(int64_t)255; };
  char* ws = static_cast<char*>(workspace);
  tg_beamlet_coef* unit = reinterpret_cast<tg_beamlet_coef*>(ws);
  float2* w = reinterpret_cast<float2*>(ws + al(n * (int64_t)sizeof(tg_beamlet_coef)));
  double* moments = reinterpret_cast<double*>(reinterpret_cast<char*>(w) +
                                              al((height * width + kRun) * (int64_t)sizeof(float2)));
  const float cut = nsigma > 0.0 ? (float)(nsigma * nsigma) : 80.0f;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(beamlet_unit_kernel, dim3((unsigned)tg::cdiv(n, 256)), dim3(256), 0, s, params, n, abcd,
                     abcd_stride, mode, pix2m[0], pix2m[1], pix2m[2], pix2m[3], pix2m[4], pix2m[5], cut, unit);
  TG_LAUNCH_CHECK("beamlet_unit_kernel launch");
  const int64_t npix = height * width;
  hipLaunchKernelGGL(conj_pad_kernel, dim3((unsigned)tg::cdiv(npix + kRun, 256)), dim3(256), 0, s,
                     (const float2*)grad_field, npix, w);
  TG_LAUNCH_CHECK("conj_pad_kernel launch");
  hipLaunchKernelGGL(beamlet_moment_kernel, dim3((unsigned)tg::cdiv(n, kBeamletsPerMomentBlock)), dim3(256), 0, s,
                     unit, n, (int32_t)height, (int32_t)width, w, moments);
  TG_LAUNCH_CHECK("beamlet_moment_kernel launch");
  hipLaunchKernelGGL(beamlet_grad_kernel, dim3((unsigned)tg::cdiv(n, 64)), dim3(64), 0, s, params, n, abcd,
                     abcd_stride, mode, pix2m[0], pix2m[1], pix2m[2], pix2m[3], pix2m[4], pix2m[5], (int32_t)height,
                     (int32_t)width, unit, moments, grad_params, grad_abcd);
  TG_LAUNCH_CHECK("beamlet_grad_kernel launch");
  return TG_OK;
}
