// Batched Gaussian-decomposition fits: forward field, intensity loss and its
// gradient for many independent instances of uneven size.
//
// Reference: one instance is evaluate_gaussian_input_image
// (gaussian.py:372-427) with the loss mean((|u|^2 - T)^2) and its jax.grad
// (examples/fit_three_gaussians.ipynb, cells 2-3).
//
// Pipeline (six launches, no allocation, no atomics, fixed summation order):
//   fit_coef_kernel     one lane per Gaussian, f64: Phi(px, py) of the
//                       Gaussian with unit amplitude ("unit") and with its
//                       amplitude folded into the constant term ("coef").
//   fit_forward_kernel  one workgroup per (instance, 64x64 tile, Gaussian
//                       slice) work item, ordered heaviest first: the
//                       recurrence tile sum shared with beamlet.hip, into
//                       per-slice partial fields.
//   fit_loss_kernel     per 4096-pixel chunk: u = sum of the slices (f64),
//                       r = |u|^2 - T, loss partial sum of r^2 and the adjoint
//                       weight W = (4/N) r conj(u).
//   fit_loss_final      per instance: fixed-order sum of its partials / N.
//   fit_moment_kernel   one wavefront per Gaussian, over the pixels of its
//                       cutoff ellipse: the six complex moments
//                       sum W E {1, u, v, u^2, uv, v^2}, E = exp(Phi_unit),
//                       (u, v) = pixel offsets from the Gaussian's centre
//                       pixel.  f32 within a 16-pixel run, f64 across runs
//                       and lanes.
//   fit_grad_kernel     one lane per Gaussian, f64: the moments mapped to
//                       metre offsets d = r - m and the chain rule to the 13
//                       parameter columns,
//                         dL/dtheta = Re[a sum_p W E dPhi/dtheta].
#include <string.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "beamlet_common.hpp"

namespace {

using namespace tgb;

constexpr int kMaxSlices = 8;           // Gaussian slices per instance tile
constexpr int kSliceGauss = 256;        // target Gaussians per slice
constexpr int kLossChunk = 4096;        // pixels per loss workgroup
constexpr int kGaussPerMomentBlock = 4; // one wavefront each

struct FitInst {
  int64_t g0, n, pix0, ws0;
  double m[6];
  int32_t H, W, ntx, nslice;
  int32_t lb0, lb1, pad0, pad1;
};
struct FitItem {
  int32_t inst, tile, slice, pad;
};
struct LossBlock {
  int32_t inst, chunk;
};

static_assert(sizeof(FitInst) == 112, "FitInst layout");

__global__ __launch_bounds__(256) void fit_coef_kernel(const double* __restrict__ params, int64_t G,
                                                       const int32_t* __restrict__ gauss_inst,
                                                       const FitInst* __restrict__ insts, float cut,
                                                       tg_beamlet_coef* __restrict__ coef,
                                                       tg_beamlet_coef* __restrict__ unit) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  const int32_t i = gauss_inst[g];
  if (i < 0) return;
  const double* p = params + g * TG_BEAMLET_NPARAM;
  const double m[6] = {insts[i].m[0], insts[i].m[1], insts[i].m[2], insts[i].m[3], insts[i].m[4], insts[i].m[5]};
  tg_beamlet_coef o = make_coef(p, nullptr, TG_BEAMLET_INPUT, m, mk(1.0, 0.0), cut);
  unit[g] = o;
  // a exp(Phi) = exp(Phi + log a); a = 0 gives log|a| = -inf, i.e. a zero field
  const cd la = clog(mk(p[5], p[6]));
  o.phi[10] += la.re;
  o.phi[11] += la.im;
  coef[g] = o;
}

template <int LPR, int RUN>
__global__ __launch_bounds__(kThreads) void fit_forward_kernel(const tg_beamlet_coef* __restrict__ coef,
                                                               const FitInst* __restrict__ insts,
                                                               const FitItem* __restrict__ items,
                                                               float2* __restrict__ partial) {
  __shared__ TileLists sh;
  const FitItem it = items[blockIdx.x];
  const FitInst& I = insts[it.inst];
  const int32_t H = I.H, W = I.W;
  const int32_t tx0 = (it.tile % I.ntx) * (LPR * RUN), ty0 = (it.tile / I.ntx) * (kThreads / LPR);
  const int64_t b_begin = I.g0 + I.n * it.slice / I.nslice, b_end = I.g0 + I.n * (it.slice + 1) / I.nslice;
  tile_sum_store<LPR, RUN>(coef, b_begin, b_end, tx0, ty0, H, W, sh,
                           partial + I.ws0 + (int64_t)it.slice * H * (int64_t)W);
}

__global__ __launch_bounds__(256) void fit_loss_kernel(const FitInst* __restrict__ insts,
                                                       const LossBlock* __restrict__ blocks,
                                                       const float2* __restrict__ partial,
                                                       const float* __restrict__ target, float2* __restrict__ field,
                                                       float2* __restrict__ weight, double* __restrict__ loss_part) {
  __shared__ double wsum[4];
  const LossBlock lb = blocks[blockIdx.x];
  const FitInst& I = insts[lb.inst];
  const int64_t npix = (int64_t)I.H * I.W;
  const double scale = 4.0 / (double)npix;
  const int t = threadIdx.x;
  double acc = 0.0;
  for (int k = 0; k < kLossChunk / 256; ++k) {
    const int64_t p = (int64_t)lb.chunk * kLossChunk + k * 256 + t;
    if (p >= npix) break;
    double ur = 0.0, ui = 0.0;
    for (int s = 0; s < I.nslice; ++s) {
      const float2 v = partial[I.ws0 + (int64_t)s * npix + p];
      ur += v.x;
      ui += v.y;
    }
    if (field) field[I.pix0 + p] = make_float2((float)ur, (float)ui);
    if (target) {
      const double r = (ur * ur + ui * ui) - (double)target[I.pix0 + p];
      acc += r * r;
      weight[I.pix0 + p] = make_float2((float)(scale * r * ur), (float)(-scale * r * ui));
    }
  }
  if (!target) return;
  acc = wave_sum_f64(acc);
  if ((t & 63) == 0) wsum[t >> 6] = acc;
  __syncthreads();
  if (t == 0) loss_part[blockIdx.x] = (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
}

__global__ __launch_bounds__(256) void fit_loss_final(const FitInst* __restrict__ insts, int32_t n_inst,
                                                      const double* __restrict__ loss_part,
                                                      double* __restrict__ loss) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_inst) return;
  double s = 0.0;
  for (int b = insts[i].lb0; b < insts[i].lb1; ++b) s += loss_part[b];
  loss[i] = s / ((double)insts[i].H * (double)insts[i].W);
}

__global__ __launch_bounds__(256) void fit_moment_kernel(const tg_beamlet_coef* __restrict__ unit,
                                                         const FitInst* __restrict__ insts,
                                                         const int32_t* __restrict__ gauss_inst,
                                                         const int32_t* __restrict__ order, int64_t n_order,
                                                         const float2* __restrict__ weight,
                                                         double* __restrict__ moments) {
  const int lane = threadIdx.x & 63;
  const int64_t slot = (int64_t)blockIdx.x * kGaussPerMomentBlock + (threadIdx.x >> 6);
  if (slot >= n_order) return;  // whole wavefronts leave; no block barrier below
  // one Gaussian per wavefront: its index and records are wave-uniform
  const int32_t g = __builtin_amdgcn_readfirstlane(order[slot]);
  const FitInst& I = insts[__builtin_amdgcn_readfirstlane(gauss_inst[g])];
  double M[kMoments];
  beamlet_moments(unit[g], I.H, I.W, weight + I.pix0, lane, M);
#pragma unroll
  for (int k = 0; k < kMoments; ++k) M[k] = wave_sum_f64(M[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < kMoments; ++k) moments[(int64_t)g * kMoments + k] = M[k];
  }
}

__device__ __forceinline__ double re_mul(cd a, cd z) { return a.re * z.re - a.im * z.im; }

__global__ __launch_bounds__(256) void fit_grad_kernel(const double* __restrict__ params, int64_t G,
                                                       const int32_t* __restrict__ gauss_inst,
                                                       const FitInst* __restrict__ insts,
                                                       const tg_beamlet_coef* __restrict__ unit,
                                                       const double* __restrict__ moments,
                                                       double* __restrict__ grad) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  double* out = grad + g * TG_BEAMLET_NPARAM;
  const int32_t i = gauss_inst[g];
  if (i < 0) {
    for (int k = 0; k < TG_BEAMLET_NPARAM; ++k) out[k] = 0.0;
    return;
  }
  const FitInst& I = insts[i];
  const double* p = params + g * TG_BEAMLET_NPARAM;
  const double x = p[0], y = p[1], tx = p[2], ty = p[3], pl = p[4];
  const cd a = mk(p[5], p[6]);
  const double wx = p[7], wy = p[8], rx = p[9], ry = p[10], lam = p[11], th = p[12];
  const double k = 2.0 * kPi / lam;
  float pcx, pcy;
  centre_pixel(unit[g], I.H, I.W, pcx, pcy);
  const double* m = I.m;
  // d = r - (x, y) = (m4 u + m3 v + ox, m1 u + m0 v + oy)
  const double ox = (m[3] * pcy + m[4] * pcx + m[5]) - x;
  const double oy = (m[0] * pcy + m[1] * pcx + m[2]) - y;
  const double* S = moments + g * kMoments;
  const cd S0 = mk(S[0], S[1]), Su = mk(S[2], S[3]), Sv = mk(S[4], S[5]);
  const cd Suu = mk(S[6], S[7]), Suv = mk(S[8], S[9]), Svv = mk(S[10], S[11]);
  const cd Lx = add(scl(Su, m[4]), scl(Sv, m[3]));  // sum (m4 u + m3 v) W E
  const cd Ly = add(scl(Su, m[1]), scl(Sv, m[0]));
  const cd Dx = add(Lx, scl(S0, ox));
  const cd Dy = add(Ly, scl(S0, oy));
  const cd Dxx = add(add(add(scl(Suu, m[4] * m[4]), scl(Suv, 2.0 * m[4] * m[3])), scl(Svv, m[3] * m[3])),
                     add(scl(Lx, 2.0 * ox), scl(S0, ox * ox)));
  const cd Dxy = add(add(add(scl(Suu, m[4] * m[1]), scl(Suv, m[4] * m[0] + m[3] * m[1])), scl(Svv, m[3] * m[0])),
                     add(add(scl(Lx, oy), scl(Ly, ox)), scl(S0, ox * oy)));
  const cd Dyy = add(add(add(scl(Suu, m[1] * m[1]), scl(Suv, 2.0 * m[1] * m[0])), scl(Svv, m[0] * m[0])),
                     add(scl(Ly, 2.0 * oy), scl(S0, oy * oy)));
  // Q1^-1 = R(theta) diag(qx, qy) R(theta)^T (gaussian.py:138-177)
  const cd qx = mk(isinf(rx) ? 0.0 : -1.0 / rx, lam / (kPi * wx * wx));
  const cd qy = mk(isinf(ry) ? 0.0 : -1.0 / ry, lam / (kPi * wy * wy));
  double sn, cs;
  sincos(th, &sn, &cs);
  const cd Q00 = add(scl(qx, cs * cs), scl(qy, sn * sn));
  const cd Q01 = scl(sub(qx, qy), cs * sn);
  const cd Q11 = add(scl(qx, sn * sn), scl(qy, cs * cs));
  const cd ik = mk(0.0, k);
  // sum W E dPhi/dQ . dQ for a symmetric dQ = (d00, d01, d11)
  auto quad = [&](cd d00, cd d01, cd d11) {
    return mul(mk(0.0, 0.5 * k), add(add(mul(d00, Dxx), scl(mul(d01, Dxy), 2.0)), mul(d11, Dyy)));
  };
  const cd dqx_dwx = mk(0.0, -2.0 * lam / (kPi * wx * wx * wx));
  const cd dqy_dwy = mk(0.0, -2.0 * lam / (kPi * wy * wy * wy));
  const cd dqx_drx = mk(isinf(rx) ? 0.0 : 1.0 / (rx * rx));
  const cd dqy_dry = mk(isinf(ry) ? 0.0 : 1.0 / (ry * ry));
  const double c2 = cs * cs, s2 = sn * sn, csn = cs * sn;
  const cd dd = sub(qx, qy);
  out[0] = re_mul(a, mul(scl(ik, -1.0), add(add(mul(Q00, Dx), mul(Q01, Dy)), scl(S0, tx))));
  out[1] = re_mul(a, mul(scl(ik, -1.0), add(add(mul(Q01, Dx), mul(Q11, Dy)), scl(S0, ty))));
  out[2] = re_mul(a, mul(ik, Dx));
  out[3] = re_mul(a, mul(ik, Dy));
  out[4] = re_mul(a, mul(ik, S0));
  out[5] = S0.re;
  out[6] = -S0.im;
  out[7] = re_mul(a, quad(scl(dqx_dwx, c2), scl(dqx_dwx, csn), scl(dqx_dwx, s2)));
  out[8] = re_mul(a, quad(scl(dqy_dwy, s2), scl(dqy_dwy, -csn), scl(dqy_dwy, c2)));
  out[9] = re_mul(a, quad(scl(dqx_drx, c2), scl(dqx_drx, csn), scl(dqx_drx, s2)));
  out[10] = re_mul(a, quad(scl(dqy_dry, s2), scl(dqy_dry, -csn), scl(dqy_dry, c2)));
  // dPhi/dlambda = -(i k / lambda) [pl + d^T Re(Q) d / 2 + d.t]  (Im Q * k does not depend on lambda)
  const cd lam_terms = add(add(scl(S0, pl), scl(add(add(scl(Dxx, Q00.re), scl(Dxy, 2.0 * Q01.re)), scl(Dyy, Q11.re)),
                                                0.5)),
                           add(scl(Dx, tx), scl(Dy, ty)));
  out[11] = re_mul(a, mul(mk(0.0, -k / lam), lam_terms));
  out[12] = re_mul(a, quad(scl(dd, -2.0 * csn), scl(dd, c2 - s2), scl(dd, 2.0 * csn)));
  out[13] = 0.0;
  out[14] = 0.0;
  out[15] = 0.0;
}

int64_t align256(int64_t x) { return (x + 255) & ~(int64_t)255; }

// Host-side tables, in the order they are laid out in the workspace.
struct Tables {
  std::vector<FitInst> inst;
  std::vector<FitItem> items;
  std::vector<LossBlock> blocks;
  std::vector<int32_t> order;
  std::vector<int32_t> gauss_inst;
};

int slices_for(int64_t n, int64_t slice_gauss) {
  int64_t s = tg::cdiv(n, slice_gauss);
  if (s > kMaxSlices) s = kMaxSlices;
  if (s < 1) s = 1;
  return (int)s;
}

// Validates the instances and builds the tables and the plan.
int build(const tg_fit_instance* inst, int32_t n_inst, int64_t n_gauss, double nsigma, tg_fit_plan* plan,
          Tables* tab) {
  TG_REQUIRE(plan != nullptr, "tg_gauss_fit_plan: NULL plan");
  TG_REQUIRE(n_inst >= 1 && (inst != nullptr), "tg_gauss_fit_plan: need at least one instance");
  TG_REQUIRE(n_gauss >= 0 && n_gauss < ((int64_t)1 << 31), "tg_gauss_fit_plan: bad Gaussian count %lld",
             (long long)n_gauss);
  int64_t n_pix = 0;
  for (int32_t i = 0; i < n_inst; ++i) {
    const tg_fit_instance& s = inst[i];
    TG_REQUIRE(s.height >= 1 && s.width >= 1, "tg_gauss_fit_plan: instance %d has an empty image", i);
    TG_REQUIRE(s.n_gauss >= 0 && s.gauss_begin >= 0 && s.gauss_begin + s.n_gauss <= n_gauss,
               "tg_gauss_fit_plan: instance %d Gaussian rows outside [0, %lld)", i, (long long)n_gauss);
    TG_REQUIRE(s.pix_begin >= 0, "tg_gauss_fit_plan: instance %d negative pixel offset", i);
    const int64_t npix = (int64_t)s.height * s.width;
    TG_REQUIRE(npix <= ((int64_t)1 << 30), "tg_gauss_fit_plan: instance %d image too large", i);
    n_pix = std::max(n_pix, s.pix_begin + npix);
  }
  // pairwise disjoint pixel and Gaussian ranges
  std::vector<int32_t> idx(n_inst);
  std::iota(idx.begin(), idx.end(), 0);
  std::sort(idx.begin(), idx.end(), [&](int32_t a, int32_t b) { return inst[a].pix_begin < inst[b].pix_begin; });
  for (int32_t k = 1; k < n_inst; ++k) {
    const tg_fit_instance& a = inst[idx[k - 1]];
    TG_REQUIRE(a.pix_begin + (int64_t)a.height * a.width <= inst[idx[k]].pix_begin,
               "tg_gauss_fit_plan: instances %d and %d share pixels", idx[k - 1], idx[k]);
  }
  std::sort(idx.begin(), idx.end(), [&](int32_t a, int32_t b) { return inst[a].gauss_begin < inst[b].gauss_begin; });
  for (int32_t k = 1; k < n_inst; ++k) {
    const tg_fit_instance& a = inst[idx[k - 1]];
    TG_REQUIRE(a.n_gauss == 0 || a.gauss_begin + a.n_gauss <= inst[idx[k]].gauss_begin ||
                   inst[idx[k]].n_gauss == 0,
               "tg_gauss_fit_plan: instances %d and %d share Gaussian rows", idx[k - 1], idx[k]);
  }
  int64_t slice_gauss = kSliceGauss;
  {
    const int64_t key[4] = {n_inst, n_gauss, n_pix, 0};
    int32_t tuned = 0;
    if (tg::tune_lookup(TG_TUNE_FIT_SLICE, key, tuned)) slice_gauss = std::min<int64_t>(std::max(tuned, 16), 4096);
  }
  int32_t shape = TG_TILE_64x64_RUN16;
  {
    const int64_t key[4] = {n_inst, n_gauss, n_pix, 1};
    tg::tune_lookup(TG_TUNE_TILE_SHAPE, key, shape);
  }
  const TileDims tile = tile_dims(shape);
  Tables t;
  t.inst.resize(n_inst);
  t.gauss_inst.assign(n_gauss, -1);
  int64_t ws_elems = 0, n_items = 0, n_blocks = 0, n_assigned = 0;
  for (int32_t i = 0; i < n_inst; ++i) {
    const tg_fit_instance& s = inst[i];
    FitInst& f = t.inst[i];
    f.g0 = s.gauss_begin;
    f.n = s.n_gauss;
    f.pix0 = s.pix_begin;
    f.ws0 = ws_elems;
    for (int k = 0; k < 6; ++k) f.m[k] = s.pix2m[k];
    f.H = s.height;
    f.W = s.width;
    f.ntx = (int32_t)tg::cdiv(s.width, tile.w);
    f.nslice = slices_for(s.n_gauss, slice_gauss);
    const int64_t npix = (int64_t)s.height * s.width;
    ws_elems += (int64_t)f.nslice * npix;
    n_items += tg::cdiv(s.height, tile.h) * f.ntx * f.nslice;
    f.lb0 = (int32_t)n_blocks;
    n_blocks += tg::cdiv(npix, kLossChunk);
    f.lb1 = (int32_t)n_blocks;
    f.pad0 = f.pad1 = 0;
    for (int64_t g = s.gauss_begin; g < s.gauss_begin + s.n_gauss; ++g) t.gauss_inst[g] = i;
    n_assigned += s.n_gauss;
  }
  TG_REQUIRE(n_items < ((int64_t)1 << 31) && n_blocks < ((int64_t)1 << 31), "tg_gauss_fit_plan: too many tiles");
  // forward work items, heaviest first (work ~ Gaussians per slice); ties in
  // instance/tile/slice order
  std::vector<int32_t> by_cost(n_inst);
  std::iota(by_cost.begin(), by_cost.end(), 0);
  std::stable_sort(by_cost.begin(), by_cost.end(), [&](int32_t a, int32_t b) {
    return t.inst[a].n * t.inst[b].nslice > t.inst[b].n * t.inst[a].nslice;
  });
  t.items.reserve(n_items);
  for (int32_t i : by_cost) {
    const FitInst& f = t.inst[i];
    const int32_t ntiles = (int32_t)(tg::cdiv(f.H, tile.h) * f.ntx);
    for (int32_t tile = 0; tile < ntiles; ++tile)
      for (int32_t sl = 0; sl < f.nslice; ++sl) t.items.push_back(FitItem{i, tile, sl, 0});
  }
  t.blocks.reserve(n_blocks);
  for (int32_t i = 0; i < n_inst; ++i)
    for (int32_t ch = 0; ch < t.inst[i].lb1 - t.inst[i].lb0; ++ch) t.blocks.push_back(LossBlock{i, ch});
  // moment wavefronts: Gaussians of larger images first (larger clipped boxes)
  std::vector<int32_t> by_pix(n_inst);
  std::iota(by_pix.begin(), by_pix.end(), 0);
  std::stable_sort(by_pix.begin(), by_pix.end(), [&](int32_t a, int32_t b) {
    return (int64_t)t.inst[a].H * t.inst[a].W > (int64_t)t.inst[b].H * t.inst[b].W;
  });
  t.order.reserve(n_assigned);
  for (int32_t i : by_pix)
    for (int64_t g = t.inst[i].g0; g < t.inst[i].g0 + t.inst[i].n; ++g) t.order.push_back((int32_t)g);

  tg_fit_plan p{};
  p.n_inst = n_inst;
  p.n_items = (int32_t)n_items;
  p.n_loss_blocks = (int32_t)n_blocks;
  p.n_moment_blocks = (int32_t)tg::cdiv(n_assigned, kGaussPerMomentBlock);
  p.n_gauss = n_gauss;
  p.n_assigned = n_assigned;
  p.n_pix = n_pix;
  p.cut = nsigma > 0.0 ? (float)(nsigma * nsigma) : 80.0f;
  p.tile_shape = shape;
  int64_t off = 0;
  auto place = [&](int64_t bytes) {
    const int64_t o = off;
    off = align256(off + bytes);
    return o;
  };
  p.off_inst = place((int64_t)sizeof(FitInst) * n_inst);
  p.off_items = place((int64_t)sizeof(FitItem) * n_items);
  p.off_loss_blocks = place((int64_t)sizeof(LossBlock) * n_blocks);
  p.off_order = place((int64_t)sizeof(int32_t) * n_assigned);
  p.off_gauss_inst = place((int64_t)sizeof(int32_t) * n_gauss);
  p.table_bytes = off;
  p.off_coef = place((int64_t)sizeof(tg_beamlet_coef) * n_gauss);
  p.off_unit = place((int64_t)sizeof(tg_beamlet_coef) * n_gauss);
  p.off_partial = place((int64_t)sizeof(float2) * ws_elems);
  p.off_weight = place((int64_t)sizeof(float2) * (n_pix + kRun));  // + tail padding read by fit_moment_kernel
  p.off_loss_part = place((int64_t)sizeof(double) * n_blocks);
  p.off_moments = place((int64_t)sizeof(double) * kMoments * n_gauss);
  p.workspace_bytes = off;
  *plan = p;
  if (tab) *tab = std::move(t);
  return TG_OK;
}

}  // namespace

extern "C" int tg_gauss_fit_plan(const tg_fit_instance* inst, int32_t n_inst, int64_t n_gauss, double nsigma,
                                 tg_fit_plan* plan) {
  tg::clear_error();
  return build(inst, n_inst, n_gauss, nsigma, plan, nullptr);
}

extern "C" int tg_gauss_fit_upload(const tg_fit_plan* plan, const tg_fit_instance* inst, void* workspace,
                                   void* stream) {
  tg::clear_error();
  TG_REQUIRE(plan != nullptr && workspace != nullptr, "tg_gauss_fit_upload: NULL plan or workspace");
  tg_fit_plan check;
  Tables t;
  const int rc = build(inst, plan->n_inst, plan->n_gauss, 0.0, &check, &t);
  if (rc != TG_OK) return rc;
  TG_REQUIRE(check.table_bytes == plan->table_bytes && check.workspace_bytes == plan->workspace_bytes &&
                 check.n_items == plan->n_items && check.n_pix == plan->n_pix && check.tile_shape == plan->tile_shape,
             "tg_gauss_fit_upload: instances (or the tuned slicing or tile shape) differ from the plan's");
  std::vector<char> host((size_t)plan->table_bytes, 0);
  auto put = [&](int64_t off, const void* src, size_t bytes) {
    if (bytes) memcpy(host.data() + off, src, bytes);
  };
  put(plan->off_inst, t.inst.data(), t.inst.size() * sizeof(FitInst));
  put(plan->off_items, t.items.data(), t.items.size() * sizeof(FitItem));
  put(plan->off_loss_blocks, t.blocks.data(), t.blocks.size() * sizeof(LossBlock));
  put(plan->off_order, t.order.data(), t.order.size() * sizeof(int32_t));
  put(plan->off_gauss_inst, t.gauss_inst.data(), t.gauss_inst.size() * sizeof(int32_t));
  TG_HIP_CHECK(hipMemcpyAsync(workspace, host.data(), host.size(), hipMemcpyHostToDevice, (hipStream_t)stream));
  // the weight buffer's tail padding is read (never used) by fit_moment_kernel
  TG_HIP_CHECK(hipMemsetAsync(static_cast<char*>(workspace) + plan->off_weight + plan->n_pix * (int64_t)sizeof(float2), 0,
                              kRun * sizeof(float2), (hipStream_t)stream));
  TG_HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
  return TG_OK;
}

extern "C" int tg_gauss_fit_run(const tg_fit_plan* plan, void* workspace, const double* params, const float* target,
                                float* field, double* loss, double* grad, void* stream) {
  tg::clear_error();
  TG_REQUIRE(plan != nullptr && workspace != nullptr, "tg_gauss_fit_run: NULL plan or workspace");
  TG_REQUIRE(plan->n_gauss == 0 || params != nullptr, "tg_gauss_fit_run: NULL params");
  TG_REQUIRE(plan->tile_shape >= TG_TILE_64x64_RUN16 && plan->tile_shape <= TG_TILE_64x32_RUN8,
             "tg_gauss_fit_run: plan has an unknown tile shape %d", plan->tile_shape);
  TG_REQUIRE(target != nullptr || (loss == nullptr && grad == nullptr),
             "tg_gauss_fit_run: loss and grad need target intensities");
  TG_REQUIRE(reinterpret_cast<uintptr_t>(workspace) % 256 == 0, "tg_gauss_fit_run: workspace not 256-byte aligned");
  char* ws = static_cast<char*>(workspace);
  const FitInst* insts = reinterpret_cast<const FitInst*>(ws + plan->off_inst);
  const FitItem* items = reinterpret_cast<const FitItem*>(ws + plan->off_items);
  const LossBlock* blocks = reinterpret_cast<const LossBlock*>(ws + plan->off_loss_blocks);
  const int32_t* order = reinterpret_cast<const int32_t*>(ws + plan->off_order);
  const int32_t* ginst = reinterpret_cast<const int32_t*>(ws + plan->off_gauss_inst);
  tg_beamlet_coef* coef = reinterpret_cast<tg_beamlet_coef*>(ws + plan->off_coef);
  tg_beamlet_coef* unit = reinterpret_cast<tg_beamlet_coef*>(ws + plan->off_unit);
  float2* partial = reinterpret_cast<float2*>(ws + plan->off_partial);
  float2* weight = reinterpret_cast<float2*>(ws + plan->off_weight);
  double* loss_part = reinterpret_cast<double*>(ws + plan->off_loss_part);
  double* moments = reinterpret_cast<double*>(ws + plan->off_moments);
  hipStream_t s = (hipStream_t)stream;
  const int64_t G = plan->n_gauss;
  if (G > 0) {
    hipLaunchKernelGGL(fit_coef_kernel, dim3((unsigned)tg::cdiv(G, 256)), dim3(256), 0, s, params, G, ginst, insts,
                       plan->cut, coef, unit);
    TG_LAUNCH_CHECK("fit_coef_kernel launch");
  }
  const dim3 fgrid((unsigned)plan->n_items);
  switch (plan->tile_shape) {
    case TG_TILE_128x32_RUN16:
      hipLaunchKernelGGL((fit_forward_kernel<8, 16>), fgrid, dim3(kThreads), 0, s, coef, insts, items, partial);
      break;
    case TG_TILE_32x128_RUN16:
      hipLaunchKernelGGL((fit_forward_kernel<2, 16>), fgrid, dim3(kThreads), 0, s, coef, insts, items, partial);
      break;
    case TG_TILE_64x32_RUN8:
      hipLaunchKernelGGL((fit_forward_kernel<8, 8>), fgrid, dim3(kThreads), 0, s, coef, insts, items, partial);
      break;
    default:
      hipLaunchKernelGGL((fit_forward_kernel<kLanesPerRow, kRun>), fgrid, dim3(kThreads), 0, s, coef, insts, items,
                         partial);
  }
  TG_LAUNCH_CHECK("fit_forward_kernel launch");
  hipLaunchKernelGGL(fit_loss_kernel, dim3((unsigned)plan->n_loss_blocks), dim3(256), 0, s, insts, blocks, partial,
                     target, (float2*)field, weight, loss_part);
  TG_LAUNCH_CHECK("fit_loss_kernel launch");
  if (target == nullptr) return TG_OK;
  if (loss) {
    hipLaunchKernelGGL(fit_loss_final, dim3((unsigned)tg::cdiv(plan->n_inst, 256)), dim3(256), 0, s, insts,
                       plan->n_inst, loss_part, loss);
    TG_LAUNCH_CHECK("fit_loss_final launch");
  }
  if (grad == nullptr || G == 0) return TG_OK;
  if (plan->n_moment_blocks > 0) {
    hipLaunchKernelGGL(fit_moment_kernel, dim3((unsigned)plan->n_moment_blocks), dim3(256), 0, s, unit, insts, ginst,
                       order, plan->n_assigned, weight, moments);
    TG_LAUNCH_CHECK("fit_moment_kernel launch");
  }
  hipLaunchKernelGGL(fit_grad_kernel, dim3((unsigned)tg::cdiv(G, 256)), dim3(256), 0, s, params, G, ginst, insts,
                     unit, moments, grad);
  TG_LAUNCH_CHECK("fit_grad_kernel launch");
  return TG_OK;
}
