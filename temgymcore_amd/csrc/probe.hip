// Device bandwidth probe: measures the achievable HBM read rate of the
// current device with the access pattern the streaming kernels use
// (lane-strided 16-byte non-temporal loads, 1 KiB per wavefront instruction).
#include "tg_common.hpp"

namespace {

constexpr int kThreads = 256;
constexpr int kUnroll = 4;  // with ~64 workgroups per CU the fastest of a sweep of read kernels
typedef float f4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(kThreads) void hbm_read_probe_kernel(const f4* __restrict__ src, int64_t n4,
                                                                  float* __restrict__ block_sums) {
  __shared__ float red[kThreads];
  f4 acc = {0.f, 0.f, 0.f, 0.f};
  const int64_t stride = (int64_t)gridDim.x * kThreads * kUnroll;
  for (int64_t base = (int64_t)blockIdx.x * kThreads * kUnroll + threadIdx.x; base < n4; base += stride) {
    f4 v[kUnroll];
#pragma unroll
    for (int k = 0; k < kUnroll; ++k) {
      const int64_t i = base + (int64_t)k * kThreads;
      v[k] = i < n4 ? __builtin_nontemporal_load(src + i) : f4{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int k = 0; k < kUnroll; ++k) acc += v[k];
  }
  red[threadIdx.x] = acc.x + acc.y + acc.z + acc.w;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) block_sums[blockIdx.x] = red[0];
}

}  // namespace

extern "C" int tg_hbm_read_probe(const void* src, int64_t bytes, float* block_sums, int32_t n_blocks, void* stream) {
  tg::clear_error();
  TG_REQUIRE(src != nullptr && block_sums != nullptr, "tg_hbm_read_probe: NULL buffer");
  TG_REQUIRE(bytes >= 0 && (bytes % 16) == 0 && ((uintptr_t)src & 15) == 0,
             "tg_hbm_read_probe: bytes must be a multiple of 16 and src 16-byte aligned");
  TG_REQUIRE(n_blocks > 0, "tg_hbm_read_probe: n_blocks must be positive");
  hipLaunchKernelGGL(hbm_read_probe_kernel, dim3(n_blocks), dim3(kThreads), 0, (hipStream_t)stream,
                     (const f4*)src, bytes / 16, block_sums);
  TG_LAUNCH_CHECK("hbm_read_probe_kernel launch");
  return TG_OK;
}
