// Ray branch kernel: runs a host-compiled per-ray register program over a
// vectorised ray bundle held as SoA arrays, optionally followed (in the same
// pass) by detector binning.
//
// Replaces the eager element-wise stream of run_iter / run_to_end
// (reference run.py:47-116 with propagator.py:51-72 and components.py) and the
// serial numba scatter Grid.into_image -> inplace_sum (grid.py:231-280,
// utils.py:83-114).
//
// Layout / mapping:
//   * one lane owns R rays (stride 256 within a 256*R ray chunk, so each
//     load/store instruction of a wavefront touches 64 consecutive elements);
//   * the program is a kernel argument (kernarg segment -> scalar loads), so
//     opcode dispatch is a uniform scalar branch and never diverges;
//   * registers are named VGPRs selected through uniform switches (no scratch);
//   * binning of small detectors (<= 32768 bins of int counts) goes through
//     an LDS-privatised histogram of packed 16-bit counters (two bins per
//     32-bit word; a workgroup bins at most 65535 rays, so a counter never
//     carries into its neighbour), flushed once per workgroup with
//     device-scope atomics; larger detectors and float accumulators use
//     device-scope atomics directly.
// Arithmetic is exactly one IEEE operation per instruction, built with
// -ffp-contract=off, so results are bit-identical to the same operation
// sequence evaluated elementwise on the CPU.
#include <algorithm>

#include "tg_common.hpp"

namespace {

constexpr int kThreads = 256;
// the program and binning spec travel by value in the kernel-argument segment
static_assert(sizeof(tg_program) + sizeof(tg_bin_spec) <= 4096, "ray program exceeds the kernel-argument limit");
constexpr int64_t kLdsBins = 32768;
constexpr int64_t kMaxRaysPerGroup = 65535;
// input slots whose loads are issued together (one memory round trip);
// strided slots carry 64-bit offsets per ray, so they go two at a time
template <bool CONTIG>
constexpr int kSlotGroup = CONTIG ? 4 : 2;

template <typename T>
__device__ __forceinline__ T imm_of(const tg_instr& in);
template <>
__device__ __forceinline__ float imm_of<float>(const tg_instr& in) { return in.immf; }
template <>
__device__ __forceinline__ double imm_of<double>(const tg_instr& in) { return in.imm; }

// The register file of each ray is one 16-wide vector value.  Indexing it
// with a uniform (scalar) register number lowers to the VGPR-indexing mode
// (s_set_gpr_idx_on / v_movrel*): no scratch memory and no selection trees.
// The host only dispatches programs whose registers are all below 16 here.
template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
  typedef float type __attribute__((ext_vector_type(16)));
};
template <>
struct Vec16<double> {
  typedef double type __attribute__((ext_vector_type(16)));
};

template <typename T, int NREG>
struct RegFile;
template <typename T>
struct RegFile<T, 16> {
  typename Vec16<T>::type v;
  __device__ __forceinline__ void clear() { v = (typename Vec16<T>::type)(T(0)); }
  __device__ __forceinline__ T get(int i) const { return v[i]; }
  __device__ __forceinline__ void set(int i, T x) { v[i] = x; }
};
// Wide programs (more than 16 live values, e.g. forward-mode tangents of
// nonlinear elements) keep their TG_MAX_REG registers in a per-lane array in
// scratch memory, indexed in bounds by masking.  (A VGPR file wider than 16
// is not safe to build from vector parts: the compiler if-converts the part
// selection and performs the indexed write of every part, which for an index
// of another part lands on unrelated VGPRs.)  Registers are not cleared:
// programs only use values they have loaded or computed, and the unused
// operand of a unary instruction is read but never used.
template <typename T>
struct RegFile<T, TG_MAX_REG> {
  T v[TG_MAX_REG];
  __device__ __forceinline__ void clear() {}
  __device__ __forceinline__ T get(int i) const { return v[i & (TG_MAX_REG - 1)]; }
  __device__ __forceinline__ void set(int i, T x) { v[i & (TG_MAX_REG - 1)] = x; }
};
static_assert((TG_MAX_REG & (TG_MAX_REG - 1)) == 0, "TG_MAX_REG must be a power of two");

template <typename T>
__device__ __forceinline__ T sign_of(T a) {
  if (a != a || a == T(0)) return a;
  return a > T(0) ? T(1) : T(-1);
}

__device__ __forceinline__ int64_t slot_offset(int64_t lin, const int64_t (&shape)[TG_MAX_DIM],
                                               const int64_t (&stride)[TG_MAX_DIM]) {
  const int64_t i3 = lin % shape[3];
  lin /= shape[3];
  const int64_t i2 = lin % shape[2];
  lin /= shape[2];
  const int64_t i1 = lin % shape[1];
  const int64_t i0 = lin / shape[1];
  return i0 * stride[0] + i1 * stride[1] + i2 * stride[2] + i3 * stride[3];
}

// Pixel index of one position; -1 when it falls outside the detector.
__device__ __forceinline__ int64_t bin_index(const tg_bin_spec& B, double y, double x) {
  // Same operation order as apply_transformation's T @ [y, x, 1]
  // (reference coordinate_transforms.py:104-140), in f64.
  const double fy = __dadd_rn(__dadd_rn(__dmul_rn(B.m[0], y), __dmul_rn(B.m[1], x)), B.m[2]);
  const double fx = __dadd_rn(__dadd_rn(__dmul_rn(B.m[3], y), __dmul_rn(B.m[4], x)), B.m[5]);
  // jnp.round is round-half-to-even; rint uses the default RNE mode.
  const int32_t iy = tg::xla_f64_to_i32(rint(fy));
  const int32_t ix = tg::xla_f64_to_i32(rint(fx));
  if (iy < 0 || iy >= B.height || ix < 0 || ix >= B.width) return -1;
  return (int64_t)iy * B.width + ix;
}

template <typename T>
__device__ __forceinline__ void global_deposit(const tg_bin_spec& B, int64_t p, T w) {
  switch (B.acc_dtype) {
    case 0:
      atomicAdd((int32_t*)B.acc + p, 1);
      break;
    case 1:
      atomicAdd((unsigned long long*)B.acc + p, 1ull);
      break;
    case 2:
      unsafeAtomicAdd((float*)B.acc + p, (float)w);
      break;
    default:
      unsafeAtomicAdd((double*)B.acc + p, (double)w);
      break;
  }
}

// BIN: 0 = no binning, 1 = LDS-privatised integer histogram, 2 = global atomics.
// NT: workgroup size (LDS-histogram launches use 1024 so that 16 wavefronts
// share one private histogram and fewer flushes hit the same global bins).
template <typename T, int R, bool CONTIG, int BIN, int NT, int NREG, bool MIXED>
__global__ __launch_bounds__(NT) void ray_program_kernel(const tg_program P, const tg_bin_spec B) {
  constexpr int kThreads = NT;
  extern __shared__ int32_t lds_hist[];
  const int64_t n = P.n;
  int64_t hw = 0;
  if constexpr (BIN == 1) {
    hw = B.height * B.width;
    for (int64_t i = threadIdx.x; i < (hw + 1) / 2; i += kThreads) lds_hist[i] = 0;
    __syncthreads();
  }
  const int64_t chunk = (int64_t)kThreads * R;
  for (int64_t base = (int64_t)blockIdx.x * chunk; base < n; base += (int64_t)gridDim.x * chunk) {
    RegFile<T, NREG> r[R];
#pragma unroll
    for (int u = 0; u < R; ++u) r[u].clear();
    int64_t idx[R];
    bool ok[R];
#pragma unroll
    for (int u = 0; u < R; ++u) {
      idx[u] = base + (int64_t)u * kThreads + threadIdx.x;
      ok[u] = idx[u] < n;
    }
    // Inputs, kSlotGroup slots at a time: every load of a group is issued
    // before the first is consumed, so a group costs one memory round trip.
    // The loads are unpredicated (indices of rays past n are clamped to ray
    // n - 1, whose values are replaced by 0 on entering the register file;
    // a group's unused tail re-reads the last slot): predicated loads made
    // the compiler wait for each slot's loads before issuing the next.
    for (int s0 = 0; s0 < P.n_slot; s0 += kSlotGroup<CONTIG>) {
      T v[kSlotGroup<CONTIG>][R];
#pragma unroll
      for (int g = 0; g < kSlotGroup<CONTIG>; ++g) {
        const int s = min(s0 + g, P.n_slot - 1);
#pragma unroll
        for (int u = 0; u < R; ++u) {
          const int64_t i = ok[u] ? idx[u] : n - 1;
          int64_t off = i;
          if constexpr (!CONTIG) off = slot_offset(i, P.shape, P.slot[s].stride);
          if constexpr (MIXED) {
            // an f64 program with f32 slots: the per-slot width test is a
            // branch, and its loads are waited for one at a time
            if (P.slot[s].dtype == TG_F32)
              v[g][u] = (T)((const float*)P.slot[s].ptr)[off];
            else
              v[g][u] = ((const T*)P.slot[s].ptr)[off];
          } else {
            v[g][u] = ((const T*)P.slot[s].ptr)[off];
          }
        }
      }
      // unconditional writes (a tail entry writes the last slot's value to
      // its register again), so no load is sunk into a conditional block
#pragma unroll
      for (int g = 0; g < kSlotGroup<CONTIG>; ++g) {
        const int dst = P.slot[min(s0 + g, P.n_slot - 1)].reg;
#pragma unroll
        for (int u = 0; u < R; ++u) r[u].set(dst, ok[u] ? v[g][u] : T(0));
      }
    }
    for (int i = 0; i < P.n_instr; ++i) {
      const tg_instr in = P.instr[i];
      const int ia = in.a, ib = in.b, id = in.dst;
      T a[R], b[R], d[R];
#pragma unroll
      for (int u = 0; u < R; ++u) {
        a[u] = r[u].get(ia);
        b[u] = r[u].get(ib);
      }
      const bool f32op = sizeof(T) == 8 && (in.op & TG_OP_F32);
      const T im = f32op ? (T)in.immf : imm_of<T>(in);
      switch (in.op & 0x7f) {
        case TG_OP_ADD:
#pragma unroll
          for (int u = 0; u < R; ++u) d[u] = a[u] + b[u];
          break;
        case TG_OP_SUB:
#pragma unroll
          for (int u = 0; u < R; ++u) d[u] = a[u] - b[u];
          break;
        case TG_OP_MUL:
#pragma unroll
          for (int u = 0; u < R; ++u) d[u] = a[u] * b[u];
          break;
        case TG_OP_DIV:
#pragma unroll
          for (int u = 0; u < R; ++u) d[u] = a[u] / b[u];
          break;
        case TG_OP_ADDI:
#pragma unroll
          for (int u = 0; u < R; ++u) d[u] = a[u] + im;
          break;
        case TG_OP_SUBI:
#pragma unroll
          for (int u = 0; u < R; ++u) d[u] = a[u] - im;
          break;
        case TG_OP_RSUBI:
#pragma unroll
          for (int u = 0; u < R; ++u) d[u] = im - a[u];
          break;
        case TG_OP_MULI:
#pragma unroll
          for (int u = 0; u < R; ++u) d[u] = a[u] * im;
          break;
        case TG_OP_DIVI:
#pragma unroll
          for (int u = 0; u < R; ++u) d[u] = a[u] / im;
          break;
        case TG_OP_RDIVI:
#pragma unroll
          for (int u = 0; u < R; ++u) d[u] = im / a[u];
          break;
        case TG_OP_NEG:
#pragma unroll
          for (int u = 0; u < R; ++u) d[u] = -a[u];
          break;
        case TG_OP_MOV:
#pragma unroll
          for (int u = 0; u < R; ++u) d[u] = a[u];
          break;
        case TG_OP_SET:
#pragma unroll
          for (int u = 0; u < R; ++u) d[u] = im;
          break;
        default:  // TG_OP_SIGN (opcodes are validated on the host)
#pragma unroll
          for (int u = 0; u < R; ++u) d[u] = sign_of<T>(a[u]);
          break;
      }
      if (f32op) {
#pragma unroll
        for (int u = 0; u < R; ++u) d[u] = (T)(float)d[u];
      }
#pragma unroll
      for (int u = 0; u < R; ++u) r[u].set(id, d[u]);
    }
    for (int k = 0; k < P.n_out; ++k) {
      T v[R];
      const int src = P.out_reg[k];
#pragma unroll
      for (int u = 0; u < R; ++u) v[u] = r[u].get(src);
      if (sizeof(T) == 8 && P.out_dtype[k] == TG_F32) {
        float* __restrict__ o = (float*)P.out[k];
#pragma unroll
        for (int u = 0; u < R; ++u)
          if (ok[u]) o[idx[u]] = (float)v[u];
      } else {
        T* __restrict__ o = (T*)P.out[k];
#pragma unroll
        for (int u = 0; u < R; ++u)
          if (ok[u]) o[idx[u]] = v[u];
      }
    }
    if constexpr (BIN != 0) {
      T xv[R], yv[R], wv[R];
      if (B.x_reg >= 0) {
        const int src = B.x_reg;
#pragma unroll
        for (int u = 0; u < R; ++u) xv[u] = r[u].get(src);
      } else {
#pragma unroll
        for (int u = 0; u < R; ++u) xv[u] = (T)B.x_uniform;
      }
      if (B.y_reg >= 0) {
        const int src = B.y_reg;
#pragma unroll
        for (int u = 0; u < R; ++u) yv[u] = r[u].get(src);
      } else {
#pragma unroll
        for (int u = 0; u < R; ++u) yv[u] = (T)B.y_uniform;
      }
      if (B.w_reg >= 0) {
        const int src = B.w_reg;
#pragma unroll
        for (int u = 0; u < R; ++u) wv[u] = r[u].get(src);
      } else {
#pragma unroll
        for (int u = 0; u < R; ++u) wv[u] = T(1);
      }
#pragma unroll
      for (int u = 0; u < R; ++u) {
        if (!ok[u]) continue;
        const int64_t p = bin_index(B, (double)yv[u], (double)xv[u]);
        if (p < 0) continue;
        if constexpr (BIN == 1) {
          atomicAdd(&lds_hist[p >> 1], 1 << (16 * (int)(p & 1)));
        } else {
          global_deposit<T>(B, p, wv[u]);
        }
      }
    }
  }
  if constexpr (BIN == 1) {
    __syncthreads();
    for (int64_t i = threadIdx.x; i < (hw + 1) / 2; i += kThreads) {
      const uint32_t w = (uint32_t)lds_hist[i];
      if (w == 0) continue;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int32_t c = (int32_t)((w >> (16 * h)) & 0xffffu);
        const int64_t bin = 2 * i + h;
        if (c == 0 || bin >= hw) continue;
        if (B.acc_dtype == 0)
          atomicAdd((int32_t*)B.acc + bin, c);
        else
          atomicAdd((unsigned long long*)B.acc + bin, (unsigned long long)(int64_t)c);
      }
    }
  }
}

template <typename T, int R, bool CONTIG, int BIN, int NREG, bool MIXED>
int launch(const tg_program& P, const tg_bin_spec& B, hipStream_t stream) {
  constexpr int kThreads = (BIN == 1) ? 1024 : 256;
  const int64_t chunk = (int64_t)kThreads * R;
  const int64_t nchunks = tg::cdiv(P.n, chunk);
  const int cus = tg::cu_count();
  // LDS histograms are zeroed and flushed once per workgroup (and every flush
  // adds into the same global bins), so that grid stays at one 1024-thread
  // workgroup per CU; streaming variants get more waves.  A caller (the
  // autotuner) may override.
  int64_t cap = (BIN == 1) ? 1LL * cus : 16LL * cus;
  int32_t tuned = 0;
  if (B.grid_hint > 0) {
    cap = B.grid_hint;
  } else if (BIN != 0) {
    const int64_t key[4] = {P.n, B.height * B.width, P.dtype, 0};
    if (tg::tune_lookup(TG_TUNE_RAY_GRID, key, tuned) && tuned > 0) cap = tuned;
  }
  int64_t grid = nchunks < cap ? nchunks : cap;
  if (BIN == 1) {
    // packed 16-bit counters: at most kMaxRaysPerGroup rays per workgroup
    const int64_t min_grid = tg::cdiv(P.n, (kMaxRaysPerGroup / chunk) * chunk);
    if (grid < min_grid) grid = min_grid;
  }
  const size_t lds = (BIN == 1) ? (size_t)((B.height * B.width + 1) / 2) * sizeof(int32_t) : 0;
  hipLaunchKernelGGL((ray_program_kernel<T, R, CONTIG, BIN, kThreads, NREG, MIXED>), dim3((unsigned)grid), dim3(kThreads),
                     lds, stream, P, B);
  TG_LAUNCH_CHECK("ray_program_kernel launch");
  return TG_OK;
}

template <typename T, int R, int NREG, bool MIXED>
int dispatch_width(const tg_program& P, const tg_bin_spec& B, bool contig, int bin, hipStream_t s) {
  // wide programs bin with global atomics (one kernel fewer per file width)
  if (NREG > 16 && bin == 1) bin = 2;
  if (contig) {
    if (bin == 0) return launch<T, R, true, 0, NREG, MIXED>(P, B, s);
    if constexpr (NREG == 16)
      if (bin == 1) return launch<T, R, true, 1, NREG, MIXED>(P, B, s);
    return launch<T, R, true, 2, NREG, MIXED>(P, B, s);
  }
  if (bin == 0) return launch<T, R, false, 0, NREG, MIXED>(P, B, s);
  if constexpr (NREG == 16)
    if (bin == 1) return launch<T, R, false, 1, NREG, MIXED>(P, B, s);
  return launch<T, R, false, 2, NREG, MIXED>(P, B, s);
}

// f64 programs whose slots are all f64 load them without a per-slot width
// test; only programs that also read f32 slots take the mixed variant
template <typename T, int R, int NREG>
int dispatch(const tg_program& P, const tg_bin_spec& B, bool contig, int bin, hipStream_t s) {
  if constexpr (sizeof(T) == 8) {
    for (int k = 0; k < P.n_slot; ++k)
      if (P.slot[k].dtype == TG_F32) return dispatch_width<T, R, NREG, true>(P, B, contig, bin, s);
  }
  return dispatch_width<T, R, NREG, false>(P, B, contig, bin, s);
}

// Highest register the program (and its binning) touches.
int max_register(const tg_program& P, const tg_bin_spec* B) {
  int m = 0;
  for (int i = 0; i < P.n_instr; ++i) {
    const tg_instr& in = P.instr[i];
    m = std::max(m, (int)std::max(in.dst, std::max(in.a, in.b)));
  }
  for (int k = 0; k < P.n_slot; ++k) m = std::max(m, (int)P.slot[k].reg);
  for (int k = 0; k < P.n_out; ++k) m = std::max(m, (int)P.out_reg[k]);
  if (B != nullptr) m = std::max(m, (int)std::max(B->x_reg, std::max(B->y_reg, B->w_reg)));
  return m;
}

}  // namespace

extern "C" int tg_ray_program_run(const tg_program* prog, const tg_bin_spec* bin, void* stream) {
  tg::clear_error();
  TG_REQUIRE(prog != nullptr, "tg_ray_program_run: prog is NULL");
  const tg_program& P = *prog;
  TG_REQUIRE(P.dtype == TG_F32 || P.dtype == TG_F64, "tg_ray_program_run: bad dtype %d", P.dtype);
  TG_REQUIRE(P.n >= 0, "tg_ray_program_run: negative ray count %lld", (long long)P.n);
  TG_REQUIRE(P.n_instr >= 0 && P.n_instr <= TG_MAX_INSTR, "tg_ray_program_run: n_instr %d out of range",
             P.n_instr);
  TG_REQUIRE(P.n_slot >= 0 && P.n_slot <= TG_MAX_SLOT, "tg_ray_program_run: n_slot %d out of range", P.n_slot);
  TG_REQUIRE(P.n_out >= 0 && P.n_out <= TG_MAX_SLOT, "tg_ray_program_run: n_out %d out of range", P.n_out);
  int64_t prod = 1;
  for (int k = 0; k < TG_MAX_DIM; ++k) {
    TG_REQUIRE(P.shape[k] >= 0, "tg_ray_program_run: negative shape");
    prod *= P.shape[k];
  }
  TG_REQUIRE(prod == P.n, "tg_ray_program_run: prod(shape)=%lld != n=%lld", (long long)prod, (long long)P.n);
  for (int i = 0; i < P.n_instr; ++i) {
    const tg_instr& in = P.instr[i];
    TG_REQUIRE((in.op & 0x7f) < TG_OP_COUNT, "tg_ray_program_run: instr %d has bad opcode %d", i, (int)in.op);
    TG_REQUIRE(P.dtype == TG_F64 || !(in.op & TG_OP_F32), "tg_ray_program_run: f32 flag in an f32 program");
    TG_REQUIRE(in.dst < TG_MAX_REG && in.a < TG_MAX_REG && in.b < TG_MAX_REG,
               "tg_ray_program_run: instr %d register out of range", i);
  }
  bool contig = true;
  for (int s = 0; s < P.n_slot; ++s) {
    TG_REQUIRE(P.slot[s].reg >= 0 && P.slot[s].reg < TG_MAX_REG, "tg_ray_program_run: slot %d bad reg", s);
    TG_REQUIRE(P.n == 0 || P.slot[s].ptr != nullptr, "tg_ray_program_run: slot %d has NULL pointer", s);
    TG_REQUIRE(P.slot[s].dtype == P.dtype || (P.dtype == TG_F64 && P.slot[s].dtype == TG_F32),
               "tg_ray_program_run: slot %d dtype %d incompatible with program dtype", s, P.slot[s].dtype);
    contig = contig && (P.slot[s].contig != 0);
  }
  for (int k = 0; k < P.n_out; ++k) {
    TG_REQUIRE(P.out_reg[k] >= 0 && P.out_reg[k] < TG_MAX_REG, "tg_ray_program_run: out %d bad reg", k);
    TG_REQUIRE(P.n == 0 || P.out[k] != nullptr, "tg_ray_program_run: out %d has NULL pointer", k);
    TG_REQUIRE(P.out_dtype[k] == P.dtype || (P.dtype == TG_F64 && P.out_dtype[k] == TG_F32),
               "tg_ray_program_run: out %d dtype %d incompatible with program dtype", k, P.out_dtype[k]);
  }
  int mode = 0;
  tg_bin_spec B{};
  if (bin != nullptr && bin->enabled) {
    B = *bin;
    TG_REQUIRE(B.acc != nullptr, "tg_ray_program_run: bin accumulator is NULL");
    TG_REQUIRE(B.height > 0 && B.width > 0, "tg_ray_program_run: bad detector shape");
    TG_REQUIRE(B.acc_dtype >= 0 && B.acc_dtype <= 3, "tg_ray_program_run: bad accumulator dtype");
    TG_REQUIRE(B.x_reg >= -1 && B.x_reg < TG_MAX_REG && B.y_reg >= -1 && B.y_reg < TG_MAX_REG &&
                   B.w_reg >= -1 && B.w_reg < TG_MAX_REG,
               "tg_ray_program_run: bin register out of range");
    TG_REQUIRE(B.acc_dtype >= 2 || B.w_reg < 0, "tg_ray_program_run: weights need a float accumulator");
    TG_REQUIRE(B.grid_hint >= 0, "tg_ray_program_run: negative grid_hint");
    const bool small = B.height * B.width <= kLdsBins;
    mode = (B.acc_dtype <= 1 && small) ? 1 : 2;
  }
  if (P.n == 0) return TG_OK;
  hipStream_t s = (hipStream_t)stream;
  if (max_register(P, mode != 0 ? &B : nullptr) < 16) {
    if (P.dtype == TG_F32) return dispatch<float, 4, 16>(P, B, contig, mode, s);
    return dispatch<double, 2, 16>(P, B, contig, mode, s);
  }
  if (P.dtype == TG_F32) return dispatch<float, 1, TG_MAX_REG>(P, B, contig, mode, s);
  return dispatch<double, 1, TG_MAX_REG>(P, B, contig, mode, s);
}
