// Shared host/device helpers for libtemgym_hip.so: error reporting across the
// C ABI, device attribute caching and small device utilities.
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>

#include "../../include/temgym_hip.h"

namespace tg {

// Thread-local last-error buffer behind tg_last_error().
void set_error(const char* fmt, ...);
void clear_error();

// Compute-unit count of the current device (cached per device id).
int cu_count();

// Convert a hipError_t into TG_EHIP with a message naming `what`.
int hip_fail(hipError_t e, const char* what);

// Tuned launch choice recorded by tg_tune_set for (knob, current device, key).
bool tune_lookup(int32_t knob, const int64_t (&key)[4], int32_t& value);

}  // namespace tg

#define TG_HIP_CHECK(expr)                                   \
  do {                                                       \
    hipError_t _e = (expr);                                  \
    if (_e != hipSuccess) return tg::hip_fail(_e, #expr);    \
  } while (0)

#define TG_REQUIRE(cond, ...)                                \
  do {                                                       \
    if (!(cond)) {                                           \
      tg::set_error(__VA_ARGS__);                            \
      return TG_EINVAL;                                      \
    }                                                        \
  } while (0)

// After a kernel launch: report launch failures (bad configuration etc).
#define TG_LAUNCH_CHECK(name)                                \
  do {                                                       \
    hipError_t _e = hipGetLastError();                       \
    if (_e != hipSuccess) return tg::hip_fail(_e, name);     \
  } while (0)

namespace tg {

// Integer ceil-div for launch geometry.
static inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// The reference casts rounded pixel coordinates with jnp.astype(int32); XLA's
// float->int conversion maps NaN to 0 and saturates out-of-range values.
// Branch-free: clamp into the int32 range (fmax/fmin return the non-NaN
// operand), truncate, then select 0 for NaN.
__device__ __forceinline__ int32_t xla_f64_to_i32(double v) {
  const int32_t t = (int32_t)fmin(fmax(v, -2147483648.0), 2147483647.0);
  return v != v ? 0 : t;
}

}  // namespace tg
