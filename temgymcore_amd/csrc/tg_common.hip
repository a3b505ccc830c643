// Host-side support shared by every entry point of libtemgym_hip.so.
#include "tg_common.hpp"

#include <mutex>
#include <string.h>

namespace {
thread_local char g_err[1024] = {0};
constexpr int kMaxDevices = 64;
int g_cu[kMaxDevices] = {0};
std::mutex g_cu_mu;
}  // namespace

namespace tg {

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

void clear_error() { g_err[0] = 0; }

int hip_fail(hipError_t e, const char* what) {
  set_error("%s failed: %s (%d)", what, hipGetErrorString(e), (int)e);
  return TG_EHIP;
}

int cu_count() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return 256;
  std::lock_guard<std::mutex> lk(g_cu_mu);
  if (g_cu[dev] == 0) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0)
      v = 256;
    g_cu[dev] = v;
  }
  return g_cu[dev];
}

}  // namespace tg

extern "C" {

const char* tg_last_error(void) { return g_err; }

int32_t tg_abi_version(void) { return 3; }

int32_t tg_device_cu_count(void) { return tg::cu_count(); }

}  // extern "C"
