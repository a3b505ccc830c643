// Host-side support shared by every entry point of libtemgym_hip.so.
#include "tg_common.hpp"

#include <map>
#include <mutex>
#include <string.h>
#include <tuple>

namespace {
thread_local char g_err[1024] = {0};
constexpr int kMaxDevices = 64;
int g_cu[kMaxDevices] = {0};
std::mutex g_cu_mu;

// tuned launch choices: (knob, device, key[4]) -> value
using TuneKey = std::tuple<int32_t, int, int64_t, int64_t, int64_t, int64_t>;
std::map<TuneKey, int32_t> g_tune;
std::mutex g_tune_mu;

int current_device() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  return dev;
}
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

bool tune_lookup(int32_t knob, const int64_t (&key)[4], int32_t& value) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  const auto it = g_tune.find(TuneKey{knob, current_device(), key[0], key[1], key[2], key[3]});
  if (it == g_tune.end()) return false;
  value = it->second;
  return true;
}

}  // namespace tg

extern "C" {

const char* tg_last_error(void) { return g_err; }

int32_t tg_abi_version(void) { return 7; }

int32_t tg_device_cu_count(void) { return tg::cu_count(); }

int tg_tune_set(int32_t knob, const int64_t* key, int32_t value) {
  tg::clear_error();
  TG_REQUIRE(knob >= TG_TUNE_BEAMLET_SPLITS && knob <= TG_TUNE_TILE_SHAPE, "tg_tune_set: unknown knob %d", knob);
  TG_REQUIRE(key != nullptr, "tg_tune_set: NULL key");
  TG_REQUIRE(value > 0, "tg_tune_set: value must be positive");
  TG_REQUIRE(knob != TG_TUNE_TILE_SHAPE || value <= TG_TILE_64x32_RUN8, "tg_tune_set: unknown tile shape %d", value);
  TG_REQUIRE(knob != TG_TUNE_BEAMLET_PATH || value <= TG_BEAMLET_PATH_VALU, "tg_tune_set: unknown tile-sum path %d",
             value);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune[TuneKey{knob, current_device(), key[0], key[1], key[2], key[3]}] = value;
  return TG_OK;
}

int tg_tune_get(int32_t knob, const int64_t* key, int32_t* value) {
  tg::clear_error();
  TG_REQUIRE(key != nullptr && value != nullptr, "tg_tune_get: NULL argument");
  const int64_t k[4] = {key[0], key[1], key[2], key[3]};
  return tg::tune_lookup(knob, k, *value) ? 1 : 0;
}

int tg_tune_clear(void) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune.clear();
  return TG_OK;
}

}  // extern "C"
