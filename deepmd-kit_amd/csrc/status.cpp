// Thread-local error state behind dpk_last_error(), and the launch tunables:
// process-wide defaults, and tuning sets (one per model) that a thread makes
// current around its launches; the launchers read the calling thread's
// current set, else the defaults.
#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <new>

#include "dpk_common.h"

namespace dpk {

static thread_local char g_err[512] = {0};

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

void clear_error() { g_err[0] = 0; }

static std::atomic<int> g_tune[DPK_TUNE_COUNT] = {
    {0},    // DPK_TUNE_NLIST_MODE: 0 staged cell neighbourhoods, 1 per-centre gather
    {125},  // DPK_TUNE_NLIST_STAGE_PCT: staging capacity, % of the mean neighbourhood
    {0},    // DPK_TUNE_FORCE_MODE: 0 cell-blocked, 1 per-atom
    {100},  // DPK_TUNE_EMBED_GRID_PCT: persistent BF16 embedding grid, % of resident slots
    {0},    // DPK_TUNE_NLIST_WAVES: search wavefronts per cell workgroup (0 = from the density)
    {0},    // DPK_TUNE_FORCE_WAVES: cell-blocked force wavefronts per workgroup (0 = auto)
    {0},    // DPK_TUNE_NLIST_SORT: 0 packed keys + full-key fallback, 1 full keys only
};

static constexpr int kTuneLo[DPK_TUNE_COUNT] = {0, 10, 0, 0, 0, 0, 0};
static constexpr int kTuneHi[DPK_TUNE_COUNT] = {1, 400, 1, 1600, 16, 16, 1};

static thread_local const dpk_tuning* t_current = nullptr;

int tune(int key) {
  const dpk_tuning* t = t_current;
  return t ? t->v[key] : g_tune[key].load(std::memory_order_relaxed);
}

static int check_tune(const char* who, int key, int value) {
  if (key < 0 || key >= DPK_TUNE_COUNT) {
    set_error("%s: unknown key %d", who, key);
    return DPK_ERR_ARG;
  }
  if (value < kTuneLo[key] || value > kTuneHi[key]) {
    set_error("%s: value %d for key %d outside [%d, %d]", who, value, key, kTuneLo[key],
              kTuneHi[key]);
    return DPK_ERR_ARG;
  }
  return DPK_OK;
}

}  // namespace dpk

extern "C" const char* dpk_last_error(void) { return dpk::g_err; }

extern "C" int dpk_abi_version(void) { return 8; }

extern "C" int dpk_tune_set(int key, int value) {
  dpk::clear_error();
  const int rc = dpk::check_tune("dpk_tune_set", key, value);
  if (rc != DPK_OK) return rc;
  dpk::g_tune[key].store(value, std::memory_order_relaxed);
  return DPK_OK;
}

extern "C" int dpk_tune_get(int key) {
  return (key >= 0 && key < DPK_TUNE_COUNT) ? dpk::g_tune[key].load(std::memory_order_relaxed)
                                            : -1;
}

extern "C" dpk_tuning* dpk_tuning_create(void) {
  dpk::clear_error();
  dpk_tuning* t = new (std::nothrow) dpk_tuning;
  if (!t) {
    dpk::set_error("dpk_tuning_create: out of host memory");
    return nullptr;
  }
  for (int k = 0; k < DPK_TUNE_COUNT; ++k) t->v[k] = dpk::g_tune[k].load(std::memory_order_relaxed);
  return t;
}

extern "C" void dpk_tuning_destroy(dpk_tuning* t) {
  if (t && dpk::t_current == t) dpk::t_current = nullptr;
  delete t;
}

extern "C" int dpk_tuning_set(dpk_tuning* t, int key, int value) {
  dpk::clear_error();
  if (!t) {
    dpk::set_error("dpk_tuning_set: null tuning set");
    return DPK_ERR_ARG;
  }
  const int rc = dpk::check_tune("dpk_tuning_set", key, value);
  if (rc != DPK_OK) return rc;
  t->v[key] = value;
  return DPK_OK;
}

extern "C" int dpk_tuning_get(const dpk_tuning* t, int key) {
  if (key < 0 || key >= DPK_TUNE_COUNT) return -1;
  return t ? t->v[key] : dpk::g_tune[key].load(std::memory_order_relaxed);
}

extern "C" const dpk_tuning* dpk_tuning_use(const dpk_tuning* t) {
  const dpk_tuning* prev = dpk::t_current;
  dpk::t_current = t;
  return prev;
}
