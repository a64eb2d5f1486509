// Thread-local error state behind dpk_last_error(), and the launch tunables
// (process-wide, read by the launchers at each call).
#include <atomic>
#include <cstdarg>
#include <cstdio>

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

int tune(int key) { return g_tune[key].load(std::memory_order_relaxed); }

}  // namespace dpk

extern "C" const char* dpk_last_error(void) { return dpk::g_err; }

extern "C" int dpk_abi_version(void) { return 7; }

extern "C" int dpk_tune_set(int key, int value) {
  dpk::clear_error();
  if (key < 0 || key >= DPK_TUNE_COUNT) {
    dpk::set_error("dpk_tune_set: unknown key %d", key);
    return DPK_ERR_ARG;
  }
  const int lo[DPK_TUNE_COUNT] = {0, 10, 0, 0, 0, 0, 0},
            hi[DPK_TUNE_COUNT] = {1, 400, 1, 1600, 16, 16, 1};
  if (value < lo[key] || value > hi[key]) {
    dpk::set_error("dpk_tune_set: value %d for key %d outside [%d, %d]", value, key, lo[key],
                   hi[key]);
    return DPK_ERR_ARG;
  }
  dpk::g_tune[key].store(value, std::memory_order_relaxed);
  return DPK_OK;
}

extern "C" int dpk_tune_get(int key) {
  return (key >= 0 && key < DPK_TUNE_COUNT) ? dpk::tune(key) : -1;
}
