// Launch-shape tuning: the workgroup size a kernel family uses for blocks of
// one shape and dtype (the columns per workgroup of the column kernels; for
// the work list, the persistent grid), set by
// the autotuner (snapy_amd/autotune.py) and read by the launchers.  A tuned
// shape changes only how the items are split into workgroups, never the
// arithmetic of an item, so the results are the same bits for every value.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <tuple>

#include "common.hpp"

namespace snapy {
namespace {

using Key = std::tuple<int, int, int, int, int>;  // tunable, nx1, nx2, nx3, dtype
std::mutex g_mu;
std::map<Key, unsigned> g_table;

Key key_of(int t, snapy_grid const* g, int dtype) {
  if (!g) return Key{t, 0, 0, 0, 0};
  return Key{t, g->nx1, g->nx2, g->nx3, dtype};
}

bool allowed(int t, int v) {
  if (t == SNAPY_TUNE_WL_GRID) return v >= 64 && v <= 65536;
  if (t == SNAPY_TUNE_VIC_COLS) return v == 1 || v == 2 || v == 4 || v == 8;
  if (t == SNAPY_TUNE_REF_COLS) return v == 1 || v == 2 || v == 4;
  return v == 64 || v == 128 || v == 256;
}

}  // namespace

unsigned tuned(int t, snapy_grid const* g, int dtype, unsigned dflt) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_table.find(key_of(t, g, dtype));
  return it == g_table.end() ? dflt : it->second;
}

}  // namespace snapy

using namespace snapy;

extern "C" {

int snapy_tune_set(int tunable, const snapy_grid* g, int dtype, int value) {
  if (tunable < 0 || tunable >= SNAPY_TUNE_COUNT) return fail(-1, "tune_set: unknown tunable");
  if (tunable != SNAPY_TUNE_WL_GRID && !g) return fail(-1, "tune_set: NULL grid");
  std::lock_guard<std::mutex> lk(g_mu);
  Key k = key_of(tunable, tunable == SNAPY_TUNE_WL_GRID ? nullptr : g, dtype);
  if (value == 0) {
    g_table.erase(k);
    return 0;
  }
  if (!allowed(tunable, value)) return fail(-1, "tune_set: value outside the tunable's range");
  g_table[k] = unsigned(value);
  return 0;
}

int snapy_tune_get(int tunable, const snapy_grid* g, int dtype) {
  if (tunable < 0 || tunable >= SNAPY_TUNE_COUNT) return fail(-1, "tune_get: unknown tunable");
  unsigned dflt = tunable == SNAPY_TUNE_WL_GRID    ? 2048u
                  : tunable == SNAPY_TUNE_VIC_COLS ? 8u
                  : tunable == SNAPY_TUNE_REF_COLS ? 4u
                                                   : 256u;
  return int(tuned(tunable, tunable == SNAPY_TUNE_WL_GRID ? nullptr : g, dtype, dflt));
}

}  // extern "C"
