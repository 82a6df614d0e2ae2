// Recording plans and building and launching groups (group.hpp).
//
// The reference steps the blocks of a process on worker threads, one CUDA
// stream each (src/mesh/mesh.cpp:74-110); here the blocks that share a deck
// share each launch.
#include <hip/hip_runtime.h>

#include <memory>
#include <string>

#include "common.hpp"
#include "group.hpp"

namespace snapy {

LaunchPlan*& recording_plan() {
  static thread_local LaunchPlan* plan = nullptr;
  return plan;
}

namespace {

struct GroupLaunch {
  RecordedLaunch::Grouped grouped;
  dim3 block;
  size_t lds;
  unsigned total;
  size_t recs_off, prefix_off;  // byte offsets into the device buffer
};

struct Group {
  int nb = 0;
  void* dev = nullptr;  // the argument records and prefixes of every launch
  std::vector<GroupLaunch> launches;
  ~Group() {
    if (dev) (void)hipFree(dev);
  }
};

}  // namespace
}  // namespace snapy

using namespace snapy;

extern "C" {

int snapy_record_begin(void) {
  if (recording_plan()) return fail(-20, "record_begin: this thread is already recording");
  recording_plan() = new LaunchPlan();
  return 0;
}

int snapy_record_end(void** plan) {
  if (!plan) return fail(-1, "record_end: NULL argument");
  LaunchPlan* p = recording_plan();
  if (!p) return fail(-20, "record_end: this thread is not recording");
  recording_plan() = nullptr;
  *plan = p;
  return 0;
}

int snapy_plan_launches(const void* plan) {
  return plan ? int(static_cast<LaunchPlan const*>(plan)->launches.size()) : -1;
}

void snapy_plan_free(void* plan) { delete static_cast<LaunchPlan*>(plan); }

int snapy_group_build(void* const* plans, int nplans, void** group) {
  if (!plans || !group || nplans < 1) return fail(-1, "group_build: no plans");
  auto P = [&](int b) { return static_cast<LaunchPlan const*>(plans[b]); };
  for (int b = 0; b < nplans; ++b)
    if (!P(b)) return fail(-1, "group_build: NULL plan");
  size_t nl = P(0)->launches.size();
  for (int b = 1; b < nplans; ++b) {
    if (P(b)->launches.size() != nl)
      return fail(-21, "group_build: the plans make different numbers of launches");
    for (size_t i = 0; i < nl; ++i) {
      RecordedLaunch const &x = P(0)->launches[i], &y = P(b)->launches[i];
      if (x.grouped != y.grouped || x.rec.size() != y.rec.size() || x.flex != y.flex ||
          (x.block.x != y.block.x && !x.flex) || x.block.y != y.block.y ||
          x.block.z != y.block.z)
        return fail(-21, "group_build: launch " + std::to_string(i) +
                             " differs between the plans (a different kernel or workgroup)");
    }
  }
  auto g = std::make_unique<Group>();
  g->nb = nplans;
  // host image: per launch, the nb records (16-byte aligned) then the nb + 1
  // prefix counts
  std::vector<unsigned char> img;
  auto align = [&](size_t a) { img.resize((img.size() + a - 1) / a * a); };
  for (size_t i = 0; i < nl; ++i) {
    GroupLaunch gl;
    RecordedLaunch const& first = P(0)->launches[i];
    gl.grouped = first.grouped;
    gl.block = first.block;
    gl.lds = 0;
    align(16);
    gl.recs_off = img.size();
    for (int b = 0; b < nplans; ++b) {
      RecordedLaunch const& L = P(b)->launches[i];
      img.insert(img.end(), L.rec.begin(), L.rec.end());
      if (L.lds > gl.lds) gl.lds = L.lds;
      if (L.block.x > gl.block.x) gl.block.x = L.block.x;  // (flex launches only)
    }
    align(16);
    gl.prefix_off = img.size();
    int64_t total = 0;
    for (int b = 0; b <= nplans; ++b) {
      int32_t v = int32_t(total);
      img.insert(img.end(), reinterpret_cast<unsigned char*>(&v),
                 reinterpret_cast<unsigned char*>(&v) + sizeof(v));
      if (b < nplans) {
        RecordedLaunch const& L = P(b)->launches[i];
        total += int64_t(L.grid.x) * L.grid.y * L.grid.z;
      }
    }
    if (total > (int64_t(1) << 31) - 1) return fail(-9, "group_build: too many workgroups");
    gl.total = unsigned(total);
    g->launches.push_back(gl);
  }
  if (!img.empty()) {
    if (hipMalloc(&g->dev, img.size()) != hipSuccess)
      return fail(-22, "group_build: no device memory for the records");
    if (hipMemcpy(g->dev, img.data(), img.size(), hipMemcpyHostToDevice) != hipSuccess)
      return fail(-22, "group_build: copy of the records failed");
  }
  *group = g.release();
  return 0;
}

int snapy_group_launch(const void* group, void* stream) {
  if (!group) return fail(-1, "group_launch: NULL group");
  auto g = static_cast<Group const*>(group);
  auto base = static_cast<unsigned char const*>(g->dev);
  for (GroupLaunch const& gl : g->launches) {
    if (gl.total == 0) continue;
    gl.grouped(base + gl.recs_off, reinterpret_cast<int32_t const*>(base + gl.prefix_off), g->nb,
               gl.total, gl.block, gl.lds, as_stream(stream));
    if (int e = launch_status("group_launch")) return e;
  }
  return 0;
}

int snapy_group_launches(const void* group) {
  return group ? int(static_cast<Group const*>(group)->launches.size()) : -1;
}

void snapy_group_free(void* group) { delete static_cast<Group*>(group); }

}  // extern "C"
