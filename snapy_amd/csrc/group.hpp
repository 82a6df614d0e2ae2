// Grouped launches: the same stage of many blocks of different grids as one
// launch per kernel.
//
// A kernel that can be grouped keeps its body in a struct with a static
// __device__ run(bid, gdim, args...) that reads its workgroup index and grid
// size from bid and gdim instead of blockIdx and gridDim; its __global__
// kernel calls run(blockIdx, gridDim, args...), and its launcher goes
// through launch<Body, BOUNDS, WPE>(kernel, grid, block, lds, stream,
// args...).  Outside a recording that is the plain launch.  While a thread
// records (snapy_record_begin .. snapy_record_end), the launch is not made:
// its arguments and grid are kept in the thread's plan, with the grouped
// twin of the kernel.  Plans of K blocks that made the same launches in the
// same order (the same deck on different grids) become a group
// (snapy_group_build): per launch, the K argument records and the prefix of
// their workgroup counts in device memory.  snapy_group_launch then makes
// each launch once over all K blocks' workgroups: a workgroup finds its
// block by a binary search of the prefix and runs the body with that block's
// arguments and its index within that block's grid.  A block's cells see
// the same arithmetic as in its own launch, so a grouped step is bit-for-bit
// the blocks' own steps.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <tuple>
#include <utility>
#include <vector>

namespace snapy {

// a kernel's own workgroup index and grid size, as a body takes them
__device__ __forceinline__ dim3 block_index() { return dim3(blockIdx.x, blockIdx.y, blockIdx.z); }
__device__ __forceinline__ dim3 grid_size() { return dim3(gridDim.x, gridDim.y, gridDim.z); }

// one recorded launch: its grouped twin, the argument record and the shape
struct RecordedLaunch {
  using Grouped = void (*)(void const* recs, int32_t const* prefix, int nb, unsigned total,
                           dim3 block, size_t lds, hipStream_t s);
  Grouped grouped = nullptr;
  std::vector<unsigned char> rec;  // a GroupRec<A...>
  dim3 grid, block;
  size_t lds = 0;
  // the body masks threads past the ones its arguments ask for, so the
  // blocks of a group may ask for different workgroup sizes along x (the
  // group takes the largest)
  bool flex = false;
};

struct LaunchPlan {
  std::vector<RecordedLaunch> launches;
};

// non-null while this thread records its launches instead of making them
LaunchPlan*& recording_plan();

// the arguments of one block's launch and its grid
template <typename... A>
struct GroupRec {
  std::tuple<A...> args;
  unsigned gx, gy, gz;
};

template <typename Body, typename Tup, size_t... I>
__device__ __forceinline__ void group_call(dim3 const& bid, dim3 const& gdim, Tup const& t,
                                           std::index_sequence<I...>) {
  Body::run(bid, gdim, std::get<I>(t)...);
}

// the grouped twin: workgroup x of the launch belongs to the last block b
// with prefix[b] <= x, at index x - prefix[b] of that block's grid
template <typename Body, int BOUNDS, int WPE, typename... A>
__global__ void __launch_bounds__(BOUNDS) __attribute__((amdgpu_waves_per_eu(WPE)))
    grouped_kernel(GroupRec<A...> const* __restrict__ recs, int32_t const* __restrict__ prefix,
                   int nb) {
  unsigned const x = blockIdx.x;
  int lo = 0, hi = nb - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (unsigned(prefix[mid]) <= x)
      lo = mid;
    else
      hi = mid - 1;
  }
  GroupRec<A...> const& R = recs[lo];
  unsigned l = x - unsigned(prefix[lo]);
  dim3 bid(l % R.gx, (l / R.gx) % R.gy, l / (R.gx * R.gy));
  group_call<Body>(bid, dim3(R.gx, R.gy, R.gz), R.args, std::index_sequence_for<A...>{});
}

template <typename Body, int BOUNDS, int WPE, typename... A>
void grouped_launch(void const* recs, int32_t const* prefix, int nb, unsigned total, dim3 block,
                    size_t lds, hipStream_t s) {
  hipLaunchKernelGGL((grouped_kernel<Body, BOUNDS, WPE, A...>), dim3(total), block, lds, s,
                     static_cast<GroupRec<A...> const*>(recs), prefix, nb);
}

template <typename Body, int BOUNDS, int WPE, typename... A, typename... B>
void launch_as(bool flex, void (*kernel)(A...), dim3 grid, dim3 block, size_t lds, hipStream_t s,
               B&&... args) {
  LaunchPlan* plan = recording_plan();
  if (!plan) {
    hipLaunchKernelGGL(kernel, grid, block, lds, s, static_cast<A>(args)...);
    return;
  }
  GroupRec<A...> r{std::tuple<A...>(static_cast<A>(args)...), grid.x, grid.y, grid.z};
  RecordedLaunch L;
  L.grouped = &grouped_launch<Body, BOUNDS, WPE, A...>;
  L.rec.resize(sizeof(r));
  std::memcpy(L.rec.data(), &r, sizeof(r));
  L.grid = grid;
  L.block = block;
  L.lds = lds;
  L.flex = flex;
  plan->launches.push_back(std::move(L));
}

// the launch of a groupable kernel (see the top of this file); WPE is the
// kernel's amdgpu_waves_per_eu minimum (1 for none)
template <typename Body, int BOUNDS, int WPE, typename... A, typename... B>
void launch(void (*kernel)(A...), dim3 grid, dim3 block, size_t lds, hipStream_t s,
            B&&... args) {
  launch_as<Body, BOUNDS, WPE>(false, kernel, grid, block, lds, s, static_cast<B&&>(args)...);
}

// the launch of a kernel whose body masks the threads past the ones its
// arguments ask for (RecordedLaunch::flex)
template <typename Body, int BOUNDS, int WPE, typename... A, typename... B>
void launch_flex(void (*kernel)(A...), dim3 grid, dim3 block, size_t lds, hipStream_t s,
                 B&&... args) {
  launch_as<Body, BOUNDS, WPE>(true, kernel, grid, block, lds, s, static_cast<B&&>(args)...);
}

}  // namespace snapy
