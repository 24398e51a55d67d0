// Persistent batch kernel: the cells of several mechanisms integrated by one
// launch, each wavefront taking work items from device queues until all are
// empty.
//
// Replaces the reference's one-launch-per-mechanism batched driver
// (src/mkpp_host/include/mkpp_host/dispatcher.hpp:258-362, driven per name by
// mechanism_manager.hpp:118-209).  A work item is one cell for a wave-per-cell
// mechanism and 64*m cells (m per lane) for a lane-per-cell one, so cells of
// very different cost share the machine without a tail per mechanism.
//
// Queues.  There is one queue per XCD (workgroup b runs on XCD b % 8).  Queue q
// owns the q-th eighth of every segment's items, so each XCD works through
// contiguous cells (the cache-line sharing of the plain wave kernel) and sees
// the same mix of mechanisms.  A wavefront drains its home queue first and
// then steals from the others in order; it leaves once all eight are empty,
// which every wavefront reaches after at most one failed take per queue.
// Queue heads are 32-bit counters that batch_reset_kernel zeroes on the
// launch stream (no memset node, so the pair can be captured in a graph).
#pragma once
#include <hip/hip_runtime.h>

#include "runtime.hpp"
#include "wave_engine.hpp"

namespace mkpp {

constexpr int BATCH_MAX_SEGMENTS = 8;
constexpr int BATCH_QUEUES = 8;

struct BatchSegment {
  DeviceArgs args;
  int engine;  // case label of the unit's BatchEngines::run
  int lanes;   // 0: one cell per wavefront; m > 0: 64 * m cells per item, m per lane
  long items;
};

struct BatchArgs {
  BatchSegment seg[BATCH_MAX_SEGMENTS];
  int nseg;
  unsigned int* heads;  // BATCH_QUEUES queue heads
};

template <int Q>
__global__ void batch_reset_kernel(unsigned int* heads) {
  if (threadIdx.x < Q) heads[threadIdx.x] = 0u;
}

// First item of segment items owned by queue q (q = BATCH_QUEUES gives items).
__device__ __forceinline__ long batch_chunk(long items, int q) { return items * q / BATCH_QUEUES; }

template <class M, class Tab>
__device__ __forceinline__ void batch_lane_item(const DeviceArgs& a, long item, int m) {
  for (int j = 0; j < m; ++j) {
    int lane = threadIdx.x;  // opaque: keeps lane-dependent values inside the item
    asm volatile("" : "+v"(lane));
    const long cell = (item * m + j) * 64 + lane;
    if (cell < a.ncells) integrate_lane_cell<M, Tab>(a, cell);
  }
}

template <class Engines, class Tab>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2))) batch_kernel(BatchArgs b) {
  extern __shared__ double lds[];
  const int home = (int)(blockIdx.x % BATCH_QUEUES);
  for (int v = 0; v < BATCH_QUEUES; ++v) {
    const int q = (home + v) % BATCH_QUEUES;
    long qlen = 0;
    for (int s = 0; s < b.nseg; ++s) qlen += batch_chunk(b.seg[s].items, q + 1) - batch_chunk(b.seg[s].items, q);
    for (;;) {
      unsigned int t = 0;
      if (threadIdx.x == 0) t = atomicAdd(&b.heads[q], 1u);
      t = __builtin_amdgcn_readfirstlane(t);
      if ((long)t >= qlen) break;
      long r = (long)t;
      int s = 0;
      for (; s < b.nseg - 1; ++s) {
        const long n = batch_chunk(b.seg[s].items, q + 1) - batch_chunk(b.seg[s].items, q);
        if (r < n) break;
        r -= n;
      }
      const long item = batch_chunk(b.seg[s].items, q) + r;
      Engines::template run<Tab>(b.seg[s].engine, b.seg[s].args, item, b.seg[s].lanes, lds);
    }
  }
}

// Host side, implemented per method by the generated batch units.
struct BatchLaunch {
  int (*launch)(const BatchArgs& b, size_t lds_bytes, int workgroups, hipStream_t s);
  int (*resident_per_cu)(size_t lds_bytes);  // workgroups of this kernel one CU holds
};

}  // namespace mkpp

// Defines mkpp::batch_launch_<TAB>() for one tableau in a unit that has
// included its gen/batch_engines.hpp.
#define MKPP_BATCH_METHOD(TAB)                                                                            \
  namespace mkpp {                                                                                       \
  static int batch_launch_impl_##TAB(const BatchArgs& b, size_t lds_bytes, int workgroups, hipStream_t s) { \
    batch_reset_kernel<BATCH_QUEUES><<<1, 64, 0, s>>>(b.heads);                                          \
    batch_kernel<BatchEngines, TAB><<<workgroups, 64, lds_bytes, s>>>(b);                                \
    return hipGetLastError() == hipSuccess ? 0 : -2;                                                     \
  }                                                                                                      \
  static int batch_resident_impl_##TAB(size_t lds_bytes) {                                               \
    int n = 0;                                                                                           \
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch_kernel<BatchEngines, TAB>, 64, lds_bytes) \
        != hipSuccess)                                                                                   \
      return 0;                                                                                          \
    return n;                                                                                            \
  }                                                                                                      \
  const BatchLaunch* batch_launch_##TAB() {                                                              \
    static const BatchLaunch l = {&batch_launch_impl_##TAB, &batch_resident_impl_##TAB};                 \
    return &l;                                                                                           \
  }                                                                                                      \
  }
