// Grouped launches: one kernel over several independent problems (the members of a group of
// constraint-path runs, each with its own lattice, walkers and buffers) in a single launch.
//
// A groupable kernel is a functor F with a device body F::run(const BlockIdx3&, const Args&), its
// threads per workgroup F::NT and waves per EU F::WPE.  glaunch<F>(args, grid, block, lds, s)
// launches it on its own (plain_kernel) or, while a GroupCollector is open on the thread, records
// it.  The collector gives each record a position in its member's sequence of launches; flushing
// launches, position by position, one group_kernel per kernel and block shape over the records of
// all members at that position.  A member's launches therefore keep their order, and the launches
// of different members (which touch disjoint buffers) share workgroups.  Launches made inside the
// branches of a GroupFork (work on independent buffers of one member, such as its two spin blocks)
// restart from the fork's position, so the branches share launches too.
//
// The per-problem arguments of a grouped launch sit in a device table; the workgroup of global
// index g finds its problem p by a binary search over the start offsets passed by value, and its
// local block index g - start[p].  Problems start at multiples of 8 workgroups, so a workgroup's
// XCD (global index mod 8) equals its local index mod 8 and the kernels' XCD-aware tile orders hold.
// Tables are kept by content for the life of the group (GroupTables), so the launches of a graph
// captured from grouped calls keep valid tables.
#pragma once
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "common.hpp"

constexpr int GROUP_MAX = 64;  // problems per grouped launch (more are split over launches)

struct BlockIdx3 {
  int x, y, z;
  int gx, gy, gz;
};

struct GroupIndex {
  int n;
  int start[GROUP_MAX + 1];
};

template <class Args>
struct GroupEntry {
  Args a;
  int gx, gy, gz, pad;
};

// F::WPE is the minimum number of waves per SIMD the register allocation must allow.  A functor
// that also defines WPE_MAX bounds it from above (waves_per_eu(WPE, WPE_MAX)), which lets the
// compiler use the registers of fewer waves.
template <class F, class = void>
struct has_wpe_max : std::false_type {};
template <class F>
struct has_wpe_max<F, std::void_t<decltype(F::WPE_MAX)>> : std::true_type {};

template <class F, class Args>
__global__ __launch_bounds__(F::NT) __attribute__((amdgpu_waves_per_eu(F::WPE))) void plain_kernel(Args a) {
  const BlockIdx3 b{(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z, (int)gridDim.x, (int)gridDim.y, (int)gridDim.z};
  F::run(b, a);
}

template <class F, class Args>
__global__ __launch_bounds__(F::NT) __attribute__((amdgpu_waves_per_eu(F::WPE))) void group_kernel(
    const GroupEntry<Args>* __restrict__ tab, GroupIndex gi) {
  const int g = (int)blockIdx.x;
  int lo = 0, hi = gi.n;  // start[lo] <= g < start[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (gi.start[mid] <= g) lo = mid;
    else hi = mid;
  }
  // the problem's arguments, read before the body stores anything
  const Args a = tab[lo].a;
  const int gx = tab[lo].gx, gy = tab[lo].gy, gz = tab[lo].gz;
  const int local = g - gi.start[lo];
  if (local >= gx * gy * gz) return;  // padding up to the next multiple of 8
  const BlockIdx3 b{local % gx, (local / gx) % gy, local / (gx * gy), gx, gy, gz};
  F::run(b, a);
}

template <class F, class Args>
__global__ __launch_bounds__(F::NT) __attribute__((amdgpu_waves_per_eu(F::WPE, F::WPE_MAX))) void plain_kernel_b(
    Args a) {
  const BlockIdx3 b{(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z, (int)gridDim.x, (int)gridDim.y, (int)gridDim.z};
  F::run(b, a);
}

template <class F, class Args>
__global__ __launch_bounds__(F::NT) __attribute__((amdgpu_waves_per_eu(F::WPE, F::WPE_MAX))) void group_kernel_b(
    const GroupEntry<Args>* __restrict__ tab, GroupIndex gi) {
  const int g = (int)blockIdx.x;
  int lo = 0, hi = gi.n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (gi.start[mid] <= g) lo = mid;
    else hi = mid;
  }
  const Args a = tab[lo].a;
  const int gx = tab[lo].gx, gy = tab[lo].gy, gz = tab[lo].gz;
  const int local = g - gi.start[lo];
  if (local >= gx * gy * gz) return;
  const BlockIdx3 b{local % gx, (local / gx) % gy, local / (gx * gy), gx, gy, gz};
  F::run(b, a);
}

// ------------------------------------------------------------------ host side
struct GroupRecord;
class GroupTables;
typedef int (*GroupFlushFn)(const GroupRecord* const* recs, int n, hipStream_t s, GroupTables* t);

struct GroupRecord {
  const void* key;  // the plain kernel: records with equal keys and block shapes share a launch
  GroupFlushFn flush;
  std::vector<unsigned char> args;
  unsigned gx, gy, gz;
  dim3 block;
  size_t lds;
  int member, pos, seq;
};

// Device copies of argument tables, by content, for the life of the owner; and the streams and
// events over which a position's launches of different kernels run side by side.
class GroupTables {
 public:
  static constexpr int FORKS = 3;  // side streams for the second and later kernels of a position
  ~GroupTables();
  // a device copy of the n bytes at p (the same pointer for the same bytes); nullptr on failure
  const void* get(const void* p, size_t n);
  size_t count() const { return map_.size(); }
  // side stream i (default priority, non-blocking) and event i (FORKS + 1 of them); created on
  // first use, false on failure
  bool fork_resources(hipStream_t* streams, hipEvent_t* events);

 private:
  std::unordered_map<std::string, void*> map_;
  hipStream_t side_ = nullptr;
  hipStream_t fork_[FORKS] = {};
  hipEvent_t ev_[FORKS + 1] = {};
};

struct GroupCollector {
  GroupTables* tables = nullptr;
  int member = 0, pos = 0;
  std::vector<int> fork_start, fork_max;
  std::vector<GroupRecord> recs;
};
extern thread_local GroupCollector* g_group;
inline bool grouping() { return g_group != nullptr; }

// the collector's records, launched position by position on s; clears them.  The launches of one
// position touch disjoint buffers (different members, or the independent branches of a GroupFork):
// the second and later kernels of a position run on side streams forked from s and joined back
// before the next position, so that kernels of different shapes (the small-dense classes of
// lattices of different sizes) share the GPU.
int group_flush_all(GroupCollector& c, hipStream_t s);

// Branches over independent buffers of one member: each branch restarts at the fork's position,
// and the fork ends after its longest branch.  A no-op outside a collector.
struct GroupFork {
  GroupFork() {
    if (g_group) {
      g_group->fork_start.push_back(g_group->pos);
      g_group->fork_max.push_back(g_group->pos);
    }
  }
  void branch() {
    if (!g_group) return;
    int& m = g_group->fork_max.back();
    if (g_group->pos > m) m = g_group->pos;
    g_group->pos = g_group->fork_start.back();
  }
  ~GroupFork() {
    if (!g_group) return;
    const int m = g_group->fork_max.back();
    if (g_group->pos < m) g_group->pos = m;
    g_group->fork_start.pop_back();
    g_group->fork_max.pop_back();
  }
};

template <class F, class Args>
int group_flush_kernel(const GroupRecord* const* r, int n, hipStream_t s, GroupTables* t) {
  for (int c0 = 0; c0 < n; c0 += GROUP_MAX) {
    const int cn = n - c0 < GROUP_MAX ? n - c0 : GROUP_MAX;
    std::vector<GroupEntry<Args>> tab(cn);
    memset(tab.data(), 0, sizeof(GroupEntry<Args>) * cn);
    GroupIndex gi;
    memset(&gi, 0, sizeof(gi));
    gi.n = cn;
    long long start = 0;
    size_t lds = 0;
    for (int i = 0; i < cn; ++i) {
      const GroupRecord& q = *r[c0 + i];
      memcpy(&tab[i].a, q.args.data(), sizeof(Args));
      tab[i].gx = (int)q.gx;
      tab[i].gy = (int)q.gy;
      tab[i].gz = (int)q.gz;
      gi.start[i] = (int)start;
      start += ((long long)q.gx * q.gy * q.gz + 7) / 8 * 8;
      if (q.lds > lds) lds = q.lds;
    }
    REQUIRE(start < (1ll << 31), "grouped launch of %lld workgroups", start);
    gi.start[cn] = (int)start;
    if (start == 0) continue;
    const void* d = t->get(tab.data(), sizeof(GroupEntry<Args>) * cn);
    REQUIRE(d, "grouped launch: no device memory for its argument table");
    if constexpr (has_wpe_max<F>::value)
      hipLaunchKernelGGL((group_kernel_b<F, Args>), dim3((unsigned)start), r[c0]->block, lds, s,
                         reinterpret_cast<const GroupEntry<Args>*>(d), gi);
    else
      hipLaunchKernelGGL((group_kernel<F, Args>), dim3((unsigned)start), r[c0]->block, lds, s,
                         reinterpret_cast<const GroupEntry<Args>*>(d), gi);
    hipError_t e = hipGetLastError();
    REQUIRE(e == hipSuccess, "grouped launch -> %s", hipGetErrorString(e));
  }
  return 0;
}

// Launch F over grid (or record it while a collector is open).
template <class F, class Args>
int glaunch(const Args& a, dim3 grid, dim3 block, size_t lds, hipStream_t s) {
  if ((long long)grid.x * grid.y * grid.z == 0) return 0;
  if (g_group) {
    GroupRecord q;
    if constexpr (has_wpe_max<F>::value) q.key = reinterpret_cast<const void*>(&plain_kernel_b<F, Args>);
    else q.key = reinterpret_cast<const void*>(&plain_kernel<F, Args>);
    q.flush = &group_flush_kernel<F, Args>;
    q.args.resize(sizeof(Args));
    memcpy(q.args.data(), &a, sizeof(Args));
    q.gx = grid.x; q.gy = grid.y; q.gz = grid.z;
    q.block = block;
    q.lds = lds;
    q.member = g_group->member;
    q.pos = g_group->pos++;
    q.seq = (int)g_group->recs.size();
    g_group->recs.push_back(std::move(q));
    return 0;
  }
  if constexpr (has_wpe_max<F>::value) hipLaunchKernelGGL((plain_kernel_b<F, Args>), grid, block, lds, s, a);
  else hipLaunchKernelGGL((plain_kernel<F, Args>), grid, block, lds, s, a);
  hipError_t e = hipGetLastError();
  REQUIRE(e == hipSuccess, "launch -> %s", hipGetErrorString(e));
  return 0;
}
