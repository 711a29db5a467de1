// batch.hip — many independent instances solved by one graph launch.
//
// The reference solves one instance per SmpEngineExecutor call
// (benchmark/smp_engine_exec.h:6-32) and has no batched path.  Here one hipGraph
// (captured once, replayed) solves all instances, in one of two modes:
//   FUSED  one persistent kernel; workgroups take instances from a global
//          counter, largest first, and each solves its instance alone: it
//          inverts the women's rows (rank[w][pref_w[w][r]] = r, the
//          F9RankMatrixInitKernel step, src/smp/kernel.cu:672-686) into an
//          int16 rank matrix, rows staged in LDS, then its waves run proposal
//          chains (the La3F9CoreKernel walk, src/smp/kernel.cu:633-670) with the
//          women's partner words in LDS and 64-bit LDS atomicMin, then it writes
//          the matching (PostProc, src/smp/smp_engine_bamboo.cu:299-303);
//   POOL   two kernels: the rank matrices of all instances as one HBM-bound
//          pass (women's rows: k_rank_rows16, or men's rows: k_rankT_tiles),
//          then a persistent kernel whose workgroups each keep `slots`
//          instances open in LDS and run several chains per wave
//          (wg_pool_chains), taking instances largest first.
// The mode and its shape (block size, workgroups per CU, lanes per chain,
// slots, rank layout) are chosen by a measured sweep at creation unless the
// caller fixes them.

#include <algorithm>
#include <numeric>

#include "engine.hpp"
#include "wg_solver.hpp"

namespace pactae {

struct BatchDesc {
  long long off;   // element offset of the instance's n x n lists
  long long moff;  // offset of its result in the concatenated matching
  int n;
  int idx;         // the instance's position in the caller's order
};

// FUSED mode.  Each workgroup takes instances largest first and solves each
// alone with wg_solve_instance, so at any moment some workgroups stream lists
// (the HBM-bound rank rows) while others walk chains (latency-bound): the two
// overlap across the GPU instead of running as two passes one after the other.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK, 8) void k_batch_solve(const BatchDesc* __restrict__ desc, int count,
                                                          const int32_t* __restrict__ pm_all,
                                                          const int32_t* __restrict__ pw_all,
                                                          int16_t* __restrict__ rank_all,
                                                          int32_t* __restrict__ match_all, int* next_inst,
                                                          unsigned long long* proposals, int* error, int max_n) {
  constexpr int NW = BLOCK / 64;
  extern __shared__ unsigned long long smem64[];
  unsigned long long* partner = smem64;
  int32_t* linv = reinterpret_cast<int32_t*>(partner + max_n);
  uint16_t* stage = reinterpret_cast<uint16_t*>(linv + max_n);
  __shared__ int ctl[4];  // cursor, differs, reach (wg_solve_instance), instance
  __shared__ unsigned long long red[NW];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int bad = 0;
  for (;;) {
    __syncthreads();  // the previous instance is done with the LDS
    if (threadIdx.x == 0) ctl[3] = atomicAdd(next_inst, 1);
    __syncthreads();
    // Wave-uniform from here on: the instance's scalars and pointers live in
    // scalar registers.
    const int inst = __builtin_amdgcn_readfirstlane(ctl[3]);
    if (inst >= count) break;
    const BatchDesc d = desc[inst];
    unsigned long long p = 0;
    bad |= wg_solve_instance<BLOCK>(pm_all + d.off, pw_all + d.off, rank_all + d.off, match_all + d.moff, d.n,
                                    partner, linv, stage, ctl, &p);
    for (int off = 32; off > 0; off >>= 1) p += __shfl_down(p, off);
    if (lane == 0) red[wave] = p;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long t = 0;
      for (int i = 0; i < NW; ++i) t += red[i];
      atomicAdd(proposals, t);
    }
  }
  if (bad) atomicOr(error, bad);
}

// Pool-mode chains: at most 64 VGPRs (8 waves per SIMD), so two 1024-thread
// workgroups share a CU.
template <int BLOCK, int G, int NS, bool TR>
__global__ __launch_bounds__(BLOCK, 8) void k_batch_pool(const BatchDesc* __restrict__ desc, int count,
                                                         const int32_t* __restrict__ pm_all,
                                                         const int16_t* __restrict__ rank_all,
                                                         int32_t* __restrict__ match_all, int* next_inst,
                                                         unsigned long long* proposals, int* error, int max_n,
                                                         const int32_t* __restrict__ pw_all,
                                                         const int* __restrict__ master_flags,
                                                         const int32_t* __restrict__ linv_all) {
  extern __shared__ unsigned long long partner[];
  __shared__ PoolSlot slots[NS];
  __shared__ int qdone;
  // Windows of up to 128 positions (8 for G <= 16): a longer window was used
  // only by walks that refuse hundreds of positions in a row, and each extra
  // window slot costs every step its unrolled instructions.
  constexpr int kJmax = G >= 64 ? 2 : G == 32 ? 4 : 8;
  const int bad = wg_pool_chains<BLOCK, G, kJmax, NS, TR>(desc, count, next_inst, pm_all, rank_all, match_all,
                                                       proposals, max_n, partner, slots, &qdone, pw_all,
                                                       master_flags, linv_all);
  if (bad) atomicOr(error, bad);
}

// Master-list detection (does every man of an instance hold the same list?),
// in three passes so that no instance is compared by one workgroup alone:
//   k_master_probe   one wave per instance compares the first 64 entries of
//                    rows 1..4 with row 0; most instances differ there
//                    (differs[i] = 1) and are done;
//   k_master_rows    the remaining candidates' rows, one wave per row over the
//                    flat row space of the batch, 512 entries a step, each row
//                    stopping at its first difference or when its instance is
//                    already known to differ;
//   k_master_linv    for every instance still equal (n >= 2): linv[moff + L[k]] = k
//                    with L = row 0 checked to be a permutation.
// differs[] is indexed by the instance's position in the caller's order.
__global__ __launch_bounds__(256) void k_master_probe(const RankTask* __restrict__ tasks, int count,
                                                      const int32_t* __restrict__ pw_all,
                                                      const int32_t* __restrict__ pm_all, int* differs) {
  const int lane = threadIdx.x & 63;
  const int waves = (int)(gridDim.x * (blockDim.x >> 6));
  for (int i = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)); i < count; i += waves) {
    const int n = tasks[i].n;
    const int32_t* pm = pm_all + (tasks[i].pw - pw_all);
    bool diff = n < 2;
    for (int m = 1; m < min(n, 5) && !diff; ++m)
      diff = __ballot(lane < n && pm[(size_t)m * n + lane] != pm[lane]) != 0;
    if (lane == 0) differs[i] = diff ? 1 : 0;
  }
}

__global__ __launch_bounds__(256) void k_master_rows(const RankTask* __restrict__ tasks,
                                                     const int* __restrict__ row_task, int total_rows,
                                                     const int32_t* __restrict__ pw_all,
                                                     const int32_t* __restrict__ pm_all, int* differs) {
  const int lane = threadIdx.x & 63;
  const int waves = (int)(gridDim.x * (blockDim.x >> 6));
  for (int g = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)); g < total_rows; g += waves) {
    const int i = row_task[g];
    if (__hip_atomic_load(&differs[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) continue;
    const int n = tasks[i].n, m = g - tasks[i].row0;
    if (m == 0) continue;
    const int32_t* pm = pm_all + (tasks[i].pw - pw_all);
    const int32_t* row = pm + (size_t)m * n;
    for (int k0 = 0; k0 < n; k0 += 512) {
      bool diff = false;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int k = k0 + 64 * j + lane;
        if (k < n && row[k] != pm[k]) diff = true;
      }
      if (__ballot(diff)) {
        if (lane == 0) atomicOr(&differs[i], 1);
        break;
      }
      if (__hip_atomic_load(&differs[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
    }
  }
}

__global__ __launch_bounds__(256) void k_master_linv(const RankTask* __restrict__ tasks, int count,
                                                     const long long* __restrict__ moffs,
                                                     const int32_t* __restrict__ pw_all,
                                                     const int32_t* __restrict__ pm_all, const int* differs,
                                                     int32_t* linv_all, int* error) {
  int bad = 0;
  for (int i = blockIdx.x; i < count; i += gridDim.x) {
    if (differs[i]) continue;
    const int n = tasks[i].n;
    const int32_t* L = pm_all + (tasks[i].pw - pw_all);
    int32_t* linv = linv_all + moffs[i];
    for (int k = threadIdx.x; k < n; k += blockDim.x) linv[k] = -1;
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
      const int w = L[k];
      if ((unsigned)w >= (unsigned)n) bad |= kErrEntryRange;
      else if (atomicCAS(&linv[w], -1, k) != -1) bad |= kErrNotPerm;
    }
    __syncthreads();
  }
  if (bad) atomicOr(error, bad);
}

__global__ void k_batch_reset(int* next_inst, unsigned long long* proposals, int* error) {
  if (threadIdx.x == 0) {
    *next_inst = 0;
    *proposals = 0;
    *error = 0;
  }
}

template <int BLOCK>
static hipError_t launch_fused(int grid, size_t lds, hipStream_t s, const BatchDesc* desc, int count,
                               const int32_t* pm, const int32_t* pw, int16_t* rank, int32_t* match,
                               int* next_inst, unsigned long long* props, int* err, int max_n) {
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_batch_solve<BLOCK>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_batch_solve<BLOCK>, dim3(grid), dim3(BLOCK), lds, s, desc, count, pm, pw, rank,
                     match, next_inst, props, err, max_n);
  return hipGetLastError();
}

struct PoolArgs {
  int grid, max_n, count;
  hipStream_t s;
  const BatchDesc* desc;
  const int32_t* pm;
  const int16_t* rank;
  int32_t* match;
  int* next_inst;
  unsigned long long* props;
  int* err;
  const int32_t* pw;
  const int* master;  // per-instance master-list flags, or null
  const int32_t* linv;
};

template <int BLOCK, int G, int NS, bool TR>
static hipError_t launch_pool4(const PoolArgs& a) {
  const size_t lds = (size_t)NS * a.max_n * 8;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_batch_pool<BLOCK, G, NS, TR>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_batch_pool<BLOCK, G, NS, TR>), dim3(a.grid), dim3(BLOCK), lds, a.s, a.desc, a.count,
                     a.pm, a.rank, a.match, a.next_inst, a.props, a.err, a.max_n, a.pw, a.master, a.linv);
  return hipGetLastError();
}

template <int BLOCK, int G, bool TR>
static hipError_t launch_pool3(int ns, const PoolArgs& a) {
  switch (ns) {
    case 1: return launch_pool4<BLOCK, G, 1, TR>(a);
    case 2: return launch_pool4<BLOCK, G, 2, TR>(a);
    case 4: return launch_pool4<BLOCK, G, 4, TR>(a);
    default: return hipErrorInvalidValue;
  }
}

template <int BLOCK, bool TR>
static hipError_t launch_pool2(int g, int ns, const PoolArgs& a) {
  switch (g) {
    case 8: return launch_pool3<BLOCK, 8, TR>(ns, a);
    case 16: return launch_pool3<BLOCK, 16, TR>(ns, a);
    case 64: return launch_pool3<BLOCK, 64, TR>(ns, a);
    default: return hipErrorInvalidValue;
  }
}

// Pool shapes built into the library: block 512 or 1024, 8/16/64 lanes per
// chain, 1/2/4 slots, either rank layout.
static bool pool_shape_ok(int blk, int g, int ns) {
  return (blk == 512 || blk == 1024) && (g == 8 || g == 16 || g == 64) && (ns == 1 || ns == 2 || ns == 4);
}

static hipError_t launch_pool(int blk, int g, int ns, bool transposed, const PoolArgs& a) {
  if (blk == 512) return transposed ? launch_pool2<512, true>(g, ns, a) : launch_pool2<512, false>(g, ns, a);
  if (blk == 1024) return transposed ? launch_pool2<1024, true>(g, ns, a) : launch_pool2<1024, false>(g, ns, a);
  return hipErrorInvalidValue;
}

}  // namespace pactae

using pactae::BatchDesc;

struct smp_batch {
  // One measured configuration of the batched solve.
  struct Cfg {
    int mode = 0, block = 0, wpc = 0;
    int group = 0, slots = 0;   // pool: lanes per chain, instances open per workgroup
    int layout = 0;             // pool: SMP_RANK_WOMEN_ROWS or SMP_RANK_MEN_ROWS
  };

  smp_batch_opts opts{};
  pactae::DeviceInfo dev;
  int count = 0;
  int max_n = 0;
  long long total_elems = 0, total_men = 0;
  std::vector<int> sizes;
  std::vector<long long> offsets, moffsets;

  int32_t* d_pm = nullptr;
  int32_t* d_pw = nullptr;
  int16_t* d_rank = nullptr;
  int32_t* d_match = nullptr;
  BatchDesc* d_desc = nullptr;
  pactae::RankTask* d_tasks = nullptr;  // pool: one per instance, instance order
  int2* d_tiles = nullptr;              // pool, men's-rows layout: rankT tiles
  int ntiles = 0;
  int* d_row_task = nullptr;            // pool, women's-rows layout: flat row -> instance
  int* d_master = nullptr;              // pool: per-instance master-list flags
  int32_t* d_linv = nullptr;            // pool: Linv of master-list instances, at matching offsets
  long long* d_moffs = nullptr;         // pool: matching offsets in the caller's order
  int* d_ctl = nullptr;                 // [0] next instance, [1] error
  unsigned long long* d_props = nullptr;

  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;   // pool: the master-list detection, beside the rank pass
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_mid = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  hipEvent_t mid_ev = nullptr;  // recorded between the rank pass and the chains when set
  float rank_ms = 0;            // rank pass of the chosen configuration, measured while tuning
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  Cfg cfg;
  bool have_result = false;
  smp_batch_times times{};
  std::unique_ptr<smp_engine> looped;  // single-instance engine for the looped baseline

  ~smp_batch() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    looped.reset();
    (void)hipFree(d_pm);
    (void)hipFree(d_pw);
    (void)hipFree(d_rank);
    (void)hipFree(d_match);
    (void)hipFree(d_desc);
    (void)hipFree(d_tasks);
    (void)hipFree(d_tiles);
    (void)hipFree(d_row_task);
    (void)hipFree(d_master);
    (void)hipFree(d_linv);
    (void)hipFree(d_moffs);
    (void)hipFree(d_ctl);
    (void)hipFree(d_props);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (ev_mid) (void)hipEventDestroy(ev_mid);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (side) (void)hipStreamDestroy(side);
    if (stream) (void)hipStreamDestroy(stream);
  }

  hipError_t enqueue(const Cfg& c) {
    hipLaunchKernelGGL(pactae::k_batch_reset, dim3(1), dim3(64), 0, stream, d_ctl, d_props, d_ctl + 1);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (c.mode == SMP_BATCH_POOL) {
      const bool tr = c.layout == SMP_RANK_MEN_ROWS;
      // The master-list detection (men's lists) runs on the side stream while
      // the rank pass (women's lists) runs here; neither writes what the other
      // reads, and the chains wait for both.
      if ((e = hipEventRecord(ev_fork, stream)) != hipSuccess) return e;
      if ((e = hipStreamWaitEvent(side, ev_fork, 0)) != hipSuccess) return e;
      const int cgrid = std::max(1, std::min((count + 3) / 4, dev.cus * 8));
      hipLaunchKernelGGL(pactae::k_master_probe, dim3(cgrid), dim3(256), 0, side, d_tasks, count, d_pw, d_pm,
                         d_master);
      hipLaunchKernelGGL(pactae::k_master_rows, dim3(dev.cus * 8), dim3(256), 0, side, d_tasks, d_row_task,
                         (int)total_men, d_pw, d_pm, d_master);
      hipLaunchKernelGGL(pactae::k_master_linv, dim3(std::max(1, std::min(count, dev.cus * 8))), dim3(256), 0,
                         side, d_tasks, count, d_moffs, d_pw, d_pm, d_master, d_linv, d_ctl + 1);
      if ((e = hipGetLastError()) != hipSuccess) return e;
      if ((e = hipEventRecord(ev_join, side)) != hipSuccess) return e;
      e = tr ? pactae::launch_rankT(stream, d_tasks, d_tiles, ntiles, max_n, d_ctl + 1, 0, dev.cus)
             : pactae::launch_rank_rows16(stream, d_tasks, count, d_row_task, (int)total_men, max_n, d_ctl + 1,
                                          0, dev.cus, dev.max_lds);
      if (e != hipSuccess) return e;
      if (mid_ev && (e = hipEventRecord(mid_ev, stream)) != hipSuccess) return e;
      if ((e = hipStreamWaitEvent(stream, ev_join, 0)) != hipSuccess) return e;
      pactae::PoolArgs a{std::max(1, std::min((count + c.slots - 1) / c.slots, dev.cus * c.wpc)),
                         max_n, count, stream, d_desc, d_pm, d_rank, d_match, d_ctl, d_props, d_ctl + 1,
                         d_pw, d_master, d_linv};
      return pactae::launch_pool(c.block, c.group, c.slots, tr, a);
    }
    const int grid = std::max(1, std::min(count, dev.cus * c.wpc));
    const size_t lds = pactae::fused_lds(c.block, max_n);
    switch (c.block) {
      case 256:
        return pactae::launch_fused<256>(grid, lds, stream, d_desc, count, d_pm, d_pw, d_rank, d_match, d_ctl,
                                         d_props, d_ctl + 1, max_n);
      case 512:
        return pactae::launch_fused<512>(grid, lds, stream, d_desc, count, d_pm, d_pw, d_rank, d_match, d_ctl,
                                         d_props, d_ctl + 1, max_n);
      default:
        return pactae::launch_fused<1024>(grid, lds, stream, d_desc, count, d_pm, d_pw, d_rank, d_match,
                                          d_ctl, d_props, d_ctl + 1, max_n);
    }
  }

  // Whether a configuration fits this batch (LDS per workgroup and per CU).
  bool fits(Cfg& c) const {
    if (c.block * c.wpc > 2048) return false;
    if (c.mode == SMP_BATCH_FUSED) {
      // a whole row per staging round, and the workgroups of a CU in its LDS
      const size_t lds = pactae::fused_lds(c.block, max_n) + 8 * (c.block / 64) + 64;
      return c.block * pactae::fused_per(c.block) >= max_n && lds * c.wpc <= 160 * 1024 &&
             (int)lds <= dev.max_lds;
    }
    if (!pactae::pool_shape_ok(c.block, c.group, c.slots)) return false;
    const size_t lds = (size_t)max_n * 8 * c.slots + 64 * (c.slots + 1);
    if (lds * c.wpc > 160 * 1024 || lds > (size_t)dev.max_lds) return false;
    if (c.layout == SMP_RANK_MEN_ROWS &&
        pactae::rankT_tile_lds(pactae::rankT_tile_width(max_n), max_n) > (size_t)dev.max_lds)
      return false;
    if (c.layout == SMP_RANK_WOMEN_ROWS && pactae::rank_rows16_lds(64, max_n) > (size_t)dev.max_lds)
      return false;
    return true;
  }

  // One timed run of a configuration (events around the enqueued work; *rms
  // receives the rank pass alone in the pool mode).
  int time_config(const Cfg& c, float* ms, float* rms) {
    const bool two = c.mode == SMP_BATCH_POOL;
    mid_ev = two ? ev_mid : nullptr;
    SMP_HIP_TRY(hipEventRecord(ev0, stream));
    hipError_t e = enqueue(c);
    mid_ev = nullptr;
    SMP_HIP_TRY(e);
    SMP_HIP_TRY(hipEventRecord(ev1, stream));
    SMP_HIP_TRY(hipStreamSynchronize(stream));
    SMP_HIP_TRY(hipEventElapsedTime(ms, ev0, ev1));
    *rms = 0;
    if (two) SMP_HIP_TRY(hipEventElapsedTime(rms, ev0, ev_mid));
    return SMP_OK;
  }

  // Autotuner: time each candidate configuration that fits this batch (best of
  // two runs) and keep the fastest.  The default search covers the shapes that
  // have won on uniform, congested, correlated and mixed batches; any fixed
  // option restricts the search to the full grid of shapes matching it.
  int tune() {
    std::vector<Cfg> cands;
    const bool fixed = opts.mode || opts.block_threads > 0 || opts.wgs_per_cu > 0 || opts.group_lanes > 0 ||
                       opts.pool_slots > 0 || opts.rank_layout > 0;
    if (fixed) {
      for (int blk : {256, 512, 1024})
        for (int w : {1, 2, 4, 8}) {
          Cfg f;
          f.mode = SMP_BATCH_FUSED, f.block = blk, f.wpc = w, f.group = 64;
          cands.push_back(f);
          for (int g : {8, 16, 64})
            for (int ns : {1, 2, 4})
              for (int lay : {SMP_RANK_MEN_ROWS, SMP_RANK_WOMEN_ROWS}) {
                Cfg p;
                p.mode = SMP_BATCH_POOL, p.block = blk, p.wpc = w, p.group = g, p.slots = ns, p.layout = lay;
                cands.push_back(p);
              }
        }
    } else {
      const int shapes[][6] = {
          // mode, block, wgs/CU, group, slots, layout
          {SMP_BATCH_FUSED, 1024, 2, 64, 0, 0},
          {SMP_BATCH_FUSED, 512, 4, 64, 0, 0},
          {SMP_BATCH_POOL, 1024, 2, 64, 1, SMP_RANK_MEN_ROWS},
          {SMP_BATCH_POOL, 1024, 2, 64, 2, SMP_RANK_MEN_ROWS},
          {SMP_BATCH_POOL, 1024, 2, 16, 1, SMP_RANK_MEN_ROWS},
          {SMP_BATCH_POOL, 512, 4, 16, 1, SMP_RANK_MEN_ROWS},
          {SMP_BATCH_POOL, 1024, 2, 64, 1, SMP_RANK_WOMEN_ROWS},
          {SMP_BATCH_POOL, 1024, 2, 16, 1, SMP_RANK_WOMEN_ROWS},
      };
      for (const auto& sh : shapes) {
        Cfg c;
        c.mode = sh[0], c.block = sh[1], c.wpc = sh[2], c.group = sh[3], c.slots = sh[4], c.layout = sh[5];
        cands.push_back(c);
      }
    }
    float best = 1e30f;
    Cfg bc;
    for (Cfg c : cands) {
      if (opts.mode && c.mode != opts.mode) continue;
      if (opts.block_threads > 0 && c.block != opts.block_threads) continue;
      if (opts.wgs_per_cu > 0 && c.wpc != opts.wgs_per_cu) continue;
      if (c.mode == SMP_BATCH_POOL) {
        if (opts.group_lanes > 0 && c.group != opts.group_lanes) continue;
        if (opts.pool_slots > 0 && c.slots != opts.pool_slots) continue;
        if (opts.rank_layout > 0 && c.layout != opts.rank_layout) continue;
      }
      if (!fits(c)) continue;
      float t1 = 0, t2 = 0, r1 = 0, r2 = 0;
      if (int st = time_config(c, &t1, &r1)) return st;
      if (int st = time_config(c, &t2, &r2)) return st;
      const float t = std::min(t1, t2);
      if (t < best) best = t, bc = c, rank_ms = std::min(r1, r2);
    }
    if (!bc.block) {
      pactae::set_error("batched solve: no configuration fits the options and the LDS budget for n=" +
                        std::to_string(max_n));
      return SMP_ERR_INVALID;
    }
    cfg = bc;
    return SMP_OK;
  }

  int build_graph() {
    if (!opts.use_graph) return SMP_OK;
    SMP_HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = enqueue(cfg);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(stream, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      pactae::set_error(std::string("graph capture failed: ") + hipGetErrorString(e != hipSuccess ? e : e2));
      return SMP_ERR_HIP;
    }
    graph = g;
    SMP_HIP_TRY(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
    return SMP_OK;
  }

  int check_error() {
    int err = 0;
    SMP_HIP_TRY(hipMemcpy(&err, d_ctl + 1, sizeof(int), hipMemcpyDeviceToHost));
    if (err) {
      pactae::set_error("batched solve: invalid preference lists (device flags " + std::to_string(err) + ")");
      return SMP_ERR_INVALID;
    }
    return SMP_OK;
  }
};

extern "C" {

void smp_batch_default_opts(smp_batch_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->device = 0;
  o->use_graph = 1;
  o->block_threads = 0;
  o->wgs_per_cu = 0;
  o->mode = SMP_BATCH_AUTO;
  o->group_lanes = 0;
  o->pool_slots = 0;
  o->rank_layout = 0;
}

int smp_batch_create(int32_t count, const int32_t* sizes, const int64_t* offsets,
                     const int32_t* pref_m_all, const int32_t* pref_w_all, const smp_batch_opts* opts,
                     smp_batch** out) {
  if (!out || count <= 0 || !sizes || !offsets || !pref_m_all || !pref_w_all) {
    pactae::set_error("smp_batch_create: bad arguments");
    return SMP_ERR_INVALID;
  }
  *out = nullptr;
  auto b = std::make_unique<smp_batch>();
  if (opts)
    b->opts = *opts;
  else
    smp_batch_default_opts(&b->opts);
  b->count = count;
  b->sizes.assign(sizes, sizes + count);
  b->offsets.assign(offsets, offsets + count);
  long long end = 0;
  for (int i = 0; i < count; ++i) {
    const int n = sizes[i];
    if (n <= 0 || n > 32767) {
      pactae::set_error("smp_batch_create: instance sizes must be in [1, 32767] (int16 ranks)");
      return SMP_ERR_INVALID;
    }
    if (offsets[i] < 0) {
      pactae::set_error("smp_batch_create: negative offset");
      return SMP_ERR_INVALID;
    }
    end = std::max(end, (long long)offsets[i] + (long long)n * n);
    b->max_n = std::max(b->max_n, n);
  }
  b->total_elems = end;
  b->moffsets.resize(count);
  long long mo = 0;
  for (int i = 0; i < count; ++i) {
    b->moffsets[i] = mo;
    mo += sizes[i];
  }
  b->total_men = mo;
  if (int st = pactae::query_device(b->opts.device, &b->dev)) return st;
  const size_t eb = (size_t)end * sizeof(int32_t);
  if (hipMalloc(&b->d_pm, eb) != hipSuccess || hipMalloc(&b->d_pw, eb) != hipSuccess ||
      hipMalloc(&b->d_rank, (size_t)end * sizeof(int16_t)) != hipSuccess ||
      hipMalloc(&b->d_match, (size_t)mo * sizeof(int32_t)) != hipSuccess ||
      hipMalloc(&b->d_desc, (size_t)count * sizeof(BatchDesc)) != hipSuccess ||
      hipMalloc(&b->d_ctl, 4 * sizeof(int)) != hipSuccess ||
      hipMalloc(&b->d_props, sizeof(unsigned long long)) != hipSuccess) {
    (void)hipGetLastError();
    pactae::set_error("smp_batch_create: device allocation failed");
    return SMP_ERR_NOMEM;
  }
  SMP_HIP_TRY(hipMemcpy(b->d_pm, pref_m_all, eb, hipMemcpyHostToDevice));
  SMP_HIP_TRY(hipMemcpy(b->d_pw, pref_w_all, eb, hipMemcpyHostToDevice));
  // Largest instances first (longest-processing-time order for the work stealing).
  std::vector<int> order(count);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return sizes[a] > sizes[c]; });
  std::vector<BatchDesc> desc(count);
  for (int i = 0; i < count; ++i) {
    const int j = order[i];
    desc[i] = BatchDesc{offsets[j], b->moffsets[j], sizes[j], j};
  }
  SMP_HIP_TRY(hipMemcpy(b->d_desc, desc.data(), desc.size() * sizeof(BatchDesc), hipMemcpyHostToDevice));
  // Pool mode: the instances as rank tasks, their transposed-rank tiles and the
  // flat row -> instance map of the women's-rows pass.
  if (b->opts.mode != SMP_BATCH_FUSED) {
    std::vector<pactae::RankTask> tasks(count);
    long long row0 = 0;
    for (int i = 0; i < count; ++i) {
      tasks[i] = pactae::RankTask{b->d_pw + offsets[i], b->d_rank + offsets[i], sizes[i], (int)row0};
      row0 += sizes[i];
    }
    if (row0 > INT32_MAX) {
      pactae::set_error("smp_batch_create: more than 2^31 rows in one batch");
      return SMP_ERR_INVALID;
    }
    std::vector<int> row_task((size_t)row0);
    for (int i = 0; i < count; ++i)
      std::fill(row_task.begin() + tasks[i].row0, row_task.begin() + tasks[i].row0 + sizes[i], i);
    std::vector<int2> tiles;
    pactae::rankT_tiles(sizes, count, b->max_n, &tiles);
    b->ntiles = (int)tiles.size();
    if (hipMalloc(&b->d_tasks, tasks.size() * sizeof(pactae::RankTask)) != hipSuccess ||
        hipMalloc(&b->d_tiles, tiles.size() * sizeof(int2)) != hipSuccess ||
        hipMalloc(&b->d_row_task, row_task.size() * sizeof(int)) != hipSuccess ||
        hipMalloc(&b->d_master, (size_t)count * sizeof(int)) != hipSuccess ||
        hipMalloc(&b->d_linv, (size_t)b->total_men * sizeof(int32_t)) != hipSuccess ||
        hipMalloc(&b->d_moffs, (size_t)count * sizeof(long long)) != hipSuccess) {
      (void)hipGetLastError();
      pactae::set_error("smp_batch_create: device allocation failed");
      return SMP_ERR_NOMEM;
    }
    SMP_HIP_TRY(hipMemcpy(b->d_tasks, tasks.data(), tasks.size() * sizeof(pactae::RankTask),
                          hipMemcpyHostToDevice));
    SMP_HIP_TRY(hipMemcpy(b->d_tiles, tiles.data(), tiles.size() * sizeof(int2), hipMemcpyHostToDevice));
    SMP_HIP_TRY(hipMemcpy(b->d_row_task, row_task.data(), row_task.size() * sizeof(int),
                          hipMemcpyHostToDevice));
    SMP_HIP_TRY(hipMemcpy(b->d_moffs, b->moffsets.data(), (size_t)count * sizeof(long long),
                          hipMemcpyHostToDevice));
  }
  SMP_HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  SMP_HIP_TRY(hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking));
  SMP_HIP_TRY(hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming));
  SMP_HIP_TRY(hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming));
  SMP_HIP_TRY(hipEventCreate(&b->ev0));
  SMP_HIP_TRY(hipEventCreate(&b->ev1));
  SMP_HIP_TRY(hipEventCreate(&b->ev_mid));
  if (int st = b->tune()) return st;
  if (int st = b->build_graph()) return st;
  *out = b.release();
  return SMP_OK;
}

int smp_batch_exec(smp_batch* b) {
  if (!b) return SMP_ERR_INVALID;
  SMP_HIP_TRY(hipSetDevice(b->opts.device));
  auto t0 = pactae::Clock::now();
  SMP_HIP_TRY(hipEventRecord(b->ev0, b->stream));
  if (b->graph_exec)
    SMP_HIP_TRY(hipGraphLaunch(b->graph_exec, b->stream));
  else
    SMP_HIP_TRY(b->enqueue(b->cfg));
  SMP_HIP_TRY(hipEventRecord(b->ev1, b->stream));
  SMP_HIP_TRY(hipStreamSynchronize(b->stream));
  b->times = smp_batch_times{};
  b->times.exec_ms = pactae::ms_since(t0);
  float ms = 0;
  if (hipEventElapsedTime(&ms, b->ev0, b->ev1) == hipSuccess) b->times.kernel_ms = ms;
  if (int st = b->check_error()) return st;
  unsigned long long p = 0;
  SMP_HIP_TRY(hipMemcpy(&p, b->d_props, sizeof(p), hipMemcpyDeviceToHost));
  b->times.proposals = (int64_t)p;
  b->times.count = b->count;
  b->times.block_threads = b->cfg.block;
  b->times.wgs_per_cu = b->cfg.wpc;
  b->times.mode = b->cfg.mode;
  b->times.group_lanes = b->cfg.group;
  b->times.pool_slots = b->cfg.slots;
  b->times.rank_layout = b->cfg.layout;
  b->times.rank_ms = b->rank_ms;
  b->times.solves_per_s = b->count / (b->times.exec_ms / 1e3);
  b->have_result = true;
  return SMP_OK;
}

int smp_batch_exec_looped(smp_batch* b) {
  if (!b) return SMP_ERR_INVALID;
  SMP_HIP_TRY(hipSetDevice(b->opts.device));
  if (!b->looped) {
    smp_opts o;
    smp_default_opts(&o);
    o.device = b->opts.device;
    o.tail_mode = SMP_TAIL_GPU;
    auto e = std::make_unique<smp_engine>();
    e->n = b->max_n;
    e->opts = o;
    if (int st = e->setup()) return st;
    b->looped = std::move(e);
  }
  auto t0 = pactae::Clock::now();
  long long props = 0;
  for (int i = 0; i < b->count; ++i) {
    const int n = b->sizes[i];
    int st = b->looped->run_view(b->d_pm + b->offsets[i], b->d_pw + b->offsets[i], n,
                                 b->d_match + b->moffsets[i]);
    if (st) return st;
    props += b->looped->times.proposals;
  }
  b->times = smp_batch_times{};
  b->times.exec_ms = pactae::ms_since(t0);
  b->times.kernel_ms = 0;
  b->times.proposals = props;
  b->times.count = b->count;
  b->times.solves_per_s = b->count / (b->times.exec_ms / 1e3);
  b->have_result = true;
  return SMP_OK;
}

int smp_batch_result(smp_batch* b, int32_t* match_all) {
  if (!b || !match_all) return SMP_ERR_INVALID;
  if (!b->have_result) {
    pactae::set_error("smp_batch_result before a successful exec");
    return SMP_ERR_STATE;
  }
  SMP_HIP_TRY(hipMemcpy(match_all, b->d_match, (size_t)b->total_men * sizeof(int32_t),
                        hipMemcpyDeviceToHost));
  return SMP_OK;
}

int smp_batch_profile(const smp_batch* b, smp_batch_times* t) {
  if (!b || !t) return SMP_ERR_INVALID;
  *t = b->times;
  return SMP_OK;
}

void smp_batch_destroy(smp_batch* b) { delete b; }

}  // extern "C"
