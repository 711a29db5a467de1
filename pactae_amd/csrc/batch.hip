// batch.hip — many independent instances solved by one persistent kernel.
//
// The reference solves one instance per SmpEngineExecutor call
// (benchmark/smp_engine_exec.h:6-32) and has no batched path.  Here one launch
// (captured once in a hipGraph and replayed) solves all instances:
//   * workgroups take instances from a global counter (work stealing), largest
//     first, so the long instances start early and the short ones fill the end;
//   * a workgroup solves its instance alone: it inverts the women's rows
//     (rank[w][pref_w[w][r]] = r, the F9RankMatrixInitKernel step,
//     src/smp/kernel.cu:672-686) into an int16 rank matrix, rows staged in LDS;
//     then its waves run proposal chains (the La3F9CoreKernel walk,
//     src/smp/kernel.cu:633-670) with the women's partner words in LDS and
//     64-bit LDS atomicMin; then it writes the matching (PostProc,
//     src/smp/smp_engine_bamboo.cu:299-303);
//   * nothing leaves the workgroup except the rank rows (HBM) and the result.
// Block size and workgroups per CU are chosen by a measured sweep at creation
// (autotune) unless the caller fixes them.

#include <algorithm>
#include <numeric>

#include "engine.hpp"
#include "wg_solver.hpp"

namespace pactae {

struct BatchDesc {
  long long off;   // element offset of the instance's n x n lists
  long long moff;  // offset of its result in the concatenated matching
  int n;
  int pad;
};

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_batch_solve(const BatchDesc* __restrict__ desc, int count,
                                                       const int32_t* __restrict__ pm_all,
                                                       const int32_t* __restrict__ pw_all,
                                                       int16_t* __restrict__ rank_all,
                                                       int32_t* __restrict__ match_all,
                                                       int* next_inst, unsigned long long* proposals,
                                                       int* error, int max_n, int rows) {
  extern __shared__ unsigned char smem[];
  __shared__ int ctl[2];
  __shared__ unsigned long long red[BLOCK / 64];
  int bad = 0;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) ctl[0] = atomicAdd(next_inst, 1);
    __syncthreads();
    const int inst = ctl[0];
    if (inst >= count) break;
    const BatchDesc d = desc[inst];
    InstanceRef ref{pm_all + d.off, pw_all + d.off, rank_all + d.off, match_all + d.moff, d.n};
    unsigned long long props = 0;
    bad |= wg_solve<BLOCK>(ref, smem, rows, max_n, ctl, red, &props);
    if (threadIdx.x == 0) atomicAdd(proposals, props);
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
static hipError_t launch_batch_t(int grid, size_t lds, hipStream_t s, const BatchDesc* desc, int count,
                                 const int32_t* pm, const int32_t* pw, int16_t* rank, int32_t* match,
                                 int* next_inst, unsigned long long* props, int* err, int max_n,
                                 int rows) {
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_batch_solve<BLOCK>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_batch_solve<BLOCK>, dim3(grid), dim3(BLOCK), lds, s, desc, count, pm, pw, rank,
                     match, next_inst, props, err, max_n, rows);
  return hipGetLastError();
}

}  // namespace pactae

using pactae::BatchDesc;

struct smp_batch {
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
  int* d_ctl = nullptr;  // [0] next instance, [1] error
  unsigned long long* d_props = nullptr;

  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int block = 0, wpc = 0, rows = 0;
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
    (void)hipFree(d_ctl);
    (void)hipFree(d_props);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }

  hipError_t enqueue(int blk, int wgs_per_cu, int rows) {
    hipLaunchKernelGGL(pactae::k_batch_reset, dim3(1), dim3(64), 0, stream, d_ctl, d_props, d_ctl + 1);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int grid = std::max(1, std::min(count, dev.cus * wgs_per_cu));
    const size_t lds = pactae::wg_lds_bytes(rows, max_n);
    switch (blk) {
      case 256:
        return pactae::launch_batch_t<256>(grid, lds, stream, d_desc, count, d_pm, d_pw, d_rank, d_match,
                                           d_ctl, d_props, d_ctl + 1, max_n, rows);
      case 512:
        return pactae::launch_batch_t<512>(grid, lds, stream, d_desc, count, d_pm, d_pw, d_rank, d_match,
                                           d_ctl, d_props, d_ctl + 1, max_n, rows);
      default:
        return pactae::launch_batch_t<1024>(grid, lds, stream, d_desc, count, d_pm, d_pw, d_rank,
                                            d_match, d_ctl, d_props, d_ctl + 1, max_n, rows);
    }
  }

  // Most rows in flight (one per wave at most, a power of two) whose LDS fits
  // `wgs_per_cu` workgroups in a CU's 160 KiB; 0 if none fits.
  int rows_for(int blk, int wgs_per_cu) const {
    for (int rows = blk / 64; rows >= 1; rows >>= 1) {
      const size_t lds = pactae::wg_lds_bytes(rows, max_n) + 8 * (blk / 64) + 64;
      if (lds * wgs_per_cu <= 160 * 1024 && (int)lds <= dev.max_lds) return rows;
    }
    return 0;
  }

  // One timed run of a configuration (events around the enqueued work).
  int time_config(int blk, int wgs, int rows, float* ms) {
    SMP_HIP_TRY(hipEventRecord(ev0, stream));
    SMP_HIP_TRY(enqueue(blk, wgs, rows));
    SMP_HIP_TRY(hipEventRecord(ev1, stream));
    SMP_HIP_TRY(hipStreamSynchronize(stream));
    SMP_HIP_TRY(hipEventElapsedTime(ms, ev0, ev1));
    return SMP_OK;
  }

  // Autotuner: measure each (block, workgroups per CU) candidate that fits the
  // LDS budget on this batch and keep the fastest (best of two runs each).
  int tune() {
    const int blocks[] = {256, 512, 1024};
    const int wgss[] = {1, 2, 4, 8};
    float best = 1e30f;
    int bb = 0, bw = 0, br = 0;
    for (int blk : blocks) {
      if (opts.block_threads > 0 && blk != opts.block_threads) continue;
      for (int w : wgss) {
        if (opts.wgs_per_cu > 0 && w != opts.wgs_per_cu) continue;
        const int rows = rows_for(blk, w);
        if (blk * w > 2048 || rows == 0) continue;
        float t1 = 0, t2 = 0;
        if (int st = time_config(blk, w, rows, &t1)) return st;
        if (int st = time_config(blk, w, rows, &t2)) return st;
        const float t = std::min(t1, t2);
        if (t < best) best = t, bb = blk, bw = w, br = rows;
      }
    }
    if (!bb) {
      pactae::set_error("batched kernel: no block configuration fits the LDS budget for n=" +
                        std::to_string(max_n));
      return SMP_ERR_INVALID;
    }
    block = bb;
    wpc = bw;
    rows = br;
    return SMP_OK;
  }

  int build_graph() {
    if (!opts.use_graph) return SMP_OK;
    SMP_HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = enqueue(block, wpc, rows);
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
    desc[i] = BatchDesc{offsets[j], b->moffsets[j], sizes[j], 0};
  }
  SMP_HIP_TRY(hipMemcpy(b->d_desc, desc.data(), desc.size() * sizeof(BatchDesc), hipMemcpyHostToDevice));
  SMP_HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  SMP_HIP_TRY(hipEventCreate(&b->ev0));
  SMP_HIP_TRY(hipEventCreate(&b->ev1));
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
    SMP_HIP_TRY(b->enqueue(b->block, b->wpc, b->rows));
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
  b->times.block_threads = b->block;
  b->times.wgs_per_cu = b->wpc;
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
