// engine.hpp — the single-instance engine (AbsSmpEngine-shaped) shared by the C ABI
// and by the looped baseline of the batched path.
//
// Host structure follows the reference engine framework:
//   AbsSmpEngine::FindStableMatching   include/smp/smp_engine_abs.h:10-36
//     precheck (IsPerfect) -> InitProc -> CoreProc, each timed with a steady clock
//   SmpEngineBamboo                    src/smp/smp_engine_bamboo.cu:1-307
// with these MI355X-specific changes:
//   * every phase runs on one HIP stream of the engine; exec returns only after all
//     device work and the host tail have finished (the reference detaches its
//     kernel thread, smp_engine_bamboo.cu:108-112);
//   * the hand-off to the host is decided by a device counter of free men that the
//     proposal kernel maintains, not by 10 ms polls of a racing snapshot
//     (smp_engine_bamboo.cu:232-287); the kernel parks its chains and exits, so the
//     host works on a consistent state;
//   * the host tail reads a compact men-major PR prefix copied while the kernel
//     runs (the reference copies all 8n^2 bytes of its PR matrix, untimed,
//     smp_engine_abs.h:23); rows beyond the prefix are gathered on demand.
#pragma once

#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "kernels.hpp"

namespace pactae {

using Clock = std::chrono::steady_clock;
inline double ms_since(Clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

// Pinned host buffers the host tail reads are placed by the calling thread's
// NUMA policy (its local node) rather than next to the GPU: the host tail's
// random reads then stay local, while the copies into them still run at full
// PCIe rate.
constexpr unsigned kHostAllocFlags = hipHostMallocNumaUser;

struct DeviceInfo {
  int cus = 256;
  int max_lds = 65536;  // bytes a workgroup may allocate
};

inline int query_device(int dev, DeviceInfo* info) {
  int count = 0;
  SMP_HIP_TRY(hipGetDeviceCount(&count));
  if (dev < 0 || dev >= count) {
    set_error("device ordinal " + std::to_string(dev) + " not present (" + std::to_string(count) +
              " devices)");
    return SMP_ERR_HIP;
  }
  SMP_HIP_TRY(hipSetDevice(dev));
  hipDeviceProp_t p;
  SMP_HIP_TRY(hipGetDeviceProperties(&p, dev));
  info->cus = p.multiProcessorCount;
  // gfx950 offers 160 KiB of LDS per CU; sharedMemPerBlock reports what a single
  // workgroup may take.
  info->max_lds = (int)std::max<size_t>(p.sharedMemPerBlock, 65536);
  return SMP_OK;
}

// ---------------------------------------------------------------------------
// AbsSmpEngine template method (smp_engine_abs.h:10-36), with status codes.
class AbsSmpEngine {
 public:
  virtual ~AbsSmpEngine() = default;

  int FindStableMatching() {
    auto t0 = Clock::now();
    bool perfect = false;
    int st = IsPerfect(&perfect);
    precheck_time_ = ms_since(t0);
    if (st) return st;
    init_time_ = core_time_ = 0;
    if (!perfect) {
      t0 = Clock::now();
      st = InitProc();
      init_time_ = ms_since(t0);
      if (st) return st;
      t0 = Clock::now();
      st = CoreProc();
      core_time_ = ms_since(t0);
      if (st) return st;
    }
    t0 = Clock::now();
    st = PostProc(perfect);
    post_time_ = ms_since(t0);
    return st;
  }

 protected:
  virtual int IsPerfect(bool* perfect) = 0;
  virtual int InitProc() = 0;
  virtual int CoreProc() = 0;
  virtual int PostProc(bool perfect) = 0;

  double precheck_time_ = 0, init_time_ = 0, core_time_ = 0, post_time_ = 0;
};

}  // namespace pactae

using pactae::Clock;
using pactae::ms_since;

// ---------------------------------------------------------------------------
struct smp_engine : public pactae::AbsSmpEngine {
  using Scalars = pactae::Scalars;

  int n = 0;        // size of the instance being solved
  int n_alloc = 0;  // size the buffers were allocated for (>= n)
  smp_opts opts{};
  pactae::DeviceInfo dev;
  int K = 0;  // PR prefix columns
  int T = 0;  // hand-off threshold (0 = GPU finishes)
  int K_alloc = 0;               // PR columns the buffers hold
  size_t pr_cap = 0;             // entries d_pr holds (grown for the whole matrix)
  // Lists whose chains walked far past the prefix on the GPU take the whole PR
  // matrix from their next solve on (see deep_lists()).
  bool full_pr_next = false;
  // Master-list path: the women share woman 0's list too (closed form).
  bool master_both = false;
  // A proposal launch that resumes parked chains (a restart over the whole PR
  // matrix) instead of starting every man, and the deep-step limit of the
  // launch (0 = off).
  const int2* resume_starts = nullptr;
  int resume_n = 0;
  long long deep_limit = 0;
  int tuned_K = 0, tuned_T = -1;  // measured choice for AUTO mode (smp_tune)
  bool owns_lists = false;
  bool have_result = false;

  const int32_t* d_pref_m = nullptr;
  const int32_t* d_pref_w = nullptr;  // null when the ranks are given (d_rank_in)
  uint16_t* d_rank_w = nullptr;  // women's ranks, 16-bit (n <= SMP_MAX_N = 65535)
  // A women's rank matrix given as input (int32, rank_in[w * n + m] = position of
  // man m in woman w's list) instead of the women's lists: init is then the
  // prefix gather alone.  Owned with the lists when owns_lists.
  const int32_t* d_rank_in = nullptr;
  uint32_t* d_pr = nullptr;
  unsigned long long* d_partner = nullptr;
  int32_t* d_match = nullptr;
  int* d_seen = nullptr;     // precheck counts; the master-list path's Linv afterwards
  bool master = false;       // this run takes the master-list path
  int* d_cas_next = nullptr; // thread-per-man schedules: list position after each man's last claim
  uint32_t* d_front = nullptr;     // frontier schedule: two lists of free men (n each)
  pactae::FrontierCtl* d_fctl = nullptr;
  bool thread_walk() const { return opts.schedule >= SMP_SCHEDULE_MW_CAS && opts.schedule <= SMP_SCHEDULE_LA_MIN; }
  bool matrices_valid = false;  // d_rank_w / d_pr hold the last run's matrices
  Scalars* d_sc = nullptr;
  int2* d_parked = nullptr;
  unsigned long long* d_wstats = nullptr;  // per proposal wave: atomics, steps
  uint32_t* d_bucket = nullptr;  // sparse prefix pass: per-woman buckets of (m, k)
  unsigned long long* d_bins = nullptr;  // sparse prefix pass: bins of 256 women
  int* d_bin_cnt = nullptr;      // entries per bin
  int* d_bucket_cnt = nullptr;   // entries per woman's bucket (inside the d_bin_cnt allocation)
  uint16_t* d_bucket_rank = nullptr;  // the row pass's ranks of the bucket entries, bucket order
  bool bins_pending = false;     // the bucket kernels of this run are queued (after the precheck)
  int K_sparse_alloc = 0;        // widest prefix the sparse buffers hold
  bool sparse_overflowed = false;  // a sparse pass on these lists overflowed a bin or bucket
  bool sparse = false;           // this run's init took the sparse prefix pass
  bool rank_built = false;       // d_rank_w holds this run's women's rank matrix

  uint32_t* h_pr = nullptr;  // pinned copy of the PR prefix for the host tail
  unsigned long long* h_partner = nullptr;
  Scalars* h_sc = nullptr;
  int2* h_parked = nullptr;
  // The same three buffers as seen from the device (k_export_state writes them).
  unsigned long long* h_partner_dev = nullptr;
  Scalars* h_sc_dev = nullptr;
  int2* h_parked_dev = nullptr;
  // Pinned flag words the device writes last when it hands data to the host:
  // [kFlagPrecheck] by k_publish_scalars, [kFlagExport] by k_export_state.  The
  // host waits for the sequence number it queued (wait_flag) instead of for an
  // event: a flag is seen about as soon as it is written, and unlike a copy
  // plus an event it leaves no gap on the queue.
  static constexpr int kFlagPrecheck = 0, kFlagExport = 1;
  unsigned* h_flags = nullptr;
  unsigned* h_flags_dev = nullptr;
  unsigned flag_seq = 0;      // last sequence number queued (either flag)
  unsigned export_seq = 0;    // the one the current export writes
  bool match_on_host_only = false;  // the host tail built the matching; d_match not yet mirrored
  int32_t* h_match = nullptr;  // pinned result (man -> woman)
  int2* d_starts = nullptr;     // chains handed back by the host tail
  bool host_finished = false;   // the host tail completed every chain
  // Extension windows: when a host chain walks past the host's PR prefix, the
  // GPU computes the next entries (woman, rank) of that man's list into pinned
  // host memory and the host continues the chain itself.  A man's window
  // replaces his previous one (his list position only grows).  Window storage
  // (win_cap entries) is handed out in order during a solve; when it runs out
  // the chains go back to the GPU.
  static constexpr int kFlagWindow = 2;
  size_t win_cap = 0;
  size_t win_used = 0;
  uint32_t* h_win = nullptr;   // win_cap entries
  uint32_t* h_win_dev = nullptr;
  int4* h_winreq = nullptr;    // (man, first position, offset in the round, length); n of them
  int4* h_winreq_dev = nullptr;
  std::vector<int> win_off, win_start, win_len, win_men;
  // A man's first window holds kWinFirst entries and each later one twice the
  // last, up to kWinMax: most chains that pass the prefix end a few dozen
  // positions past it, and without a rank matrix each entry costs a search of
  // a row.
  static constexpr int kWinFirst = 64;
  static constexpr int kWinMax = 4096;
  // When a rank matrix exists (a window entry is then one gather), each round
  // also gives a window to every man matched at most kWinAhead positions before
  // the end of the entries the host holds for him, so that a chain that
  // displaces him continues without another round trip.  deep_w collects the
  // women whose partner may be such a man.
  static constexpr int kWinAhead = 8;
  size_t window_cap() const {
    return opts.tail_window_entries > 0 ? std::min(win_cap, (size_t)opts.tail_window_entries) : win_cap;
  }
  std::vector<int> deep_w, deep_stamp;
  int deep_round = 0;
  // Bands: when chains keep leaving what the host holds one man at a time, each
  // a GPU round trip, the host takes a band of columns of every man's list
  // instead (row-major, gathered on the GPU into d_band and copied into h_band,
  // both grown on demand).  A band starts at the end of the prefix and reaches
  // 64 positions past the deepest position a chain has stopped at, and is at
  // least as wide as the prefix, 256 columns and twice the last band; one wider
  // than n / 4 columns is the whole matrix instead.  A band of c columns is
  // taken once the rounds since the last one (for the whole matrix: all rounds
  // of the solve) have cost about what copying it over PCIe would (~50 GB/s
  // against ~50 us per round: 1.6e-6 n c rounds, and at least 8).
  int band_rounds(int cols) const {
    return std::max(8, (int)std::min<long long>(1 << 20, 16ll * n * cols / 10000000));
  }
  uint32_t* d_band = nullptr;
  uint32_t* h_band = nullptr;
  size_t band_cap = 0;  // entries of both

  hipStream_t stream = nullptr, copy_stream = nullptr, copy_stream2 = nullptr;
  static constexpr int kMaxBlocks = 8;  // PR prefix blocks copied to the host (K <= 256)
  hipEvent_t ev_blk[kMaxBlocks] = {};
  int blocks_copied = 0, blocks_ready = 0;
  int requeues = 0;  // copy-stream replacements made by check_copy_overlap
  bool state_copied = false;  // control block, partner words and parked list are on the host
  // Events on the engine's stream.  Each one recorded between two kernels costs
  // a gap of several microseconds on the queue, so the general path records
  // only ev[0] (init's main pass starts), ev[2] (init done; the prefix copies
  // wait for it) and ev[5] (behind the state export, for timing only: the host
  // waits for the export's flag word); ev[1], ev[4] and ev[6] serve the full
  // init, the baseline and master-list paths and buckets queued without the
  // precheck.  The spans below name the phases they bound in this run (null:
  // not timed).
  hipEvent_t ev[9] = {};
  struct Span {
    hipEvent_t* a = nullptr;
    hipEvent_t* b = nullptr;
  };
  Span span_bucket, span_rank, span_gather, span_propose;
  hipEvent_t* bucket_ev = nullptr;
  // Elapsed time of a span into *out (when both events completed); clears it.
  void take_span(Span& sp, double* out) {
    float ms = 0;
    if (sp.a && sp.b && hipEventElapsedTime(&ms, *sp.a, *sp.b) == hipSuccess) *out = ms;
    (void)hipGetLastError();
    sp = Span{};
  }

  smp_times times{};
  volatile uint64_t touch_sink = 0;
  long long host_proposals_total = 0;
  // Host tail state, one word per woman: rank in bits 0-15 (0xFFFF = free), man
  // in 16-31, list position in 32-47 (all < n <= 65535).  One word keeps each
  // proposal on one cache line.
  std::vector<uint64_t> hst;
  // The rank field of hst alone (2 bytes per woman: 40 KB at n = 20,000, inside
  // a core's L1), read by the walk's eight-position scans; kept equal to hst's.
  std::vector<uint16_t> hrank;
  static constexpr uint32_t kHostFree = 0xFFFFu;
  static constexpr int kTailSlots = 4;  // chains the host walks interleaved
  static uint64_t hs_pack(uint32_t r, uint32_t m, uint32_t p) {
    return (uint64_t)r | ((uint64_t)m << 16) | ((uint64_t)p << 32);
  }
  static int tail_error(const char* what) {
    pactae::set_error(what);
    return SMP_ERR_INVALID;
  }

  ~smp_engine() override { release(); }

  void release() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (copy_stream) (void)hipStreamSynchronize(copy_stream);
    if (copy_stream2) (void)hipStreamSynchronize(copy_stream2);
    if (owns_lists) {
      (void)hipFree(const_cast<int32_t*>(d_pref_m));
      (void)hipFree(const_cast<int32_t*>(d_pref_w));
      (void)hipFree(const_cast<int32_t*>(d_rank_in));
    }
    d_pref_m = d_pref_w = d_rank_in = nullptr;
    (void)hipFree(d_rank_w);
    (void)hipFree(d_pr);
    (void)hipFree(d_partner);
    (void)hipFree(d_match);
    (void)hipFree(d_seen);
    (void)hipFree(d_cas_next);
    (void)hipFree(d_front);
    (void)hipFree(d_fctl);
    d_front = nullptr;
    d_fctl = nullptr;
    (void)hipFree(d_sc);
    (void)hipFree(d_parked);
    (void)hipFree(d_bucket);
    d_bucket = nullptr;
    (void)hipFree(d_bins);
    d_bins = nullptr;
    (void)hipFree(d_bin_cnt);
    d_bin_cnt = nullptr;
    d_bucket_cnt = nullptr;  // inside d_bin_cnt
    (void)hipFree(d_bucket_rank);
    d_bucket_rank = nullptr;
    (void)hipFree(d_wstats);
    d_wstats = nullptr;
    (void)hipFree(d_starts);
    d_starts = nullptr;
    (void)hipHostFree(h_match);
    h_match = nullptr;
    (void)hipHostFree(h_pr);
    (void)hipHostFree(h_partner);
    (void)hipHostFree(h_sc);
    (void)hipHostFree(h_parked);
    (void)hipHostFree(h_flags);
    h_flags = h_flags_dev = nullptr;
    (void)hipFree(d_band);
    (void)hipHostFree(h_band);
    d_band = h_band = nullptr;
    band_cap = 0;
    (void)hipHostFree(h_win);
    (void)hipHostFree(h_winreq);
    h_win = h_win_dev = nullptr;
    h_winreq = h_winreq_dev = nullptr;
    d_rank_w = nullptr;
    d_pr = nullptr;
    d_partner = nullptr;
    d_match = nullptr;
    d_seen = nullptr;
    d_cas_next = nullptr;
    d_sc = nullptr;
    d_parked = nullptr;
    h_pr = nullptr;
    h_partner = nullptr;
    h_sc = nullptr;
    h_parked = nullptr;
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e), e = nullptr;
    for (auto& e : ev_blk)
      if (e) (void)hipEventDestroy(e), e = nullptr;
    if (stream) (void)hipStreamDestroy(stream), stream = nullptr;
    if (copy_stream) (void)hipStreamDestroy(copy_stream), copy_stream = nullptr;
    if (copy_stream2) (void)hipStreamDestroy(copy_stream2), copy_stream2 = nullptr;
  }

  int setup() {
    if (int st = pactae::query_device(opts.device, &dev)) return st;
    const size_t nn = (size_t)n * n;
    n_alloc = n;
    K_alloc = std::min(n, std::max(prefix_for(n), 256));
    configure(n);
    SMP_HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    SMP_HIP_TRY(hipStreamCreateWithFlags(&copy_stream, hipStreamNonBlocking));
    SMP_HIP_TRY(hipStreamCreateWithFlags(&copy_stream2, hipStreamNonBlocking));
    for (auto& e : ev) SMP_HIP_TRY(hipEventCreate(&e));
    for (auto& e : ev_blk) SMP_HIP_TRY(hipEventCreate(&e));
    if (hipMalloc(&d_rank_w, nn * sizeof(uint16_t)) != hipSuccess ||
        hipMalloc(&d_pr, (pr_cap = pactae::pr_entries(n, K_alloc)) * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&d_partner, (size_t)n * sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(&d_match, (size_t)n * sizeof(int32_t)) != hipSuccess ||
        hipMalloc(&d_seen, (size_t)n * sizeof(int)) != hipSuccess ||
        hipMalloc(&d_cas_next, (size_t)n * sizeof(int)) != hipSuccess ||
        hipMalloc(&d_front, 2 * (size_t)n * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&d_fctl, sizeof(pactae::FrontierCtl)) != hipSuccess ||
        hipMalloc(&d_sc, sizeof(Scalars)) != hipSuccess ||
        hipMalloc(&d_parked, (size_t)n * sizeof(int2)) != hipSuccess ||
        hipMalloc(&d_wstats, (size_t)dev.cus * 32 * 2 * sizeof(unsigned long long)) != hipSuccess) {
      (void)hipGetLastError();
      pactae::set_error("device allocation failed for n=" + std::to_string(n));
      return SMP_ERR_NOMEM;
    }
    if (hipHostMalloc(&h_pr, pactae::pr_entries(n, K_alloc) * sizeof(uint32_t), pactae::kHostAllocFlags) !=
            hipSuccess ||
        hipHostMalloc(&h_partner, (size_t)n * sizeof(unsigned long long), pactae::kHostAllocFlags) !=
            hipSuccess ||
        hipHostMalloc(&h_sc, sizeof(Scalars), pactae::kHostAllocFlags) != hipSuccess ||
        hipHostMalloc(&h_parked, (size_t)n * sizeof(int2), pactae::kHostAllocFlags) != hipSuccess ||
        hipHostMalloc(&h_flags, 64, pactae::kHostAllocFlags) != hipSuccess) {
      (void)hipGetLastError();
      pactae::set_error("pinned host allocation failed for n=" + std::to_string(n));
      return SMP_ERR_NOMEM;
    }
    for (int i = 0; i < 16; ++i) __atomic_store_n(&h_flags[i], 0u, __ATOMIC_RELAXED);
    if (hipError_t e = pactae::set_invert_lds_limit(((invert_chunk() + 7) & ~7) * 2); e != hipSuccess) {
      pactae::set_error(std::string("hipFuncSetAttribute(LDS) failed: ") + hipGetErrorString(e));
      return SMP_ERR_HIP;
    }
    if (hipHostGetDevicePointer(reinterpret_cast<void**>(&h_partner_dev), h_partner, 0) != hipSuccess ||
        hipHostGetDevicePointer(reinterpret_cast<void**>(&h_sc_dev), h_sc, 0) != hipSuccess ||
        hipHostGetDevicePointer(reinterpret_cast<void**>(&h_parked_dev), h_parked, 0) != hipSuccess ||
        hipHostGetDevicePointer(reinterpret_cast<void**>(&h_flags_dev), h_flags, 0) != hipSuccess) {
      (void)hipGetLastError();
      pactae::set_error("pinned hand-off buffers are not visible to the device");
      return SMP_ERR_HIP;
    }
    if (hipHostMalloc(&h_match, (size_t)n * sizeof(int32_t), pactae::kHostAllocFlags) != hipSuccess ||
        hipMalloc(&d_starts, (size_t)n * sizeof(int2)) != hipSuccess) {
      (void)hipGetLastError();
      pactae::set_error("allocation of the result buffers failed");
      return SMP_ERR_NOMEM;
    }
    win_cap = std::min<size_t>(1u << 22, std::max<size_t>(1u << 16, (size_t)8 * n * kWinFirst));
    if (hipHostMalloc(&h_win, win_cap * sizeof(uint32_t), pactae::kHostAllocFlags) != hipSuccess ||
        hipHostMalloc(&h_winreq, (size_t)std::max(n, 1) * sizeof(int4), pactae::kHostAllocFlags) != hipSuccess ||
        hipHostGetDevicePointer(reinterpret_cast<void**>(&h_win_dev), h_win, 0) != hipSuccess ||
        hipHostGetDevicePointer(reinterpret_cast<void**>(&h_winreq_dev), h_winreq, 0) != hipSuccess) {
      (void)hipGetLastError();
      pactae::set_error("allocation of the host tail's extension windows failed");
      return SMP_ERR_NOMEM;
    }
    win_off.assign(n, -1);
    win_start.assign(n, 0);
    win_len.assign(n, 0);
    deep_stamp.assign(n, 0);
    // Sparse prefix pass buffers, for prefixes up to 256 columns (any width when
    // n <= 2048, where n x n buckets are small).
    K_sparse_alloc = n <= 2048 ? K_alloc : std::min(K_alloc, 256);
    if (hipMalloc(&d_bucket, (size_t)n * bucket_cap(n, K_sparse_alloc) * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&d_bucket_rank, (size_t)n * bucket_cap(n, K_sparse_alloc) * sizeof(uint16_t)) != hipSuccess ||
        hipMalloc(&d_bins, (size_t)pactae::bin_count(n) * bin_cap(K_sparse_alloc) * sizeof(unsigned long long)) !=
            hipSuccess ||
        hipMalloc(&d_bin_cnt, counter_words() * sizeof(int)) != hipSuccess) {
      (void)hipGetLastError();
      pactae::set_error("allocation of the prefix buckets failed for n=" + std::to_string(n));
      return SMP_ERR_NOMEM;
    }
    d_bucket_cnt = d_bin_cnt + 256 * pactae::kBinCounterStride;
    return SMP_OK;
  }
  // Bin counters, then the per-woman bucket counts: one range the reset clears.
  size_t counter_words() const { return 256 * (size_t)pactae::kBinCounterStride + (size_t)n_alloc; }

  // Sparse prefix pass: bucket capacity per woman for prefix width k.  A woman's
  // suitor count over the first k positions of n uniform lists is about
  // Poisson(k); k + max(64, k/2) is more than 5.6 standard deviations above it
  // (overflow only falls back to the full rank matrix).  With k = ni every man
  // names every woman, so ni slots always suffice.
  static int bucket_cap(int ni, int k) { return std::min(ni, k + std::max(64, k / 2)); }
  // Entries per bin of 256 women: 256 k on average for uniform lists (exactly
  // that at most when k = n), plus 12.5% and 1024 of slack.
  static int bin_cap(int k) { return 256 * k + 32 * k + 1024; }
  // AUTO also stops taking the sparse pass on lists where it has overflowed once
  // (men whose prefixes name the same few women, e.g. correlated lists): its
  // buckets would only be built to be thrown away.
  bool use_sparse() const {
    if (d_rank_in || opts.init_mode == SMP_INIT_FULL || K > K_sparse_alloc) return false;
    if (opts.init_mode == SMP_INIT_SPARSE) return true;
    return 4ll * K <= n && !sparse_overflowed;
  }
  // The device result buffer, mirrored from the host's matching if the host tail
  // built it.
  int device_match(const int32_t** out) {
    if (match_on_host_only) {
      SMP_HIP_TRY(hipMemcpyAsync(d_match, h_match, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, stream));
      SMP_HIP_TRY(wait_stream(stream));
      match_on_host_only = false;
    }
    *out = d_match;
    return SMP_OK;
  }
  // Queues the sparse pass's bucket kernels on the engine's stream.  Queued
  // behind the precheck (gated) they are not timed (an event there would delay
  // them); otherwise from ev[6] to the init's ev[0].
  // stages as in launch_bucket_prefix (1: bin the prefix, 2: sort the bins);
  // publish_seq != 0: the binning kernel also publishes the precheck's flags.
  int launch_buckets(int gated, int stages = 3, unsigned publish_seq = 0) {
    if (stages & 1) {
      bucket_ev = nullptr;
      if (!gated) {
        SMP_HIP_TRY(hipEventRecord(ev[6], stream));
        bucket_ev = &ev[6];
      }
    }
    SMP_HIP_TRY(pactae::launch_bucket_prefix(d_pref_m, n, K, d_bin_cnt, d_bins, bin_cap(K), d_bucket_cnt, d_bucket,
                                             bucket_cap(n, K), gated, d_sc, dev.cus, stream, stages, h_sc_dev,
                                             publish_seq ? h_flags_dev + kFlagPrecheck : nullptr, publish_seq));
    bins_pending = true;
    return SMP_OK;
  }
  // The women's rank matrix, if this run has not built it (sparse prefix pass).
  int ensure_rank_matrix() {
    if (rank_built || d_rank_in) return SMP_OK;
    SMP_HIP_TRY(pactae::launch_invert_rows(d_pref_w, d_rank_w, n, invert_chunk(), opts.validate, d_sc, stream,
                                           /*nontemporal=*/1));
    rank_built = true;
    times.rank_matrix = 1;
    return SMP_OK;
  }

  // Host waits.  With opts.host_spin the host polls the stream / event instead
  // of blocking in the runtime, which takes the wake-up latency out of every
  // hand-off on the critical path (the waits here are short and few).
  hipError_t wait_stream(hipStream_t s) {
    if (!opts.host_spin) return hipStreamSynchronize(s);
    for (;;) {
      hipError_t e = hipStreamQuery(s);
      if (e != hipErrorNotReady) return e;
    }
  }
  hipError_t wait_event(hipEvent_t e) {
    if (!opts.host_spin) return hipEventSynchronize(e);
    for (;;) {
      hipError_t q = hipEventQuery(e);
      if (q != hipErrorNotReady) return q;
    }
  }
  bool flag_seen(int f, unsigned seq) const { return __atomic_load_n(&h_flags[f], __ATOMIC_ACQUIRE) == seq; }
  // Polls flag f for seq (a read of host memory).  Now and then it asks the
  // runtime about the stream, so that a failed launch or a device error ends
  // the wait; a stream that drained without the flag is an error too.
  hipError_t wait_flag(int f, unsigned seq) {
    for (unsigned i = 1;; ++i) {
      if (flag_seen(f, seq)) return hipSuccess;
      __builtin_ia32_pause();  // leaves the core's other hardware thread its resources
      if ((i & 4095) == 0) {
        const hipError_t q = hipStreamQuery(stream);
        if (q == hipErrorNotReady) continue;
        if (q != hipSuccess) return q;
        return flag_seen(f, seq) ? hipSuccess : hipErrorUnknown;
      }
    }
  }

  // Prefix width and hand-off threshold for an instance of size ni <= n_alloc.
  int prefix_for(int ni) const {
    return opts.prefix_cols <= 0 ? std::min(ni, 128) : std::min(ni, opts.prefix_cols);
  }
  void configure(int ni) {
    n = ni;
    K = prefix_for(ni);
    full_pr_next = false;
    // Default from measured sweeps (uniform lists): the host takes over at 48
    // free men, which the tuner picks at n = 20,000 (untuned solves at 64 were
    // 0.04-0.10 ms slower there); at n = 2,000 to 8,000, 48 and 64 are within
    // 2%.  smp_tune replaces this with a measurement.
    if (opts.tail_mode == SMP_TAIL_GPU)
      T = 0;
    else
      T = opts.tail_threshold > 0 ? std::min(opts.tail_threshold, ni) : std::min(ni, 48);
    if (ni == n_alloc && opts.tail_mode == SMP_TAIL_AUTO && tuned_T >= 0) {
      if (opts.prefix_cols <= 0) K = tuned_K;
      if (opts.tail_threshold <= 0) T = tuned_T;
    }
    // The prefix buffers (d_pr, h_pr) hold K_alloc columns; a width set after
    // they were sized is cut to them (the whole matrix goes through grow_pr).
    if (K_alloc > 0) K = std::min(K, K_alloc);
  }

  // Measured choice of the hand-off threshold T and the prefix width K: every
  // candidate pair is run `trials` times on this instance and ranked by its
  // median wall time; the four best, and the untuned defaults, are then run
  // 3 * trials times more and the lowest median of those wins (one noisy median
  // among ~33 candidates must not decide).  AUTO mode then uses it.  Results
  // are identical for every pair (the matching is unique); only the time differs.
  int tune(int trials, smp_tune_result* out) {
    if (trials <= 0) trials = 5;
    // Slow instances (lists that share their top women: contention) first pick
    // the schedule and the proposal kernel's occupancy; the threshold / prefix
    // sweep below only runs where one solve is cheap.
    int best_schedule = opts.schedule, best_waves = opts.waves_per_cu;
    double sched_ms = 0;
    if (int st = tune_schedule(trials, &best_schedule, &best_waves, &sched_ms)) return st;
    if (sched_ms > kSweepLimitMs) {
      if (out) {
        *out = smp_tune_result{};
        out->prefix_cols = K;
        out->tail_threshold = T;
        out->best_ms = sched_ms;
        out->gpu_only_ms = 0;
        out->candidates = 0;
        out->schedule = best_schedule;
        out->waves_per_cu = best_waves;
      }
      return SMP_OK;
    }
    const int Ks[] = {64, 96, 128, 256};
    const int Ts[] = {0, 1, 2, 4, 8, 16, 24, 32, 48, 64, 96, 128, 256, 1024};
    const smp_opts saved = opts;
    struct Cand {
      int k, t;
      double med;
    };
    std::vector<Cand> cands;
    double gpu_only = 1e300;
    // Median of `runs` solves with prefix k and threshold t.
    auto measure = [&](int k, int t, int runs, double* med) -> int {
      opts.prefix_cols = k;
      opts.tail_mode = t == 0 ? SMP_TAIL_GPU : SMP_TAIL_HOST;
      opts.tail_threshold = t;
      configure(n);
      std::vector<double> w(runs);
      for (int i = 0; i < runs; ++i) {
        auto t0 = Clock::now();
        if (int st = exec()) return st;
        w[i] = ms_since(t0);
      }
      std::sort(w.begin(), w.end());
      *med = w[runs / 2];
      return SMP_OK;
    };
    auto fail = [&](int st) {
      opts = saved;
      configure(n);
      return st;
    };
    for (int k : Ks) {
      if (k > K_alloc) continue;
      for (int t : Ts) {
        if (t > n) continue;
        double med = 0;
        if (int st = measure(k, t, trials, &med)) return fail(st);
        cands.push_back({k, t, med});
        if (t == 0) gpu_only = std::min(gpu_only, med);
      }
    }
    std::sort(cands.begin(), cands.end(), [](const Cand& x, const Cand& y) { return x.med < y.med; });
    std::vector<Cand> finals(cands.begin(), cands.begin() + std::min<size_t>(4, cands.size()));
    {
      opts = saved;
      opts.tail_mode = SMP_TAIL_AUTO;
      opts.prefix_cols = opts.tail_threshold = 0;
      const int def_k = prefix_for(n), def_t = std::min(n, 48);
      bool have = false;
      for (const Cand& c : finals) have = have || (c.k == def_k && c.t == def_t);
      if (!have) finals.push_back({def_k, def_t, 0});
    }
    Cand best{prefix_for(n), 0, 1e300};
    for (Cand& c : finals) {
      if (int st = measure(c.k, c.t, 3 * trials, &c.med)) return fail(st);
      if (c.med < best.med) best = c;
    }
    // The prefix width that wins is fitted to how deep this instance's chains
    // end; other lists of the same kind end deeper (uniform n = 20,000: deepest
    // final position 91 to 131 over six seeds), and a chain past the prefix
    // costs a GPU round trip.  So among the finals within 3% of the fastest,
    // the widest prefix is kept (a wider prefix costs a few microseconds).
    {
      const double limit = best.med * 1.03;
      for (const Cand& c : finals)
        if (c.med <= limit && (c.k > best.k || (c.k == best.k && c.med < best.med))) best = c;
    }
    opts = saved;
    tuned_K = best.k;
    tuned_T = best.t;
    configure(n);
    if (out) {
      *out = smp_tune_result{};
      out->prefix_cols = best.k;
      out->tail_threshold = best.t;
      out->best_ms = best.med;
      out->gpu_only_ms = gpu_only;
      out->candidates = (int)cands.size();
      out->schedule = opts.schedule;
      out->waves_per_cu = opts.waves_per_cu;
    }
    return SMP_OK;
  }

  // Instances slower than this per solve skip the threshold / prefix sweep.
  static constexpr double kSweepLimitMs = 5.0;
  // The schedule and occupancy, measured on this instance when its default
  // solve is slow (kSweepLimitMs): this build's pipeline at 16, 8 and 4 waves per
  // CU, the frontier rounds and the reference's four thread-per-man walks, each
  // solved `trials` times (after one untimed solve); the lowest median is kept
  // in the engine's options for later solves.  Only where the caller left the
  // schedule and occupancy at their defaults, and not with ranks given (the
  // other schedules need the women's lists).
  int tune_schedule(int trials, int* best_schedule, int* best_waves, double* best_ms) {
    const smp_opts saved = opts;
    auto median_of = [&](int runs, double* med) -> int {
      std::vector<double> w(runs);
      for (int i = 0; i < runs; ++i) {
        auto t0 = Clock::now();
        if (int st = exec()) return st;
        w[i] = ms_since(t0);
      }
      std::sort(w.begin(), w.end());
      *med = w[runs / 2];
      return SMP_OK;
    };
    configure(n);
    double base = 0;
    if (int st = median_of(std::max(1, trials / 2), &base)) return st;
    *best_ms = base;
    *best_schedule = opts.schedule;
    *best_waves = opts.waves_per_cu;
    if (base <= kSweepLimitMs || d_rank_in || saved.schedule != SMP_SCHEDULE_BAMBOO || saved.waves_per_cu != 0)
      return SMP_OK;
    struct C {
      int schedule, waves;
    };
    const C cands[] = {{SMP_SCHEDULE_BAMBOO, 8},  {SMP_SCHEDULE_BAMBOO, 4}, {SMP_SCHEDULE_FRONTIER, 0},
                       {SMP_SCHEDULE_MW_CAS, 0},  {SMP_SCHEDULE_LA_CAS, 0}, {SMP_SCHEDULE_MW_MIN, 0},
                       {SMP_SCHEDULE_LA_MIN, 0}};
    for (const C& c : cands) {
      opts.schedule = c.schedule;
      opts.waves_per_cu = c.waves;
      configure(n);
      double med = 0;
      int st = exec();  // untimed: first run of the schedule
      if (st == SMP_OK) st = median_of(trials, &med);
      if (st != SMP_OK) {
        opts = saved;
        configure(n);
        return st;
      }
      if (med < *best_ms) {
        *best_ms = med;
        *best_schedule = c.schedule;
        *best_waves = c.waves;
      }
    }
    opts = saved;
    opts.schedule = *best_schedule;
    opts.waves_per_cu = *best_waves;
    configure(n);
    return SMP_OK;
  }

  // Solve another instance (lists already in HBM, size <= n_alloc) with this
  // engine's buffers: the plain single-instance pipeline, used by the looped
  // baseline of the batched path.
  int run_view(const int32_t* pm, const int32_t* pw, int ni, int32_t* d_out) {
    if (ni <= 0 || ni > n_alloc || d_rank_in) return SMP_ERR_INVALID;
    d_pref_m = pm;
    d_pref_w = pw;
    sparse_overflowed = false;  // other lists
    configure(ni);
    if (int st = exec()) return st;
    if (d_out) {
      const int32_t* dm = nullptr;
      if (int st = device_match(&dm)) return st;
      SMP_HIP_TRY(hipMemcpyAsync(d_out, dm, (size_t)ni * sizeof(int32_t), hipMemcpyDeviceToDevice, stream));
      SMP_HIP_TRY(wait_stream(stream));
    }
    return SMP_OK;
  }

  // Columns per LDS pass of the rank inversion: the whole row when its 16-bit
  // staging fits in a workgroup's LDS (n up to 81,920 on gfx950), else a multiple
  // of 4 that does.
  int invert_chunk() const {
    const int cap = (dev.max_lds / 2) & ~7;
    return n <= cap ? n : cap;
  }

  // Workgroups (4 waves each) of the persistent proposal kernel.
  int propose_grid() const {
    const int wpc = std::min(32, opts.waves_per_cu > 0 ? opts.waves_per_cu : 16);
    const long long waves = std::min<long long>((long long)dev.cus * wpc, (long long)n);
    return (int)std::max<long long>(1, (waves + 3) / 4);
  }

  int read_scalars() {
    SMP_HIP_TRY(hipMemcpyAsync(h_sc, d_sc, sizeof(Scalars), hipMemcpyDeviceToHost, stream));
    SMP_HIP_TRY(wait_stream(stream));
    return SMP_OK;
  }

  int check_device_error(const char* phase) {
    if (h_sc->error == 0) return SMP_OK;
    std::string what;
    if (h_sc->error & pactae::kErrEntryRange) what += " entry outside [0,n);";
    if (h_sc->error & pactae::kErrNotPerm) what += " a woman's row is not a permutation;";
    if (h_sc->error & pactae::kErrListEnd) what += " a man exhausted his list;";
    if (h_sc->error & pactae::kErrUnmatched) what += " a woman was left unmatched;";
    pactae::set_error(std::string("invalid preference lists (") + phase + "):" + what);
    return SMP_ERR_INVALID;
  }

  // -- precheck (smp_engine_bamboo.h:25-33) ---------------------------------
  int IsPerfect(bool* perfect) override {
    times = smp_times{};
    span_bucket = span_rank = span_gather = span_propose = Span{};
    have_result = false;
    master = false;
    master_both = false;
    if (full_pr_next) {
      full_pr_next = false;
      if (grow_pr(n)) K = n;
    }
    matrices_valid = false;
    sparse = false;
    rank_built = false;
    bins_pending = false;
    times.schedule = opts.schedule;
    // One launch resets the control block, clears the precheck's counts and the
    // sparse pass's bin counters, and marks every woman unclaimed (the
    // proposal kernel's start state; the master-list path refills the words
    // for its men).
    const bool run_precheck = opts.precheck && !thread_walk();
    const bool may_sparse = opts.schedule == SMP_SCHEDULE_BAMBOO && use_sparse();
    SMP_HIP_TRY(pactae::launch_reset_scalars(
        d_sc, std::min(n, propose_grid() * 4), stream, run_precheck ? d_seen : nullptr, (size_t)n,
        may_sparse ? d_bin_cnt : nullptr, counter_words(), thread_walk() ? nullptr : d_partner, (size_t)n));
    if (!run_precheck) {
      *perfect = false;
      return SMP_OK;
    }
    SMP_HIP_TRY(pactae::launch_first_choices(d_pref_m, n, d_seen, d_sc, stream));
    // The sparse pass's bucket kernels go right behind the precheck, gated on its
    // flags on the device, so they run while the host reads the flags; the
    // binning kernel publishes those flags as it starts (one launch fewer for
    // the host to queue ahead of the kernels it feeds).
    const unsigned seq = ++flag_seq;
    if (may_sparse) {
      if (int st = launch_buckets(/*gated=*/1, /*stages=*/3, seq)) return st;
    } else {
      SMP_HIP_TRY(pactae::launch_publish_scalars(d_sc, h_sc_dev, h_flags_dev + kFlagPrecheck, seq, stream));
    }
    SMP_HIP_TRY(wait_flag(kFlagPrecheck, seq));
    if (int st = check_device_error("precheck")) return st;
    *perfect = h_sc->dup_first == 0;
    times.is_perfect = *perfect ? 1 : 0;
    // Every man names the same first choice: is every list the same permutation?
    master = false;
    // (With ranks given there are no women's lists to walk: the general path.)
    if (!*perfect && n > 1 && h_sc->first_differs == 0 && !d_rank_in) {
      bins_pending = false;  // the gated bucket kernels found the path closed and did nothing
      // Speculatively the master-list path: the inverse of man 0's list (into
      // d_seen, free after the precheck), then one launch that walks the
      // women's lists as if every man held man 0's list, beside the scans that
      // decide whether he does (and check the women's rows).  The men's partner
      // words start unclaimed (the reset).
      SMP_HIP_TRY(hipEventRecord(ev[0], stream));
      SMP_HIP_TRY(pactae::launch_master_inverse(d_pref_m, n, d_seen, d_sc, stream));
      SMP_HIP_TRY(pactae::launch_master_fused(d_pref_m, d_pref_w, d_seen, d_partner, n, opts.validate, d_sc,
                                              dev.cus, stream));
      SMP_HIP_TRY(hipEventRecord(ev[2], stream));
      if (int st = read_scalars()) return st;
      master = h_sc->not_master == 0;
      master_both = master && h_sc->both_master != 0;
      if (master) {
        span_propose = {&ev[0], &ev[2]};
      } else {
        // Not a master list: the walk's partner words, counters and flags are
        // void; back to the state the reset left, and on to the general path.
        SMP_HIP_TRY(pactae::launch_reset_scalars(
            d_sc, std::min(n, propose_grid() * 4), stream, nullptr, 0, may_sparse ? d_bin_cnt : nullptr,
            counter_words(), thread_walk() ? nullptr : d_partner, (size_t)n));
      }
    }
    times.master_list = master ? 1 : 0;
    times.master_both = master_both ? 1 : 0;
    return SMP_OK;
  }

  // -- init: rank matrix + PR prefix gather (smp_engine_bamboo.cu:33-106) -----
  int InitProc() override {
    if (thread_walk() || (opts.schedule == SMP_SCHEDULE_FRONTIER && !master)) {
      // The reference's walks and the frontier: the women's rank matrix, and for
      // the PR-entry (LA) walks and the frontier the PR prefix gathered from it.
      blocks_copied = 0;
      matrices_valid = true;
      SMP_HIP_TRY(hipEventRecord(ev[0], stream));
      SMP_HIP_TRY(pactae::launch_invert_rows(d_pref_w, d_rank_w, n, invert_chunk(), opts.validate, d_sc, stream,
                                             /*nontemporal=*/1));
      rank_built = true;
      times.rank_matrix = 1;
      if (opts.schedule == SMP_SCHEDULE_LA_CAS || opts.schedule == SMP_SCHEDULE_LA_MIN ||
          opts.schedule == SMP_SCHEDULE_FRONTIER) {
        SMP_HIP_TRY(hipEventRecord(ev[1], stream));
        SMP_HIP_TRY(pactae::launch_gather_pr(d_pref_m, d_rank_w, d_pr, n, K, d_sc, stream));
        span_rank = {&ev[0], &ev[1]};
        span_gather = {&ev[1], &ev[2]};
      } else {
        span_rank = {&ev[0], &ev[2]};
      }
      SMP_HIP_TRY(hipEventRecord(ev[2], stream));
      return SMP_OK;
    }
    if (master) {
      // Master list: the precheck's launch has walked it already; neither the
      // rank matrix nor the PR prefix is needed.
      blocks_copied = 0;
      return SMP_OK;
    }
    const int chunk = invert_chunk();
    matrices_valid = true;
    sparse = use_sparse();
    times.sparse_init = sparse ? 1 : 0;
    if (d_rank_in) {
      // Ranks given: the prefix gather from them is the whole init.
      SMP_HIP_TRY(hipEventRecord(ev[1], stream));
      SMP_HIP_TRY(pactae::launch_gather_pr(d_pref_m, d_rank_in, d_pr, n, K, d_sc, stream));
      span_gather = {&ev[1], &ev[2]};
      return end_init();
    }
    // (The women's partner words were set to unclaimed by the precheck's reset.)
    if (sparse) {
      // Buckets of (man, position) per woman over the first K positions (queued
      // by the precheck unless it was off or saw all first choices equal), then
      // one pass over the women's rows that writes the PR prefix directly.  If a
      // bucket overflowed, that pass writes the full rank matrix instead; the
      // proposal kernel then does nothing, and CoreProc gathers the prefix from
      // the matrix and runs the proposal phase again.
      const int cap = bucket_cap(n, K);
      if (!bins_pending) {
        if (int st = launch_buckets(/*gated=*/0)) return st;
      }
      bins_pending = false;
      SMP_HIP_TRY(hipEventRecord(ev[0], stream));
      span_bucket = {bucket_ev, &ev[0]};
      SMP_HIP_TRY(pactae::launch_rank_prefix(d_pref_w, d_rank_w, n, chunk, opts.validate, d_bucket_cnt, d_bucket, cap,
                                             d_bucket_rank, d_pr, d_sc, stream));
      span_rank = {&ev[0], &ev[2]};
    } else {
      SMP_HIP_TRY(hipEventRecord(ev[0], stream));
      SMP_HIP_TRY(pactae::launch_invert_rows(d_pref_w, d_rank_w, n, chunk, opts.validate, d_sc, stream,
                                             /*nontemporal=*/1));
      rank_built = true;
      times.rank_matrix = 1;
      SMP_HIP_TRY(hipEventRecord(ev[1], stream));
      SMP_HIP_TRY(pactae::launch_gather_pr(d_pref_m, d_rank_w, d_pr, n, K, d_sc, stream));
      span_rank = {&ev[0], &ev[1]};
      span_gather = {&ev[1], &ev[2]};
    }
    // No wait here: the proposal kernel does nothing when init flagged malformed
    // lists, and the error is reported after it.
    return end_init();
  }

  // The end of init (ev[2]) and, for the host tail, the PR prefix copied to the
  // host beside the proposal kernel: one 32-position block per copy (block 0,
  // where nearly all late proposals fall, first; the others alternate between
  // two streams), each with an event so the host can cache a block as soon as it
  // lands.
  int end_init() {
    SMP_HIP_TRY(hipEventRecord(ev[2], stream));
    blocks_copied = 0;
    blocks_ready = 0;
    return SMP_OK;
  }

  // The PR prefix copies to the host on the copy streams, each waiting for the
  // end of init (ev[2]).  Queued after the proposal kernel and the state export
  // (not before them): a copy stream that shares a hardware queue with the
  // engine's stream then waits behind the kernel, which check_copy_overlap sees
  // and repairs, instead of holding the kernel back until its copy is done.
  int queue_prefix_copies() {
    if (T > 0 && opts.prefix_copy == 0) {
      const size_t blk = (size_t)n * pactae::kPrBlock;
      const int nb = (K + pactae::kPrBlock - 1) / pactae::kPrBlock;
      SMP_HIP_TRY(hipStreamWaitEvent(copy_stream, ev[2], 0));
      SMP_HIP_TRY(hipStreamWaitEvent(copy_stream2, ev[2], 0));
      // When each copy stream starts its first copy (check_copy_overlap).
      SMP_HIP_TRY(hipEventRecord(ev[7], copy_stream));
      SMP_HIP_TRY(hipEventRecord(ev[8], copy_stream2));
      for (int b = 0; b < nb && b < kMaxBlocks; ++b) {
        hipStream_t cs = b == 0 ? copy_stream : copy_stream2;
        SMP_HIP_TRY(hipMemcpyAsync(h_pr + b * blk, d_pr + b * blk, blk * sizeof(uint32_t),
                                   hipMemcpyDeviceToHost, cs));
        SMP_HIP_TRY(hipEventRecord(ev_blk[b], cs));
        blocks_copied = b + 1;
      }
    }
    return SMP_OK;
  }

  // -- core: persistent proposal kernel + host tail --------------------------
  pactae::ProposeArgs propose_args(int T_, const int2* starts, int nstarts) const {
    pactae::ProposeArgs a;
    a.pr = d_pr;
    a.pref_m = d_pref_m;
    a.rank_w = d_rank_w;
    a.rank32 = d_rank_in;
    a.partner = d_partner;
    a.parked = d_parked;
    a.sc = d_sc;
    a.wstats = d_wstats;
    a.n = n;
    a.K = K;
    a.T = T_;
    a.coherent_hint = opts.coherent_hint;
    a.poll = opts.poll_interval > 0 ? opts.poll_interval : 16;
    a.starts = starts;
    a.nstarts = nstarts;
    a.rank_sparse = rank_built || d_rank_in ? 0 : 1;
    a.deep_limit = deep_limit;
    // Men start in the order of woman 0's list (opts.start_order 0): when the
    // women's lists share a common order (the reference's RANDOM lists share
    // one in groups of 5), the men the women prefer then propose first and
    // displace fewer men; for independent lists the order is as good as any.
    // The row is a permutation (the init validates it), and the matching does
    // not depend on the order.
    a.start_order = opts.start_order == 0 && !d_rank_in ? d_pref_w : nullptr;
    return a;
  }

  int CoreProc() override {
    host_finished = false;
    const int grid = propose_grid();
    if (thread_walk()) {
      SMP_HIP_TRY(pactae::launch_walk_thread(opts.schedule, d_pref_m, d_pref_w, d_rank_w, d_pr, K, d_seen, d_cas_next,
                                             n, d_sc, stream));
      SMP_HIP_TRY(hipEventRecord(ev[4], stream));
      span_propose = {&ev[2], &ev[4]};
      if (int st = read_scalars()) return st;
      if (int st = check_device_error("thread-per-man walk")) return st;
      times.handoff_free = 0;
      return SMP_OK;
    }
    if (opts.schedule == SMP_SCHEDULE_FRONTIER && !master) {
      pactae::FrontierArgs fa;
      fa.pr = d_pr;
      fa.pref_m = d_pref_m;
      fa.rank_w = d_rank_w;
      fa.partner = d_partner;
      fa.f0 = d_front;
      fa.f1 = d_front + n_alloc;
      fa.ctl = d_fctl;
      fa.sc = d_sc;
      fa.n = n;
      fa.K = K;
      SMP_HIP_TRY(pactae::launch_frontier(fa, dev.cus, stream));
      SMP_HIP_TRY(hipEventRecord(ev[4], stream));
      span_propose = {&ev[2], &ev[4]};
      pactae::FrontierCtl fc{};
      SMP_HIP_TRY(hipMemcpyAsync(&fc, d_fctl, sizeof(fc), hipMemcpyDeviceToHost, stream));
      if (int st = read_scalars()) return st;
      if (fc.abort) {
        pactae::set_error("frontier schedule: a grid barrier wait ran out");
        return SMP_ERR_TIMEOUT;
      }
      if (int st = check_device_error("frontier walk")) return st;
      times.steps = fc.rounds;
      times.handoff_free = 0;
      return SMP_OK;
    }
    if (master) {  // walked in the precheck's launch
      if (int st = check_device_error("women-proposing walk")) return st;
      times.handoff_free = 0;
      return SMP_OK;
    }
    // Deep chains: with the automatic width and K < n, a launch whose waves
    // make more than n^2 / 128 steps past the prefix parks every chain; the
    // host then gathers the whole PR matrix and resumes them from the shared
    // partner words (lists that share their top women, such as the
    // reference's RANDOM with groups of 5: its first solve at n = 20,000 takes
    // the whole matrix a few percent into the walk).
    resume_starts = nullptr;
    resume_n = 0;
    deep_limit = opts.prefix_cols <= 0 && opts.schedule == SMP_SCHEDULE_BAMBOO && K < n && n <= 32768
                     ? std::max<long long>(1, (long long)n * n / 128)
                     : 0;
    for (int run = 0;; ++run) {
      if (int st = propose_phase(grid)) return st;
      if (h_sc->restart && resume_starts == nullptr) {
        const int parked = h_sc->parked_count;
        times.deep_restart = 1;
        deep_limit = 0;
        SMP_HIP_TRY(hipMemcpyAsync(d_starts, d_parked, (size_t)parked * sizeof(int2), hipMemcpyDeviceToDevice,
                                   stream));
        if (grow_pr(n)) K = n;
        if (!d_rank_in) {
          if (int rc = ensure_rank_matrix()) return rc;
        }
        SMP_HIP_TRY(hipEventRecord(ev[1], stream));
        if (d_rank_in)
          SMP_HIP_TRY(pactae::launch_gather_pr(d_pref_m, d_rank_in, d_pr, n, K, d_sc, stream));
        else
          SMP_HIP_TRY(pactae::launch_gather_pr(d_pref_m, d_rank_w, d_pr, n, K, d_sc, stream));
        span_gather = {&ev[1], &ev[2]};
        SMP_HIP_TRY(pactae::launch_fill_i32(&d_sc->parked_count, 1, 0, stream));
        SMP_HIP_TRY(pactae::launch_fill_i32(&d_sc->restart, 1, 0, stream));
        SMP_HIP_TRY(pactae::launch_fill_i32(&d_sc->active, 1, std::min(parked, grid * 4), stream));
        if (int st = end_init()) return st;
        resume_starts = d_starts;
        resume_n = parked;
        if (parked > 0) continue;
      }
      if (run == 0 && sparse && !rank_built && h_sc->prefix_overflow != 0) {
        // A bucket overflowed: the sparse pass wrote the rank matrix instead of
        // the prefix and the proposal kernel did nothing.  Gather the prefix
        // from the matrix and run the phase again.
        rank_built = true;
        times.rank_matrix = 1;
        sparse_overflowed = true;
        // Lists that overflow a bucket name the same women near the top of many
        // men's lists (shared or correlated lists), and their chains end deeper
        // in the lists: with the automatic width, the prefix is widened to the
        // 256 columns the buffers hold, so that fewer chains walk past it and
        // are resumed (correlated lists, n = 1,500 to 20,000: 1.1 to 2.6 times
        // faster; below n = 1,024 the wider host copy costs more than it saves).
        // The engine keeps the width for its later solves of the same lists.
        if (opts.prefix_cols <= 0 && n >= 1024) K = std::min(K_alloc, std::max(K, 256));
        // Their tail is also long: the host walks each of its proposals, while
        // the GPU's many chains still find candidates fast.  With the automatic
        // threshold the host takes over at 16 free men instead of 64 (correlated
        // lists, alternating runs on one box: n = 20,000 three seeds 4.3-4.4 ->
        // 3.1, 1.8-1.9 -> 1.5-1.6 and 4.0 -> 2.9 ms, n = 8,000 0.79-0.88 -> 0.71
        // ms, n = 2,000 to 5,000 the same or a little faster; 8 was slower).
        if (T > 16 && opts.tail_threshold <= 0 && !(opts.tail_mode == SMP_TAIL_AUTO && tuned_T >= 0)) T = 16;
        take_span(span_rank, &times.rank_kernel_ms);
        SMP_HIP_TRY(hipEventRecord(ev[1], stream));
        SMP_HIP_TRY(pactae::launch_gather_pr(d_pref_m, d_rank_w, d_pr, n, K, d_sc, stream));
        span_gather = {&ev[1], &ev[2]};
        if (int st = end_init()) return st;
        continue;
      }
      break;
    }
    const int handoff = h_sc->parked_count;
    times.handoff_free = handoff;
    if (handoff > 0) {
      if (int st = host_tail()) return st;
    }
    return SMP_OK;
  }

  // The persistent proposal kernel from the start state (or resume list) and
  // what follows it on the engine's stream, up to the point where the host has
  // the control block (and, with a hand-off, the partner words and parked
  // chains).  ev[5] marks the end of the phase on the device.
  int propose_phase(int grid) {
    SMP_HIP_TRY(pactae::launch_propose(propose_args(T, resume_starts, resume_n), grid, stream));
    span_propose = {&ev[2], &ev[5]};
    state_copied = false;
    if (T == 0) {
      SMP_HIP_TRY(pactae::launch_fold_wstats(d_wstats, grid * 4, d_sc, stream));
      SMP_HIP_TRY(hipEventRecord(ev[5], stream));
    }
    if (T > 0 && opts.prefix_copy >= 1) {
      // Everything the host tail needs follows the kernel on the engine's own
      // stream: first the small state (control block, partner words, parked
      // chains, written to host memory by k_export_state), then the PR prefix
      // block by block, each with an event.
      export_seq = ++flag_seq;
      SMP_HIP_TRY(pactae::launch_export_state(d_sc, d_wstats, grid * 4, d_partner, d_parked, n, h_sc_dev,
                                              h_partner_dev, h_parked_dev, h_flags_dev + kFlagExport,
                                              export_seq, stream));
      SMP_HIP_TRY(hipEventRecord(ev[5], stream));
      const size_t blk = (size_t)n * pactae::kPrBlock;
      const int nb = (K + pactae::kPrBlock - 1) / pactae::kPrBlock;
      for (int b = 0; b < nb && b < kMaxBlocks; ++b) {
        // prefix_copy 2: blocks past the first go on the copy stream, so that any
        // the tail never reads cannot delay the engine's stream.
        hipStream_t cs = (b > 0 && opts.prefix_copy == 2) ? copy_stream : stream;
        if (b == 1 && cs == copy_stream) SMP_HIP_TRY(hipStreamWaitEvent(copy_stream, ev_blk[0], 0));
        SMP_HIP_TRY(hipMemcpyAsync(h_pr + b * blk, d_pr + b * blk, blk * sizeof(uint32_t),
                                   hipMemcpyDeviceToHost, cs));
        SMP_HIP_TRY(hipEventRecord(ev_blk[b], cs));
        blocks_copied = b + 1;
      }
      auto t0 = Clock::now();
      SMP_HIP_TRY(wait_flag(kFlagExport, export_seq));
      times.copy_wait_ms = ms_since(t0);
      state_copied = true;
      if (int st = check_device_error("proposal kernel")) return st;
      if (h_sc->parked_count > 0) {
        t0 = Clock::now();
        SMP_HIP_TRY(wait_event(ev_blk[0]));
        times.copy_wait_ms += ms_since(t0);
        blocks_ready = 1;
        t0 = Clock::now();
        touch_block(0, false);
        times.warm_ms = ms_since(t0);
      }
    } else if (T > 0) {
      // The state the host tail reads (control block, partner words, parked
      // chains) is written to host memory by k_export_state right behind the
      // kernel on its own stream, so that one wait covers all of it.
      export_seq = ++flag_seq;
      SMP_HIP_TRY(pactae::launch_export_state(d_sc, d_wstats, grid * 4, d_partner, d_parked, n, h_sc_dev,
                                              h_partner_dev, h_parked_dev, h_flags_dev + kFlagExport,
                                              export_seq, stream));
      SMP_HIP_TRY(hipEventRecord(ev[5], stream));
      if (int st = queue_prefix_copies()) return st;
      // While the kernel runs: pull the PR prefix blocks into the host caches as
      // they land (a DMA-written buffer starts out in DRAM only).  Block 0 always;
      // later blocks only while the kernel is still running.
      bool kernel_done = false;
      for (int b = 0; b < blocks_copied && !kernel_done; ++b) {
        if (b > 0 && flag_seen(kFlagExport, export_seq)) break;
        auto t0 = Clock::now();
        SMP_HIP_TRY(wait_event(ev_blk[b]));
        times.copy_wait_ms += ms_since(t0);
        blocks_ready = b + 1;
        t0 = Clock::now();
        kernel_done = !touch_block(b, /*stop_when_done=*/b > 0);
        times.warm_ms += ms_since(t0);
      }
      auto t0 = Clock::now();
      SMP_HIP_TRY(wait_flag(kFlagExport, export_seq));
      times.copy_wait_ms += ms_since(t0);
      state_copied = true;
      if (int st = check_device_error("proposal kernel")) return st;
    }
    if (!state_copied) {
      if (int st = read_scalars()) return st;
      if (int st = check_device_error("proposal kernel")) return st;
    }
    return SMP_OK;
  }

  // Reads one line in every 64 bytes of block b.  With stop_when_done it checks
  // every 256 KB whether the proposal kernel has finished and then stops (the
  // host tail must not wait for a warm-up); returns false if it stopped early.
  bool touch_block(int b, bool stop_when_done) {
    const size_t blk = (size_t)n * pactae::kPrBlock;
    const uint64_t* p = reinterpret_cast<const uint64_t*>(h_pr + b * blk);
    const size_t words = blk / 2, chunk = 32768;
    uint64_t acc = 0;
    for (size_t c = 0; c < words; c += chunk) {
      if (stop_when_done && flag_seen(kFlagExport, export_seq)) {
        touch_sink = acc;
        return false;
      }
      const size_t e = std::min(words, c + chunk);
      for (size_t i = c; i < e; i += 8) acc += p[i];
    }
    touch_sink = acc;
    return true;
  }

  // Columns c0 .. c0 + cols - 1 of the PR matrix (those positions of every
  // man's list with their ranks, row after row: a man's band is contiguous)
  // gathered on the GPU into d_band and copied into h_band, for a host tail
  // whose chains walk deep into the lists; c0 = 0, cols = n is the whole matrix
  // (the reference copies its whole PR matrix to the host for the same
  // purpose, smp_engine_bamboo.h:41-61).  Sets *ok to false, and returns
  // SMP_OK, if the buffers cannot be allocated: the caller then keeps to
  // windows.
  int host_band(int c0, int cols, bool* ok) {
    *ok = false;
    const size_t entries = (size_t)n * cols;
    if (entries > band_cap) {
      (void)hipFree(d_band);
      (void)hipHostFree(h_band);
      d_band = h_band = nullptr;
      band_cap = 0;
      if (hipMalloc(&d_band, entries * sizeof(uint32_t)) != hipSuccess ||
          hipHostMalloc(&h_band, entries * sizeof(uint32_t), pactae::kHostAllocFlags) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(d_band);
        d_band = nullptr;
        h_band = nullptr;
        return SMP_OK;
      }
      band_cap = entries;
    }
    if (d_rank_in) {
      SMP_HIP_TRY(pactae::launch_gather_pr(d_pref_m, d_rank_in, d_band, n, cols, d_sc, stream, /*row_major=*/1, c0));
    } else {
      if (int rc = ensure_rank_matrix()) return rc;
      SMP_HIP_TRY(pactae::launch_gather_pr(d_pref_m, d_rank_w, d_band, n, cols, d_sc, stream, /*row_major=*/1, c0));
    }
    SMP_HIP_TRY(hipMemcpyAsync(h_band, d_band, entries * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    SMP_HIP_TRY(wait_stream(stream));
    if (cols == n) times.tail_full_matrix = 1;
    ++times.tail_bands;
    *ok = true;
    return SMP_OK;
  }

  // Host tail: McVitie-Wilson chains (smp_engine_mw_1_f9.cu:15-79, and the
  // reference's host LAProcedure smp_engine_bamboo.cu:197-230) over the partner
  // state the kernel left, kept as one packed word per woman (hst) so that it
  // stays in the host caches.  kTailSlots chains are walked interleaved: a chain
  // runs until its next acceptance, prefetches the row of the man it displaced
  // and yields to the next chain, so the row fetches of several chains overlap
  // (the stable matching does not depend on the order of proposals).  A chain
  // that walks past the host's PR prefix (or past its man's extension window)
  // stops and is collected; once every other chain is done, the GPU computes an
  // extension window for each collected man (the next entries of his list, by
  // searching the women's rows, or from the rank matrix when one exists), and
  // for the men matched near the end of what the host holds for them, and the
  // host continues the chains, in rounds until every chain has ended.  Only if
  // the window storage runs out are the collected chains handed back to the
  // GPU, which resumes them from the shared partner words with the full rank
  // matrix.
  int host_tail() {
    auto t0 = Clock::now();
    int parked = h_sc->parked_count;
    if (!state_copied) {
      SMP_HIP_TRY(hipMemcpyAsync(h_partner, d_partner, (size_t)n * sizeof(unsigned long long),
                                 hipMemcpyDeviceToHost, stream));
      SMP_HIP_TRY(hipMemcpyAsync(h_parked, d_parked, (size_t)parked * sizeof(int2),
                                 hipMemcpyDeviceToHost, stream));
      SMP_HIP_TRY(wait_stream(stream));
    }
    // The GPU finished alone (no prefix copy beside the kernel) but parked chains
    // at the end of a sparse prefix: bring the prefix over now, so that the host
    // walks the displaced men's chains from it instead of asking for a window
    // for each of them.
    {
      const int nb = std::min(kMaxBlocks, (std::min(K, n) + pactae::kPrBlock - 1) / pactae::kPrBlock);
      if (blocks_copied < nb) {
        for (; blocks_ready < blocks_copied; ++blocks_ready) SMP_HIP_TRY(wait_event(ev_blk[blocks_ready]));
        const size_t blk = (size_t)n * pactae::kPrBlock;
        SMP_HIP_TRY(hipMemcpyAsync(h_pr + blocks_copied * blk, d_pr + blocks_copied * blk,
                                   (size_t)(nb - blocks_copied) * blk * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                   stream));
        SMP_HIP_TRY(wait_stream(stream));
        blocks_copied = blocks_ready = nb;
      }
    }
    times.handoff_ms = ms_since(t0);

    t0 = Clock::now();
    int kcap = std::min(std::min(K, n), blocks_copied * pactae::kPrBlock);  // host-side prefix
    hst.resize(n);
    hrank.resize(n);
    deep_w.clear();
    for (int w = 0; w < n; ++w) {
      const unsigned long long v = h_partner[w];
      if (v == pactae::kEmpty) {
        hst[w] = kHostFree;
        hrank[w] = (uint16_t)kHostFree;
      } else {
        hst[w] = hs_pack(pactae::pw_rank(v), pactae::pw_man(v), pactae::pw_pos(v));
        hrank[w] = (uint16_t)pactae::pw_rank(v);
        if ((int)pactae::pw_pos(v) + 1 + kWinAhead >= kcap) deep_w.push_back(w);
      }
    }
    for (int m : win_men) win_off[m] = -1, win_len[m] = 0;  // windows of an earlier solve
    win_men.clear();
    win_used = 0;
    long long props = 0;
    int bailed = 0;
    const uint32_t* pr = h_pr;  // the prefix, column-blocked (pr_index)
    uint64_t* st = hst.data();
    uint16_t* sr = hrank.data();
    // A band the host holds: positions kb0 .. kb1 - 1 of every man, row-major
    // with bw = kb1 - kb0 entries per man (none while kb1 = 0).
    const uint32_t* band = nullptr;
    int kb0 = 0, kb1 = 0, bw = 0;
    int hold = kcap;  // positions below it are held for every man (prefix or band)
    int2* bail = h_parked;  // stopped chains overwrite the consumed parked list (bailed <= taken)
    int cm[kTailSlots], ck[kTailSlots];
    // Walks the `parked` chains of h_parked; the chains that leave the entries
    // the host holds are collected in bail[0..bailed).  MODE 0: the prefix and
    // windows only; 1: a band beside them; 2: the band is the whole matrix (the
    // loop then has one source of entries).
    auto walk_impl = [&](auto mode_tag) -> int {
      constexpr int MODE = decltype(mode_tag)::value;
      constexpr bool FULL = MODE == 2, BAND = MODE == 1;
      int act = 0, taken = 0;
      bailed = 0;
      // What the host holds does not change during a walk: keep it in locals.
      const uint32_t* const bd = band;
      const int b0 = kb0, b1 = kb1, bs = bw, kc = kcap, hd = hold;
      // Puts parked chain `taken` into slot s; false if it names a man outside [0,n).
      auto take = [&](int sl) {
        cm[sl] = h_parked[taken].x;
        ck[sl] = h_parked[taken].y;
        ++taken;
        return (unsigned)cm[sl] < (unsigned)n;
      };
      for (; act < kTailSlots && taken < parked; ++act)
        if (!take(act)) return tail_error("host tail: parked man outside [0,n)");
      int i = 0;
      while (act > 0) {
        if (i >= act) i = 0;
        int m = cm[i], k = ck[i];
        bool ended = false;  // chain reached a free woman, or left the entries the host holds
        for (;;) {
          // Eight positions of one 32-entry block of the prefix (or of the band)
          // at a time, without branches: the women's words are eight
          // independent loads, and the first position where the man beats the
          // woman's partner is found from a bit mask.  Positions before it are
          // refusals; the candidate itself goes through the step below.
          const bool in_band = FULL ? k + 8 <= n : BAND && k >= b0 && k + 8 <= b1;
          if (in_band || (!FULL && k + 8 <= kc && (k & 31) <= 24 && (k >> 5) < blocks_ready)) {
            const uint32_t* p8 = FULL      ? bd + (size_t)m * n + k
                                 : in_band ? bd + (size_t)m * bs + (k - b0)
                                           : pr + pactae::pr_index(m, k, n);
            uint32_t cand = 0, range = 0;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
              const uint32_t e8 = p8[j];
              const uint32_t w8 = pactae::pr_woman(e8);
              range |= (uint32_t)(w8 >= (uint32_t)n) << j;
              const uint32_t wc = w8 < (uint32_t)n ? w8 : 0u;
              cand |= (uint32_t)(pactae::pr_rank(e8) < (uint32_t)sr[wc]) << j;
            }
            const uint32_t stop = cand | range;  // a malformed entry goes to the step below
            if (stop == 0) {
              props += 8;
              k += 8;
              continue;
            }
            const int j0 = __builtin_ctz(stop);
            props += j0;
            k += j0;
          }
          uint32_t e;
          if (FULL) {
            if (k >= n) return tail_error("host tail: a man exhausted his list");
            e = bd[(size_t)m * n + k];
          } else if (BAND && k >= b0 && k < b1) {
            e = bd[(size_t)m * bs + (k - b0)];
          } else if (k < kc) {
            if ((k >> 5) >= blocks_ready) {  // a block the warm-up did not reach
              while (blocks_ready <= (k >> 5)) SMP_HIP_TRY(wait_event(ev_blk[blocks_ready++]));
            }
            e = pr[pactae::pr_index(m, k, n)];
          } else if (k >= n) {
            return tail_error("host tail: a man exhausted his list");
          } else {
            const int wo = win_off[m];
            const int off = k - win_start[m];
            if (wo < 0 || off < 0 || off >= win_len[m]) {
              bail[bailed++] = make_int2(m, k);
              ended = true;
              break;
            }
            e = h_win[(size_t)wo + off];
            if (e == pactae::kWinInvalid) return tail_error("host tail: malformed list entry");
          }
          ++props;
          const uint32_t w = pactae::pr_woman(e), r = pactae::pr_rank(e);
          if (w >= (uint32_t)n) return tail_error("host tail: preference entry outside [0,n)");
          const uint64_t v = st[w];
          if (r < (uint32_t)(v & 0xFFFFu)) {
            st[w] = hs_pack(r, (uint32_t)m, (uint32_t)k);
            sr[w] = (uint16_t)r;
            if (!FULL && k + 1 + kWinAhead >= hd) deep_w.push_back((int)w);  // he may soon leave what the host holds
            if ((v & 0xFFFFu) == kHostFree) {
              ended = true;
              break;
            }
            m = (int)((v >> 16) & 0xFFFFu);
            k = (int)(v >> 32) + 1;
            if (FULL) __builtin_prefetch(bd + (size_t)m * n + k);
            else if (BAND && k >= b0 && k < b1) __builtin_prefetch(bd + (size_t)m * bs + (k - b0));
            else if (k < kc) __builtin_prefetch(pr + pactae::pr_index(m, k, n));
            break;
          }
          ++k;
        }
        if (!ended) {
          cm[i] = m;
          ck[i] = k;
          ++i;
        } else if (taken < parked) {
          if (!take(i)) return tail_error("host tail: parked man outside [0,n)");
        } else {
          --act;
          cm[i] = cm[act];
          ck[i] = ck[act];
        }
      }
      return SMP_OK;
    };
    auto walk = [&]() -> int {
      if (kb1 == 0) return walk_impl(std::integral_constant<int, 0>{});
      if (kb0 == 0 && kb1 == n) return walk_impl(std::integral_constant<int, 2>{});
      return walk_impl(std::integral_constant<int, 1>{});
    };
    int rounds = 0, band_round = 0, max_stop = 0;
    long long searched = 0;  // window entries found by searching rows
    for (;;) {
      if (int rc = walk()) return rc;
      if (bailed == 0) break;
      for (int b = 0; b < bailed; ++b) max_stop = std::max(max_stop, bail[b].y);
      // Chains that keep leaving what the host holds, one man at a time, cost a
      // GPU round trip each: after enough rounds the host takes a band of every
      // man's list instead, wide enough for the deepest position a chain has
      // stopped at and wider each time, and finally the whole matrix.
      if (kb1 < n) {
        int cols = std::max(bw ? 2 * bw : std::max(kcap, 256), (max_stop + 64 - kcap + 255) / 256 * 256);
        int c0 = kcap;
        if (cols > n / 4 || kcap + cols >= n) c0 = 0, cols = n;  // chains that walk this deep: the whole matrix
        // A partial band against the rounds since the last one, the whole
        // matrix against every round of this solve.
        if ((c0 > 0 ? rounds - band_round : rounds) < band_rounds(cols)) c0 = -1;
        if (c0 >= 0) {
          bool ok = false;
          if (int rc = host_band(c0, cols, &ok)) return rc;
          band = ok ? h_band : nullptr;
          kb0 = ok ? c0 : 0;
          kb1 = ok ? c0 + cols : 0;
          bw = ok ? cols : 0;
          hold = std::max(kcap, kb1);
          band_round = rounds;
          if (ok) {
            parked = bailed;  // the stopped chains go round again, over the band
            continue;
          }
        }
      }
      // Windows for the stopped chains, each starting where its chain stopped;
      // out of window storage, the chains go back to the GPU.
      const bool have_rank = rank_built || d_rank_in;
      int count = 0;
      long long total = 0;
      auto next_len = [&](int m, int start) {
        const int len = win_len[m] == 0 ? kWinFirst : std::min(kWinMax, 2 * win_len[m]);
        return std::max(1, std::min(len, n - start));
      };
      auto request = [&](int m, int start, int len) {
        if (win_len[m] == 0) win_men.push_back(m);
        h_winreq[count++] = make_int4(m, start, (int)total, len);
        win_off[m] = (int)(win_used + total);
        win_start[m] = start;
        win_len[m] = len;
        total += len;
      };
      {
        long long need = 0;
        for (int b = 0; b < bailed; ++b) need += next_len(bail[b].x, bail[b].y);
        if (win_used + need > window_cap()) break;
        for (int b = 0; b < bailed; ++b) request(bail[b].x, bail[b].y, next_len(bail[b].x, bail[b].y));
      }
      // Men who may soon leave the entries the host holds for them, while the
      // storage lasts (each man is matched to one woman, and a stopped chain's
      // man to none, so no man is asked for twice); only with a rank matrix,
      // where a window entry is one gather rather than a search of a row.
      if (have_rank) {
        ++deep_round;
        for (int w : deep_w) {
          if (deep_stamp[w] == deep_round) continue;
          deep_stamp[w] = deep_round;
          const uint64_t v = st[w];
          if ((v & 0xFFFFu) == kHostFree) continue;
          const int m = (int)((v >> 16) & 0xFFFFu);
          const int start = std::max((int)(v >> 32) + 1, hold);
          if (start >= n) continue;
          if (win_len[m] > 0 && start >= win_start[m] && start + kWinAhead < win_start[m] + win_len[m]) continue;
          const int len = next_len(m, start);
          if (win_used + total + len > window_cap()) break;
          request(m, start, len);
        }
        deep_w.clear();
      }
      // Search the women's rows (4 n bytes per entry) while the searches of
      // this solve together read less than building the rank matrix would
      // (6 n^2 bytes); then build it once and gather.
      if (!have_rank) {
        if ((searched + total) * 4 > 6ll * n) {
          if (int rc = ensure_rank_matrix()) return rc;
        } else {
          searched += total;
        }
      }
      const bool gather = rank_built || d_rank_in;
      const unsigned seq = ++flag_seq;
      SMP_HIP_TRY(pactae::launch_window(d_pref_m, d_pref_w, gather && !d_rank_in ? d_rank_w : nullptr, d_rank_in, n,
                                        h_winreq_dev, count, (int)total, h_win_dev + win_used, d_sc,
                                        h_flags_dev + kFlagWindow, seq, stream));
      SMP_HIP_TRY(wait_flag(kFlagWindow, seq));
      win_used += (size_t)total;
      ++rounds;
      times.tail_windows += count;
      parked = bailed;  // the stopped chains, in h_parked[0..bailed), go round again
    }
    times.tail_rounds += rounds;
    times.host_proposals = props;
    times.bailed = bailed;
    if (bailed == 0) {
      // Every chain finished on the host: assemble the matching here.
      long long total = 0;
      for (int w = 0; w < n; ++w) {
        const uint64_t v = st[w];
        if ((v & 0xFFFFu) == kHostFree) return tail_error("host tail: a woman was left unmatched");
        h_match[(v >> 16) & 0xFFFFu] = w;
        total += (long long)(v >> 32) + 1;
      }
      host_proposals_total = total;
      host_finished = true;
      times.host_tail_ms = ms_since(t0);
      return SMP_OK;
    }
    for (int w = 0; w < n; ++w) {
      const uint64_t v = st[w];
      h_partner[w] = (v & 0xFFFFu) == kHostFree
                         ? pactae::kEmpty
                         : pactae::pw_pack((uint32_t)(v & 0xFFFFu), (uint32_t)((v >> 16) & 0xFFFFu),
                                           (uint32_t)(v >> 32));
    }
    times.host_tail_ms = ms_since(t0);

    // Resume the handed-back chains on the GPU (no hand-off this time), with the
    // rank matrix for the positions past the prefix.
    t0 = Clock::now();
    if (int rc = ensure_rank_matrix()) return rc;
    SMP_HIP_TRY(hipMemcpyAsync(d_partner, h_partner, (size_t)n * sizeof(unsigned long long),
                               hipMemcpyHostToDevice, stream));
    SMP_HIP_TRY(hipMemcpyAsync(d_starts, bail, (size_t)bailed * sizeof(int2), hipMemcpyHostToDevice,
                               stream));
    const int grid = std::max(1, std::min(propose_grid(), (bailed + 3) / 4));
    SMP_HIP_TRY(pactae::launch_fill_i32(&d_sc->active, 1, std::min(bailed, grid * 4), stream));
    // The resumed chains must all finish here: no hand-off and no deep-list
    // restart (a restart would park chains that nothing resumes).
    pactae::ProposeArgs ra = propose_args(0, d_starts, bailed);
    ra.deep_limit = 0;
    SMP_HIP_TRY(pactae::launch_propose(ra, grid, stream));
    SMP_HIP_TRY(pactae::launch_fold_wstats(d_wstats, grid * 4, d_sc, stream));
    SMP_HIP_TRY(wait_stream(stream));
    times.resume_ms = ms_since(t0);
    return SMP_OK;
  }

  // -- post (smp_engine_bamboo.cu:289-306) ----------------------------------
  int PostProc(bool perfect) override {
    match_on_host_only = false;
    if (host_finished) {
      // The host holds the matching (the result the reference returns); the
      // device copy is made only if a caller asks for it (device_match()).
      match_on_host_only = true;
      times.proposals = host_proposals_total;
      times.atomics = (int64_t)h_sc->atomics;
      times.steps = (int64_t)h_sc->steps;
      return SMP_OK;
    }
    if (perfect) {
      SMP_HIP_TRY(pactae::launch_first_match(d_pref_m, n, d_match, stream));
    } else if (thread_walk()) {
      SMP_HIP_TRY(pactae::launch_post_cas(d_pref_w, d_seen, d_cas_next, n, d_match, d_sc, stream));
    } else if (master_both) {
      SMP_HIP_TRY(pactae::launch_master_both(d_pref_m, d_pref_w, n, d_match, d_sc, stream));
    } else if (master) {
      SMP_HIP_TRY(pactae::launch_post_master(d_partner, d_seen, n, d_match, d_sc, stream));
    } else {
      SMP_HIP_TRY(pactae::launch_post(d_partner, n, d_match, d_sc, stream));
    }
    SMP_HIP_TRY(hipMemcpyAsync(h_match, d_match, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost,
                               stream));
    SMP_HIP_TRY(hipMemcpyAsync(h_sc, d_sc, sizeof(Scalars), hipMemcpyDeviceToHost, stream));
    SMP_HIP_TRY(wait_stream(stream));
    if (int st = check_device_error("post")) return st;
    if (perfect) {
      times.proposals = n;  // every man proposes once
    } else {
      times.proposals = (int64_t)h_sc->proposals;
      times.atomics = (int64_t)h_sc->atomics;
      if (opts.schedule != SMP_SCHEDULE_FRONTIER || master) times.steps = (int64_t)h_sc->steps;  // frontier: rounds
    }
    return SMP_OK;
  }

  // With the automatic width, lists on which the proposal kernel made more
  // than n^2 / 8 proposals (men walking on average more than n / 8 positions:
  // lists that share their top women, such as the reference's RANDOM with
  // groups of 5) take the whole PR matrix, the reference's layout, from their
  // next solve on.  Past the prefix each step of a chain otherwise gathers a
  // rank from the rank matrix after loading the woman from the list, two
  // dependent loads instead of one; the gather of the whole matrix (4 n^2
  // bytes written) is one parallel pass.  The reference's RANDOM, n = 20,000:
  // 108 -> 83 ms per solve; n = 4,000: 9.9 -> 7.8 ms.  Up to n = 32,768 (4 GiB
  // of PR matrix).
  bool deep_lists() const {
    if (opts.prefix_cols > 0 || opts.schedule != SMP_SCHEDULE_BAMBOO || master || K >= n || n > 32768)
      return false;
    const long long gpu_props = (long long)times.proposals - (long long)times.host_proposals;
    return gpu_props * 8 > (long long)n * n;
  }
  // d_pr room for `cols` columns; false (the old buffer kept) if it cannot be had.
  // The old buffer is freed only after the engine's streams are idle: prefix
  // copies the host tail did not wait for may still read it.
  bool grow_pr(int cols) {
    const size_t need = pactae::pr_entries(n_alloc, cols);
    if (need <= pr_cap) return true;
    uint32_t* p = nullptr;
    if (hipMalloc(&p, need * sizeof(uint32_t)) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    if (hipStreamSynchronize(stream) != hipSuccess || hipStreamSynchronize(copy_stream) != hipSuccess ||
        hipStreamSynchronize(copy_stream2) != hipSuccess) {
      (void)hipGetLastError();
      (void)hipFree(p);
      return false;
    }
    (void)hipFree(d_pr);
    d_pr = p;
    pr_cap = need;
    return true;
  }

  // The prefix copies must run beside the proposal kernel.  Streams share the
  // device's few hardware queues; when a copy stream lands on the compute
  // stream's queue, its copies wait for the kernel instead.  Seen from the
  // events, the copy streams are then replaced, which moves them to other
  // queues (at most 8 times per engine).  Both copy streams are checked.
  void check_copy_overlap() {
    // A copy stream that shares the engine's hardware queue starts its copies
    // only once the kernel and the export ahead of it in that queue are done:
    // its start event then comes at or after the end of the phase (ev[5]).  A
    // copy that merely outlasts the kernel starts right after init and is fine.
    float kern_ms = 0, s0 = 0, s1 = 0;
    if (hipEventQuery(ev[5]) != hipSuccess || hipEventQuery(ev[7]) != hipSuccess ||
        hipEventElapsedTime(&kern_ms, ev[2], ev[5]) != hipSuccess ||
        hipEventElapsedTime(&s0, ev[2], ev[7]) != hipSuccess) {
      (void)hipGetLastError();
      return;
    }
    if (blocks_copied > 1 && hipEventQuery(ev[8]) == hipSuccess &&
        hipEventElapsedTime(&s1, ev[2], ev[8]) == hipSuccess)
      s0 = std::max(s0, s1);
    (void)hipGetLastError();
    const bool waited = kern_ms > 0.05f && s0 >= 0.9f * kern_ms;
    times.copy_serialized = waited && requeues >= 8 ? 1 : 0;
    if (waited && requeues < 8) {
      (void)hipStreamSynchronize(copy_stream);
      (void)hipStreamSynchronize(copy_stream2);
      hipStream_t a = nullptr, b = nullptr;
      if (hipStreamCreateWithFlags(&a, hipStreamNonBlocking) == hipSuccess &&
          hipStreamCreateWithFlags(&b, hipStreamNonBlocking) == hipSuccess) {
        (void)hipStreamDestroy(copy_stream);
        (void)hipStreamDestroy(copy_stream2);
        copy_stream = a;
        copy_stream2 = b;
        ++requeues;
      } else {
        if (a) (void)hipStreamDestroy(a);
        (void)hipGetLastError();
      }
    }
  }

  // Init-only run: rank matrix, optionally the PR matrix, then the D2H copy of
  // the last matrix built (AbsSmpInitEngine::Init, smp_init_abs.h:16-28).
  int init_only(bool init_pr, smp_init_times* t) {
    *t = smp_init_times{};
    if (d_rank_in) {
      pactae::set_error("smp_init_exec builds the rank matrix from the women's lists; this engine was given ranks");
      return SMP_ERR_INVALID;
    }
    matrices_valid = true;
    sparse = false;
    rank_built = true;
    SMP_HIP_TRY(pactae::launch_reset_scalars(d_sc, std::min(n, propose_grid() * 4), stream));
    const int chunk = invert_chunk();
    auto t0 = Clock::now();
    SMP_HIP_TRY(hipEventRecord(ev[0], stream));
    SMP_HIP_TRY(pactae::launch_invert_rows(d_pref_w, d_rank_w, n, chunk, opts.validate, d_sc, stream,
                                           /*nontemporal=*/1));
    SMP_HIP_TRY(hipEventRecord(ev[1], stream));
    SMP_HIP_TRY(wait_stream(stream));
    t->rank_ms = ms_since(t0);
    if (init_pr) {
      t0 = Clock::now();
      SMP_HIP_TRY(pactae::launch_gather_pr(d_pref_m, d_rank_w, d_pr, n, K, d_sc, stream));
      SMP_HIP_TRY(hipEventRecord(ev[2], stream));
      SMP_HIP_TRY(wait_stream(stream));
      t->pr_ms = ms_since(t0);
    }
    if (int st = read_scalars()) return st;
    if (int st = check_device_error("init")) return st;
    t0 = Clock::now();
    std::vector<uint32_t> host;
    if (init_pr) {
      host.resize(pactae::pr_entries(n, K));
      SMP_HIP_TRY(hipMemcpy(host.data(), d_pr, host.size() * 4, hipMemcpyDeviceToHost));
    } else {
      host.resize(((size_t)n * n + 1) / 2);
      SMP_HIP_TRY(hipMemcpy(host.data(), d_rank_w, (size_t)n * n * sizeof(uint16_t), hipMemcpyDeviceToHost));
    }
    t->transfer_ms = ms_since(t0);
    t->total_ms = t->rank_ms + t->pr_ms + t->transfer_ms;
    float ms = 0;
    if (hipEventElapsedTime(&ms, ev[0], ev[1]) == hipSuccess) t->rank_kernel_ms = ms;
    if (init_pr && hipEventElapsedTime(&ms, ev[1], ev[2]) == hipSuccess) t->pr_kernel_ms = ms;
    t->pr_cols = K;
    return SMP_OK;
  }

  int exec() {
    int st = FindStableMatching();
    if (st) return st;
    times.precheck_ms = precheck_time_;
    times.init_ms = init_time_;
    times.core_ms = core_time_;
    times.post_ms = post_time_;
    times.total_ms = precheck_time_ + init_time_ + core_time_ + post_time_;
    times.prefix_cols = K;
    times.tail_threshold = T;
    full_pr_next = deep_lists();
    if (!times.is_perfect && T > 0 && blocks_copied > 0 && opts.prefix_copy == 0) check_copy_overlap();
    times.copy_requeues = requeues;
    take_span(span_bucket, &times.bucket_kernel_ms);
    take_span(span_rank, &times.rank_kernel_ms);
    take_span(span_gather, &times.gather_kernel_ms);
    take_span(span_propose, &times.propose_kernel_ms);
    have_result = true;
    return SMP_OK;
  }
};
