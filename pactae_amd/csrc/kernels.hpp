// kernels.hpp — host-side declarations of the kernel launchers in kernels.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "smp_internal.hpp"

namespace pactae {

struct ProposeArgs {
  const uint32_t* pr;
  const int32_t* pref_m;
  const uint16_t* rank_w;     // women's ranks, rank_w[w * n + m]
  const int32_t* rank32;      // or, when not null, a rank matrix given as input
  unsigned long long* partner;
  int2* parked;
  Scalars* sc;
  unsigned long long* wstats;  // per wave: [2*g] atomics, [2*g+1] segment steps
  int n, K, T;
  int coherent_hint;           // 1: partner hints bypass non-coherent caches
  int poll;                    // steps between reads of the shared `active` counter
  const int2* starts;          // resume mode: (man, position) pairs to start from;
  int nstarts;                 //   null = every man from position 0
  int rank_sparse;             // 1: after a sparse prefix pass, rank_w does not exist;
                               //   a chain reaching position K is parked at K for the
                               //   host, which builds rank_w.  If the pass overflowed
                               //   (sc->prefix_overflow), the kernel does nothing.
  const int32_t* start_order;  // not null: work item i is man start_order[i] (a
                               //   permutation of 0..n-1), else man i
  long long deep_limit;        // > 0: once the waves have made this many steps past
                               //   position K, every chain is parked and sc->restart
                               //   set (the host then resumes them over the whole PR
                               //   matrix); 0 = off
};

// Control block of the frontier schedule (device memory).
struct FrontierCtl {
  int cnt[3];     // list sizes, rotating by round % 3
  int bar_count;  // grid barrier: arrivals
  int bar_gen;    // grid barrier: generation
  int rounds;     // rounds run
  int abort;      // a bounded barrier wait ran out
  int pad;
};

struct FrontierArgs {
  const uint32_t* pr;       // PR prefix (positions < K)
  const int32_t* pref_m;    // positions >= K: pref_m and the rank matrix
  const uint16_t* rank_w;
  unsigned long long* partner;
  uint32_t* f0;             // free-man lists, n entries each: k << 16 | m
  uint32_t* f1;
  FrontierCtl* ctl;
  Scalars* sc;
  int n, K;
};

hipError_t launch_frontier(const FrontierArgs& a, int cus, hipStream_t s);

hipError_t launch_copy16(void* dst, const void* src, size_t bytes, int cus, hipStream_t s);
hipError_t launch_fill_u64(unsigned long long* p, size_t count, unsigned long long v, hipStream_t s);
hipError_t launch_fill_i32(int* p, size_t count, int v, hipStream_t s);
// Also zeroes z0[0..n0) and z1[0..n1) and sets partner[0..np) to kEmpty (null = none).
hipError_t launch_reset_scalars(Scalars* sc, int active, hipStream_t s, int* z0 = nullptr, size_t n0 = 0,
                                int* z1 = nullptr, size_t n1 = 0, unsigned long long* partner = nullptr,
                                size_t np = 0);
hipError_t launch_first_choices(const int32_t* pref_m, int n, int* seen, Scalars* sc, hipStream_t s);
// The master-list path in one launch: the women-proposing walk over linv (the
// inverse of man 0's list, from launch_master_inverse) into partner_m, beside the
// comparison of every men's list with man 0's (not_master set if one differs:
// the walk's result is then void) and, when check != 0, the permutation check of
// every women's row.
hipError_t launch_master_fused(const int32_t* pref_m, const int32_t* pref_w, const int32_t* linv,
                               unsigned long long* partner_m, int n, int check, Scalars* sc, int cus, hipStream_t s);
hipError_t launch_master_inverse(const int32_t* L, int n, int32_t* linv, Scalars* sc, hipStream_t s);
// Both sides hold master lists (every man man 0's list, every woman woman 0's):
// match[pref_w[k]] = pref_m[k], and sc->proposals += sum of (k + 1).
hipError_t launch_master_both(const int32_t* pref_m, const int32_t* pref_w, int n, int32_t* match, Scalars* sc,
                              hipStream_t s);
hipError_t launch_post_master(const unsigned long long* partner_m, const int32_t* linv, int n, int32_t* match,
                              Scalars* sc, hipStream_t s);
// The reference's thread-per-man walks (schedule SMP_SCHEDULE_MW_CAS, _LA_CAS,
// _MW_MIN or _LA_MIN) over 32-bit partner ranks and next[]; pr/K: the PR prefix
// the LA variants read (positions >= K from pref_m and rank_w).
hipError_t launch_walk_thread(int schedule, const int32_t* pref_m, const int32_t* pref_w, const uint16_t* rank_w,
                              const uint32_t* pr, int K, int* partner_rank, int* next, int n, Scalars* sc,
                              hipStream_t s);
hipError_t launch_post_cas(const int32_t* pref_w, const int* partner_rank, const int* next, int n, int32_t* match,
                           Scalars* sc, hipStream_t s);
hipError_t launch_first_match(const int32_t* pref_m, int n, int32_t* match, hipStream_t s);
hipError_t launch_invert_rows(const int32_t* pref, uint16_t* rank, int n, int chunk, int validate,
                              Scalars* sc, hipStream_t s, int nontemporal);
hipError_t set_invert_lds_limit(int bytes);
// row_major = 1: out[m * K + k] instead of the column-blocked pr_index layout,
// for list positions col0 + k (a band of K columns from col0; col0 = 0 unless
// row_major).
hipError_t launch_gather_pr(const int32_t* pref_m, const uint16_t* rank_w, uint32_t* pr, int n, int K,
                            Scalars* sc, hipStream_t s, int row_major = 0, int col0 = 0);
// The same from an int32 rank matrix given as input.
hipError_t launch_gather_pr(const int32_t* pref_m, const int32_t* rank_w, uint32_t* pr, int n, int K,
                            Scalars* sc, hipStream_t s, int row_major = 0, int col0 = 0);
// Sparse prefix pass: buckets of (m, k < K) per woman (bins of 256 women with
// bin_cap entries each and counters kBinCounterStride ints apart in bin_cnt,
// then per-woman buckets of cap entries and their counts in cnt; bin_cnt and
// cnt must be zero: the engine keeps them in one allocation that
// launch_reset_scalars clears), then the PR prefix from the women's rows without
// the n x n rank matrix (or the full rank matrix if a bin or bucket overflowed).
int bin_count(int n);
constexpr int kBinCounterStride = 32;  // bin_cnt holds 256 * kBinCounterStride ints
// gated = 1: the kernels do nothing unless the precheck (k_first_choices, queued
// before them) found two men sharing a first choice and not all first choices equal.
// stages: 1 = k_bin_prefix, 2 = k_bin_sort, 3 = both (in that order).  With
// h_flag, k_bin_prefix first publishes the control block to the host (h_sc, then
// seq into *h_flag), as launch_publish_scalars does.
hipError_t launch_bucket_prefix(const int32_t* pref_m, int n, int K, int* bin_cnt, unsigned long long* bins,
                                int bin_cap, int* cnt, uint32_t* bucket, int cap, int gated, Scalars* sc, int cus,
                                hipStream_t s, int stages = 3, Scalars* h_sc = nullptr, unsigned* h_flag = nullptr,
                                unsigned seq = 0);
// The row pass (ranks of each woman's bucket entries into brank, in bucket
// order; or the full rank matrix after an overflow), then k_scatter_prefix
// (brank into the PR prefix).
hipError_t launch_rank_prefix(const int32_t* pref_w, uint16_t* rank, int n, int chunk, int validate,
                              const int* cnt, const uint32_t* bucket, int cap, uint16_t* brank, uint32_t* pr,
                              Scalars* sc, hipStream_t s);
hipError_t launch_propose(const ProposeArgs& a, int grid, hipStream_t s);
// Extension windows for the host tail: req[i].w list entries (woman, rank) from
// position req[i].y of man req[i].x into out[req[i].z ..] (pinned host memory,
// device address), kWinInvalid past the list's end.  The requests' entries are
// back to back: req[0].z = 0, req[i + 1].z = req[i].z + req[i].w, and total is
// the sum of the lengths.  Ranks from rank32 or rank16 when given, else by
// searching the woman's row (pref_w).  The kernel then stores seq into *h_flag
// (pinned) once every entry is visible to the host.
hipError_t launch_window(const int32_t* pref_m, const int32_t* pref_w, const uint16_t* rank16, const int32_t* rank32,
                         int n, const int4* req, int count, int total, uint32_t* out, Scalars* sc, unsigned* h_flag,
                         unsigned seq, hipStream_t s);
hipError_t launch_fold_wstats(const unsigned long long* wstats, int waves, Scalars* sc, hipStream_t s);
// Folds the proposal waves' counters, then writes the control block, the partner
// words and the parked chains into host memory (device-visible pinned pointers)
// and, once all of it is out, seq into *h_flag (pinned).
hipError_t launch_export_state(Scalars* sc, const unsigned long long* wstats, int waves,
                               const unsigned long long* partner, const int2* parked, int n, Scalars* h_sc,
                               unsigned long long* h_partner, int2* h_parked, unsigned* h_flag, unsigned seq,
                               hipStream_t s);
// The control block into pinned host memory, then seq into *h_flag (pinned).
hipError_t launch_publish_scalars(const Scalars* sc, Scalars* h_sc, unsigned* h_flag, unsigned seq, hipStream_t s);
hipError_t launch_post(const unsigned long long* partner, int n, int32_t* match, Scalars* sc,
                       hipStream_t s);

}  // namespace pactae
