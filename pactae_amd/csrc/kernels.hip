// kernels.hip — gfx950 kernels of the single-instance proposal engine.
//
// Phases (reference counterparts, paths relative to the reference repository root):
//   k_first_choices / k_first_match   IsPerfect + perfect PostProc
//                                     include/smp/smp_engine_bamboo.h:25-33,
//                                     src/smp/smp_engine_bamboo.cu:291-297
//   k_invert_rows                     F9RankMatrixInitKernel src/smp/kernel.cu:672-686
//   k_gather_pr                       F9PRMatrixInitKernel   src/smp/kernel.cu:688-709
//   k_propose                         La3F9CoreKernel        src/smp/kernel.cu:633-670
//   k_post                            PostProc               src/smp/smp_engine_bamboo.cu:299-303
//   k_master_fused .. k_post_master   master-list path (every man holds the same list):
//                                     detection, women-proposing walk, matching
//   k_walk_thread / k_post_cas        the reference's thread-per-man schedules:
//                                     MW-Par-GPU-CAS src/smp/smp_engine_mw_3_f9.cu:37-97,
//                                     src/smp/kernel.cu:5-55, 122-165, 468-512, 515-552
//   k_frontier_init, k_frontier       frontier schedule (rounds, LDS pre-reduction)
//   k_window_search, k_window_gather  the host tail's extension windows
//   k_bin_prefix, k_bin_sort,         sparse prefix pass: F9RankMatrixInitKernel and
//   k_invert_rows<.., SP=1>           F9PRMatrixInitKernel restricted to the first K list
//                                     positions, without writing the n x n rank matrix
//
// All arithmetic is integer.  Every loop is bounded by n (list length) or by the
// grid size; preference entries are range-checked before they index anything.

#include <algorithm>

#include "kernels.hpp"

namespace pactae {

// ---------------------------------------------------------------------------
// Fills (used instead of hipMemsetAsync so the same sequence can be captured
// into a hipGraph).
__global__ void k_fill_u64(unsigned long long* p, size_t count, unsigned long long v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_fill_i32(int* p, size_t count, int v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Resets the control block, zeroes up to two int arrays (the precheck's
// first-choice counts, the sparse pass's bin counters) and sets the women's
// partner words to unclaimed, all in one launch.
__global__ void k_reset_scalars(Scalars* sc, int active, int* z0, size_t n0, int* z1, size_t n1,
                                unsigned long long* partner, size_t np) {
  const size_t nth = (size_t)gridDim.x * blockDim.x;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n0 + n1 + np; i += nth) {
    if (i < n0) z0[i] = 0;
    else if (i < n0 + n1) z1[i - n0] = 0;
    else partner[i - n0 - n1] = kEmpty;
  }
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    sc->active = active;
    sc->parked_count = 0;
    sc->prefix_overflow = 0;
    sc->error = 0;
    sc->dup_first = 0;
    sc->first_differs = 0;
    sc->not_master = 0;
    sc->export_blocks = 0;
    sc->restart = 0;
    sc->women_differ = 0;
    sc->scans_done = 0;
    sc->both_master = 0;
    sc->deep_steps = 0;
    sc->atomics = 0;
    sc->steps = 0;
    sc->proposals = 0;
  }
}

// ---------------------------------------------------------------------------
// Precheck: do all men have distinct first choices?  seen[] must be zero.
// The flags are combined per wave (ballot) before one atomic each: on uniform
// input thousands of men repeat a first choice, and one same-address atomic per
// man would serialize the kernel.  The lanes that name the same woman as lane 0
// count themselves with one atomic (the rest one each): when most men share a
// first choice (shared or correlated lists), 20,000 same-address atomics took
// 0.2-0.4 ms.
__global__ void k_first_choices(const int32_t* __restrict__ pref_m, int n, int* seen, Scalars* sc) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  bool range = false, dup = false, diff = false;
  const int w = m < n ? pref_m[(size_t)m * n] : -1;
  const bool valid = m < n && (unsigned)w < (unsigned)n;
  const int w0 = __shfl(w, 0);
  const unsigned long long grp = __ballot(valid && w == w0);  // lane 0 is in it when valid
  int grp_old = 0;
  if (lane == 0 && grp) grp_old = atomicAdd(&seen[w0], __popcll(grp));
  grp_old = __shfl(grp_old, 0);
  if (m < n) {
    if (!valid) {
      range = true;
    } else {
      const bool in_grp = (grp >> lane) & 1;
      dup = in_grp ? (grp_old != 0 || __popcll(grp) >= 2) : atomicAdd(&seen[w], 1) != 0;
      diff = w != pref_m[0];
    }
  }
  const unsigned long long br = __ballot(range), bd = __ballot(dup || range), bf = __ballot(diff);
  if ((threadIdx.x & 63) == 0) {
    if (br) atomicOr(&sc->error, kErrEntryRange);
    if (bd) atomicOr(&sc->dup_first, 1);
    if (bf) atomicOr(&sc->first_differs, 1);
  }
}

// Perfect case: every man keeps his first choice.
__global__ void k_first_match(const int32_t* __restrict__ pref_m, int n, int32_t* match) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < n) match[m] = pref_m[(size_t)m * n];
}

// ---------------------------------------------------------------------------
// Rank inversion, one workgroup per row: rank[w][pref[w][r]] = r.
// The row's inverse is assembled in LDS as 16-bit positions (n <= 65535, so a
// position fits and 0xFFFF is free to mark an empty slot), `chunk` columns per
// pass (one pass when the row fits: 40 KB at n = 20,000).  Each thread first
// issues up to kBlk 16-byte loads of the row and only then scatters them, so
// many loads are in flight per thread; the read of pref and the write of rank
// are both coalesced 16-byte accesses and only the LDS scatter is random.  With
// validate != 0 a slot left at 0xFFFF marks a row that is not a permutation.
//
// SP = 1 is the sparse prefix pass: instead of writing the row out, the
// workgroup looks up only the men whose first K list positions name this woman
// (her bucket, filled by k_bucket_prefix) and writes their PR prefix entries
// pr[m, k] = (w, rank) straight into the column-blocked prefix.  The n x n rank
// matrix is then never written (2n^2 bytes saved) and the separate gather is not
// needed.  If some bucket overflowed (sc->prefix_overflow), every row is written
// out in full as with SP = 0, and the engine builds the prefix with k_gather_pr
// once the host has seen the flag.
// Measured at n = 20,000: 1024 > 512 ≈ 256 threads per row for the full
// inversion; for the sparse pass 1024 x 8 loads ≈ 768 x 7 < 512 x 10 < 640 x 8.
constexpr int kInvBlock = 1024;
constexpr int kBlk = 8;
constexpr int kBlkShort = 5;  // k_rank_prefix_row: rows of at most 20,480 entries


struct PrefixBuckets {
  const int* cnt;          // suitors per woman (may exceed cap: overflow)
  const uint32_t* bucket;  // bucket[w * cap + j] = k << 16 | m
  int cap;
  uint16_t* brank;         // brank[w * cap + j] = rank of bucket[w * cap + j]'s man
};

// Two 1024-thread workgroups per CU (at most 64 VGPRs): with one, half the row
// loads the CU could keep in flight are missing.
template <int VEC, int NT, int SP, int BS = kInvBlock, int KB = kBlk>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_invert_rows(const int32_t* __restrict__ pref,
                                                           uint16_t* __restrict__ rank, int n,
                                                           int chunk, int validate, Scalars* sc,
                                                           PrefixBuckets pb) {
  extern __shared__ uint16_t lds16[];
  uint32_t* lds32 = reinterpret_cast<uint32_t*>(lds16);
  const int tid = threadIdx.x;
  int bad = 0;
  // Sparse pass: look up the bucket entries unless a bucket overflowed.
  // (The flag is read here and tested after the scatter, so the row's loads do
  // not wait for it.)
  int overflow = SP ? __hip_atomic_load(&sc->prefix_overflow, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 1;
  {  // one row per workgroup: the grid is n workgroups
    const int row = blockIdx.x;
    const int32_t* src = pref + (size_t)row * n;
    uint16_t* dst = rank + (size_t)row * n;

    for (int c0 = 0; c0 < n; c0 += chunk) {
      const int clen = min(chunk, n - c0);
      if (validate)
        for (int c = tid; c < (clen + 1) / 2; c += BS) lds32[c] = 0xFFFFFFFFu;
      __syncthreads();
      if constexpr (VEC == 4) {
        const int4* src4 = reinterpret_cast<const int4*>(src);
        const int nq = n >> 2;
        for (int q0 = tid; q0 < nq; q0 += BS * KB) {
          int4 v[KB];
          // Loads past the row's end are clamped to its last vector (and not
          // scattered), so every load is unconditional and all KB are issued
          // before the first wait.
#pragma unroll
          for (int j = 0; j < KB; ++j) {
            // A 32-bit byte offset from the row's (uniform) base: one VGPR per
            // address instead of two (a row is at most 256 KB).
            const uint32_t off = (uint32_t)min(q0 + j * BS, nq - 1) << 4;
            const int4* a = reinterpret_cast<const int4*>(reinterpret_cast<const char*>(src4) + off);
            if constexpr (NT) {
              v[j].x = __builtin_nontemporal_load(&a->x);
              v[j].y = __builtin_nontemporal_load(&a->y);
              v[j].z = __builtin_nontemporal_load(&a->z);
              v[j].w = __builtin_nontemporal_load(&a->w);
            } else {
              v[j] = *a;
            }
          }
#pragma unroll
          for (int j = 0; j < KB; ++j) {
            const int q = q0 + j * BS;
            if (q < nq) {
              const int r = q << 2;
              const int e[4] = {v[j].x, v[j].y, v[j].z, v[j].w};
#pragma unroll
              for (int t = 0; t < 4; ++t) {
                const unsigned u = (unsigned)(e[t] - c0);
                if (u < (unsigned)clen) lds16[u] = (uint16_t)(r + t);
                else bad |= (unsigned)e[t] >= (unsigned)n;
              }
            }
          }
        }
      } else {
        for (int r0 = tid; r0 < n; r0 += BS * KB) {
          int v[KB];
#pragma unroll
          for (int j = 0; j < KB; ++j) {
            const int r = r0 + j * BS;
            if (r < n) v[j] = src[r];
          }
#pragma unroll
          for (int j = 0; j < KB; ++j) {
            const int r = r0 + j * BS;
            if (r >= n) break;
            const unsigned u = (unsigned)(v[j] - c0);
            if (u < (unsigned)clen) lds16[u] = (uint16_t)r;
            else bad |= (unsigned)v[j] >= (unsigned)n;
          }
        }
      }
      __syncthreads();
      // Opaque here, or the compiler tests the flag ahead of the row's loads.
      if constexpr (SP) asm volatile("" : "+v"(overflow));
      const bool lookup = SP && overflow == 0;
      if (lookup) {
        // Prefix entries of this woman's suitors whose man falls in this chunk.
        const int cntw = min(pb.cnt[row], pb.cap);
        const uint32_t* bk = pb.bucket + (size_t)row * pb.cap;
        for (int j = tid; j < cntw; j += BS) {
          const uint32_t e = bk[j];
          const unsigned u = (e & 0xFFFFu) - (unsigned)c0;
          if (u < (unsigned)clen) pb.brank[(size_t)row * pb.cap + j] = lds16[u];
        }
        if (validate) {  // a hole in the LDS row marks a row that is not a permutation
          for (int c = tid; c < (clen + 1) / 2; c += BS) {
            const uint32_t x = lds32[c];
            if ((x & 0xFFFFu) == 0xFFFFu || (2 * c + 1 < clen && (x >> 16) == 0xFFFFu)) bad |= 2;
          }
        }
      } else if constexpr (VEC == 4) {
        // The LDS row is already the 16-bit output: 16-byte stores (8 ranks) when
        // n is a multiple of 8 (rows then start 16-byte aligned), 8-byte stores
        // otherwise (n is a multiple of 4 here).
        typedef unsigned int v4u __attribute__((ext_vector_type(4)));
        typedef unsigned int v2u __attribute__((ext_vector_type(2)));
        auto has_hole = [](unsigned x) { return (x & 0xFFFFu) == 0xFFFFu || (x >> 16) == 0xFFFFu; };
        if ((n & 7) == 0) {
          const v4u* l4 = reinterpret_cast<const v4u*>(lds16);
          v4u* dst8 = reinterpret_cast<v4u*>(dst + c0);
          for (int q = tid; q < (clen >> 3); q += BS) {
            const v4u p = l4[q];
            if (validate && (has_hole(p.x) || has_hole(p.y) || has_hole(p.z) || has_hole(p.w))) bad |= 2;
            if constexpr (NT) __builtin_nontemporal_store(p, dst8 + q);
            else dst8[q] = p;
          }
        } else {
          const v2u* l2 = reinterpret_cast<const v2u*>(lds16);
          v2u* dst4 = reinterpret_cast<v2u*>(dst + c0);
          for (int q = tid; q < (clen >> 2); q += BS) {
            const v2u p = l2[q];
            if (validate && (has_hole(p.x) || has_hole(p.y))) bad |= 2;
            if constexpr (NT) __builtin_nontemporal_store(p, dst4 + q);
            else dst4[q] = p;
          }
        }
      } else {
        for (int c = tid; c < clen; c += BS) {
          const uint16_t v = lds16[c];
          if (validate && v == 0xFFFFu) bad |= 2;
          dst[c0 + c] = v;
        }
      }
      __syncthreads();
    }
  }
  if (bad & 1) atomicOr(&sc->error, kErrEntryRange);
  if (bad & 2) atomicOr(&sc->error, kErrNotPerm);
}

// ---------------------------------------------------------------------------
// Sparse prefix pass for rows that one round of loads covers and the LDS holds
// whole (n <= BS * KB * 4; 20,480 with 1024 threads x 5 loads): the work of
// k_invert_rows<4, 1, 1> for one row per workgroup, as straight-line code.
// Every load is issued first (the row's, then the woman's bucket count and
// first BS entries), so they are in flight while the LDS row is cleared and
// during the barrier; the overflow flag read at the start is tested only after
// the scatter.  Bucket entries at or past the count, or all of them after an
// overflow, are read but not used.
template <int BS, int KB>
__device__ __forceinline__ void load_row_vec4(const int32_t* __restrict__ src, int nq, int tid, int4 (&v)[KB]) {
#pragma unroll
  for (int j = 0; j < KB; ++j) {
    // 32-bit byte offsets from the row's uniform base (a row is at most 256 KB);
    // past the row's end the last vector is loaded again.
    const uint32_t off = (uint32_t)min(tid + j * BS, nq - 1) << 4;
    const int4* a = reinterpret_cast<const int4*>(reinterpret_cast<const char*>(src) + off);
    v[j].x = __builtin_nontemporal_load(&a->x);
    v[j].y = __builtin_nontemporal_load(&a->y);
    v[j].z = __builtin_nontemporal_load(&a->z);
    v[j].w = __builtin_nontemporal_load(&a->w);
  }
}

template <int BS, int KB>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_rank_prefix_row(
    const int32_t* __restrict__ pref, uint16_t* __restrict__ rank, int n, int validate, Scalars* sc,
    PrefixBuckets pb) {
  extern __shared__ uint16_t lds16[];
  uint32_t* lds32 = reinterpret_cast<uint32_t*>(lds16);
  const int tid = threadIdx.x;
  const int row = blockIdx.x;
  const int nq = n >> 2;
  int bad = 0;
  int overflow = __hip_atomic_load(&sc->prefix_overflow, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  int4 v[KB];
  load_row_vec4<BS, KB>(pref + (size_t)row * n, nq, tid, v);
  int cntw = pb.cnt[row];
  uint32_t e_first = tid < pb.cap ? pb.bucket[(size_t)row * pb.cap + tid] : 0u;
  if (validate) {
    for (int c = tid; c < (n + 1) / 2; c += BS) lds32[c] = 0xFFFFFFFFu;
    __syncthreads();
  }
  // The clamped vectors rewrite the row's last four slots with the same values,
  // so the scatter has no branch.
#pragma unroll
  for (int j = 0; j < KB; ++j) {
    const int q = min(tid + j * BS, nq - 1);
    const int e[4] = {v[j].x, v[j].y, v[j].z, v[j].w};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      if ((unsigned)e[t] < (unsigned)n) lds16[e[t]] = (uint16_t)((q << 2) + t);
      else bad |= 1;
    }
  }
  __syncthreads();
  // Opaque from here: the compiler may not compute with these values (and so
  // wait for their loads) ahead of this point.
  asm volatile("" : "+v"(overflow), "+v"(cntw), "+v"(e_first));
  if (overflow == 0) {
    cntw = min(cntw, pb.cap);
    const uint32_t* bk = pb.bucket + (size_t)row * pb.cap;
    for (int j = tid; j < cntw; j += BS) {
      const uint32_t e = j == tid ? e_first : bk[j];
      const int m = (int)(e & 0xFFFFu);
      if (m < n) pb.brank[(size_t)row * pb.cap + j] = lds16[m];
    }
    if (validate) {  // a hole in the LDS row marks a row that is not a permutation
      for (int c = tid; c < (n + 1) / 2; c += BS) {
        const uint32_t x = lds32[c];
        if ((x & 0xFFFFu) == 0xFFFFu || (2 * c + 1 < n && (x >> 16) == 0xFFFFu)) bad |= 2;
      }
    }
  } else {
    // A bucket overflowed: the whole rank row, as the full inversion writes it.
    typedef unsigned int v4u __attribute__((ext_vector_type(4)));
    typedef unsigned int v2u __attribute__((ext_vector_type(2)));
    auto has_hole = [](unsigned x) { return (x & 0xFFFFu) == 0xFFFFu || (x >> 16) == 0xFFFFu; };
    uint16_t* dst = rank + (size_t)row * n;
    if ((n & 7) == 0) {
      const v4u* l4 = reinterpret_cast<const v4u*>(lds16);
      v4u* dst8 = reinterpret_cast<v4u*>(dst);
      for (int q = tid; q < (n >> 3); q += BS) {
        const v4u p = l4[q];
        if (validate && (has_hole(p.x) || has_hole(p.y) || has_hole(p.z) || has_hole(p.w))) bad |= 2;
        __builtin_nontemporal_store(p, dst8 + q);
      }
    } else {
      const v2u* l2 = reinterpret_cast<const v2u*>(lds16);
      v2u* dst4 = reinterpret_cast<v2u*>(dst);
      for (int q = tid; q < (n >> 2); q += BS) {
        const v2u p = l2[q];
        if (validate && (has_hole(p.x) || has_hole(p.y))) bad |= 2;
        __builtin_nontemporal_store(p, dst4 + q);
      }
    }
  }
  if (bad & 1) atomicOr(&sc->error, kErrEntryRange);
  if (bad & 2) atomicOr(&sc->error, kErrNotPerm);
}

// ---------------------------------------------------------------------------
// Sparse prefix pass, last step: the ranks the row pass looked up, stored in
// bucket order (coalesced, so that the row pass does not wait for scattered
// writes), go to the men's PR prefix entries pr[m, k] = (w, rank).  One wave
// per woman.  Nothing to do after a bucket overflow (the row pass wrote the
// rank matrix instead).
__global__ __launch_bounds__(256) void k_scatter_prefix(const int* __restrict__ cnt,
                                                        const uint32_t* __restrict__ bucket,
                                                        const uint16_t* __restrict__ brank, int cap, int n,
                                                        uint32_t* __restrict__ pr, Scalars* sc) {
  if (__hip_atomic_load(&sc->prefix_overflow, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
  const int w = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  if (w >= n) return;
  const int cntw = min(cnt[w], cap);
  const uint32_t* bk = bucket + (size_t)w * cap;
  const uint16_t* br = brank + (size_t)w * cap;
  // Bound by the scattered 4-byte stores (2.56 M at n = 20,000, K = 128: ~34 us);
  // loading a wave's slots ahead of the count changed nothing.
  for (int j = (int)(threadIdx.x & 63); j < cntw; j += 64) {
    const uint32_t e = bk[j];
    const int m = (int)(e & 0xFFFFu);
    if (m < n) pr[pr_index(m, (int)(e >> 16), n)] = pr_pack((uint32_t)w, br[j]);
  }
}

// ---------------------------------------------------------------------------
// PR gather: pr[m, k] = (w = pref_m[m][k], rank_w[w][m]) for k < K, stored
// column-blocked (pr_index).
// Tiles are 32 men x 256 columns.  The rank reads rank_w[w][m0..m0+32) of one
// men block share 128-byte lines, so tiles of the same men block are kept on one
// XCD (workgroup b runs on XCD b % 8) and consecutive in its dispatch order,
// which keeps those lines in that XCD's L2 while the block is processed.
constexpr int kGatherMen = 32;
constexpr int kGatherCols = 256;

// RankT: uint16_t (the rank matrix built on the device) or int32_t (a rank
// matrix given as input); a rank outside [0, n) is reported as a bad entry.
// row_major: out[m * K + k] (a whole man's list contiguous, for a host walking
// long chains) instead of the column-blocked pr_index layout.
template <typename RankT>
__global__ __launch_bounds__(256) void k_gather_pr(const int32_t* __restrict__ pref_m,
                                                   const RankT* __restrict__ rank_w,
                                                   uint32_t* __restrict__ pr, int n, int K,
                                                   int tiles, int per_xcd, Scalars* sc, int row_major, int col0) {
  const int b = blockIdx.x;
  const int t = (b & 7) * per_xcd + (b >> 3);
  if (t >= tiles) return;
  const int col_tiles = (K + kGatherCols - 1) / kGatherCols;
  const int mb = t / col_tiles, ct = t - mb * col_tiles;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int chunks = kGatherCols / 64;
  int bad = 0;
  for (int i = wave; i < kGatherMen * chunks; i += 4) {
    const int m = mb * kGatherMen + (i / chunks);
    const int k = ct * kGatherCols + (i % chunks) * 64 + lane;
    if (m >= n || k >= K) continue;
    const int w = pref_m[(size_t)m * n + col0 + k];
    uint32_t e = pr_pack(0, 0xFFFFu);
    if ((unsigned)w < (unsigned)n) {
      const uint32_t r = (uint32_t)rank_w[(size_t)w * n + m];
      if (r < (uint32_t)n) e = pr_pack((uint32_t)w, r);
      else bad = 1;
    } else {
      bad = 1;
    }
    pr[row_major ? (size_t)m * K + k : pr_index(m, k, n)] = e;
  }
  if (bad) atomicOr(&sc->error, kErrEntryRange);
}

// Buckets of the sparse prefix pass: every entry (m, k < K) of the men's lists
// goes to the bucket of woman w = pref_m[m][k] as k << 16 | m (both fit 16 bits:
// n <= 65535).  A transpose of the n x K prefix, done as a two-level partition so
// that global atomics are few:
//   k_bin_prefix  each workgroup takes kBinEntries consecutive prefix entries
//                 (coalesced along the men's rows), counts them per bin of 256
//                 women in LDS, reserves room in each bin with one global
//                 atomicAdd per bin, and writes (w << 32 | k << 16 | m) there;
//   k_bin_sort    kSortSplit workgroups per bin, each taking spans of 8192 of
//                 its entries: LDS atomics count them per woman, one global
//                 atomicAdd per woman reserves their slots in her bucket (the
//                 per-woman counts, zeroed by the reset launch), then they are
//                 placed.
// The order inside a bucket varies from run to run, which changes nothing (each
// entry names its own PR slot).  A bin or bucket that runs out of room sets
// sc->prefix_overflow, and the pass falls back to the full rank matrix.
constexpr int kBinWomen = 256;      // women per bin: bin = w >> 8 (at most 256 bins)
constexpr int kBinThreads = 1024;   // k_bin_prefix workgroup
constexpr int kBinEntries = 8192;   // prefix entries per k_bin_prefix workgroup (8 per thread)
constexpr int kSortSplit = 4;       // k_bin_sort workgroups per bin
// Bin counters sit one 128-byte line apart: every k_bin_prefix workgroup adds to
// each of them once, and same-line atomics would serialize.
constexpr int kBinPad = kBinCounterStride;

// gated: queued right behind the precheck, before the host has read it; the
// kernels do nothing unless the precheck left the general path open (two men
// share a first choice, and not all first choices are the same).
// A 64-bit value of one lane, the lane index being wave-uniform (a ballot's
// first set bit): two readlanes into scalar registers, where __shfl would go
// through the LDS crossbar and sit on the chain's dependent path.
__device__ inline unsigned long long lane_u64(unsigned long long v, int lane) {
  return ((unsigned long long)(uint32_t)__builtin_amdgcn_readlane((int)(v >> 32), lane) << 32) |
         (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
}

__device__ inline bool general_path_open(const Scalars* sc) {
  return __hip_atomic_load(&sc->dup_first, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 &&
         __hip_atomic_load(&sc->first_differs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

// With h_flag, workgroup 0 first publishes the control block (the precheck's
// flags, final at this kernel's start) to the host, as k_publish_scalars does.
__global__ __launch_bounds__(kBinThreads) void k_bin_prefix(const int32_t* __restrict__ pref_m, int n, int K, int nbins,
                                                    int* __restrict__ bin_cnt, unsigned long long* __restrict__ bins,
                                                    int bin_cap, int gated, Scalars* sc, Scalars* h_sc,
                                                    unsigned* h_flag, unsigned seq) {
  if (h_flag && blockIdx.x == 0 && threadIdx.x == 0) {
    *h_sc = *sc;
    __hip_atomic_store(h_flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
  if (gated && !general_path_open(sc)) return;
  __shared__ int hist[kBinWomen];
  __shared__ int base[kBinWomen];
  constexpr int kPer = kBinEntries / kBinThreads;
  const int tid = threadIdx.x;
  const unsigned long long total = (unsigned long long)n * (unsigned long long)K;  // <= 65535^2 < 2^32
  int bad = 0, over = 0;
  for (unsigned long long e0 = (unsigned long long)blockIdx.x * kBinEntries; e0 < total;
       e0 += (unsigned long long)gridDim.x * kBinEntries) {
    if (tid < kBinWomen) hist[tid] = 0;
    __syncthreads();
    uint32_t wv[kPer], mk[kPer];
    int loc[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const unsigned long long e64 = e0 + (unsigned long long)(j * kBinThreads + tid);
      wv[j] = 0xFFFFFFFFu;
      mk[j] = 0;
      if (e64 < total) {
        const unsigned e = (unsigned)e64;
        const unsigned m = e / (unsigned)K, k = e - m * (unsigned)K;
        const int w = pref_m[(size_t)m * n + k];
        if ((unsigned)w < (unsigned)n) {
          wv[j] = (uint32_t)w;
          mk[j] = (k << 16) | m;
        } else {
          bad = 1;
        }
      }
    }
#pragma unroll
    for (int j = 0; j < kPer; ++j) loc[j] = wv[j] != 0xFFFFFFFFu ? atomicAdd(&hist[wv[j] >> 8], 1) : 0;
    __syncthreads();
    if (tid < nbins && hist[tid] > 0) base[tid] = atomicAdd(&bin_cnt[tid * kBinPad], hist[tid]);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      if (wv[j] == 0xFFFFFFFFu) continue;
      const int bi = (int)(wv[j] >> 8);
      const int pos = base[bi] + loc[j];
      if (pos < bin_cap) bins[(size_t)bi * bin_cap + pos] = ((unsigned long long)wv[j] << 32) | mk[j];
      else over = 1;
    }
    __syncthreads();  // hist and base are reused by the next chunk
  }
  if (__ballot(bad) && (tid & 63) == 0) atomicOr(&sc->error, kErrEntryRange);
  if (__ballot(over) && (tid & 63) == 0) atomicOr(&sc->prefix_overflow, 1);
}

__global__ __launch_bounds__(1024) void k_bin_sort(const unsigned long long* __restrict__ bins,
                                                   const int* __restrict__ bin_cnt, int bin_cap,
                                                   int* __restrict__ cnt, uint32_t* __restrict__ bucket, int cap,
                                                   int gated, Scalars* sc) {
  if (gated && !general_path_open(sc)) return;
  __shared__ int c[kBinWomen];
  __shared__ int base[kBinWomen];
  const int tid = threadIdx.x;
  const int b = (int)blockIdx.x / kSortSplit, part = (int)blockIdx.x % kSortSplit;
  const int total = min(bin_cnt[b * kBinPad], bin_cap);
  const unsigned long long* src = bins + (size_t)b * bin_cap;
  int over = 0;
  constexpr int kLoads = 8;  // entries loaded per thread before any is placed
  constexpr int kSpan = 1024 * kLoads;
  for (int i0 = part * kSpan; i0 < total; i0 += kSortSplit * kSpan) {
    if (tid < kBinWomen) c[tid] = 0;
    __syncthreads();
    unsigned long long v[kLoads];
    int loc[kLoads];
#pragma unroll
    for (int j = 0; j < kLoads; ++j) {
      const int i = i0 + j * 1024 + tid;
      v[j] = i < total ? src[i] : ~0ull;
    }
#pragma unroll
    for (int j = 0; j < kLoads; ++j)
      loc[j] = v[j] != ~0ull ? atomicAdd(&c[(uint32_t)(v[j] >> 32) & (kBinWomen - 1)], 1) : 0;
    __syncthreads();
    // One global reservation per woman of the bin for this span's entries.
    if (tid < kBinWomen && c[tid] > 0) base[tid] = atomicAdd(&cnt[b * kBinWomen + tid], c[tid]);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kLoads; ++j) {
      if (v[j] == ~0ull) continue;
      const uint32_t w = (uint32_t)(v[j] >> 32);
      const int slot = base[w & (kBinWomen - 1)] + loc[j];
      if (slot < cap) bucket[(size_t)w * cap + slot] = (uint32_t)v[j];
      else over = 1;
    }
    __syncthreads();  // c and base are reused by the next span
  }
  if (__ballot(over) && (tid & 63) == 0) atomicOr(&sc->prefix_overflow, 1);
}

// ---------------------------------------------------------------------------
// Proposal kernel.  Persistent: each wave owns one free man ("chain") at a time.
// Wave g starts the men g, g + W, g + 2W, ... (W waves in the grid) in turn, so
// starting a man needs no shared counter.  A step loads the next 16 to 32
// entries of the man's list (one coalesced PR load, or pref_m + rank_w beyond the
// gathered prefix), reads those women's partner words and ballots the lanes where the
// man beats the current partner.  A partner word only ever decreases, so a stale
// read can only make the wave try an atomic that then fails; coherent_hint picks
// whether those reads bypass the non-coherent caches.  Candidates are tried in
// list order with a 64-bit atomicMin; the first success engages the man.  If the
// woman was taken the wave continues as the displaced man at his stored position
// + 1; if she was free the chain ends.
//
// Hand-off: `active` counts waves that still hold work; a wave decrements it once,
// when its chain ends and it has no unstarted man left.  With T > 0, a wave that
// sees active <= T after a step parks its chain (man, next position) and its
// unstarted men for the host, and exits.  Near the end every unstarted man has
// been started, so `active` equals the number of free men.  After a sparse prefix
// pass (no rank matrix), a chain that reaches position K is parked at K in the
// same list, and its wave starts its next man.
// A step examines the next `win` positions of the man's list: kProposeWin0 for a
// man just started or displaced, doubled up to kProposeWinMax after a step that
// found no candidate.  Each examined position costs a load of a woman's word
// from its own cache line, and the positions past the one that accepts him are
// wasted.  Measured against a fixed 64: the proposal phase of n = 20,000 uniform
// 0.219 -> 0.193 ms (median of five seeds), the reference's RANDOM lists (groups
// of 5, n = 20,000, whole PR matrix) 65.1 -> 43.7 ms per solve; a maximum of 16
// was 1.5 ms faster still on RANDOM but slower on correlated lists.
constexpr int kProposeWin0 = 16;
constexpr int kProposeWinMax = 32;

__global__ __launch_bounds__(256) void k_propose(ProposeArgs a) {
  const int lane = threadIdx.x & 63;
  const int n = a.n, K = a.K, T = a.T;
  const int waves = (int)(gridDim.x * (blockDim.x >> 6));
  const int gwave = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  const int total = a.starts ? a.nstarts : n;  // work items: men, or resumed chains
  int next_init = gwave;
  int m = -1, k = 0;
  int win = kProposeWin0;  // positions examined per step (lanes < win)
  // Speculation: while a candidate's atomicMin is in flight, the segment of the
  // man it would displace is loaded too; (spec_m, spec_k) says whose it is.
  int spec_m = -1, spec_k = -1;
  uint32_t spec_e = 0;
  unsigned long long n_atomics = 0, n_steps = 0;
  unsigned long long deep_local = 0;  // steps past the prefix not yet flushed
  int since_poll = 0;
  // Steps between polls: a.poll while many waves are busy, then every step once
  // the count is near the threshold, so that the last waves see the hand-off
  // within one step instead of up to a.poll steps (the kernel ends with them).
  int poll_every = a.poll;
  if (__hip_atomic_load(&a.sc->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
    next_init = total;  // init found malformed lists: do nothing
  }
  // After a sparse prefix pass only positions < K have ranks.  If a bucket
  // overflowed, the pass wrote the rank matrix instead of the prefix, which the
  // gather has not built yet: the kernel then does nothing at all (not even the
  // `active` count), and the host gathers and launches it again.
  if (a.rank_sparse && __hip_atomic_load(&a.sc->prefix_overflow, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
    if (lane == 0) a.wstats[2 * gwave] = a.wstats[2 * gwave + 1] = 0;
    return;
  }
  const bool have_rank = !a.rank_sparse;
  for (;;) {
    if (m < 0) {
      if (next_init >= total) {
        if (lane == 0 && gwave < total) atomicSub(&a.sc->active, 1);
        break;
      }
      if (a.starts) {
        const int2 st = a.starts[next_init];
        m = st.x;
        k = st.y;
      } else {
        m = a.start_order ? a.start_order[next_init] : next_init;
        k = 0;
      }
      next_init += waves;
      win = kProposeWin0;
      if ((unsigned)m >= (unsigned)n) {  // a resume list holds only men of this instance
        if (lane == 0) atomicOr(&a.sc->error, kErrEntryRange);
        m = -1;
        continue;
      }
    }
    if (k >= n) {  // only possible when rows are not permutations
      if (lane == 0) atomicOr(&a.sc->error, kErrListEnd);
      break;
    }
    int act = 0;
    const bool poll = T > 0 && ++since_poll >= poll_every;
    if (poll && lane == 0)
      act = __hip_atomic_load(&a.sc->active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    ++n_steps;
    // Deep chains: every 256 steps a wave adds its steps past the prefix to the
    // shared count; past deep_limit the first wave to see it sets `restart`,
    // and every wave that sees the flag parks its chain and its unstarted men
    // (as for the hand-off) so that the host gathers the whole PR matrix and
    // resumes them.
    bool restart = false;
    if (a.deep_limit > 0) {
      deep_local += (unsigned long long)(k >= K);
      if ((n_steps & 255) == 0) {
        int rs = 0;
        if (lane == 0) {
          if (deep_local) {
            const unsigned long long tot = atomicAdd(&a.sc->deep_steps, deep_local) + deep_local;
            if (tot > (unsigned long long)a.deep_limit) atomicOr(&a.sc->restart, 1);
          }
          rs = __hip_atomic_load(&a.sc->restart, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        deep_local = 0;
        restart = __builtin_amdgcn_readfirstlane(rs) != 0;
      }
    }
    if (restart) {
      if (lane == 0) {
        a.parked[atomicAdd(&a.sc->parked_count, 1)] = make_int2(m, k);
        for (; next_init < total; next_init += waves)
          a.parked[atomicAdd(&a.sc->parked_count, 1)] =
              a.starts ? a.starts[next_init] : make_int2(a.start_order ? a.start_order[next_init] : next_init, 0);
      }
      break;
    }
    const int kk = k + lane;
    bool valid = lane < win && kk < n && (have_rank || kk < K);
    uint32_t w = 0, r = 0;
    if (valid) {
      if (kk < K) {
        const uint32_t e = (spec_m == m && spec_k == k) ? spec_e : a.pr[pr_index(m, kk, n)];
        w = pr_woman(e);
        r = pr_rank(e);
      } else {
        const int wi = a.pref_m[(size_t)m * n + kk];
        if ((unsigned)wi < (unsigned)n) {
          w = (uint32_t)wi;
          r = a.rank32 ? (uint32_t)a.rank32[(size_t)w * n + m] : (uint32_t)a.rank_w[(size_t)w * n + m];
        }
        if ((unsigned)wi >= (unsigned)n || r >= (uint32_t)n) {
          valid = false;
          atomicOr(&a.sc->error, kErrEntryRange);
        }
      }
      valid = valid && w < (uint32_t)n;
    }
    spec_m = -1;
    const unsigned long long mine = pw_pack(r, (uint32_t)m, (uint32_t)kk);
    unsigned long long hint = 0;
    if (valid) {
      hint = a.coherent_hint
                 ? __hip_atomic_load(&a.partner[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                 : a.partner[w];
    }
    unsigned long long cand = __ballot(valid && mine < hint);
    bool moved = false;
    while (cand) {
      const int j = __ffsll((long long)cand) - 1;
      unsigned long long old = 0;
      if (lane == j) old = atomicMin(&a.partner[w], mine);
      // The man this attempt would displace, by the hint; his next segment is
      // loaded while the atomic is in flight.
      const unsigned long long hj = lane_u64(hint, j);
      int sm = -1, sk = -1;
      if (hj != kEmpty) {
        sm = (int)pw_man(hj);
        sk = (int)pw_pos(hj) + 1;
        if (sk + lane < K && sk + lane < n) spec_e = a.pr[pr_index(sm, sk + lane, n)];
      }
      old = lane_u64(old, j);
      const unsigned long long mj = lane_u64(mine, j);
      ++n_atomics;
      if (old > mj) {
        if (old == kEmpty) {
          m = -1;
        } else {
          m = (int)pw_man(old);
          k = (int)pw_pos(old) + 1;
          if (m == sm && k == sk) spec_m = sm, spec_k = sk;
        }
        win = kProposeWin0;
        moved = true;
        break;
      }
      cand &= cand - 1;
    }
    if (!moved) {
      k += win;
      win = min(kProposeWinMax, 2 * win);
      if (!have_rank && k >= K) {
        // Positions below K are settled; the rest needs the rank matrix, which the
        // host builds before it resumes the chain at K.
        if (lane == 0) a.parked[atomicAdd(&a.sc->parked_count, 1)] = make_int2(m, K);
        m = -1;
      }
    }
    if (poll) {
      since_poll = 0;
      act = __builtin_amdgcn_readfirstlane(act);
      if (act <= T) {
        if (lane == 0) {
          if (m >= 0) a.parked[atomicAdd(&a.sc->parked_count, 1)] = make_int2(m, k);
          for (; next_init < total; next_init += waves)
            a.parked[atomicAdd(&a.sc->parked_count, 1)] =
                a.starts ? a.starts[next_init] : make_int2(a.start_order ? a.start_order[next_init] : next_init, 0);
        }
        break;
      }
      if (act <= 4 * T) poll_every = 1;
    }
  }
  if (lane == 0) {
    a.wstats[2 * gwave] = n_atomics;
    a.wstats[2 * gwave + 1] = n_steps;
  }
}

// Adds the proposal kernel's per-wave counters into the control block.
__global__ __launch_bounds__(256) void k_fold_wstats(const unsigned long long* __restrict__ wstats,
                                                     int waves, Scalars* sc) {
  __shared__ unsigned long long red[2][4];
  unsigned long long at = 0, st = 0;
  for (int g = threadIdx.x; g < waves; g += blockDim.x) {
    at += wstats[2 * g];
    st += wstats[2 * g + 1];
  }
  for (int off = 32; off > 0; off >>= 1) {
    at += __shfl_down(at, off);
    st += __shfl_down(st, off);
  }
  if ((threadIdx.x & 63) == 0) red[0][threadIdx.x >> 6] = at, red[1][threadIdx.x >> 6] = st;
  __syncthreads();
  if (threadIdx.x == 0) {
    sc->atomics += red[0][0] + red[0][1] + red[0][2] + red[0][3];
    sc->steps += red[1][0] + red[1][1] + red[1][2] + red[1][3];
  }
}

// Hand-off export after the proposal kernel: folds the per-wave counters into
// the control block (block 0), then writes the control block, the partner words
// and the parked chains straight into pinned host memory (device-visible
// pointers of hipHostMalloc buffers).  One kernel on the engine's queue instead
// of three copies, each of which would wait for a switch between the compute
// queue and a copy engine.  The host reads the buffers after the kernel's event.
__global__ __launch_bounds__(256) void k_export_state(Scalars* sc, const unsigned long long* __restrict__ wstats,
                                                      int waves, const unsigned long long* __restrict__ partner,
                                                      const int2* __restrict__ parked, int n, Scalars* h_sc,
                                                      unsigned long long* h_partner, int2* h_parked,
                                                      unsigned* h_flag, unsigned seq) {
  if (blockIdx.x == 0) {
    __shared__ unsigned long long red[2][4];
    unsigned long long at = 0, st = 0;
    for (int g = threadIdx.x; g < waves; g += blockDim.x) {
      at += wstats[2 * g];
      st += wstats[2 * g + 1];
    }
    for (int off = 32; off > 0; off >>= 1) {
      at += __shfl_down(at, off);
      st += __shfl_down(st, off);
    }
    if ((threadIdx.x & 63) == 0) red[0][threadIdx.x >> 6] = at, red[1][threadIdx.x >> 6] = st;
    __syncthreads();
    if (threadIdx.x == 0) {
      sc->atomics += red[0][0] + red[0][1] + red[0][2] + red[0][3];
      sc->steps += red[1][0] + red[1][1] + red[1][2] + red[1][3];
      *h_sc = *sc;
    }
  }
  const int tid = blockIdx.x * blockDim.x + threadIdx.x, nth = gridDim.x * blockDim.x;
  for (int i = tid; i < n; i += nth) h_partner[i] = partner[i];
  const int pc = min(sc->parked_count, n);
  for (int i = tid; i < pc; i += nth) h_parked[i] = parked[i];
  // Publish: each wave's host writes are released at system scope, then the
  // last block to finish writes seq into the host's flag word, so a host that
  // sees seq sees everything above.
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0 && atomicAdd(&sc->export_blocks, 1) == (int)gridDim.x - 1) {
    sc->export_blocks = 0;
    __hip_atomic_store(h_flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// Extension windows for the host tail.  Request i = (m, k0, off, len) asks for
// positions k0 .. k0 + len - 1 of man m's list; the requests' entries are laid
// out back to back (off of request i + 1 = off + len of request i, off of
// request 0 = 0), so entry e < total of the launch belongs to the last request
// whose off <= e.  Entry (w, rank) of position k0 + j goes to out[off + j]
// (pinned host memory, written through its device-visible address), or
// kWinInvalid for a position past the list's end or a malformed entry.  The rank
// comes from the rank matrix when one exists (rank16, or rank32 given as input),
// else from a search of woman w's row for m.  The last workgroup to finish
// releases the writes at system scope and stores seq into *h_flag.
__device__ inline void publish_blocks(Scalars* sc, unsigned* h_flag, unsigned seq) {
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0 && atomicAdd(&sc->export_blocks, 1) == (int)gridDim.x - 1) {
    sc->export_blocks = 0;
    __hip_atomic_store(h_flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// The request that entry e belongs to (binary search over the ascending offsets).
__device__ inline int window_request(const int4* __restrict__ req, int count, int e) {
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (req[mid].z <= e) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Four entries per 1024-thread workgroup, one per 256-thread quarter (fewer
// workgroups: each one's system-scope release before the flag costs more than
// its search).
__global__ __launch_bounds__(1024) void k_window_search(const int32_t* __restrict__ pref_m,
                                                        const int32_t* __restrict__ pref_w, int n,
                                                        const int4* __restrict__ req, int count, int total,
                                                        uint32_t* __restrict__ out, Scalars* sc, unsigned* h_flag,
                                                        unsigned seq) {
  __shared__ int found[4];
  const int quarter = (int)threadIdx.x >> 8, tq = (int)threadIdx.x & 255;
  const int entry = (int)blockIdx.x * 4 + quarter;
  int m = -1, k = -1;
  bool used = false;
  if (entry < total) {
    const int4 q = req[window_request(req, count, entry)];
    const int j = entry - q.z;
    used = j >= 0 && j < q.w;
    m = q.x;
    k = q.y + j;
  }
  if (tq == 0) found[quarter] = -1;
  __syncthreads();
  int w = -1;
  if (used && (unsigned)m < (unsigned)n && (unsigned)k < (unsigned)n) w = pref_m[(size_t)m * n + k];
  if ((unsigned)w < (unsigned)n) {
    const int32_t* row = pref_w + (size_t)w * n;
    if ((n & 3) == 0) {
      // 16 loads per thread in flight: one round for rows up to 16,384 entries.
      const int4* row4 = reinterpret_cast<const int4*>(row);
      const int nq = n >> 2;
      for (int c0 = tq; c0 < nq; c0 += 256 * 16) {
        int4 v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) v[u] = row4[min(c0 + 256 * u, nq - 1)];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          const int c = c0 + 256 * u;
          if (c < nq) {
            if (v[u].x == m) found[quarter] = 4 * c;
            if (v[u].y == m) found[quarter] = 4 * c + 1;
            if (v[u].z == m) found[quarter] = 4 * c + 2;
            if (v[u].w == m) found[quarter] = 4 * c + 3;
          }
        }
      }
    } else {
      for (int r = tq; r < n; r += 256)
        if (row[r] == m) found[quarter] = r;
    }
  }
  __syncthreads();
  if (tq == 0 && used) {
    uint32_t* dst = out + entry;
    const int r = found[quarter];
    if ((unsigned)w < (unsigned)n && r >= 0) {
      *dst = pr_pack((uint32_t)w, (uint32_t)r);
    } else {
      *dst = kWinInvalid;
      if ((unsigned)k < (unsigned)n) atomicOr(&sc->error, (unsigned)w < (unsigned)n ? kErrNotPerm : kErrEntryRange);
    }
  }
  publish_blocks(sc, h_flag, seq);
}

template <typename RankT>
__global__ __launch_bounds__(256) void k_window_gather(const int32_t* __restrict__ pref_m,
                                                       const RankT* __restrict__ rank_w, int n,
                                                       const int4* __restrict__ req, int count, int total,
                                                       uint32_t* __restrict__ out, Scalars* sc, unsigned* h_flag,
                                                       unsigned seq) {
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (t < total) {
    const int4 q = req[window_request(req, count, t)];
    const int j = t - q.z;
    if (j >= 0 && j < q.w) {
      const int m = q.x, k = q.y + j;
      uint32_t e = kWinInvalid;
      if ((unsigned)m < (unsigned)n && (unsigned)k < (unsigned)n) {
        const int w = pref_m[(size_t)m * n + k];
        if ((unsigned)w < (unsigned)n) {
          const uint32_t r = (uint32_t)rank_w[(size_t)w * n + m];
          if (r < (uint32_t)n) e = pr_pack((uint32_t)w, r);
        }
        if (e == kWinInvalid) atomicOr(&sc->error, kErrEntryRange);
      }
      out[t] = e;
    }
  }
  publish_blocks(sc, h_flag, seq);
}

// Precheck publish: the control block into pinned host memory, then seq into
// the host's flag word with a system-scope release.  Queued behind the
// precheck, it stands in for a copy and an event, each of which would leave
// the queue idle for several microseconds before the kernels behind it.
__global__ void k_publish_scalars(const Scalars* __restrict__ sc, Scalars* h_sc, unsigned* h_flag, unsigned seq) {
  if (threadIdx.x == 0) {
    *h_sc = *sc;
    __hip_atomic_store(h_flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// ---------------------------------------------------------------------------
// Post: match[man] = woman from the partner words; also sums (position + 1)
// over all women, which equals the total number of proposals of any
// Gale-Shapley run (each man proposes down to his final partner exactly once).
__global__ __launch_bounds__(256) void k_post(const unsigned long long* __restrict__ partner, int n,
                                              int32_t* __restrict__ match, Scalars* sc) {
  __shared__ unsigned long long red[4];
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long props = 0;
  if (w < n) {
    const unsigned long long v = partner[w];
    if (v == kEmpty) {
      atomicOr(&sc->error, kErrUnmatched);
    } else {
      const uint32_t man = pw_man(v);
      if (man < (uint32_t)n) match[man] = w;
      props = pw_pos(v) + 1ull;
    }
  }
  for (int off = 32; off > 0; off >>= 1) props += __shfl_down(props, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = props;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(&sc->proposals, red[0] + red[1] + red[2] + red[3]);
}

// ---------------------------------------------------------------------------
// Master list (every man's list is the same permutation L).  With one side
// ranking the other identically the stable matching is unique, so the
// women-proposing deferred acceptance returns the man-optimal matching the
// reference's proposal walk computes (smp_engine_bamboo.cu:108-306), in about
// n*H_n proposals instead of n(n+1)/2.  A man ranks woman w at Linv[w], the same
// for every man, so a proposal needs no rank gather.

// One wave per man m >= 1 compares his list with man 0's and stops at the first
// difference (or when another wave has found one).  With n % 4 == 0 a step reads
// 2,048 entries as 16-byte streaming loads (8 per lane in flight); man 0's row is
// read through the caches, where every wave finds it.
// Workgroups blk = 0..nblk-1 of a launch (256 threads each).
__device__ void rows_equal_blocks(const int32_t* __restrict__ pref_m, int n, Scalars* sc, int blk, int nblk) {
  const int lane = threadIdx.x & 63;
  const int waves = nblk * (int)(blockDim.x >> 6);
  const bool vec = (n & 3) == 0;
  for (int m = 1 + (int)(blk * (blockDim.x >> 6) + (threadIdx.x >> 6)); m < n; m += waves) {
    if (__hip_atomic_load(&sc->not_master, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    const int32_t* row = pref_m + (size_t)m * n;
    if (vec) {
      const int4* r4 = reinterpret_cast<const int4*>(row);
      const int4* l4 = reinterpret_cast<const int4*>(pref_m);
      const int nq = n >> 2;
      for (int q0 = 0; q0 < nq; q0 += 512) {
        int4 a[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int q = min(q0 + 64 * j + lane, nq - 1);
          a[j].x = __builtin_nontemporal_load(&r4[q].x);
          a[j].y = __builtin_nontemporal_load(&r4[q].y);
          a[j].z = __builtin_nontemporal_load(&r4[q].z);
          a[j].w = __builtin_nontemporal_load(&r4[q].w);
        }
        bool diff = false;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int q = min(q0 + 64 * j + lane, nq - 1);
          const int4 b = l4[q];
          diff |= a[j].x != b.x || a[j].y != b.y || a[j].z != b.z || a[j].w != b.w;
        }
        if (__ballot(diff)) {
          if (lane == 0 && !__hip_atomic_load(&sc->not_master, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicOr(&sc->not_master, 1);
          return;
        }
      }
      continue;
    }
    for (int k0 = 0; k0 < n; k0 += 512) {
      bool diff = false;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int k = k0 + 64 * j + lane;
        if (k < n && row[k] != pref_m[k]) diff = true;
      }
      if (__ballot(diff)) {
        if (lane == 0 && !__hip_atomic_load(&sc->not_master, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
          atomicOr(&sc->not_master, 1);
        return;
      }
    }
  }
}

// linv[L[k]] = k over the shared list L (linv pre-filled with -1).
__global__ void k_master_inverse(const int32_t* __restrict__ L, int n, int32_t* linv, Scalars* sc) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int w = L[k];
  if ((unsigned)w >= (unsigned)n) {
    atomicOr(&sc->error, kErrEntryRange);
  } else if (atomicCAS(&linv[w], -1, k) != -1) {
    atomicOr(&sc->error, kErrNotPerm);
  }
}

// Marks entry v of a row in the row's n-bit LDS set; a value out of range or
// seen before is an error.
__device__ __forceinline__ void perm_mark(int v, int n, uint32_t* bits, int& bad) {
  if ((unsigned)v >= (unsigned)n) {
    bad |= kErrEntryRange;
  } else {
    const uint32_t bit = 1u << (v & 31);
    if (atomicOr(&bits[v >> 5], bit) & bit) bad |= kErrNotPerm;
  }
}

// Whether row `src` equals row 0 (`row0`), decided by the whole workgroup (the
// result is the same in every thread).  4096 entries a round; a round that
// differs ends the comparison.
__device__ bool row_equals_row0(const int32_t* __restrict__ src, const int32_t* __restrict__ row0, int n) {
  if ((n & 3) == 0) {
    const int4* s4 = reinterpret_cast<const int4*>(src);
    const int4* l4 = reinterpret_cast<const int4*>(row0);
    const int nq = n >> 2;
    for (int q0 = 0; q0 < nq; q0 += 256 * 4) {
      bool diff = false;
      int4 a[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int q = min(q0 + 256 * j + (int)threadIdx.x, nq - 1);
        a[j].x = __builtin_nontemporal_load(&s4[q].x);
        a[j].y = __builtin_nontemporal_load(&s4[q].y);
        a[j].z = __builtin_nontemporal_load(&s4[q].z);
        a[j].w = __builtin_nontemporal_load(&s4[q].w);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int q = min(q0 + 256 * j + (int)threadIdx.x, nq - 1);
        const int4 b = l4[q];
        diff |= a[j].x != b.x || a[j].y != b.y || a[j].z != b.z || a[j].w != b.w;
      }
      if (__syncthreads_or(diff)) return false;
    }
    return true;
  }
  for (int r0 = 0; r0 < n; r0 += 256 * 8) {
    bool diff = false;
    int v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int r = min(r0 + 256 * j + (int)threadIdx.x, n - 1);
      v[j] = __builtin_nontemporal_load(src + r);
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) diff |= v[j] != row0[min(r0 + 256 * j + (int)threadIdx.x, n - 1)];
    if (__syncthreads_or(diff)) return false;
  }
  return true;
}

// The women's rows of the master-list launch, by workgroups blk = 0..nblk-1.
// Every row is compared with row 0 until some row differs (that row sets
// sc->women_differ); a row equal to row 0 is a permutation exactly when row 0
// is, so only row 0 and the rows that differ from it get the permutation check
// (an n-bit LDS set per row).  The scan stops once another workgroup has found
// that the men's lists differ (it then ran ahead of a decision that went the
// other way; the general path checks the rows itself).
__device__ void check_rows_blocks(const int32_t* __restrict__ pref, int n, Scalars* sc, int blk, int nblk,
                                  uint32_t* bits) {
  const int words = (n + 31) >> 5;
  int bad = 0;
  bool compare = true;
  for (int row = blk; row < n; row += nblk) {
    // Thread 0 reads the flags and the workgroup follows it (the barrier of
    // __syncthreads_or hands its value to every thread), so every thread leaves
    // the loop, or stops comparing, together.
    const bool lead = threadIdx.x == 0;
    if (__syncthreads_or(lead && __hip_atomic_load(&sc->not_master, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)))
      break;
    if (compare &&
        __syncthreads_or(lead && __hip_atomic_load(&sc->women_differ, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)))
      compare = false;
    const int32_t* src = pref + (size_t)row * n;
    if (compare && row != 0) {
      if (row_equals_row0(src, pref, n)) continue;
      if (lead) atomicOr(&sc->women_differ, 1);
      compare = false;
    }
    for (int i = threadIdx.x; i < words; i += blockDim.x) bits[i] = 0;
    __syncthreads();
    if ((n & 3) == 0) {
      const int4* s4 = reinterpret_cast<const int4*>(src);
      const int nq = n >> 2;
      for (int q0 = threadIdx.x; q0 < nq; q0 += 256 * 4) {
        int4 a[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int q = min(q0 + 256 * j, nq - 1);
          a[j].x = __builtin_nontemporal_load(&s4[q].x);
          a[j].y = __builtin_nontemporal_load(&s4[q].y);
          a[j].z = __builtin_nontemporal_load(&s4[q].z);
          a[j].w = __builtin_nontemporal_load(&s4[q].w);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          if (q0 + 256 * j < nq) {
            perm_mark(a[j].x, n, bits, bad);
            perm_mark(a[j].y, n, bits, bad);
            perm_mark(a[j].z, n, bits, bad);
            perm_mark(a[j].w, n, bits, bad);
          }
        }
      }
    } else {
      for (int r0 = threadIdx.x; r0 < n; r0 += 256 * 8) {
        int v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int r = r0 + 256 * j;
          v[j] = r < n ? __builtin_nontemporal_load(src + r) : 0;
        }
#pragma unroll
        for (int j = 0; j < 8; ++j)
          if (r0 + 256 * j < n) perm_mark(v[j], n, bits, bad);
      }
    }
    __syncthreads();
  }
  if (bad) atomicOr(&sc->error, bad);
}

// Women-proposing walk: one wave per chain (woman), 64 list positions a step.
// Woman w's proposal to the man at position k of her list is the packed word
// (Linv[w], w, k); the man keeps the smallest (atomicMin on his partner word), so
// he holds the woman who comes first in L.  A displaced woman continues at her
// stored position + 1.  Hints read before the atomic can only be stale upward
// (words only decrease), so they never hide a candidate.
// Wave `wave` of `waves` walks the women wave, wave + waves, ...  Run beside the
// scan that decides whether the lists are a master list, it is speculative: it
// stops once the scan has found two men's lists that differ (sc->not_master),
// checked every 16 steps, and its result is then thrown away.
// With wait_scans != 0 the launch also compares every woman's list with woman
// 0's (wait_scans scan workgroups, each counted in sc->scans_done when it ends):
// the walk starts only once a woman's list has differed (sc->women_differ) or
// every scan has ended, since when all women share a list the walk is not needed
// (sc->both_master) and, on such lists a serial chain of n^2 / 2 proposals, would
// only take bandwidth from the scans.  The scans wait on nothing, so the wait ends.
__device__ void master_walk(const int32_t* __restrict__ pref_w, const int32_t* __restrict__ linv,
                            unsigned long long* partner_m, int n, Scalars* sc, int wave, int waves,
                            int wait_scans) {
  const int lane = threadIdx.x & 63;
  int next_init = wave;
  int w = -1, k = 0;
  // Speculation: while a claim is in flight, the next segment of the woman it
  // would displace (by the hint) is loaded too; (spec_w, spec_k) says whose.
  int spec_w = -1, spec_k = -1, spec_m = 0;
  unsigned long long n_atomics = 0, n_steps = 0;
  if (wait_scans) {
    for (;;) {
      int go = 0;
      if (lane == 0)
        go = __hip_atomic_load(&sc->not_master, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) |
             __hip_atomic_load(&sc->women_differ, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) |
             (__hip_atomic_load(&sc->scans_done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= wait_scans);
      if (__shfl(go, 0)) break;
      __builtin_amdgcn_s_sleep(16);
    }
  }
  if (__hip_atomic_load(&sc->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
      __hip_atomic_load(&sc->not_master, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
      __hip_atomic_load(&sc->both_master, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != 0)
    next_init = n;
  for (;;) {
    if (w < 0) {
      if (next_init >= n) break;
      w = next_init;
      k = 0;
      next_init += waves;
    }
    if (k >= n) {
      if (lane == 0) atomicOr(&sc->error, kErrListEnd);
      break;
    }
    if ((n_steps & 15) == 15 &&
        (__hip_atomic_load(&sc->not_master, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
         __hip_atomic_load(&sc->both_master, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0))
      break;
    const uint32_t rw = (uint32_t)linv[w];
    const int kk = k + lane;
    bool valid = kk < n;
    int m = 0;
    if (valid) {
      m = (spec_w == w && spec_k == k) ? spec_m : pref_w[(size_t)w * n + kk];
      if ((unsigned)m >= (unsigned)n) {
        valid = false;
        atomicOr(&sc->error, kErrEntryRange);
      }
    }
    spec_w = -1;
    const unsigned long long mine = pw_pack(rw, (uint32_t)w, (uint32_t)kk);
    // Plain (cached) reads: a stale hint is only ever too large (words only
    // decrease), which can cost a failed atomic but never hides a candidate.
    const unsigned long long hint = valid ? partner_m[m] : 0ull;
    unsigned long long cand = __ballot(valid && mine < hint);
    ++n_steps;
    bool moved = false;
    while (cand) {
      const int j = __ffsll((long long)cand) - 1;
      unsigned long long old = 0;
      if (lane == j) old = atomicMin(&partner_m[m], mine);
      const unsigned long long hj = lane_u64(hint, j);
      int sw = -1, sk = -1;
      if (hj != kEmpty) {
        sw = (int)pw_man(hj);
        sk = (int)pw_pos(hj) + 1;
        if ((unsigned)sw < (unsigned)n && sk + lane < n) spec_m = pref_w[(size_t)sw * n + sk + lane];
      }
      old = lane_u64(old, j);
      ++n_atomics;
      if (old > lane_u64(mine, j)) {
        if (old == kEmpty) {
          w = -1;
        } else {
          w = (int)pw_man(old);
          k = (int)pw_pos(old) + 1;
          if (w == sw && k == sk) spec_w = sw, spec_k = sk;
        }
        moved = true;
        break;
      }
      cand &= cand - 1;
    }
    if (!moved) k += 64;
  }
  if (lane == 0) {
    atomicAdd(&sc->atomics, n_atomics);
    atomicAdd(&sc->steps, n_steps);
  }
}

// The master-list path in one launch, three kinds of workgroup:
//   [0, g_walk)                 the women-proposing walk, speculatively (it needs
//                               only man 0's list, inverted into linv beforehand);
//   [g_walk, g_walk + g_rows)   every men's list compared with man 0's;
//   the rest (when check != 0)  the women's rows checked for being permutations.
// The scans read 8n^2 bytes and set the HBM rate; the walk, latency-bound,
// overlaps them instead of following them.  The walk's workgroups come first so
// that they are dispatched first; nothing in the launch waits for another part.
__global__ __launch_bounds__(256) void k_master_fused(const int32_t* __restrict__ pref_m,
                                                      const int32_t* __restrict__ pref_w,
                                                      const int32_t* __restrict__ linv,
                                                      unsigned long long* partner_m, int n, int check, int g_walk,
                                                      int g_rows, Scalars* sc) {
  extern __shared__ uint32_t bits[];
  const int b = (int)blockIdx.x;
  if (b < g_walk) {
    master_walk(pref_w, linv, partner_m, n, sc, b * 4 + (int)(threadIdx.x >> 6), g_walk * 4,
                check ? (int)gridDim.x - g_walk : 0);
    return;
  }
  const int r = b - g_walk;
  if (r < g_rows) {
    rows_equal_blocks(pref_m, n, sc, r, g_rows);
  } else if (check) {
    check_rows_blocks(pref_w, n, sc, r - g_rows, (int)gridDim.x - g_walk - g_rows, bits);
  }
  // With the women's rows checked too: the last scan workgroup to finish, if
  // no man's list differs from man 0's and no woman's from woman 0's, sets
  // both_master, which stops the walk (the matching then has a closed form).
  if (check) {
    __syncthreads();
    if (threadIdx.x == 0) {
      __threadfence();
      const int scans = (int)gridDim.x - g_walk;
      const int done = __hip_atomic_fetch_add(&sc->scans_done, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) + 1;
      if (done == scans && __hip_atomic_load(&sc->not_master, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == 0 &&
          __hip_atomic_load(&sc->women_differ, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == 0)
        __hip_atomic_store(&sc->both_master, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// Every man holds man 0's list L and every woman woman 0's list P: the stable
// matching is the serial dictatorship match[P[k]] = L[k] (the man all women put
// at position k gets the woman all men put there), and the man-proposing count
// is sum over k of (k + 1).  Both rows were checked to be permutations.
__global__ __launch_bounds__(256) void k_master_both(const int32_t* __restrict__ pref_m,
                                                     const int32_t* __restrict__ pref_w, int n,
                                                     int32_t* __restrict__ match, Scalars* sc) {
  __shared__ unsigned long long red[4];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long props = 0;
  if (k < n) {
    const int m = pref_w[k], w = pref_m[k];
    if ((unsigned)m < (unsigned)n && (unsigned)w < (unsigned)n) {
      match[m] = w;
      props = (unsigned long long)k + 1ull;
    } else {
      atomicOr(&sc->error, kErrEntryRange);
    }
  }
  for (int off = 32; off > 0; off >>= 1) props += __shfl_down(props, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = props;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(&sc->proposals, red[0] + red[1] + red[2] + red[3]);
}

// match[m] = the woman in man m's partner word; proposals (the man-proposing
// count, the same for every schedule) = sum over men of Linv[match[m]] + 1.
__global__ __launch_bounds__(256) void k_post_master(const unsigned long long* __restrict__ partner_m,
                                                     const int32_t* __restrict__ linv, int n,
                                                     int32_t* __restrict__ match, Scalars* sc) {
  __shared__ unsigned long long red[4];
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long props = 0;
  if (m < n) {
    const unsigned long long v = partner_m[m];
    if (v == kEmpty) {
      atomicOr(&sc->error, kErrUnmatched);
    } else {
      const uint32_t w = pw_man(v);
      if (w < (uint32_t)n) {
        match[m] = (int32_t)w;
        props = (unsigned long long)linv[w] + 1ull;
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) props += __shfl_down(props, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = props;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(&sc->proposals, red[0] + red[1] + red[2] + red[3]);
}

// ---------------------------------------------------------------------------
// The reference's thread-per-man GPU walks as schedules (SmpEngineMw3F9
// src/smp/smp_engine_mw_3_f9.cu:37-82 and the variants of src/smp/kernel.cu:
// 5-55, 122-165, 468-512, 515-552): one thread per man walks his chain; a
// woman's partner is the 32-bit rank of her current man (n = free); the
// displaced man is pref_w[w][old rank] and resumes at the list position recorded
// in next[] when he was accepted.  The variants differ in
//   LA    entries (woman, rank) from the PR prefix for positions < K (the
//         reference's PR matrix, types.h:58-61), else pref_m + rank_w;
//   MIN   claim with atomicMin (the woman keeps the smaller rank) instead of a
//         compare-and-swap retry loop;
//   SKIP  read the partner rank first and skip women who hold a better man.
// next[m] is written before the claim and fenced, so a displacer that reads it
// after its own claim sees that value or an earlier one (an earlier one only
// repeats refused proposals); the reference writes it after the claim.
template <bool LA, bool MIN, bool SKIP>
__global__ __launch_bounds__(256) void k_walk_thread(const int32_t* __restrict__ pref_m,
                                                     const int32_t* __restrict__ pref_w,
                                                     const uint16_t* __restrict__ rank_w,
                                                     const uint32_t* __restrict__ pr, int K, int* partner_rank,
                                                     int* next, int n, Scalars* sc) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long tries = 0;
  int bad = 0;
  if (m < n) {
    int k = 0;
    for (;;) {
      if (k >= n) {
        bad |= kErrListEnd;
        break;
      }
      int w;
      int r;
      if (LA && k < K) {
        const uint32_t e = pr[pr_index(m, k, n)];
        w = (int)pr_woman(e);
        r = (int)pr_rank(e);
      } else {
        w = pref_m[(size_t)m * n + k];
        if ((unsigned)w >= (unsigned)n) {
          bad |= kErrEntryRange;
          break;
        }
        r = rank_w[(size_t)w * n + m];
      }
      if ((unsigned)w >= (unsigned)n || (unsigned)r >= (unsigned)n) {
        bad |= kErrEntryRange;
        break;
      }
      ++k;
      int cur = n + 1;  // only tested when SKIP or in the CAS loop
      if (SKIP || !MIN) {
        cur = __hip_atomic_load(&partner_rank[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur < r) continue;  // she holds a better man
      }
      int prev;
      bool won = false;
      if (MIN) {
        __hip_atomic_store(&next[m], k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        ++tries;
        prev = atomicMin(&partner_rank[w], r);
        won = prev > r;
      } else {
        prev = cur;
        while (prev > r) {
          __hip_atomic_store(&next[m], k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __threadfence();
          ++tries;
          const int seen = atomicCAS(&partner_rank[w], prev, r);
          if (seen == prev) {
            won = true;
            break;
          }
          prev = seen;
        }
      }
      if (!won) continue;
      if (prev >= n) break;  // she was free
      __threadfence();
      m = pref_w[(size_t)w * n + prev];
      if ((unsigned)m >= (unsigned)n) {
        bad |= kErrEntryRange;
        break;
      }
      k = __hip_atomic_load(&next[m], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  for (int off = 32; off > 0; off >>= 1) tries += __shfl_down(tries, off);
  const unsigned long long bb = __ballot(bad != 0);
  if ((threadIdx.x & 63) == 0) {
    if (tries) atomicAdd(&sc->atomics, tries);
    if (bb) atomicOr(&sc->error, kErrEntryRange | kErrListEnd);
  }
}

// ---------------------------------------------------------------------------
// Frontier schedule: rounds over a compacted list of free men (entries
// k << 16 | m in f0 / f1, the sizes in ctl->cnt rotating by round % 3).  In a
// round every free man is taken by a group of G lanes, which scans up to
// kFrontierScan * G positions of his list from his position for the first woman
// whose partner word he beats (a plain read: words only decrease, so a woman
// whose word is already smaller would refuse him, and skipping her is her
// refusal).  The workgroup's candidates for one woman are reduced in an LDS hash
// table (64-bit atomicMin on the packed (rank, man, position) word) and only the
// smallest goes to the global word, with one atomicMin.  A man who loses in LDS
// or to the global word, or who found no candidate in his window, goes to the
// next round's list; a man displaced by a successful claim goes there at his
// stored position + 1; a claim of a free woman ends the chain.  The next list is
// compacted by a ballot and prefix sum per wave and one atomicAdd per wave.  A
// grid barrier (all workgroups co-resident: a cooperative launch) separates the
// rounds; its waits are bounded, and a wait that runs out sets ctl->abort, which
// ends every workgroup at its next barrier or round.
constexpr int kFrontierScan = 4;  // loads per lane of a round's window (G * 4 positions)
constexpr long long kBarrierSpinLimit = 1ll << 27;

__device__ bool frontier_barrier(FrontierCtl* c, int nblocks) {
  __shared__ int ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    ok = 1;
    __threadfence();
    const int gen = __hip_atomic_load(&c->bar_gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    if (atomicAdd(&c->bar_count, 1) == nblocks - 1) {
      __hip_atomic_store(&c->bar_count, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&c->bar_gen, gen + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      long long spins = 0;
      while (__hip_atomic_load(&c->bar_gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
        __builtin_amdgcn_s_sleep(1);
        if (++spins > kBarrierSpinLimit ||
            __hip_atomic_load(&c->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
          atomicOr(&c->abort, 1);
          ok = 0;
          break;
        }
      }
    }
    __threadfence();
  }
  __syncthreads();
  return ok != 0;
}

__global__ void k_frontier_init(uint32_t* f0, int n, FrontierCtl* c) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f0[i] = (uint32_t)i;  // man i at position 0
  if (i == 0) {
    c->cnt[0] = n;
    c->cnt[1] = c->cnt[2] = 0;
    c->bar_count = c->bar_gen = 0;
    c->rounds = 0;
    c->abort = 0;
  }
}

template <int BS, int G>
__global__ __launch_bounds__(BS) void k_frontier(FrontierArgs a) {
  constexpr int MPW = BS / G;   // men per workgroup per chunk
  constexpr int TBL = 2 * MPW;  // LDS hash slots (power of two)
  static_assert((TBL & (TBL - 1)) == 0, "table size");
  __shared__ uint32_t keys[TBL];
  __shared__ unsigned long long vals[TBL];
  const int n = a.n, K = a.K;
  const int lane = threadIdx.x & 63, sub = lane & (G - 1), gbase = lane & ~(G - 1);
  const int grp = threadIdx.x / G;
  const unsigned long long gmask = G == 64 ? ~0ull : ((1ull << G) - 1) << gbase;
  unsigned long long n_atomics = 0;
  int bad = 0;
  if (__hip_atomic_load(&a.sc->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;  // uniform
  for (int round = 0;; ++round) {
    const int size = __hip_atomic_load(&a.ctl->cnt[round % 3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (size == 0 || __hip_atomic_load(&a.ctl->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      a.ctl->cnt[(round + 2) % 3] = 0;
      a.ctl->rounds = round + 1;
    }
    const uint32_t* cur = (round & 1) ? a.f1 : a.f0;
    uint32_t* nxt = (round & 1) ? a.f0 : a.f1;
    int* nxt_cnt = &a.ctl->cnt[(round + 1) % 3];
    for (int base = (int)blockIdx.x * MPW; base < size; base += (int)gridDim.x * MPW) {
      for (int i = threadIdx.x; i < TBL; i += BS) {
        keys[i] = 0xFFFFFFFFu;
        vals[i] = kEmpty;
      }
      __syncthreads();
      const int item = base + grp;
      const bool have = item < size;
      int m = 0, k = 0;
      if (have) {
        const uint32_t e = cur[item];
        m = (int)(e & 0xFFFFu);
        k = (int)(e >> 16);
      }
      // The window: positions k + sub + G * j, j < kFrontierScan, all loaded at
      // once (entries, then the women's words); the first candidate in list
      // order is the lowest j with a candidate lane, then the lowest lane.
      bool cand = false;
      uint32_t cw = 0;
      unsigned long long cmine = 0;
      int ck = 0;
      uint32_t wv[kFrontierScan];
      unsigned long long mv[kFrontierScan];
#pragma unroll
      for (int j = 0; j < kFrontierScan; ++j) {
        const int pos = k + sub + G * j;
        wv[j] = 0xFFFFFFFFu;
        mv[j] = 0;
        if (have && pos < n) {
          uint32_t w = 0, r = 0;
          if (pos < K) {
            const uint32_t e = a.pr[pr_index(m, pos, n)];
            w = pr_woman(e);
            r = pr_rank(e);
          } else {
            const int wi = a.pref_m[(size_t)m * n + pos];
            if ((unsigned)wi < (unsigned)n) {
              w = (uint32_t)wi;
              r = a.rank_w[(size_t)w * n + m];
            } else {
              w = (uint32_t)n;
            }
          }
          if (w < (uint32_t)n && r < (uint32_t)n) {
            wv[j] = w;
            mv[j] = pw_pack(r, (uint32_t)m, (uint32_t)pos);
          } else {
            bad |= kErrEntryRange;
          }
        }
      }
      unsigned long long hv[kFrontierScan];
#pragma unroll
      for (int j = 0; j < kFrontierScan; ++j) hv[j] = wv[j] != 0xFFFFFFFFu ? a.partner[wv[j]] : 0ull;
      int jf = kFrontierScan, lf = gbase;
#pragma unroll
      for (int j = kFrontierScan - 1; j >= 0; --j) {
        const unsigned long long bm = __ballot(wv[j] != 0xFFFFFFFFu && mv[j] < hv[j]) & gmask;
        if (bm) {
          jf = j;
          lf = __ffsll((long long)bm) - 1;
        }
      }
      // Every lane takes part in the shuffles; jf and lf are the same in a group.
      uint32_t wsel = wv[0];
      unsigned long long msel = mv[0];
#pragma unroll
      for (int j = 1; j < kFrontierScan; ++j)
        if (jf == j) wsel = wv[j], msel = mv[j];
      const uint32_t wj = (uint32_t)__shfl((int)wsel, lf);
      const unsigned long long mj = __shfl(msel, lf);
      bool list_end = false;
      if (have) {
        if (jf < kFrontierScan) {
          cand = true;
          cw = wj;
          cmine = mj;
          ck = k + G * jf + (lf - gbase);
        } else {
          k += G * kFrontierScan;
          if (k >= n) list_end = true;
        }
      }
      if (list_end && !cand) bad |= kErrListEnd;
      // LDS pre-reduction: the group's leader inserts its candidate.
      int slot = -1;
      if (cand && sub == 0) {
        int h = (int)((cw * 2654435761u) >> 16) & (TBL - 1);
        for (int probe = 0; probe < TBL; ++probe) {
          const uint32_t prev = atomicCAS(&keys[h], 0xFFFFFFFFu, cw);
          if (prev == 0xFFFFFFFFu || prev == cw) {
            slot = h;
            break;
          }
          h = (h + 1) & (TBL - 1);
        }
        if (slot >= 0) atomicMin(&vals[slot], cmine);
      }
      __syncthreads();
      bool push = false;
      uint32_t out = 0;
      if (sub == 0 && have) {
        if (cand) {
          if (slot >= 0 && vals[slot] == cmine) {
            const unsigned long long old = atomicMin(&a.partner[cw], cmine);
            ++n_atomics;
            if (old > cmine) {
              if (old != kEmpty) {
                push = true;
                out = ((pw_pos(old) + 1) << 16) | pw_man(old);
              }
            } else {
              push = true;
              out = ((uint32_t)(ck + 1) << 16) | (uint32_t)m;
            }
          } else {
            push = true;
            out = ((uint32_t)(ck + 1) << 16) | (uint32_t)m;
          }
        } else if (!list_end) {
          push = true;
          out = ((uint32_t)k << 16) | (uint32_t)m;
        }
      }
      // Compaction: one reservation per wave, each pushing lane at its prefix.
      const unsigned long long pm_ = __ballot(push);
      int wbase = 0;
      if (lane == 0 && pm_) wbase = atomicAdd(nxt_cnt, __popcll(pm_));
      wbase = __shfl(wbase, 0);
      if (push) {
        const int at = wbase + __popcll(pm_ & ((1ull << lane) - 1));
        if (at < n) nxt[at] = out;
        else bad |= kErrListEnd;
      }
      __syncthreads();  // the LDS table is cleared for the next chunk
    }
    if (!frontier_barrier(a.ctl, (int)gridDim.x)) break;
  }
  for (int off = 32; off > 0; off >>= 1) n_atomics += __shfl_down(n_atomics, off);
  const unsigned long long bb = __ballot(bad != 0);
  if (lane == 0) {
    if (n_atomics) atomicAdd(&a.sc->atomics, n_atomics);
    if (bb) atomicOr(&a.sc->error, bad);
  }
}

// match[pref_w[w][partner_rank[w]]] = w (the reference's PostProc,
// smp_engine_mw_3_f9.cu:85-97); proposals = sum over men of next[m], the list
// position after each man's final proposal.
__global__ __launch_bounds__(256) void k_post_cas(const int32_t* __restrict__ pref_w,
                                                  const int* __restrict__ partner_rank,
                                                  const int* __restrict__ next, int n,
                                                  int32_t* __restrict__ match, Scalars* sc) {
  __shared__ unsigned long long red[4];
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long props = 0;
  bool unmatched = false;
  if (w < n) {
    const int r = partner_rank[w];
    if ((unsigned)r >= (unsigned)n) {
      unmatched = true;
    } else {
      const int m = pref_w[(size_t)w * n + r];
      if ((unsigned)m < (unsigned)n) {
        match[m] = w;
        props = (unsigned long long)next[m];
      }
    }
  }
  const unsigned long long bu = __ballot(unmatched);
  for (int off = 32; off > 0; off >>= 1) props += __shfl_down(props, off);
  if ((threadIdx.x & 63) == 0) {
    red[threadIdx.x >> 6] = props;
    if (bu) atomicOr(&sc->error, kErrUnmatched);
  }
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(&sc->proposals, red[0] + red[1] + red[2] + red[3]);
}

// Streaming copy, 16 bytes per thread per iteration.
__global__ void k_copy16(int4* __restrict__ dst, const int4* __restrict__ src, size_t count) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// Host-callable launchers.

hipError_t launch_copy16(void* dst, const void* src, size_t bytes, int cus, hipStream_t s) {
  const size_t count = bytes / 16;
  int grid = (int)std::min<size_t>((count + 255) / 256, (size_t)cus * 16);
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, s, reinterpret_cast<int4*>(dst),
                     reinterpret_cast<const int4*>(src), count);
  return hipGetLastError();
}

hipError_t launch_fill_u64(unsigned long long* p, size_t count, unsigned long long v, hipStream_t s) {
  int grid = (int)std::min<size_t>((count + 255) / 256, 4096);
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(k_fill_u64, dim3(grid), dim3(256), 0, s, p, count, v);
  return hipGetLastError();
}

hipError_t launch_fill_i32(int* p, size_t count, int v, hipStream_t s) {
  int grid = (int)std::min<size_t>((count + 255) / 256, 4096);
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(k_fill_i32, dim3(grid), dim3(256), 0, s, p, count, v);
  return hipGetLastError();
}

hipError_t launch_reset_scalars(Scalars* sc, int active, hipStream_t s, int* z0, size_t n0, int* z1, size_t n1,
                                unsigned long long* partner, size_t np) {
  if (!z0) n0 = 0;
  if (!z1) n1 = 0;
  if (!partner) np = 0;
  const int grid = (int)std::max<size_t>(1, std::min<size_t>(128, (n0 + n1 + np + 255) / 256));
  hipLaunchKernelGGL(k_reset_scalars, dim3(grid), dim3(256), 0, s, sc, active, z0, n0, z1, n1, partner, np);
  return hipGetLastError();
}

hipError_t launch_master_fused(const int32_t* pref_m, const int32_t* pref_w, const int32_t* linv,
                               unsigned long long* partner_m, int n, int check, Scalars* sc, int cus, hipStream_t s) {
  // Two walk waves per CU: measured 788 us at n = 20,000 against 822 (four),
  // 843 (eight), 905 (one) and 870 for the scans and the walk one after the other.
  const int g_walk = std::max(1, std::min((n + 3) / 4, std::max(1, cus / 2)));
  const int g_rows = std::max(1, std::min((n + 3) / 4, cus * 4));
  const int g_check = check ? std::max(1, std::min(n, cus * 4)) : 0;
  const size_t lds = (size_t)((n + 31) / 32) * 4;
  hipLaunchKernelGGL(k_master_fused, dim3(g_walk + g_rows + g_check), dim3(256), lds, s, pref_m, pref_w, linv,
                     partner_m, n, check, g_walk, g_rows, sc);
  return hipGetLastError();
}

hipError_t launch_master_inverse(const int32_t* L, int n, int32_t* linv, Scalars* sc, hipStream_t s) {
  hipError_t e = launch_fill_i32(linv, n, -1, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_master_inverse, dim3((n + 255) / 256), dim3(256), 0, s, L, n, linv, sc);
  return hipGetLastError();
}

hipError_t launch_master_both(const int32_t* pref_m, const int32_t* pref_w, int n, int32_t* match, Scalars* sc,
                              hipStream_t s) {
  hipLaunchKernelGGL(k_master_both, dim3((n + 255) / 256), dim3(256), 0, s, pref_m, pref_w, n, match, sc);
  return hipGetLastError();
}

hipError_t launch_post_master(const unsigned long long* partner_m, const int32_t* linv, int n, int32_t* match,
                              Scalars* sc, hipStream_t s) {
  hipLaunchKernelGGL(k_post_master, dim3((n + 255) / 256), dim3(256), 0, s, partner_m, linv, n, match, sc);
  return hipGetLastError();
}

hipError_t launch_walk_thread(int schedule, const int32_t* pref_m, const int32_t* pref_w, const uint16_t* rank_w,
                              const uint32_t* pr, int K, int* partner_rank, int* next, int n, Scalars* sc,
                              hipStream_t s) {
  hipError_t e = launch_fill_i32(partner_rank, n, n, s);
  if (e != hipSuccess) return e;
  if ((e = launch_fill_i32(next, n, 0, s)) != hipSuccess) return e;
  const dim3 g((n + 255) / 256), b(256);
  switch (schedule) {
    case SMP_SCHEDULE_MW_CAS:
      hipLaunchKernelGGL((k_walk_thread<false, false, true>), g, b, 0, s, pref_m, pref_w, rank_w, pr, K, partner_rank,
                         next, n, sc);
      break;
    case SMP_SCHEDULE_LA_CAS:
      hipLaunchKernelGGL((k_walk_thread<true, false, true>), g, b, 0, s, pref_m, pref_w, rank_w, pr, K, partner_rank,
                         next, n, sc);
      break;
    case SMP_SCHEDULE_MW_MIN:
      hipLaunchKernelGGL((k_walk_thread<false, true, true>), g, b, 0, s, pref_m, pref_w, rank_w, pr, K, partner_rank,
                         next, n, sc);
      break;
    case SMP_SCHEDULE_LA_MIN:
      hipLaunchKernelGGL((k_walk_thread<true, true, false>), g, b, 0, s, pref_m, pref_w, rank_w, pr, K, partner_rank,
                         next, n, sc);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_post_cas(const int32_t* pref_w, const int* partner_rank, const int* next, int n, int32_t* match,
                           Scalars* sc, hipStream_t s) {
  hipLaunchKernelGGL(k_post_cas, dim3((n + 255) / 256), dim3(256), 0, s, pref_w, partner_rank, next, n, match, sc);
  return hipGetLastError();
}

hipError_t launch_first_choices(const int32_t* pref_m, int n, int* seen, Scalars* sc, hipStream_t s) {
  hipLaunchKernelGGL(k_first_choices, dim3((n + 255) / 256), dim3(256), 0, s, pref_m, n, seen, sc);
  return hipGetLastError();
}

hipError_t launch_first_match(const int32_t* pref_m, int n, int32_t* match, hipStream_t s) {
  hipLaunchKernelGGL(k_first_match, dim3((n + 255) / 256), dim3(256), 0, s, pref_m, n, match);
  return hipGetLastError();
}

// chunk: columns of one LDS pass (<= max LDS bytes / 4).
// chunk: columns of one LDS pass (a multiple of 4, <= max LDS bytes / 2).
hipError_t launch_invert_rows(const int32_t* pref, uint16_t* rank, int n, int chunk, int validate,
                              Scalars* sc, hipStream_t s, int nontemporal) {
  const size_t lds = (size_t)((chunk + 7) & ~7) * sizeof(uint16_t);
  const bool vec = (n % 4 == 0) && (chunk % 8 == 0 || chunk == n);
  const PrefixBuckets none{nullptr, nullptr, 0, nullptr};
  if (vec && nontemporal)
    hipLaunchKernelGGL((k_invert_rows<4, 1, 0>), dim3(n), dim3(kInvBlock), lds, s, pref, rank, n, chunk,
                       validate, sc, none);  // streaming: rank rows are written once, read much later
  else if (vec)
    hipLaunchKernelGGL((k_invert_rows<4, 0, 0>), dim3(n), dim3(kInvBlock), lds, s, pref, rank, n, chunk,
                       validate, sc, none);
  else
    hipLaunchKernelGGL((k_invert_rows<1, 0, 0>), dim3(n), dim3(kInvBlock), lds, s, pref, rank, n, chunk,
                       validate, sc, none);
  return hipGetLastError();
}

int bin_count(int n) { return (n + kBinWomen - 1) / kBinWomen; }

hipError_t launch_bucket_prefix(const int32_t* pref_m, int n, int K, int* bin_cnt, unsigned long long* bins,
                                int bin_cap, int* cnt, uint32_t* bucket, int cap, int gated, Scalars* sc, int cus,
                                hipStream_t s, int stages, Scalars* h_sc, unsigned* h_flag, unsigned seq) {
  const int nbins = bin_count(n);
  if (nbins > kBinWomen) return hipErrorInvalidValue;
  if (stages & 1) {
    const unsigned long long total = (unsigned long long)n * (unsigned long long)K;
    const long long chunks = (long long)((total + kBinEntries - 1) / kBinEntries);
    const int grid = (int)std::max<long long>(1, std::min<long long>(chunks, (long long)cus * 16));
    hipLaunchKernelGGL(k_bin_prefix, dim3(grid), dim3(kBinThreads), 0, s, pref_m, n, K, nbins, bin_cnt, bins,
                       bin_cap, gated, sc, h_sc, h_flag, seq);
  }
  if (stages & 2)
    hipLaunchKernelGGL(k_bin_sort, dim3(nbins * kSortSplit), dim3(1024), 0, s, bins, bin_cnt, bin_cap, cnt, bucket,
                       cap, gated, sc);
  return hipGetLastError();
}

hipError_t launch_rank_prefix(const int32_t* pref_w, uint16_t* rank, int n, int chunk, int validate,
                              const int* cnt, const uint32_t* bucket, int cap, uint16_t* brank, uint32_t* pr,
                              Scalars* sc, hipStream_t s) {
  const size_t lds = (size_t)((chunk + 7) & ~7) * sizeof(uint16_t);
  const bool vec = (n % 4 == 0) && (chunk % 8 == 0 || chunk == n);
  const PrefixBuckets pb{cnt, bucket, cap, brank};
  // Five loads per thread cover rows up to 20,480 entries in one round.
  if (vec && chunk == n && n / 4 <= kInvBlock * kBlkShort)
    hipLaunchKernelGGL((k_rank_prefix_row<kInvBlock, kBlkShort>), dim3(n), dim3(kInvBlock), lds, s, pref_w, rank,
                       n, validate, sc, pb);
  else if (vec)
    hipLaunchKernelGGL((k_invert_rows<4, 1, 1>), dim3(n), dim3(kInvBlock), lds, s, pref_w, rank, n, chunk,
                       validate, sc, pb);
  else
    hipLaunchKernelGGL((k_invert_rows<1, 0, 1>), dim3(n), dim3(kInvBlock), lds, s, pref_w, rank, n, chunk,
                       validate, sc, pb);
  hipLaunchKernelGGL(k_scatter_prefix, dim3((n + 3) / 4), dim3(256), 0, s, cnt, bucket, brank, cap, n, pr, sc);
  return hipGetLastError();
}

hipError_t set_invert_lds_limit(int bytes) {
  const void* fns[] = {reinterpret_cast<const void*>(&k_invert_rows<4, 1, 0>),
                       reinterpret_cast<const void*>(&k_invert_rows<4, 0, 0>),
                       reinterpret_cast<const void*>(&k_invert_rows<1, 0, 0>),
                       reinterpret_cast<const void*>(&k_invert_rows<4, 1, 1>),
                       reinterpret_cast<const void*>(&k_rank_prefix_row<kInvBlock, kBlkShort>),
                       reinterpret_cast<const void*>(&k_invert_rows<1, 0, 1>)};
  for (const void* f : fns) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

template <typename RankT>
static hipError_t launch_gather_pr_t(const int32_t* pref_m, const RankT* rank_w, uint32_t* pr, int n, int K,
                                     Scalars* sc, hipStream_t s, int row_major, int col0) {
  if (col0 < 0 || K < 0 || col0 > n - K || (col0 != 0 && !row_major)) return hipErrorInvalidValue;
  const int men_blocks = (n + kGatherMen - 1) / kGatherMen;
  const int col_tiles = (K + kGatherCols - 1) / kGatherCols;
  const long long tiles = (long long)men_blocks * col_tiles;
  if (tiles > (1ll << 30)) return hipErrorInvalidValue;
  const int per_xcd = (int)((tiles + 7) / 8);
  hipLaunchKernelGGL((k_gather_pr<RankT>), dim3(per_xcd * 8), dim3(256), 0, s, pref_m, rank_w, pr, n, K,
                     (int)tiles, per_xcd, sc, row_major, col0);
  return hipGetLastError();
}

hipError_t launch_gather_pr(const int32_t* pref_m, const uint16_t* rank_w, uint32_t* pr, int n, int K,
                            Scalars* sc, hipStream_t s, int row_major, int col0) {
  return launch_gather_pr_t(pref_m, rank_w, pr, n, K, sc, s, row_major, col0);
}

hipError_t launch_gather_pr(const int32_t* pref_m, const int32_t* rank_w, uint32_t* pr, int n, int K,
                            Scalars* sc, hipStream_t s, int row_major, int col0) {
  return launch_gather_pr_t(pref_m, rank_w, pr, n, K, sc, s, row_major, col0);
}

hipError_t launch_export_state(Scalars* sc, const unsigned long long* wstats, int waves,
                               const unsigned long long* partner, const int2* parked, int n, Scalars* h_sc,
                               unsigned long long* h_partner, int2* h_parked, unsigned* h_flag, unsigned seq,
                               hipStream_t s) {
  // Few blocks: each one's wave-level system-scope fences write back the L2, so
  // their number, not the copy, sets the kernel's time.
  const int grid = std::max(1, std::min(16, (n + 255) / 256));
  hipLaunchKernelGGL(k_export_state, dim3(grid), dim3(256), 0, s, sc, wstats, waves, partner, parked, n, h_sc,
                     h_partner, h_parked, h_flag, seq);
  return hipGetLastError();
}

hipError_t launch_publish_scalars(const Scalars* sc, Scalars* h_sc, unsigned* h_flag, unsigned seq, hipStream_t s) {
  hipLaunchKernelGGL(k_publish_scalars, dim3(1), dim3(64), 0, s, sc, h_sc, h_flag, seq);
  return hipGetLastError();
}

hipError_t launch_fold_wstats(const unsigned long long* wstats, int waves, Scalars* sc, hipStream_t s) {
  hipLaunchKernelGGL(k_fold_wstats, dim3(1), dim3(256), 0, s, wstats, waves, sc);
  return hipGetLastError();
}

hipError_t launch_window(const int32_t* pref_m, const int32_t* pref_w, const uint16_t* rank16, const int32_t* rank32,
                         int n, const int4* req, int count, int total, uint32_t* out, Scalars* sc, unsigned* h_flag,
                         unsigned seq, hipStream_t s) {
  if (count <= 0 || total <= 0) return hipErrorInvalidValue;
  if (rank32) {
    hipLaunchKernelGGL(k_window_gather<int32_t>, dim3((total + 255) / 256), dim3(256), 0, s, pref_m, rank32, n, req,
                       count, total, out, sc, h_flag, seq);
  } else if (rank16) {
    hipLaunchKernelGGL(k_window_gather<uint16_t>, dim3((total + 255) / 256), dim3(256), 0, s, pref_m, rank16, n, req,
                       count, total, out, sc, h_flag, seq);
  } else {
    hipLaunchKernelGGL(k_window_search, dim3((total + 3) / 4), dim3(1024), 0, s, pref_m, pref_w, n, req, count, total,
                       out, sc, h_flag, seq);
  }
  return hipGetLastError();
}

// The frontier schedule: initial list (every man at position 0), then one
// cooperative launch of k_frontier with as many workgroups as fit on the device
// at once (at most one per MPW men).
hipError_t launch_frontier(const FrontierArgs& a, int cus, hipStream_t s) {
  constexpr int BS = 1024, G = 8;
  hipLaunchKernelGGL(k_frontier_init, dim3((a.n + 255) / 256), dim3(256), 0, s, a.f0, a.n, a.ctl);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  int per_cu = 0;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&k_frontier<BS, G>), BS, 0);
  if (e != hipSuccess) return e;
  if (per_cu < 1) return hipErrorInvalidConfiguration;
  const int grid = std::max(1, std::min(cus * std::min(per_cu, 2), (a.n + BS / G - 1) / (BS / G)));
  FrontierArgs args = a;
  void* params[] = {&args};
  return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_frontier<BS, G>), dim3(grid), dim3(BS), params,
                                    0, s);
}

hipError_t launch_propose(const ProposeArgs& a, int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_propose, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_post(const unsigned long long* partner, int n, int32_t* match, Scalars* sc,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_post, dim3((n + 255) / 256), dim3(256), 0, s, partner, n, match, sc);
  return hipGetLastError();
}

}  // namespace pactae
