// wg_solver.hpp — one workgroup solves one whole instance (device code).
//
// Used by the batched kernel and by the serving kernel.  Steps:
//   1. women's rank rows (rank[w][pref_w[w][r]] = r, the F9RankMatrixInitKernel
//      step, src/smp/kernel.cu:672-686) as int16, `rows` rows in flight, each row
//      assembled in LDS by a team of BLOCK/rows threads, then written coalesced;
//   2. proposal chains (the La3F9CoreKernel walk, src/smp/kernel.cu:633-670): each
//      wave owns one free man, scans 64 list positions per step, ballots the lanes
//      where he beats the current partner and claims with a 64-bit LDS atomicMin on
//      the packed (rank, man, position) word; displaced men continue in the wave;
//   3. the matching (PostProc, src/smp/smp_engine_bamboo.cu:299-303) and the
//      proposal count (sum over women of position + 1).
// LDS: partner words (8 * max_n) + row buffers (2 * rows * max_n) + 2 ints.
#pragma once

#include <hip/hip_runtime.h>

#include "smp_internal.hpp"

namespace pactae {

struct InstanceRef {
  const int32_t* pm;
  const int32_t* pw;
  int16_t* rank;
  int32_t* match;
  int n;
};

inline size_t wg_lds_bytes(int rows, int max_n) {
  return (size_t)max_n * 8 + (size_t)rows * max_n * 2 + 16;
}

// Step 1: the instance's women's rank rows, `rows` rows in flight.  All threads
// of the workgroup call it together.  Returns this thread's error bits.
template <int BLOCK>
__device__ int wg_rank_rows(const InstanceRef& d, int16_t* rowbuf, int rows, int max_n) {
  const int n = d.n;
  int bad = 0;
  const int team_size = BLOCK / rows;
  const int team = threadIdx.x / team_size, tl = threadIdx.x - team * team_size;
  int16_t* buf = rowbuf + (size_t)team * max_n;
  for (int base = 0; base < n; base += rows) {
    const int row = base + team;
    if (row < n) {
      for (int c = tl; c < n; c += team_size) buf[c] = -1;
    }
    __syncthreads();
    if (row < n) {
      const int32_t* src = d.pw + (size_t)row * n;
      for (int r = tl; r < n; r += team_size) {
        const int m = src[r];
        if ((unsigned)m < (unsigned)n) buf[m] = (int16_t)r;
        else bad |= kErrEntryRange;
      }
    }
    __syncthreads();
    if (row < n) {
      int16_t* dst = d.rank + (size_t)row * n;
      for (int c = tl; c < n; c += team_size) {
        const int16_t v = buf[c];
        if (v < 0) bad |= kErrNotPerm;
        dst[c] = v;
      }
    }
    __syncthreads();  // a team may span several waves: finish reading before the next reset
  }
  return bad;
}

// Steps 2 and 3: proposal chains over partner words in LDS, then the matching.
// `partner` holds n words of LDS, `ctl` one int of LDS, `red` BLOCK/64 words.
// The rank rows must be complete and visible to the workgroup.  Thread 0
// receives the proposal count in *props.  Returns this thread's error bits.
template <int BLOCK>
__device__ int wg_chains(const InstanceRef& d, unsigned long long* partner, int* cursor,
                         unsigned long long* red, unsigned long long* props) {
  constexpr int NW = BLOCK / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int n = d.n;
  int bad = 0;
  for (int w = threadIdx.x; w < n; w += BLOCK) partner[w] = kEmpty;
  if (threadIdx.x == 0) *cursor = 0;
  __syncthreads();

  const int16_t* rk = d.rank;
  int m = -1, k = 0;
  for (;;) {
    if (m < 0) {
      int got = 0;
      if (lane == 0) got = atomicAdd(cursor, 1);
      got = __shfl(got, 0);
      if (got >= n) break;
      m = got;
      k = 0;
    }
    if (k >= n) {
      bad |= kErrListEnd;
      break;
    }
    const int kk = k + lane;
    bool valid = kk < n;
    uint32_t w = 0, r = 0;
    if (valid) {
      const int wi = d.pm[(size_t)m * n + kk];
      if ((unsigned)wi < (unsigned)n) {
        w = (uint32_t)wi;
        r = (uint32_t)(uint16_t)rk[(size_t)w * n + m];
      } else {
        valid = false;
        bad |= kErrEntryRange;
      }
    }
    const unsigned long long mine = pw_pack(r, (uint32_t)m, (uint32_t)kk);
    unsigned long long hint = 0;
    if (valid) hint = __hip_atomic_load(&partner[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    unsigned long long cand = __ballot(valid && mine < hint);
    bool moved = false;
    while (cand) {
      const int j = __ffsll((long long)cand) - 1;
      unsigned long long old = 0;
      if (lane == j) old = atomicMin(&partner[w], mine);
      old = __shfl(old, j);
      const unsigned long long mj = __shfl(mine, j);
      if (old > mj) {
        if (old == kEmpty) {
          m = -1;
        } else {
          m = (int)pw_man(old);
          k = (int)pw_pos(old) + 1;
        }
        moved = true;
        break;
      }
      cand &= cand - 1;
    }
    if (!moved) k += 64;
  }
  __syncthreads();

  unsigned long long p = 0;
  for (int w = threadIdx.x; w < n; w += BLOCK) {
    const unsigned long long v = partner[w];
    if (v == kEmpty) {
      bad |= kErrUnmatched;
    } else {
      const uint32_t man = pw_man(v);
      if (man < (uint32_t)n) d.match[man] = w;
      p += pw_pos(v) + 1ull;
    }
  }
  for (int off = 32; off > 0; off >>= 1) p += __shfl_down(p, off);
  if (lane == 0) red[wave] = p;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long s = 0;
    for (int i = 0; i < NW; ++i) s += red[i];
    *props = s;
  }
  return bad;
}

// Whole instance in one workgroup: rank rows, then chains.  `ctl` points to two
// ints of LDS (ctl[1] is the chain cursor).
template <int BLOCK>
__device__ int wg_solve(const InstanceRef& d, unsigned char* smem, int rows, int max_n, int* ctl,
                        unsigned long long* red, unsigned long long* props) {
  unsigned long long* partner = reinterpret_cast<unsigned long long*>(smem);
  int16_t* rowbuf = reinterpret_cast<int16_t*>(partner + max_n);
  int bad = wg_rank_rows<BLOCK>(d, rowbuf, rows, max_n);
  bad |= wg_chains<BLOCK>(d, partner, &ctl[1], red, props);
  return bad;
}

}  // namespace pactae
