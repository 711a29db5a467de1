// wg_solver.hpp — device code shared by the batched and the serving paths.
//
//   wg_chains_g     one workgroup walks one instance's proposal chains (the
//                   La3F9CoreKernel walk, src/smp/kernel.cu:633-670) over partner
//                   words in LDS, claimed with a 64-bit LDS atomicMin on the
//                   packed (rank, man, position) word; displaced men continue in
//                   the same group of lanes; then the matching (PostProc,
//                   src/smp/smp_engine_bamboo.cu:299-303) and the proposal count
//                   (the serving lanes' chain kernel, for launches above
//                   2,048 men);
//   wg_pool_chains  the same walk over a pool of instances open in LDS at once
//                   (the batched pool mode), with a window per step that follows
//                   each instance's running mean reach;
//   wg_solve_instance  one workgroup solves one instance alone: shared men's
//                   list check, women's rank rows staged in LDS, the walk, the
//                   matching (the batched fused mode and the serving lanes up to
//                   2,048 men);
//   k_rank_rows16   the women's rank rows (rank[w][pref_w[w][r]] = r, the
//                   F9RankMatrixInitKernel step, src/smp/kernel.cu:672-686) of many
//                   instances in one flat pass, one wave per row;
//   k_rankT_tiles   the transposed rank matrix (men's rows) of many instances.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "smp_internal.hpp"

namespace pactae {

struct InstanceRef {
  const int32_t* pm;
  const int32_t* pw;
  int16_t* rank;
  int32_t* match;
  int n;
};

// Rank of man m in woman w's list: rank[w][m] in the row-major (women's rows)
// layout, rankT[m][w] in the transposed (men's rows) layout, where a chain's
// lookups all fall inside its man's own row.
template <bool kTransposed>
__device__ inline uint32_t rank_of(const int16_t* rk, int n, uint32_t w, uint32_t m) {
  return (uint32_t)(uint16_t)(kTransposed ? rk[(size_t)m * n + w] : rk[(size_t)w * n + m]);
}

// Steps 2 and 3: proposal chains over partner words in LDS, then the matching.
// Each group of G lanes walks one chain (G = 64: one chain per wave), so a wave
// carries 64 / G chains.  A step scans G * J consecutive positions of the man's
// list (lane gl covers positions k + gl + G * j, j < J): one list load and one
// rank lookup per position, one LDS hint read per position, then the group's
// first position in list order where the man beats the woman's current partner
// claims her with a 64-bit LDS atomicMin on the packed (rank, man, position)
// word.  J starts at 1 for every new or displaced man and doubles (up to JMAX)
// after each step without a candidate, so short uniform walks read few
// positions per step and long congested ones scan G * JMAX.
// A claim that loses the race resumes the scan just after that position: the
// woman's partner word only decreases, so every position the man has passed
// stays one she would refuse.
// `partner` holds n words of LDS, `cursor` one int of LDS, `red` BLOCK/64 words.
// The rank matrix must be complete and visible to the workgroup.  Thread 0
// receives the proposal count in *props.  Returns this thread's error bits.
template <int BLOCK, int G, int JMAX, bool kTransposed>
__device__ int wg_chains_g(const InstanceRef& d, unsigned long long* partner, int* cursor,
                           unsigned long long* red, unsigned long long* props) {
  static_assert(G >= 1 && G <= 64 && (G & (G - 1)) == 0, "group size");
  static_assert(JMAX >= 1 && JMAX <= 8, "window");
  constexpr int NW = BLOCK / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int gl = lane & (G - 1), gbase = lane & ~(G - 1);
  const int n = d.n;
  int bad = 0;
  for (int w = threadIdx.x; w < n; w += BLOCK) partner[w] = kEmpty;
  if (threadIdx.x == 0) *cursor = 0;
  __syncthreads();

  const int16_t* rk = d.rank;
  const int32_t* pm = d.pm;
  int m = -1, k = 0, J = 1;  // the same in all lanes of a group
  bool done = false;
  for (;;) {
    // Idle groups take the next man; a group that finds none is done.
    int got = 0;
    if (!done && m < 0 && gl == 0) got = atomicAdd(cursor, 1);
    got = __shfl(got, gbase);
    if (!done && m < 0) {
      if (got >= n) {
        done = true;
      } else {
        m = got;
        k = 0;
        J = 1;
      }
    }
    if (!done && k >= n) {
      bad |= kErrListEnd;
      done = true;
    }
    if (__ballot(!done) == 0) break;

    // Scan G * J positions; keep this lane's first candidate.
    int jf = 8;
    uint32_t wf = 0;
    unsigned long long mf = 0;
    if (!done) {
      int wv[JMAX];
      uint32_t rv[JMAX];
#pragma unroll
      for (int j = 0; j < JMAX; ++j) {
        const int pos = k + gl + G * j;
        wv[j] = -1;
        rv[j] = 0;
        if (j < J && pos < n) {
          wv[j] = pm[(size_t)m * n + pos];
          if ((unsigned)wv[j] >= (unsigned)n) {
            bad |= kErrEntryRange;
            wv[j] = -1;
          }
        }
      }
#pragma unroll
      for (int j = 0; j < JMAX; ++j)
        if (wv[j] >= 0) rv[j] = rank_of<kTransposed>(rk, n, (uint32_t)wv[j], (uint32_t)m);
#pragma unroll
      for (int j = 0; j < JMAX; ++j) {
        if (wv[j] >= 0 && jf == 8) {
          const unsigned long long mine = pw_pack(rv[j], (uint32_t)m, (uint32_t)(k + gl + G * j));
          const unsigned long long hint =
              __hip_atomic_load(&partner[wv[j]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          if (mine < hint) {
            jf = j;
            wf = (uint32_t)wv[j];
            mf = mine;
          }
        }
      }
    }
    // The group's first candidate in list order: smallest (j, gl).
    int key = jf * G + gl;
#pragma unroll
    for (int off = 1; off < G; off <<= 1) key = min(key, __shfl_xor(key, off));
    const bool owner = !done && jf < 8 && key == jf * G + gl;
    unsigned long long old = 0;
    if (owner) old = atomicMin(&partner[wf], mf);
    const int src = gbase + (key & (G - 1));
    old = __shfl(old, src);
    const int won = __shfl((int)(owner && old > mf), src);
    if (!done) {
      if (key >= 8 * G) {  // no candidate in the window
        k += G * J;
        J = min(2 * J, JMAX);
      } else if (won) {
        if (old == kEmpty) {
          m = -1;
        } else {
          m = (int)pw_man(old);
          k = (int)pw_pos(old) + 1;
        }
        J = 1;
      } else {  // lost the race for that woman
        k += key + 1;  // key = j * G + gl is the position's offset from k
        J = 1;
      }
    }
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

// ---------------------------------------------------------------------------
// Instance pool: one workgroup keeps NS instances open at once, each with its
// partner words in an LDS slot, and its chain groups take men from any open
// slot.  A group whose chain ends starts another man of any open instance, so
// no group idles while one instance finishes its last long displacement chain.
// Slot protocol (LDS atomics, workgroup scope):
//   state   FREE -> OPENING (one group won the slot, takes the next instance
//           from the global queue and sets its partner words to kEmpty)
//           -> OPEN -> FINAL (one group writes the matching) -> FREE;
//   active  counts the groups inside the slot: every chain in progress, and
//           a group that is probing for a man.  It is incremented before the
//           group checks that the slot is OPEN and before it takes a man, and
//           decremented when that group leaves, so while active > 0 the slot
//           cannot be finalized or reopened.  The decrement that brings it to 0
//           with every man handed out (cursor >= n) finalizes the instance, at
//           most once (CAS OPEN -> FINAL);
//   cursor  the next unstarted man, reset only while the slot is OPENING.
// A chain ends when its man claims a free woman (nobody displaced) or runs off
// the end of his list (an error).  All n men handed out and no chain running
// means every woman holds her final partner.
struct PoolSlot {
  int state, n, cursor, active;
  long long off, moff;  // element offset of the lists / of the matching
  int master;           // 1: every man holds the same list; the women propose
  int reach16;          // running mean, x16, of how far into a step's window its
                        // candidate lay (a step without one counts its whole
                        // window): sets the window a chain starts a step with
};
// The window (list positions) a pool chain scans at the start of a man's walk
// and after each claim: about twice the instance's running mean reach, 16 to 64.
// Independent lists find a candidate within a few positions, and scanning 64
// then fetched 4x the rank lines it used (each lookup is a line of its own);
// nearly shared lists refuse dozens of positions in a row and keep 64.
__device__ inline int pool_window0(int reach16) {
  const int want = (2 * reach16 + 15) >> 4;  // 2 x mean reach, rounded up
  if (want <= 16) return 16;
  if (want <= 32) return 32;
  return 64;
}
// r16 is the mean the caller read at the start of its step: r16 += (16 * reach -
// r16) / 8.  Racing groups may lose an update, which only delays the mean.
__device__ inline void pool_reach_note(PoolSlot& sl, int r16, int reach) {
  __hip_atomic_store(&sl.reach16, r16 + ((16 * reach - r16) >> 3), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
constexpr int kReach16Init = 8 * 16;  // a first window of 16 positions
enum { kSlotFree = 0, kSlotOpening = 1, kSlotOpen = 2, kSlotFinal = 3 };

__device__ inline int lds_ld(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ inline void lds_st(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Leaves the slot; true if the caller must finalize the instance.
__device__ inline bool pool_leave(PoolSlot& sl) {
  const int v = atomicSub(&sl.active, 1);
  if (v == 1 && lds_ld(&sl.cursor) >= lds_ld(&sl.n))
    return atomicCAS(&sl.state, kSlotOpen, kSlotFinal) == kSlotOpen;
  return false;
}

enum { kActNone = 0, kActMan = 1, kActInit = 2, kActFinal = 3, kActExit = 4 };

// One lane per group: find work.  Returns kActMan with (*os, *ov) = (slot,
// man), kActInit with (slot, instance) for a slot this group must open,
// kActFinal with the slot this group must finalize, kActExit when no work is
// left or can appear, else kActNone (try again next step).
template <int NS>
__device__ int pool_acquire(PoolSlot* slots, int* qdone, int* next_inst, int count, int start, int* os,
                            int* ov) {
  for (int i = 0; i < NS; ++i) {
    const int ss = (start + i) & (NS - 1);
    PoolSlot& sl = slots[ss];
    if (lds_ld(&sl.state) != kSlotOpen || lds_ld(&sl.cursor) >= lds_ld(&sl.n)) continue;
    atomicAdd(&sl.active, 1);
    if (lds_ld(&sl.state) != kSlotOpen) {
      atomicSub(&sl.active, 1);  // not inside an OPEN slot: never finalizes
      continue;
    }
    const int c = atomicAdd(&sl.cursor, 1);
    if (c < lds_ld(&sl.n)) {
      *os = ss;
      *ov = c;
      return kActMan;
    }
    if (pool_leave(sl)) {
      *os = ss;
      return kActFinal;
    }
  }
  if (!lds_ld(qdone)) {
    for (int i = 0; i < NS; ++i) {
      const int ss = (start + i) & (NS - 1);
      if (lds_ld(&slots[ss].state) == kSlotFree &&
          atomicCAS(&slots[ss].state, kSlotFree, kSlotOpening) == kSlotFree) {
        const int inst = atomicAdd(next_inst, 1);
        if (inst >= count) {
          lds_st(qdone, 1);
          lds_st(&slots[ss].state, kSlotFree);
          break;
        }
        *os = ss;
        *ov = inst;
        return kActInit;
      }
    }
  }
  if (lds_ld(qdone)) {
    for (int ss = 0; ss < NS; ++ss) {
      const int st = lds_ld(&slots[ss].state);
      if (st == kSlotOpening || (st == kSlotOpen && lds_ld(&slots[ss].cursor) < lds_ld(&slots[ss].n)))
        return kActNone;
    }
    return kActExit;
  }
  return kActNone;
}

// The batched chain walk over an instance pool (see wg_chains_g for the step).
// Desc has fields off, moff, n, idx.  `partner` holds NS * max_n words of LDS,
// `slots` NS slots, `qdone` one int.  Returns this thread's error bits;
// proposal counts go to *proposals.
// Master-list instances (n >= 2 and master_flags[d.idx] = 0, i.e. no man's list
// differs from man 0's: every man holds the same list L)
// are solved from the women's side: the stable matching is then unique, the
// proposers are the women with their lists in pw_all, a proposal of woman w
// carries the rank Linv[w] (linv_all at the instance's matching offset) that
// every man gives her, and the partner words belong to the men.  The reported
// count is the men's Σ(Linv[match[m]] + 1), what the man-proposing walk makes.
template <int BLOCK, int G, int JMAX, int NS, bool kTransposed, typename Desc>
__device__ int wg_pool_chains(const Desc* __restrict__ desc, int count, int* next_inst,
                              const int32_t* __restrict__ pm_all, const int16_t* __restrict__ rank_all,
                              int32_t* __restrict__ match_all, unsigned long long* proposals, int max_n,
                              unsigned long long* partner, PoolSlot* slots, int* qdone,
                              const int32_t* __restrict__ pw_all, const int* __restrict__ master_flags,
                              const int32_t* __restrict__ linv_all) {
  static_assert(G >= 1 && G <= 64 && (G & (G - 1)) == 0, "group size");
  static_assert(NS >= 1 && (NS & (NS - 1)) == 0, "slots");
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int gl = lane & (G - 1), gbase = lane & ~(G - 1);
  int bad = 0;

  // The first NS instances, opened by the whole workgroup.
  if (threadIdx.x == 0) {
    *qdone = 0;
    for (int ss = 0; ss < NS; ++ss) {
      const int inst = atomicAdd(next_inst, 1);
      PoolSlot& sl = slots[ss];
      sl.cursor = 0;
      sl.active = 0;
      if (inst < count) {
        const Desc d = desc[inst];
        sl.n = d.n;
        sl.off = d.off;
        sl.moff = d.moff;
        sl.master = master_flags && d.n >= 2 && !master_flags[d.idx];
        sl.reach16 = kReach16Init;
        sl.state = kSlotOpen;
      } else {
        sl.n = 0;
        sl.state = kSlotFree;
        *qdone = 1;
      }
    }
  }
  __syncthreads();
  for (int ss = 0; ss < NS; ++ss) {
    const int n0 = slots[ss].n;
    for (int w = threadIdx.x; w < n0; w += BLOCK) partner[(size_t)ss * max_n + w] = kEmpty;
  }
  __syncthreads();

  // Group-wide pieces of the protocol.
  auto finalize = [&](int ss) {
    PoolSlot& sl = slots[ss];
    const int n = sl.n;
    const unsigned long long* pw_ = partner + (size_t)ss * max_n;
    int32_t* match = match_all + sl.moff;
    unsigned long long p = 0;
    if (sl.master) {
      const int32_t* linv = linv_all + sl.moff;
      for (int a = gl; a < n; a += G) {  // the words belong to the men
        const unsigned long long v = pw_[a];
        if (v == kEmpty) {
          bad |= kErrUnmatched;
        } else {
          const uint32_t w = pw_man(v);
          if (w < (uint32_t)n) {
            match[a] = (int32_t)w;
            p += (unsigned long long)linv[w] + 1ull;
          }
        }
      }
    } else {
      for (int w = gl; w < n; w += G) {
        const unsigned long long v = pw_[w];
        if (v == kEmpty) {
          bad |= kErrUnmatched;
        } else {
          const uint32_t man = pw_man(v);
          if (man < (uint32_t)n) match[man] = w;
          p += pw_pos(v) + 1ull;
        }
      }
    }
#pragma unroll
    for (int off = 1; off < G; off <<= 1) p += __shfl_xor(p, off);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    if (gl == 0) {
      atomicAdd(proposals, p);
      lds_st(&sl.state, kSlotFree);
    }
  };
  auto open = [&](int ss, int inst) {
    const Desc d = desc[inst];
    unsigned long long* pw_ = partner + (size_t)ss * max_n;
    for (int w = gl; w < d.n; w += G) pw_[w] = kEmpty;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    if (gl == 0) {
      PoolSlot& sl = slots[ss];
      sl.n = d.n;
      sl.off = d.off;
      sl.moff = d.moff;
      sl.master = master_flags && d.n >= 2 && !master_flags[d.idx];
      sl.reach16 = kReach16Init;
      lds_st(&sl.cursor, 0);
      lds_st(&sl.state, kSlotOpen);
    }
  };

  const int start = (wave * (64 / G) + lane / G) & (NS - 1);
  int s = 0, n = 0, m = -1, k = 0, W = 16;  // W: positions scanned by the next step
  long long off = 0, moff = 0;
  bool exited = false, mst = false;
  uint32_t rprop = 0;  // master-list slot: the proposing woman's rank, Linv[m]
  for (;;) {
    int act = kActNone, a_s = 0, a_v = 0;
    if (!exited && m < 0 && gl == 0) act = pool_acquire<NS>(slots, qdone, next_inst, count, start, &a_s, &a_v);
    if constexpr (G == 64) {
      // One chain per wave: the leader's values are wave-uniform; read them into
      // scalar registers instead of an exchange through the LDS crossbar.
      act = __builtin_amdgcn_readfirstlane(act);
      a_s = __builtin_amdgcn_readfirstlane(a_s);
      a_v = __builtin_amdgcn_readfirstlane(a_v);
    } else {
      act = __shfl(act, gbase);
      a_s = __shfl(a_s, gbase);
      a_v = __shfl(a_v, gbase);
    }
    if (act == kActInit) {
      open(a_s, a_v);
    } else if (act == kActFinal) {
      finalize(a_s);
    } else if (act == kActExit) {
      exited = true;
    } else if (act == kActMan) {
      s = a_s;
      m = a_v;
      k = 0;
      W = pool_window0(__hip_atomic_load(&slots[s].reach16, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
      n = slots[s].n;
      off = slots[s].off;
      moff = slots[s].moff;
      mst = slots[s].master != 0;
      if (mst) rprop = (uint32_t)linv_all[moff + m];
    }
    if (__ballot(!exited) == 0) break;

    const bool live = !exited && m >= 0;
    bool ended = false;
    if (live && k >= n) {
      bad |= kErrListEnd;
      ended = true;
    }
    const bool scan = live && !ended;
    unsigned long long* pw_ = partner + (size_t)s * max_n;
    // The instance's running mean reach, read now so that the window after this
    // step's claim does not wait for an LDS load.
    const int reach16 =
        scan ? __hip_atomic_load(&slots[s].reach16, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) : 0;
    int jf = 8;
    uint32_t wf = 0;
    unsigned long long mf = 0;
    if (scan) {
      const int32_t* pm = (mst ? pw_all : pm_all) + off;  // the proposers' lists
      const int16_t* rk = rank_all + off;
      int wv[JMAX];
#pragma unroll
      for (int j = 0; j < JMAX; ++j) {
        const int pos = k + gl + G * j;
        wv[j] = -1;
        if (gl + G * j < W && pos < n) {
          wv[j] = pm[(size_t)m * n + pos];
          if ((unsigned)wv[j] >= (unsigned)n) {
            bad |= kErrEntryRange;
            wv[j] = -1;
          }
        }
      }
      uint32_t rv[JMAX];
#pragma unroll
      for (int j = 0; j < JMAX; ++j)
        rv[j] = wv[j] < 0 ? 0u : mst ? rprop : rank_of<kTransposed>(rk, n, (uint32_t)wv[j], (uint32_t)m);
#pragma unroll
      for (int j = 0; j < JMAX; ++j) {
        if (wv[j] >= 0 && jf == 8) {
          const unsigned long long mine = pw_pack(rv[j], (uint32_t)m, (uint32_t)(k + gl + G * j));
          const unsigned long long hint =
              __hip_atomic_load(&pw_[wv[j]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          if (mine < hint) {
            jf = j;
            wf = (uint32_t)wv[j];
            mf = mine;
          }
        }
      }
    }
    int key = jf * G + gl;
    if constexpr (G == 64) {
      // One chain per wave: the first candidate in list order from ballots, one
      // per window slot, instead of six rounds of lane exchanges.
      key = 8 * G;
#pragma unroll
      for (int j = 0; j < JMAX; ++j) {
        const unsigned long long b = __ballot(jf == j);
        if (b) {
          key = j * G + __ffsll((long long)b) - 1;
          break;
        }
      }
    } else {
#pragma unroll
      for (int o = 1; o < G; o <<= 1) key = min(key, __shfl_xor(key, o));
    }
    const bool owner = scan && jf < 8 && key == jf * G + gl;
    unsigned long long old = 0;
    if (owner) old = atomicMin(&pw_[wf], mf);
    const int src = gbase + (key & (G - 1));
    int won;
    if constexpr (G == 64) {
      // key came from ballots, so src is wave-uniform: read the owner's words
      // with readlane.
      const unsigned long long mo = ((unsigned long long)(uint32_t)__builtin_amdgcn_readlane((int)(mf >> 32), src) << 32) |
                                    (uint32_t)__builtin_amdgcn_readlane((int)mf, src);
      old = ((unsigned long long)(uint32_t)__builtin_amdgcn_readlane((int)(old >> 32), src) << 32) |
            (uint32_t)__builtin_amdgcn_readlane((int)old, src);
      won = key < 8 * G && old > mo;
    } else {
      old = __shfl(old, src);
      won = __shfl((int)(owner && old > mf), src);
    }
    if (scan) {
      if (gl == 0) pool_reach_note(slots[s], reach16, key >= 8 * G ? W : key + 1);
      if (key >= 8 * G) {
        k += W;
        W = min(2 * W, G * JMAX);
      } else if (won) {
        if (old == kEmpty) {
          ended = true;
        } else {
          m = (int)pw_man(old);
          k = (int)pw_pos(old) + 1;
          if (mst) rprop = (uint32_t)linv_all[moff + m];
        }
        W = pool_window0(reach16);
      } else {
        k += key + 1;
        W = pool_window0(reach16);
      }
    }
    if (ended) {
      m = -1;
      int fin = 0;
      if (gl == 0) fin = pool_leave(slots[s]) ? 1 : 0;
      if constexpr (G == 64) fin = __builtin_amdgcn_readfirstlane(fin);
      else fin = __shfl(fin, gbase);
      if (fin) finalize(s);
    }
  }
  return bad;
}

// ---------------------------------------------------------------------------
// One workgroup solves one instance from its lists in global memory (the batched
// fused mode and the serving lanes):
//   1. master-list check: rows 1..4 against row 0 on their first 64 entries,
//      then, for a candidate, every row (the workgroup's waves, stopping at the
//      first difference);
//   2. the women's rank rows (rank[w][pref_w[w][r]] = r, the F9RankMatrixInitKernel
//      step, src/smp/kernel.cu:672-686): rounds of whole rows filling the LDS
//      staging, every load of a round issued before the staging is cleared;
//      written as int16 with plain stores, so the walk finds them in this XCD's
//      L2 (a master-list instance only validates its rows);
//   3. for a master-list instance, Linv[L[k]] = k in LDS;
//   4. the walk (the La3F9CoreKernel proposal, src/smp/kernel.cu:633-670): one
//      chain per wave, the pool walk's step (a window of about twice the
//      instance's running mean reach, doubling to at most 64 positions after an
//      empty step, first candidate from ballots, claim with a 64-bit LDS
//      atomicMin, wave-uniform values through readlane); a master-list
//      instance is walked from the women's side with the rank Linv[w] every man
//      gives woman w;
//   5. the matching (PostProc, src/smp/smp_engine_bamboo.cu:299-303) written to
//      `match`, and this thread's share of the proposal count (the men's sum of
//      list position + 1) in *props.
// LDS: `partner` and `linv` hold n words each, `stage` BLOCK * fused_per(BLOCK)
// entries (which must be >= n), `ctl` three ints.  All threads of the workgroup
// call it together.  Returns this thread's error bits.
__host__ __device__ constexpr int fused_per(int) { return 8; }  // 64 VGPRs with no spill
inline size_t fused_lds(int block, int max_n) {
  return (size_t)max_n * 12 + (size_t)block * fused_per(block) * 2;
}

template <int BLOCK>
__device__ int wg_solve_instance(const int32_t* __restrict__ pm, const int32_t* __restrict__ pw,
                                 int16_t* __restrict__ rank, int32_t* __restrict__ match, int n,
                                 unsigned long long* partner, int32_t* linv, uint16_t* stage, int* ctl,
                                 unsigned long long* props) {
  constexpr int NW = BLOCK / 64;
  constexpr int kPer = fused_per(BLOCK);
  constexpr int kStage = BLOCK * kPer;
  // At most 64 positions per step: 4,096 mixed instances 6.5 -> 6.0 ms against
  // 128 (each examined position looks up a rank on a line of its own), 7.0 ms
  // at 32 (nearly shared lists refuse dozens of positions in a row).
  constexpr int G = 64, JMAX = 1;
  int& ctl_cursor = ctl[0];
  int& ctl_differs = ctl[1];
  int& ctl_reach16 = ctl[2];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int bad = 0;
  for (int w = threadIdx.x; w < n; w += BLOCK) partner[w] = kEmpty;
  if (threadIdx.x == 0) {
    ctl_cursor = 0;
    ctl_differs = n < 2 ? 1 : 0;
    ctl_reach16 = kReach16Init;
  }
  __syncthreads();
  // 1. Master-list check.
  if (wave == 0 && n >= 2) {
    bool diff = false;
    for (int m = 1; m < min(n, 5) && !diff; ++m)
      diff = __ballot(lane < n && pm[(size_t)m * n + lane] != pm[lane]) != 0;
    if (diff && lane == 0) ctl_differs = 1;
  }
  __syncthreads();
  if (!ctl_differs) {
    for (int m = 1 + wave; m < n; m += NW) {
      if (__hip_atomic_load(&ctl_differs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
      const int32_t* row = pm + (size_t)m * n;
      bool diff = false;
      for (int k0 = 0; k0 < n && !diff; k0 += 512) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int k = k0 + 64 * j + lane;
          if (k < n && row[k] != pm[k]) diff = true;
        }
        diff = __ballot(diff) != 0;
      }
      if (diff) {
        if (lane == 0) atomicOr(&ctl_differs, 1);
        break;
      }
    }
  }
  __syncthreads();
  const bool mst = __builtin_amdgcn_readfirstlane(ctl_differs) == 0;
  // 2. Rank rows (validated always, written unless the instance is a master list).
  const int rows_per = kStage / n;  // >= 1: the host keeps kStage >= max_n
  // Entry e = threadIdx.x + j * BLOCK of a round lies in row e / n, column
  // e % n; from one step to the next both advance by BLOCK / n and BLOCK % n.
  const int step_r = BLOCK / n, step_c = BLOCK % n;
  const int row0 = (int)threadIdx.x / n, col0 = (int)threadIdx.x - row0 * n;
  for (int r0 = 0; r0 < n; r0 += rows_per) {
    const int cnt = min(rows_per, n - r0) * n;
    const int32_t* src = pw + (size_t)r0 * n;
    int v[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const int e = (int)threadIdx.x + j * BLOCK;
      v[j] = e < cnt ? src[e] : 0;
    }
    for (int e = threadIdx.x; e < cnt; e += BLOCK) stage[e] = 0xFFFFu;
    __syncthreads();
    int row = row0, col = col0;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const int e = (int)threadIdx.x + j * BLOCK;
      if (e < cnt) {
        if ((unsigned)v[j] < (unsigned)n) stage[row * n + v[j]] = (uint16_t)col;
        else bad |= kErrEntryRange;
      }
      row += step_r;
      col += step_c;
      if (col >= n) col -= n, ++row;
    }
    __syncthreads();
    int16_t* dst = rank + (size_t)r0 * n;
    for (int e = threadIdx.x; e < cnt; e += BLOCK) {
      const uint16_t x = stage[e];
      if (x == 0xFFFFu) bad |= kErrNotPerm;
      if (!mst) dst[e] = (int16_t)x;
    }
    __syncthreads();
  }
  // 3. Linv of the shared men's list L (row 0).
  if (mst) {
    for (int k = threadIdx.x; k < n; k += BLOCK) linv[k] = -1;
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += BLOCK) {
      const int w = pm[k];
      if ((unsigned)w >= (unsigned)n) bad |= kErrEntryRange;
      else if (atomicCAS(&linv[w], -1, k) != -1) bad |= kErrNotPerm;
    }
    __syncthreads();
  }
  // 4. The walk, one chain per wave.
  {
    const int32_t* P = mst ? pw : pm;  // the proposers' lists
    int m = -1, k = 0, W = 16;
    uint32_t rprop = 0;
    for (;;) {
      if (m < 0) {
        int got = 0;
        if (lane == 0) got = atomicAdd(&ctl_cursor, 1);
        got = __builtin_amdgcn_readfirstlane(got);
        if (got >= n) break;
        m = got;
        k = 0;
        W = pool_window0(__hip_atomic_load(&ctl_reach16, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
        if (mst) rprop = (uint32_t)max(linv[m], 0);
      }
      if (k >= n) {  // rows that are not permutations
        bad |= kErrListEnd;
        m = -1;
        continue;
      }
      const int reach16 = __hip_atomic_load(&ctl_reach16, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      int wv[JMAX];
      uint32_t rv[JMAX];
#pragma unroll
      for (int j = 0; j < JMAX; ++j) {
        const int pos = k + lane + G * j;
        wv[j] = -1;
        if (lane + G * j < W && pos < n) {
          wv[j] = P[(size_t)m * n + pos];
          if ((unsigned)wv[j] >= (unsigned)n) {
            bad |= kErrEntryRange;
            wv[j] = -1;
          }
        }
      }
#pragma unroll
      for (int j = 0; j < JMAX; ++j)
        rv[j] = wv[j] < 0 ? 0u : mst ? rprop : (uint32_t)(uint16_t)rank[(size_t)wv[j] * n + m];
      int jf = 8;
      uint32_t wf = 0;
      unsigned long long mf = 0;
#pragma unroll
      for (int j = 0; j < JMAX; ++j) {
        if (wv[j] >= 0 && jf == 8) {
          const unsigned long long mine = pw_pack(rv[j], (uint32_t)m, (uint32_t)(k + lane + G * j));
          const unsigned long long hint =
              __hip_atomic_load(&partner[wv[j]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          if (mine < hint) {
            jf = j;
            wf = (uint32_t)wv[j];
            mf = mine;
          }
        }
      }
      int key = 8 * G;
#pragma unroll
      for (int j = 0; j < JMAX; ++j) {
        const unsigned long long b = __ballot(jf == j);
        if (b) {
          key = j * G + __ffsll((long long)b) - 1;
          break;
        }
      }
      const bool owner = jf < 8 && key == jf * G + lane;
      unsigned long long old = 0;
      if (owner) old = atomicMin(&partner[wf], mf);
      const int src = key & (G - 1);
      const unsigned long long mo =
          ((unsigned long long)(uint32_t)__builtin_amdgcn_readlane((int)(mf >> 32), src) << 32) |
          (uint32_t)__builtin_amdgcn_readlane((int)mf, src);
      old = ((unsigned long long)(uint32_t)__builtin_amdgcn_readlane((int)(old >> 32), src) << 32) |
            (uint32_t)__builtin_amdgcn_readlane((int)old, src);
      const bool won = key < 8 * G && old > mo;
      if (lane == 0) {
        const int reach = key >= 8 * G ? W : key + 1;
        __hip_atomic_store(&ctl_reach16, reach16 + ((16 * reach - reach16) >> 3), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
      }
      if (key >= 8 * G) {
        k += W;
        W = min(2 * W, G * JMAX);
      } else if (won) {
        if (old == kEmpty) {
          m = -1;
        } else {
          m = (int)pw_man(old);
          k = (int)pw_pos(old) + 1;
          if (mst) rprop = (uint32_t)max(linv[m], 0);
        }
        W = pool_window0(reach16);
      } else {
        k += key + 1;
        W = pool_window0(reach16);
      }
    }
  }
  __syncthreads();
  // 5. The matching and the proposal count (the men's sum of list position + 1).
  unsigned long long p = 0;
  for (int a = threadIdx.x; a < n; a += BLOCK) {
    const unsigned long long v = partner[a];
    if (v == kEmpty) {
      bad |= kErrUnmatched;
    } else if (mst) {  // the words belong to the men
      const uint32_t w = pw_man(v);
      if (w < (uint32_t)n) {
        match[a] = (int32_t)w;
        p += (unsigned long long)max(linv[w], 0) + 1ull;
      }
    } else {
      const uint32_t man = pw_man(v);
      if (man < (uint32_t)n) match[man] = a;
      p += pw_pos(v) + 1ull;
    }
  }
  *props = p;
  return bad;
}

// ---------------------------------------------------------------------------
// Flat rank rows for many instances at once.  Row g of the launch is row
// g - row0 of the task t with row0 <= g < row0 + n (t = row_task[g] when the
// map is given, else a binary search over the tasks' row0).  Each wave owns a
// 16-bit LDS row buffer and takes rows on its own (grid-stride over waves, no
// workgroup barrier); each lane keeps kIn loads in flight before scattering
// them, so the launch streams the women's lists at HBM rate whatever the mix of
// instance sizes.  Errors go to errs[t * err_stride].  The rank[w][pref_w[w][r]] = r step of F9RankMatrixInitKernel
// (src/smp/kernel.cu:672-686), int16 as every n here is below 32768.
struct RankTask {
  const int32_t* pw;
  int16_t* rank;
  int n;
  int row0;
};

__device__ inline void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int BLOCK, typename Task>
__global__ __launch_bounds__(BLOCK) void k_rank_rows16(const Task* __restrict__ tasks, int ntasks,
                                                       const int* __restrict__ row_task, int total_rows,
                                                       int max_n, int* errs, int err_stride) {
  constexpr int NW = BLOCK / 64;
  constexpr int kIn = 8;
  extern __shared__ uint16_t rbuf16[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint16_t* buf = rbuf16 + (size_t)wave * max_n;
  for (int g = blockIdx.x * NW + wave; g < total_rows; g += gridDim.x * NW) {
    int ti;
    if (row_task) {
      ti = row_task[g];
    } else {
      int lo = 0, hi = ntasks - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tasks[mid].row0 <= g) lo = mid;
        else hi = mid - 1;
      }
      ti = lo;
    }
    const int n = tasks[ti].n;
    const int row = g - tasks[ti].row0;
    int bad = 0;
    for (int c = lane; c < n; c += 64) buf[c] = 0xFFFFu;
    wave_sync_lds();
    // The row as 16-byte loads between a scalar head (up to 16-byte alignment:
    // rows of odd-sized instances start anywhere) and a scalar tail, kIn vectors
    // per lane in flight.
    const int32_t* src = tasks[ti].pw + (size_t)row * n;
    const int head = min(n, (int)((4 - (((uintptr_t)src >> 2) & 3)) & 3));
    const int nvec = (n - head) >> 2;
    const int tail0 = head + 4 * nvec;
    if (lane < head) {
      const int v = __builtin_nontemporal_load(src + lane);
      if ((unsigned)v < (unsigned)n) buf[v] = (uint16_t)lane;
      else bad |= kErrEntryRange;
    }
    if (lane < n - tail0) {
      const int v = __builtin_nontemporal_load(src + tail0 + lane);
      if ((unsigned)v < (unsigned)n) buf[v] = (uint16_t)(tail0 + lane);
      else bad |= kErrEntryRange;
    }
    const int4* src4 = reinterpret_cast<const int4*>(src + head);
    for (int q0 = lane; q0 < nvec; q0 += 64 * kIn) {
      int4 v[kIn];
#pragma unroll
      for (int j = 0; j < kIn; ++j) {
        const int q = min(q0 + 64 * j, nvec - 1);
        v[j].x = __builtin_nontemporal_load(&src4[q].x);
        v[j].y = __builtin_nontemporal_load(&src4[q].y);
        v[j].z = __builtin_nontemporal_load(&src4[q].z);
        v[j].w = __builtin_nontemporal_load(&src4[q].w);
      }
#pragma unroll
      for (int j = 0; j < kIn; ++j) {
        const int q = q0 + 64 * j;
        if (q < nvec) {
          const int r = head + 4 * q;
          const int e[4] = {v[j].x, v[j].y, v[j].z, v[j].w};
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            if ((unsigned)e[t] < (unsigned)n) buf[e[t]] = (uint16_t)(r + t);
            else bad |= kErrEntryRange;
          }
        }
      }
    }
    wave_sync_lds();
    // The ranks as 16-byte stores (eight per store) between a scalar head and
    // tail, the same way.
    int16_t* dst = tasks[ti].rank + (size_t)row * n;
    const int whead = min(n, (int)((8 - (((uintptr_t)dst >> 1) & 7)) & 7));
    const int wvec = (n - whead) >> 3;
    const int wtail0 = whead + 8 * wvec;
    if (lane < whead) {
      const uint16_t v = buf[lane];
      if (v == 0xFFFFu) bad |= kErrNotPerm;
      dst[lane] = (int16_t)v;
    }
    if (lane < n - wtail0) {
      const uint16_t v = buf[wtail0 + lane];
      if (v == 0xFFFFu) bad |= kErrNotPerm;
      dst[wtail0 + lane] = (int16_t)v;
    }
    typedef unsigned int v4u __attribute__((ext_vector_type(4)));
    v4u* dst8 = reinterpret_cast<v4u*>(dst + whead);
    for (int q = lane; q < wvec; q += 64) {
      const uint16_t* b = buf + whead + 8 * q;
      uint32_t wd[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const uint32_t lo = b[2 * t], hi = b[2 * t + 1];
        if (lo == 0xFFFFu || hi == 0xFFFFu) bad |= kErrNotPerm;
        wd[t] = lo | (hi << 16);
      }
      __builtin_nontemporal_store(v4u{wd[0], wd[1], wd[2], wd[3]}, dst8 + q);
    }
    if (bad) atomicOr(&errs[(size_t)ti * err_stride], bad);
    wave_sync_lds();  // every lane has read the row before the next reset
  }
}

inline size_t rank_rows16_lds(int block, int max_n) { return (size_t)(block / 64) * max_n * 2; }

template <int BLK, typename Task>
hipError_t launch_rank_rows16_t(int grid, size_t lds, hipStream_t s, const Task* tasks, int ntasks,
                                const int* row_task, int total_rows, int max_n, int* errs, int err_stride) {
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_rank_rows16<BLK, Task>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_rank_rows16<BLK, Task>), dim3(grid), dim3(BLK), lds, s, tasks, ntasks, row_task,
                     total_rows, max_n, errs, err_stride);
  return hipGetLastError();
}

// k_rank_rows16 over `total_rows` rows: 4 waves per workgroup (fewer when the
// row buffers of the largest instance do not fit), as many workgroups per CU as
// the buffers allow, at most 8, never more workgroups than row groups.
template <typename Task>
hipError_t launch_rank_rows16(hipStream_t s, const Task* tasks, int ntasks, const int* row_task,
                              int total_rows, int max_n, int* errs, int err_stride, int cus, int max_lds) {
  int blk = 256;
  while (blk > 64 && rank_rows16_lds(blk, max_n) > (size_t)max_lds) blk >>= 1;
  const size_t lds = rank_rows16_lds(blk, max_n);
  const int per_cu = std::max(1, std::min<int>(8, (int)((160 * 1024) / std::max<size_t>(lds, 1))));
  const long long groups = ((long long)total_rows + blk / 64 - 1) / (blk / 64);
  const int grid = (int)std::max(1LL, std::min<long long>(groups, (long long)cus * per_cu));
  switch (blk) {
    case 256:
      return launch_rank_rows16_t<256>(grid, lds, s, tasks, ntasks, row_task, total_rows, max_n, errs,
                                       err_stride);
    case 128:
      return launch_rank_rows16_t<128>(grid, lds, s, tasks, ntasks, row_task, total_rows, max_n, errs,
                                       err_stride);
    default:
      return launch_rank_rows16_t<64>(grid, lds, s, tasks, ntasks, row_task, total_rows, max_n, errs,
                                      err_stride);
  }
}

// ---------------------------------------------------------------------------
// Transposed rank matrix, rankT[m][w] = r where pref_w[w][r] = m, for many
// instances at once.  A tile is T consecutive women (w0 .. w0 + T - 1) of one
// instance; their rows are tw * n consecutive ints.  The workgroup reads them
// as 16-byte vectors (scalar head and tail up to 16-byte alignment, kIn vectors
// in flight per lane), scatters each entry into an LDS tile laid out [m][t]
// with a row stride of T + 2 halfwords (so lanes that scatter to random m
// spread over the LDS banks), then writes, for every man m, the T consecutive
// ranks rankT[m][w0 .. w0 + T - 1] (2T contiguous bytes).  Reads and writes
// are one pass over HBM each.  The workgroup barriers order LDS only, so one
// tile's global stores drain while the next tile loads.  Tiles are (task, w0)
// pairs; errors go to errs[task * err_stride].
__device__ inline void wg_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <int BLOCK, int T, int KIN, int MINW, typename Task>
__global__ __launch_bounds__(BLOCK, MINW) void k_rankT_tiles(const Task* __restrict__ tasks,
                                                             const int2* __restrict__ tiles, int ntiles,
                                                             int* errs, int err_stride) {
  static_assert(BLOCK % T == 0 && T % 2 == 0, "tile shape");
  constexpr int S = T + 2;   // LDS row stride in halfwords
  constexpr int kIn = KIN;   // 16-byte loads per lane per round
  typedef int v4i __attribute__((ext_vector_type(4)));
  extern __shared__ uint32_t tileT32[];
  uint16_t* tile = reinterpret_cast<uint16_t*>(tileT32);
  const int t_wr = threadIdx.x % T, m_wr = threadIdx.x / T;

  // One tile's geometry and its first round of loads, held in registers so the
  // next tile's loads are in flight while the current tile is written out.
  struct Round {
    int task, n, w0, tw, total, head, nvec, tail0;
    const int32_t* src;
    int hv, tv;
    v4i v[kIn];
  };
  auto issue = [&](int ti, Round& R) {
    const int2 tl = tiles[ti];
    R.task = tl.x;
    R.n = tasks[tl.x].n;
    R.w0 = tl.y;
    R.tw = min(T, R.n - R.w0);
    R.total = R.tw * R.n;
    R.src = tasks[tl.x].pw + (size_t)R.w0 * R.n;
    R.head = min(R.total, (int)((4 - (((uintptr_t)R.src >> 2) & 3)) & 3));
    R.nvec = (R.total - R.head) >> 2;
    R.tail0 = R.head + 4 * R.nvec;
    R.hv = (int)threadIdx.x < R.head ? __builtin_nontemporal_load(R.src + threadIdx.x) : 0;
    R.tv = (int)threadIdx.x < R.total - R.tail0 ? __builtin_nontemporal_load(R.src + R.tail0 + threadIdx.x) : 0;
    const v4i* vsrc = reinterpret_cast<const v4i*>(R.src + R.head);
#pragma unroll
    for (int j = 0; j < kIn; ++j) {
      const int c = (int)threadIdx.x + BLOCK * j;
      R.v[j] = c < R.nvec ? __builtin_nontemporal_load(vsrc + c) : v4i{0, 0, 0, 0};
    }
  };

  // XCD-aware order: workgroup b runs on XCD b % 8, so the workgroups of one
  // XCD take consecutive tiles, and the neighbouring tiles of an instance (which
  // write the two halves of the same rankT lines) meet in that XCD's L2 instead
  // of reaching HBM as partial lines from two XCDs.
  const int xcds = (gridDim.x % 8 == 0) ? 8 : 1;
  const int per_xcd = gridDim.x / xcds;
  Round R;
  int ti = (blockIdx.x % xcds) * per_xcd + blockIdx.x / xcds;
  if (ti < ntiles) issue(ti, R);
  while (ti < ntiles) {
    const int n = R.n, tw = R.tw, task = R.task;
    int bad = 0;
    for (int i = threadIdx.x; i < n * (S / 2); i += BLOCK) tileT32[i] = 0xFFFFFFFFu;
    wg_sync_lds();
    // Entry e of the tile is woman t = e / n, rank r = e % n; the quotient comes
    // from a float reciprocal and one correction step (exact for e < 2^22).
    const float inv_n = 1.0f / (float)n;
    auto put = [&](int e, int m) {
      int t = (int)(((float)e + 0.5f) * inv_n);
      int r = e - t * n;
      if (r < 0) r += n, --t;
      else if (r >= n) r -= n, ++t;
      if ((unsigned)m < (unsigned)n) tile[m * S + t] = (uint16_t)r;
      else bad |= kErrEntryRange;
    };
    if ((int)threadIdx.x < R.head) put(threadIdx.x, R.hv);
    if ((int)threadIdx.x < R.total - R.tail0) put(R.tail0 + threadIdx.x, R.tv);
#pragma unroll
    for (int j = 0; j < kIn; ++j) {
      const int c = (int)threadIdx.x + BLOCK * j;
      if (c < R.nvec) {
        const int e = R.head + 4 * c;
        put(e, R.v[j].x);
        put(e + 1, R.v[j].y);
        put(e + 2, R.v[j].z);
        put(e + 3, R.v[j].w);
      }
    }
    // Tiles larger than one round: the remaining vectors, kIn per lane at a
    // time, in R's registers.
    const v4i* vsrc = reinterpret_cast<const v4i*>(R.src + R.head);
    for (int c0 = (int)threadIdx.x + BLOCK * kIn; c0 < R.nvec; c0 += BLOCK * kIn) {
#pragma unroll
      for (int j = 0; j < kIn; ++j) {
        const int c = c0 + BLOCK * j;
        R.v[j] = c < R.nvec ? __builtin_nontemporal_load(vsrc + c) : v4i{0, 0, 0, 0};
      }
#pragma unroll
      for (int j = 0; j < kIn; ++j) {
        const int c = c0 + BLOCK * j;
        if (c < R.nvec) {
          const int e = R.head + 4 * c;
          put(e, R.v[j].x);
          put(e + 1, R.v[j].y);
          put(e + 2, R.v[j].z);
          put(e + 3, R.v[j].w);
        }
      }
    }
    wg_sync_lds();
    int16_t* dst = tasks[task].rank + R.w0 + t_wr;
    const int next = ti + gridDim.x;
    if (next < ntiles) issue(next, R);  // R's registers are free again
    if (t_wr < tw) {
#pragma unroll 4
      for (int m = m_wr; m < n; m += BLOCK / T) {
        const uint16_t v = tile[m * S + t_wr];
        if (v == 0xFFFFu) bad |= kErrNotPerm;
        dst[(size_t)m * n] = (int16_t)v;
      }
    }
    if (bad) atomicOr(&errs[(size_t)task * err_stride], bad);
    wg_sync_lds();  // the tile is read out before the next reset
    ti = next;
  }
}

inline size_t rankT_tile_lds(int T, int max_n) { return (size_t)(T + 2) * max_n * 2; }

// Widest tile (up to 64 women) whose LDS image [max_n][T + 2] fits 160 KiB.
inline int rankT_tile_width(int max_n) {
  int T = 64;
  while (T > 2 && rankT_tile_lds(T, max_n) > 160 * 1024) T >>= 1;
  return T;
}

template <int BLK, int T, int KIN = 16, int MINW = 1, typename Task>
hipError_t launch_rankT_t(int grid, hipStream_t s, const Task* tasks, const int2* tiles, int ntiles,
                          int max_n, int* errs, int err_stride) {
  const size_t lds = rankT_tile_lds(T, max_n);
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_rankT_tiles<BLK, T, KIN, MINW, Task>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_rankT_tiles<BLK, T, KIN, MINW, Task>), dim3(grid), dim3(BLK), lds, s, tasks, tiles,
                     ntiles, errs, err_stride);
  return hipGetLastError();
}

// k_rankT_tiles over `ntiles` tiles of width rankT_tile_width(max_n), 1024
// threads per workgroup, as many workgroups as fit the CUs' LDS.
template <typename Task>
hipError_t launch_rankT(hipStream_t s, const Task* tasks, const int2* tiles, int ntiles, int max_n, int* errs,
                        int err_stride, int cus) {
  const int T = rankT_tile_width(max_n);
  const size_t lds = rankT_tile_lds(T, max_n);
  const int per_cu = std::max(1, std::min<int>(2, (int)((160 * 1024) / std::max<size_t>(lds, 1))));
  const int grid = std::max(1, std::min(ntiles, cus * per_cu));
  switch (T) {
    case 64: return launch_rankT_t<1024, 64>(grid, s, tasks, tiles, ntiles, max_n, errs, err_stride);
    case 32: return launch_rankT_t<1024, 32>(grid, s, tasks, tiles, ntiles, max_n, errs, err_stride);
    case 16: return launch_rankT_t<1024, 16>(grid, s, tasks, tiles, ntiles, max_n, errs, err_stride);
    case 8: return launch_rankT_t<1024, 8>(grid, s, tasks, tiles, ntiles, max_n, errs, err_stride);
    case 4: return launch_rankT_t<1024, 4>(grid, s, tasks, tiles, ntiles, max_n, errs, err_stride);
    default: return launch_rankT_t<1024, 2>(grid, s, tasks, tiles, ntiles, max_n, errs, err_stride);
  }
}

// The tiles of a set of instances, largest instances first.
inline void rankT_tiles(const int* sizes, int count, int max_n, std::vector<int2>* out) {
  const int T = rankT_tile_width(max_n);
  std::vector<int> order(count);
  for (int i = 0; i < count; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return sizes[a] > sizes[b]; });
  out->clear();
  for (int i : order)
    for (int w0 = 0; w0 < sizes[i]; w0 += T) out->push_back(make_int2(i, w0));
}

}  // namespace pactae
