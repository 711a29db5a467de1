/*
 * smp_hip.h — C ABI of the MI355X stable-marriage (BambooSMP-shaped) proposal engine.
 *
 * Plain pointers and sizes only; no C++ or framework types cross this boundary.
 * Every call returns an int status (SMP_OK == 0) and never throws; the text of
 * the last error on the calling thread is available from smp_last_error().
 *
 * Data layout (same as the reference's flat lists, include/utils/types.h:33):
 *   pref_m[m*n + k] = man m's k-th choice of woman   (int32, row-major n x n)
 *   pref_w[w*n + k] = woman w's k-th choice of man   (int32, row-major n x n)
 *   result match[m] = woman matched to man m         (int32[n]; Matching, types.h:44)
 * The result is the man-optimal stable matching, which is unique.
 *
 * Reference interfaces each entry point replaces (paths relative to the
 * reference repository root):
 *   smp_create         SmpEngineBamboo::SmpEngineBamboo(const SmpObj&, size_t n)
 *                      src/smp/smp_engine_bamboo.cu:6-31 (device allocation + H2D of
 *                      both lists; SmpObj::CreateFromFlatLists include/smp/smp.h:25-29)
 *   smp_create_device  same constructor, for lists already resident in HBM (borrowed)
 *   smp_create_ranked, smp_create_device_ranked
 *                      same constructor with the women's inverse rank matrix as input,
 *                      as the rank_w the reference builds in InitProc
 *                      (F9RankMatrixInitKernel src/smp/kernel.cu:672-686) but derived by
 *                      the caller; init keeps only the gather (kernel.cu:688-709)
 *   smp_exec           AbsSmpEngine::FindStableMatching include/smp/smp_engine_abs.h:10-36
 *                      (precheck IsPerfect smp_engine_bamboo.h:25-33, InitProc
 *                      smp_engine_bamboo.cu:33-106, CoreProc :108-287, PostProc :289-306)
 *   smp_result         the Matching returned by FindStableMatching (abs.h:35)
 *   smp_profile        AbsSmpEngine::PrintProfilingInfo abs.h:38-52 (precheck/init/core/total ms)
 *   smp_destroy        ~SmpEngineBamboo smp_engine_bamboo.h:18-19 (RAII device arrays)
 *   smp_generate       GenerateWorkload include/utils/generate_workloads_flat.h:250-280
 *                      (seeded here; the reference seeds from std::random_device)
 *   smp_batch_*        no reference counterpart: the reference runs one instance per
 *                      SmpEngineExecutor call (benchmark/smp_engine_exec.h:6-32); this is
 *                      that executor looped over many instances, in one launch.
 *   smp_server_*       no reference counterpart: request-serving loop around the engine.
 */
#ifndef PACTAE_SMP_HIP_H
#define PACTAE_SMP_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define SMP_OK 0
#define SMP_ERR_INVALID 1  /* bad argument or malformed preference lists      */
#define SMP_ERR_HIP 2      /* HIP runtime error (no device, launch failure...) */
#define SMP_ERR_NOMEM 3    /* device or pinned host allocation failed          */
#define SMP_ERR_STATE 4    /* call out of order (e.g. result before exec)      */
#define SMP_ERR_TIMEOUT 5  /* a bounded wait expired                          */
#define SMP_PENDING 6      /* server: request not finished yet                */

/* Largest n accepted: women, men and list positions are packed into 16-bit
 * fields of the men-major PR entries. */
#define SMP_MAX_N 65535

/* ---- workloads ---------------------------------------------------------- */
enum smp_workload {
  SMP_WL_UNIFORM = 0,     /* every row an independent uniform permutation          */
  SMP_WL_CONGESTED = 1,   /* all men share one permutation L; women uniform         */
  SMP_WL_SHIFTED = 2,     /* man i's list is L rotated left by i; women uniform     */
  SMP_WL_CORRELATED = 3,  /* men: one base permutation + param transpositions each;
                             param <= 0 means n/10; women uniform                   */
  SMP_WL_REF_RANDOM = 4,  /* reference RANDOM: shuffled id order, rows shuffled in
                             groups of param (default 5); both sides
                             (generate_workloads_flat.h:85-137)                     */
  SMP_WL_REF_CONGESTED = 5, /* reference CONGESTED: men share L, women share P (:14-47) */
  SMP_WL_REF_PERFECT = 6, /* reference PERFECT: distinct first choices (:50-82)     */
  SMP_WL_REF_SOLO = 7     /* reference SOLO: long displacement chain (:140-247)     */
};

/* Fill pref_m and pref_w (each n*n int32, host memory) with workload `kind`.
 * Deterministic in (kind, n, seed, param): every row is drawn from its own
 * xoshiro256** stream seeded through SplitMix64 from (seed, side, row), with
 * Fisher-Yates and unbiased bounded draws.  threads <= 0 picks a default. */
int smp_generate(int32_t kind, int32_t n, uint64_t seed, int32_t param, int32_t* pref_m,
                 int32_t* pref_w, int32_t threads);

/* The reference's on-disk workload cache (include/utils/generate_workloads_flat.h:
 * 283-376): text, first line n, then n rows of pref_m and n rows of pref_w, each
 * value followed by a space.  smp_workload_load with pref_m == NULL only reads n;
 * otherwise capacity (elements of each output) must be >= n*n. */
int smp_workload_save(const char* path, int32_t n, const int32_t* pref_m, const int32_t* pref_w);
int smp_workload_load(const char* path, int32_t* n_out, int32_t* pref_m, int32_t* pref_w,
                      size_t capacity);
/* GetWorkloadFilename (generate_workloads_flat.h:285-300) with the reference's
 * WorkloadType numbering: PERFECT 0, RANDOM 1, CONGESTED 2, SOLO 3. */
int smp_workload_filename(int32_t ref_type, int32_t n, int32_t group, char* buf, size_t len);

/* ---- single-instance engine ---------------------------------------------- */
typedef struct smp_engine smp_engine;

enum smp_tail_mode {
  SMP_TAIL_AUTO = 0, /* hand the last free men to the host when that is faster     */
  SMP_TAIL_GPU = 1,  /* finish every chain in the persistent proposal kernel       */
  SMP_TAIL_HOST = 2  /* always hand off at tail_threshold free men                 */
};

typedef struct smp_opts {
  int32_t prefix_cols;    /* PR columns gathered at init: 0 = auto, >= n = full n x n
                             (the reference's layout); beyond it the kernel reads
                             pref_m and the rank matrix directly.  Auto is 128 (or
                             smp_tune's choice), widened to 256 for n >= 1024 when
                             the sparse pass overflows a bucket (lists that share
                             their top women)                                     */
  int32_t tail_mode;      /* enum smp_tail_mode                                    */
  int32_t tail_threshold; /* free-man count at which the host takes over; 0 = auto */
  int32_t device;         /* HIP device ordinal                                    */
  int32_t precheck;       /* 1 = IsPerfect fast path (default), 0 = skip           */
  int32_t waves_per_cu;   /* proposal-kernel occupancy; 0 = auto                   */
  int32_t validate;       /* 1 = reject rows that are not permutations (default)   */
  int32_t coherent_hint;  /* 1 = partner-word hints read past non-coherent caches;
                             0 = plain cached reads (default; a stale hint costs a
                             failed atomic, never correctness)                    */
  int32_t poll_interval;  /* proposal-kernel steps between reads of the shared
                             active-wave counter that triggers the hand-off; 0 = 16 */
  int32_t host_spin;      /* 1 = the host polls the runtime for device completion,
                             0 = it blocks in the runtime (default; polling was
                             measured slower).  The two hand-offs on the critical
                             path (after the precheck and after the proposal
                             kernel) are always polled, through flag words the
                             device writes into pinned host memory            */
  int32_t prefix_copy;    /* how the PR prefix reaches the host for the tail:
                             0 = copied beside the proposal kernel on copy streams
                             (default, measured fastest), 1 = copied after it on
                             the engine's stream, 2 = block 0 after it and later
                             blocks on a copy stream                              */
  int32_t schedule;       /* enum smp_schedule                                     */
  int32_t init_mode;      /* enum smp_init_mode                                    */
  int32_t tail_window_entries; /* host-tail extension-window storage per solve, in
                             list entries: 0 = auto (512 n, between 65,536 and
                             4,194,304); a smaller value caps it.  Chains that
                             find no storage left go back to the GPU            */
  int32_t start_order;    /* order in which the proposal kernel starts the men:
                             0 = the order of woman 0's list (default: when the
                             women's lists share a common order, the men they
                             prefer propose first and displace fewer men; the
                             reference's RANDOM lists, n = 20,000: 83 -> 64 ms),
                             1 = by index.  The matching is the same          */
  int32_t reserved[1];
} smp_opts;

enum smp_init_mode {
  SMP_INIT_AUTO = 0,   /* sparse prefix pass when prefix_cols <= n / 4, else full  */
  SMP_INIT_FULL = 1,   /* the n x n women's rank matrix, then the PR prefix gather
                          (the reference's init, kernel.cu:672-709)                */
  SMP_INIT_SPARSE = 2  /* PR prefix straight from the women's rows: the rank
                          matrix is built only if a chain walks past the prefix
                          or a woman is named in more prefixes than her bucket
                          holds; results are identical in every mode              */
};

enum smp_schedule {
  SMP_SCHEDULE_BAMBOO = 0, /* this build's pipeline (default)                       */
  SMP_SCHEDULE_MW_CAS = 1, /* the reference's GPU baseline MW-Par-GPU-CAS
                              (src/smp/smp_engine_mw_3_f9.cu:37-97, and the CAS
                              retry of src/smp/kernel.cu:5-55): one thread per man,
                              32-bit CAS on the partner rank, pref_m + rank_w      */
  SMP_SCHEDULE_LA_CAS = 2, /* the same with the PR matrix's (woman, rank) entries
                              and a read-and-skip before the CAS
                              (src/smp/kernel.cu:122-165)                          */
  SMP_SCHEDULE_MW_MIN = 3, /* pref_m + rank_w, read-and-skip, then 32-bit atomicMin
                              (src/smp/kernel.cu:468-512)                          */
  SMP_SCHEDULE_LA_MIN = 4, /* PR entries, atomicMin without the read
                              (src/smp/kernel.cu:515-552)                          */
  SMP_SCHEDULE_FRONTIER = 5 /* rounds over a compacted frontier of free men: each
                              round every free man skips the women whose partner
                              word already beats him and proposes to the next one;
                              proposals to one woman are reduced in LDS before one
                              global 64-bit atomicMin per workgroup; the free men
                              of the next round are compacted by wave ballots and
                              prefix sums.  For lists that share their top women  */
};
/* The MW/LA schedules skip the precheck and the master-list path, as the
 * reference's engines do; all schedules return the same matching. */
#define SMP_SCHEDULE_COUNT 6

typedef struct smp_times {
  /* Wall clock of the timed phases.  precheck/init/core are the three phases
   * AbsSmpEngine reports; post (result assembly and the copy of the matching to
   * the host) is untimed in the reference and is timed here.  total is all four. */
  double precheck_ms, init_ms, core_ms, post_ms, total_ms;
  /* Device time of the main kernels (HIP events on the engine's stream; the
   * engine records as few as it can, since each event between two kernels
   * delays the second by several microseconds):
   *   rank_kernel_ms     the rank pass: the full inversion, or the sparse prefix
   *                      pass (from its launch to the end of init);
   *   gather_kernel_ms   the PR prefix gather (full init, or after a bucket
   *                      overflow of the sparse pass);
   *   propose_kernel_ms  the proposal kernel and what follows it on the stream
   *                      until its state is in host memory (the state export,
   *                      or the counter fold when the GPU finishes alone). */
  double rank_kernel_ms, gather_kernel_ms, propose_kernel_ms;
  /* Host time of the hand-off (copies) and of the host tail loop. */
  double handoff_ms, host_tail_ms;
  double copy_wait_ms;     /* host wait for the PR-prefix copy after the kernel launch */
  double warm_ms;          /* host cache warm-up of that copy                          */
  double resume_ms;        /* proposal launch for chains the host handed back          */
  double bucket_kernel_ms; /* sparse prefix pass: the bucket kernels (HIP events,
                              from the precheck's flag copy to the rank pass)     */
  int64_t proposals;       /* sum over men of (final list position + 1)           */
  int64_t atomics;         /* atomicMin attempts made by the proposal kernel(s)   */
  int64_t host_proposals;  /* proposals made by the host tail                     */
  int64_t steps;           /* list-segment scans (16 to 32 entries) made by the
                              proposal kernel                                      */
  int32_t is_perfect;      /* 1 if the first-choice fast path answered            */
  int32_t handoff_free;    /* free men handed to the host (0 = GPU finished)      */
  int32_t prefix_cols;     /* PR columns actually gathered                        */
  int32_t tail_threshold;  /* threshold used                                      */
  int32_t bailed;          /* chains the host handed back to the GPU (past the PR
                              prefix with no extension window left)               */
  int32_t master_list;     /* 1 if every man's list was the same permutation and
                              the women-proposing walk answered (the stable
                              matching is then unique)                            */
  int32_t schedule;        /* enum smp_schedule of this run                       */
  int32_t sparse_init;     /* 1 if init took the sparse prefix pass               */
  int32_t rank_matrix;     /* 1 if the n x n rank matrix was built (full init, a
                              bucket overflow, or a chain past the prefix)        */
  int32_t copy_requeues;   /* copy-stream pairs the engine has replaced so far
                              because the PR-prefix copies had started only after
                              the proposal kernel (their stream shared its
                              hardware queue); at most 8                          */
  int32_t copy_serialized; /* 1 if this run's prefix copies waited for the
                              proposal kernel and no replacement was left          */
  int32_t tail_windows;    /* extension windows (the next 64 to 4,096 entries of
                              one man's list) the GPU computed for host-tail
                              chains that walked past the PR prefix, and for men
                              matched near its end                                */
  int32_t tail_full_matrix; /* 1 if the host tail's chains went so deep that the
                              whole PR matrix was gathered and copied to the host
                              (4 n^2 bytes, buffers allocated on first use)       */
  int32_t tail_rounds;     /* GPU round trips for extension windows               */
  int32_t tail_bands;      /* bands of columns of every man's list (the whole
                              matrix included) the host tail took because chains
                              kept leaving what it held                           */
  int32_t deep_restart;    /* 1 if the proposal kernel's chains walked so far past
                              the prefix that every chain was parked, the whole
                              PR matrix gathered and the chains resumed over it   */
  int32_t master_both;     /* 1 if, besides every man holding the same list, every
                              woman held the same list: the matching then is the
                              serial dictatorship match[P[k]] = L[k] (the
                              reference's CONGESTED lists)                        */
  int32_t reserved[1];
} smp_times;

/* Init-only run (AbsSmpInitEngine::Init, include/smp/smp_init_abs.h:16-28;
 * SmpInitEngine3 include/smp/smp_init_3.h:68-129): build the women's rank matrix
 * and, if init_pr != 0, the men-major PR matrix over the engine's prefix_cols
 * (set prefix_cols >= n for the full matrix), then copy the result to the host as
 * the reference does.  Times are wall-clock ms plus HIP-event kernel ms. */
typedef struct smp_init_times {
  double rank_ms, pr_ms, transfer_ms, total_ms;
  double rank_kernel_ms, pr_kernel_ms;
  int32_t pr_cols;
  int32_t reserved[7];
} smp_init_times;

void smp_default_opts(smp_opts* opts);

/* Copies both host lists into HBM (untimed, like the reference constructor). */
int smp_create(const int32_t* pref_m, const int32_t* pref_w, int32_t n, const smp_opts* opts,
               smp_engine** out);
/* Borrows lists already in HBM; they must outlive the engine and stay unchanged. */
int smp_create_device(const int32_t* d_pref_m, const int32_t* d_pref_w, int32_t n,
                      const smp_opts* opts, smp_engine** out);
/* The men's lists and the women's inverse rank matrix derived by the caller
 * (rank_w[w*n + m] = position of man m in woman w's list, int32, row-major n x n)
 * instead of the women's lists.  Init is then the PR prefix gather alone.  Each
 * rank is range-checked where it is read; the rank rows are not checked for
 * being permutations (validate covers the women's lists only).  The women's
 * lists are absent, so the master-list path is not taken (the general walk
 * answers instead) and the MW-CAS schedule is rejected. */
int smp_create_ranked(const int32_t* pref_m, const int32_t* rank_w, int32_t n, const smp_opts* opts,
                      smp_engine** out);
/* The same for arrays already in HBM (borrowed, as smp_create_device). */
int smp_create_device_ranked(const int32_t* d_pref_m, const int32_t* d_rank_w, int32_t n,
                             const smp_opts* opts, smp_engine** out);
/* One full solve: precheck + init + core + post.  Synchronous: every device and
 * host-tail step has finished when it returns.  May be called repeatedly. */
int smp_exec(smp_engine* e);
/* Copy the last result (int32[n], man -> woman) to host memory. */
int smp_result(smp_engine* e, int32_t* match_m2w);
/* Device address of the last result (int32[n] in HBM, owned by the engine). */
int smp_result_device(smp_engine* e, const int32_t** d_match);
int smp_profile(const smp_engine* e, smp_times* t);
int smp_init_exec(smp_engine* e, int32_t init_pr, smp_init_times* t);
/* Measure the configuration on this instance.  First the default solve: when
 * it takes more than 5 ms (lists that share their top women) and the caller
 * left schedule and waves_per_cu at their defaults, the schedules (this build's
 * pipeline at 16, 8 and 4 waves per CU, the frontier, and the reference's four
 * thread-per-man walks) are each solved `trials` times and the fastest is kept
 * in the engine's options; nothing more is swept then.  Otherwise the hand-off
 * threshold and PR prefix width: every candidate (threshold in {0 (GPU
 * finishes), 1, 2, 4, 8, 16, 24, 32, 48, 64, 96, 128, 256, 1024}, prefix in {64,
 * 96, 128, 256}) is solved `trials` times (0 = 5) and ranked by median wall time;
 * the four best and the untuned defaults are solved 3 * trials times more, and
 * the lowest of those medians is kept and used by later smp_exec calls in
 * SMP_TAIL_AUTO mode. */
typedef struct smp_tune_result {
  int32_t prefix_cols, tail_threshold;
  double best_ms;      /* median wall ms of the chosen configuration      */
  double gpu_only_ms;  /* best median with the GPU finishing every chain  */
  int32_t candidates;  /* threshold / prefix pairs measured               */
  int32_t schedule;    /* enum smp_schedule chosen                        */
  int32_t waves_per_cu; /* proposal-kernel occupancy chosen (0 = default) */
  int32_t reserved[5];
} smp_tune_result;
int smp_tune(smp_engine* e, int32_t trials, smp_tune_result* out);
/* Copies of the engine's device matrices after smp_exec or smp_init_exec (the
 * reference's RetrieveRankMatrixForTest / RetrievePRMatrixForTest hooks,
 * include/smp/smp_engine_hybrid.cuh:26-32).  rank_out: int32[n*n].  pr_out:
 * int32[n * cols_cap * 2], filled as n rows of *cols PRNode pairs (woman, rank),
 * include/utils/types.h:58-61, where *cols receives the PR columns the last run
 * gathered (its smp_times.prefix_cols; the automatic width can grow after a
 * sparse-pass overflow, and smp_tune can change it).  pr_out == NULL only sets
 * *cols; a cols_cap below it returns SMP_ERR_INVALID and writes nothing. */
int smp_rank_matrix(smp_engine* e, int32_t* rank_out);
int smp_pr_matrix(smp_engine* e, int32_t* pr_out, int32_t cols_cap, int32_t* cols);
void smp_destroy(smp_engine* e);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  One solve of lists already in HBM (n <= 4096) into the caller's
 * d_match (int32[n] in HBM), with *proposals (may be null) set to the proposal
 * count.  The same result as smp_create_device + smp_exec + smp_result_device,
 * by the GPU-only full-init pipeline (init_mode FULL, tail_mode GPU, precheck 0),
 * with every intermediate buffer in a fixed 34.2 MiB device array of the library
 * that each GPU it runs on holds until the process ends.  It calls no hipMalloc
 * or hipFree.  n above 4096 returns SMP_ERR_INVALID.  Synchronous (default
 * stream).  Calls that overlap on one GPU share that array and race: serialize
 * them per device. */
int smp_exec_device_noalloc(const int32_t* d_pref_m, const int32_t* d_pref_w, int32_t n, int32_t device,
                            int32_t* d_match, int64_t* proposals);

/* ---- batched: many independent instances in one persistent launch -------- */
typedef struct smp_batch smp_batch;

typedef struct smp_batch_opts {
  int32_t device;
  int32_t use_graph;      /* 1 = capture the launch sequence in a hipGraph (default) */
  int32_t block_threads;  /* workgroup size of the batched kernel; 0 = autotuned     */
  int32_t wgs_per_cu;     /* persistent workgroups per CU; 0 = autotuned            */
  int32_t mode;           /* SMP_BATCH_AUTO (autotuned), _FUSED or _POOL            */
  int32_t group_lanes;    /* pool mode: lanes per proposal chain (8, 16 or 64);
                             0 = autotuned                                           */
  int32_t pool_slots;     /* pool mode: instances open per workgroup (1, 2 or 4);
                             0 = autotuned                                           */
  int32_t rank_layout;    /* pool mode: SMP_RANK_WOMEN_ROWS or SMP_RANK_MEN_ROWS;
                             0 = autotuned                                           */
  int32_t reserved[8];
} smp_batch_opts;

enum smp_batch_mode {
  SMP_BATCH_AUTO = 0,
  SMP_BATCH_FUSED = 1, /* one kernel: each workgroup inverts its instance's rows, then
                          runs its chains                                          */
  SMP_BATCH_POOL = 2   /* two kernels in one graph: the rank matrices of all
                          instances as one HBM-bound pass, then workgroups that each
                          keep pool_slots instances open in LDS and run several
                          chains per wave                                          */
};

enum smp_rank_layout {
  SMP_RANK_WOMEN_ROWS = 1, /* rank[w][m]: row w is woman w's inverted list          */
  SMP_RANK_MEN_ROWS = 2    /* rank[m][w]: a chain's lookups stay in its man's row   */
};

typedef struct smp_batch_times {
  double exec_ms;         /* wall time of the last smp_batch_exec / _looped         */
  double kernel_ms;       /* device time of the persistent kernel(s)                */
  double solves_per_s;
  int64_t proposals;
  int32_t count;
  int32_t block_threads, wgs_per_cu;  /* of the chain (or fused) kernel           */
  int32_t mode;                       /* SMP_BATCH_FUSED or SMP_BATCH_POOL          */
  int32_t group_lanes;                /* pool mode: lanes per proposal chain        */
  int32_t pool_slots;                 /* pool mode: instances open per workgroup    */
  int32_t rank_layout;                /* pool mode: SMP_RANK_*                      */
  double rank_ms;                     /* pool mode: the launch up to the end of the
                                         rank pass (the master-list detection runs
                                         beside it on a second stream), measured
                                         while tuning (min of two runs)             */
  int32_t reserved[4];
} smp_batch_times;

void smp_batch_default_opts(smp_batch_opts* opts);
/* sizes[i] = n_i; instance i's lists start at element offset offsets[i] of
 * pref_m_all / pref_w_all (each list n_i*n_i int32).  Host buffers are copied. */
int smp_batch_create(int32_t count, const int32_t* sizes, const int64_t* offsets,
                     const int32_t* pref_m_all, const int32_t* pref_w_all,
                     const smp_batch_opts* opts, smp_batch** out);
/* Solve every instance in one persistent launch (work-stealing over instances). */
int smp_batch_exec(smp_batch* b);
/* Baseline: solve the same instances one by one with plain single-instance launches. */
int smp_batch_exec_looped(smp_batch* b);
/* Results concatenated in instance order: match_all[moff[i] + m], moff = prefix sums of sizes. */
int smp_batch_result(smp_batch* b, int32_t* match_all);
int smp_batch_profile(const smp_batch* b, smp_batch_times* t);
void smp_batch_destroy(smp_batch* b);

/* Deterministic batch plan: sizes log-uniform in [n_min, n_max], kinds drawn with
 * weights mix[0..2] over {UNIFORM, CONGESTED, CORRELATED}. */
int smp_batch_plan(int32_t count, int32_t n_min, int32_t n_max, uint64_t seed,
                   const double* mix3, int32_t* sizes, int32_t* kinds);

/* ---- serving: a request loop with continuous batching -------------------- */
typedef struct smp_server smp_server;

enum smp_server_mode {
  SMP_SERVE_BATCHED = 0, /* admitted requests are launched together, several launches
                            in flight at once; new requests join the next free lane */
  SMP_SERVE_PLAIN = 1    /* baseline: one request at a time through the plain
                            single-instance pipeline                            */
};

typedef struct smp_server_opts {
  int32_t device;
  int32_t mode;             /* enum smp_server_mode                                   */
  int32_t lanes;            /* launches in flight at once (streams); 0 = 3           */
  int32_t max_batch;        /* requests per launch; 0 = 64                           */
  int32_t max_n;            /* largest request accepted; 0 = 8192 (int16 ranks)      */
  int32_t block_threads;    /* workgroup size of the serving kernel; 0 = auto        */
  int64_t hbm_budget_bytes; /* device arena for request state (the lanes' int16 rank
                               matrices and the list copies of host-submitted
                               requests); 0 = 1 GiB, room for two host-submitted
                               n = 8192 requests.  Allocated when the server is
                               created and freed when it is destroyed            */
  int32_t large_n;          /* batched mode: requests with n >= large_n are solved by
                               the multi-workgroup single-instance pipeline on
                               workers of their own, beside the lanes;
                               0 = 512; above max_n = never                        */
  int32_t large_workers;    /* threads (each with its own pipeline) serving the large
                               requests; 0 = 2                                     */
  int32_t large_spill_us;   /* > 0: worker i > 0 takes the oldest waiting large
                               request only when more than i are waiting or it has
                               waited this long (so that at low load a request runs
                               alone on the GPU); 0 = any idle worker at once
                               (default: measured best for p50 and p99)            */
  int32_t small_to_idle_worker; /* batched mode: 1 = a request below large_n goes to
                               a large-request worker when one is idle and none
                               waits for the workers, else to the lanes;
                               0 = always to the lanes; -1 = the default          */
  int32_t fused_lanes;      /* 1 = lane launches whose requests all have n <= 2048
                               run one kernel, a workgroup per request (a shared-
                               list check, rank rows and chains in it); 0 = a flat
                               rank pass and a chain kernel; -1 = the default (1) */
  int32_t reserved[3];
} smp_server_opts;

typedef struct smp_request_info {
  double queue_ms;    /* submit -> launch                                        */
  double service_ms;  /* launch -> completion seen by the server                 */
  double latency_ms;  /* submit -> completion                                    */
  int64_t proposals;
  int32_t n;
  int32_t batch_size; /* requests in the launch that carried this one            */
  int32_t lane;       /* launch lane, or -1 for the large-request worker          */
  int32_t reserved[5];
} smp_request_info;

typedef struct smp_server_stats {
  int64_t submitted, completed, launches;
  int64_t arena_bytes, peak_arena_bytes;
  int32_t peak_inflight;  /* admitted and not completed, maximum seen            */
  int32_t peak_lanes_busy;
  int32_t large_solved;   /* requests solved by the large-request worker           */
  int32_t reserved[7];
} smp_server_stats;

void smp_server_default_opts(smp_server_opts* opts);
int smp_server_create(const smp_server_opts* opts, smp_server** out);
/* Submit one instance.  on_device = 0: host lists, copied into the arena at
 * admission (the caller may reuse them once this returns).  on_device = 1: lists
 * already in HBM, borrowed until the request completes.  Never blocks on the GPU;
 * admission happens when the arena and a lane have room.  The result must be
 * collected with smp_server_poll / smp_server_wait. */
int smp_server_submit(smp_server* s, const int32_t* pref_m, const int32_t* pref_w, int32_t n,
                      int32_t on_device, uint64_t* ticket);
/* SMP_OK and the matching if the request finished (its ticket is then retired),
 * SMP_PENDING otherwise. */
int smp_server_poll(smp_server* s, uint64_t ticket, int32_t* match, smp_request_info* info);
/* Wait up to timeout_ms (< 0: no limit) for a request. SMP_ERR_TIMEOUT on expiry. */
int smp_server_wait(smp_server* s, uint64_t ticket, int32_t* match, smp_request_info* info,
                    int32_t timeout_ms);
int smp_server_stats_get(smp_server* s, smp_server_stats* st);
/* Finishes in-flight work, then frees everything. */
void smp_server_destroy(smp_server* s);

/* ---- device memory helpers (so hosts need no other GPU runtime) ----------- */
int smp_device_count(int32_t* count);
int smp_device_alloc(size_t bytes, void** d_ptr);
int smp_device_free(void* d_ptr);
int smp_copy_h2d(void* d_dst, const void* h_src, size_t bytes);
int smp_copy_d2h(void* h_dst, const void* d_src, size_t bytes);
/* Device-to-device copy by a plain streaming kernel (16-byte loads and stores;
 * bytes must be a multiple of 16).  Synchronous. */
int smp_device_copy(void* d_dst, const void* d_src, size_t bytes);
int smp_device_sync(void);

const char* smp_last_error(void);
const char* smp_version(void);

#ifdef __cplusplus
}
#endif

#endif /* PACTAE_SMP_HIP_H */
