/*
 * bpmax.h -- C ABI of the MI355X BPMax engine (libbpmax_hip.so).
 *
 * BPMax (base-pair maximization over two RNA strands) fills the 4-D "triangle of
 * triangles" F[i1,j1,i2,j2], 0<=i1<=j1<M over seq1 and 0<=i2<=j2<N over REVERSED seq2,
 * and returns the optimal score F(0,M-1,0,N-1).  Every function here is plain C: no
 * exceptions cross it, pointers are host pointers unless a name says "device", and the
 * return value is 0 or a negative BPMAX_E* code.
 *
 * Reference interfaces replaced (chiranjeb/BPMaxCPU, paths relative to
 * TPDS/source/release/MPT_RT_V3/bpmax_complete_otf_ftable/):
 *   bpmax()                  <- src/bpmax.c:129 (prototype src/bpmax-wrapper.c:51), same
 *                               signature, same 6-D F layout, same parameter check;
 *   bpmax_engine_solve()     <- the same call as a library entry point with the weight
 *                               table passed in (the reference resolves it at link time
 *                               through e_intra_score/e_inter_score, src/external_functions.h:33-35)
 *                               and F returned packed or dense 4-D instead of the 6-D tree;
 *   bpmax_engine_score_table <- src/external_functions.c:21-68 (default weight table);
 *   bpmax_write_checking_dump <- the CHECKING dump of src/bpmax-wrapper.c:318-342;
 *   bpmax_read_sequence_file  <- ReadSequencesFromFiles, src/external_functions.c:447-468
 *                               (returns an error on a length mismatch instead of silently
 *                               substituting random bases).
 */
#ifndef BPMAX_H
#define BPMAX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BPMAX_ABI_VERSION 1

/* ---- error codes ---------------------------------------------------------------- */
#define BPMAX_OK 0
#define BPMAX_EINVAL -1      /* bad argument (lengths, codes, layout, null pointer)   */
#define BPMAX_ENOMEM -2      /* device or host allocation failed                      */
#define BPMAX_EDEVICE -3     /* HIP runtime error                                     */
#define BPMAX_EWEIGHTS -4    /* weight table not exact (see bpmax_engine_solve)       */
#define BPMAX_EBUSY -5       /* serving queue full / admission refused                */
#define BPMAX_ENOTREADY -6   /* request not finished yet                              */
#define BPMAX_ENOTFOUND -7   /* unknown request id                                    */
#define BPMAX_ETOOBIG -8     /* pair can never fit the device memory budget           */

/* ---- F output layouts -------------------------------------------------------------- */
typedef enum {
    BPMAX_LAYOUT_NONE = 0,       /* score only; F stays on the device                  */
    BPMAX_LAYOUT_PACKED_TRI_TRI, /* tri(M)*tri(N) floats, order i1, j1>=i1, i2, j2>=i2
                                    (the order of the reference's CHECKING dump)        */
    BPMAX_LAYOUT_DENSE_4D        /* [M][M][N][N] floats (IPDPSW-style 4-D table,
                                    IPDPSW/source/.../src/bpmax-wrapper.c:175); only the
                                    cells with i1<=j1 and i2<=j2 are written            */
} bpmax_layout;

typedef struct bpmax_ctx bpmax_ctx; /* one HIP stream + device arena on one GPU */

/* Sequence codes: 0..4 = A, C, G, U/T, N (src/external_functions.c:74-85). */
unsigned char bpmax_base_code(char c);

/* Weight tables are 5x5 row-major float arrays, w[5*x+y] = score of pairing code x with
 * code y.  Fills the reference's hard-coded table (GC=3, AU=1, GU=1.5, everything else 0,
 * identical for intra- and inter-strand pairs). */
void bpmax_engine_score_table(float intra[25], float inter[25]);

/* Context on HIP device `device`.  Not thread-safe: one host thread per context. */
int bpmax_ctx_create(bpmax_ctx **out, int device);
int bpmax_ctx_destroy(bpmax_ctx *ctx);
const char *bpmax_last_error(const bpmax_ctx *ctx); /* human-readable, never NULL */

/* Solve one pair.
 *   seq1[M], seq2[N]: codes 0..4, seq2 in its natural 5'->3' order (not reversed).
 *   intra, inter: 5x5 weight tables; NULL selects the reference table.  Weights must be
 *     finite, >= 0 and all multiples of 2^-p for some p <= 12; with p the smallest such
 *     exponent and wmax the largest weight, wmax * ((M+N)/2 + 1) * 2^p must stay below
 *     2^24 (a structure has at most (M+N)/2 pairs, so every partial score is an integer
 *     multiple of 2^-p below 2^24 * 2^-p, exact in FP32).  Max-plus evaluation is then
 *     exact in any order; BPMAX_EWEIGHTS otherwise.
 *   F_out: host buffer for `layout`, or NULL with BPMAX_LAYOUT_NONE.
 *   score_out: receives F(0,M-1,0,N-1).
 * M >= 1, N >= 1. */
int bpmax_engine_solve(bpmax_ctx *ctx, const int *seq1, int M, const int *seq2, int N,
                       const float *intra, const float *inter, float *F_out,
                       bpmax_layout layout, float *score_out);

/* Bytes that `layout` needs for an MxN pair on the host. */
size_t bpmax_layout_bytes(int M, int N, bpmax_layout layout);

/* Device bytes the engine holds for one resident MxN pair (F in its tiled layout plus
 * the single-strand tables). */
size_t bpmax_engine_pair_device_bytes(int M, int N);

/* ---- batched solve ----------------------------------------------------------------- */
/* Solve `npairs` independent pairs of any sizes in one scheduled pass on the device.
 * seq1s[p] / seq2s[p] point to codes of length Ms[p] / Ns[p].  scores[p] receives each
 * optimal score.  F tables stay on the device (score-only).  Weights as in
 * bpmax_engine_solve, shared by all pairs. */
int bpmax_engine_solve_batch(bpmax_ctx *ctx, int npairs, const int *const *seq1s,
                             const int *Ms, const int *const *seq2s, const int *Ns,
                             const float *intra, const float *inter, float *scores);
/* Copy the F table of pair `pair` (0-based, in call order) of the context's last
 * bpmax_engine_solve or bpmax_engine_solve_batch, still resident on the device, to F_out in
 * `layout` (PACKED_TRI_TRI or DENSE_4D).  BPMAX_EINVAL when the last solve failed, another
 * call has since reused the arena, or `pair` is out of range.  No counterpart in the
 * reference (its bpmax() writes F into the caller's table directly). */
int bpmax_engine_fetch(bpmax_ctx *ctx, int pair, bpmax_layout layout, float *F_out);

/* ---- solve without device allocation ----------------------------------------------- */
/* Size of the per-GPU device scratch array that bpmax_engine_solve_noalloc works in.  The
 * array is part of the library's code object: each GPU the library runs on holds it until
 * the process ends. */
#define BPMAX_NOALLOC_SCRATCH_BYTES (64u << 20)
/* As bpmax_engine_solve with layout BPMAX_LAYOUT_NONE or BPMAX_LAYOUT_PACKED_TRI_TRI, but it
 * allocates no device memory (no hipMalloc/hipFree, no allocating library routine): the F
 * table, codes, schedule, score and the packed copy of F all live in the scratch array
 * above.  A pair whose tables, schedule and packed output do not fit its
 * BPMAX_NOALLOC_SCRATCH_BYTES returns BPMAX_ENOMEM with a message (64 MiB holds, with the
 * packed output, e.g. 24 x 160, 32 x 128 and 16 x 256).  Every call on one GPU shares that
 * GPU's array and keeps nothing between calls (F is re-initialised by each call), so calls
 * that overlap on one GPU -- from several threads or contexts -- race and must not be issued.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). */
int bpmax_engine_solve_noalloc(bpmax_ctx *ctx, const int *seq1, int M, const int *seq2, int N,
                               const float *intra, const float *inter, float *F_out,
                               bpmax_layout layout, float *score_out);

/* ---- serving loop ---------------------------------------------------------------------- */
/* Continuous batching over one device arena of `memory_budget` bytes.  The server thread
 * advances every resident pair by one outer diagonal per "tick" -- one step of the outer
 * wavefront for all resident pairs at once, each on its own diagonal -- and enqueues the
 * ticks on two streams of its own, `slots` ticks ahead of the device.  Submitted pairs queue
 * FIFO and join at the next tick that has room: room in the arena for the pair's extent (its
 * tiled F layout, S1, S2 and codes, ~M(M+1)N(N+1)/4 x 4 B, bpmax_engine_pair_device_bytes;
 * half the bytes of F more for a pair with FP16 R0 operands, see bpmax_ctx_set_r0_half),
 * fewer than `max_batch` resident pairs, and room in the tick's work-item staging.  Requests
 * that arrive while others run therefore start while those run, and pairs leave after their
 * last diagonal.  Every device buffer is allocated by bpmax_server_start; the serving loop
 * never allocates or frees device memory.  Results are bit-identical to bpmax_engine_solve
 * of the same pair.  No counterpart in the reference (its bpmax() solves one pair per call). */
typedef struct {
    size_t memory_budget; /* arena bytes; 0 = 80% of the free device memory at start */
    int max_batch;        /* most pairs resident at once (0 = 256)                    */
    int slots;            /* ticks enqueued ahead of the device, 1..8 (0 = 3)         */
    int queue_capacity;   /* queued requests before submit returns BPMAX_EBUSY (0 = 4096) */
    int keep_tables;      /* 1: F tables stay resident until bpmax_server_release and can be
                             read with bpmax_server_fetch; 0: freed when the score is ready */
    const float *intra;   /* weight tables of every request (NULL: reference table)    */
    const float *inter;
} bpmax_server_config;

typedef struct {
    long long submitted, completed, failed;
    long long batches;            /* ticks enqueued                                 */
    long long max_batch_seen;     /* most pairs in one tick                         */
    long long max_inflight_pairs; /* most pairs resident at one time                */
    size_t budget, bytes_in_use, peak_bytes_in_use;
} bpmax_server_stats;

int bpmax_server_start(bpmax_ctx *ctx, const bpmax_server_config *cfg);
/* Queue one pair; *id receives its request id.  BPMAX_ETOOBIG if the pair can never fit
 * the budget, BPMAX_EBUSY if the queue is full, BPMAX_EWEIGHTS as bpmax_engine_solve. */
int bpmax_server_submit(bpmax_ctx *ctx, const int *seq1, int M, const int *seq2, int N, long long *id);
/* BPMAX_OK with the score and submit-to-completion latency, BPMAX_ENOTREADY while
 * queued or running, BPMAX_ENOTFOUND for an unknown or released id. */
int bpmax_server_poll(bpmax_ctx *ctx, long long id, float *score, double *latency_ms);
/* As poll, blocking up to timeout_ms (< 0: no limit). */
int bpmax_server_wait(bpmax_ctx *ctx, long long id, double timeout_ms, float *score, double *latency_ms);
/* Copy a finished request's F table (keep_tables = 1) in `layout` to F_out. */
int bpmax_server_fetch(bpmax_ctx *ctx, long long id, bpmax_layout layout, float *F_out);
/* Forget a finished request (and free its table when it was the last of its batch). */
int bpmax_server_release(bpmax_ctx *ctx, long long id);
int bpmax_server_get_stats(bpmax_ctx *ctx, bpmax_server_stats *stats);
/* Finish every queued request, then stop the server thread and free the arena.  Queued
 * requests that can never be admitted because unreleased kept tables hold the arena fail
 * with BPMAX_ENOMEM (reported by poll/wait) instead of blocking the stop. */
int bpmax_server_stop(bpmax_ctx *ctx);
/* ---- drop-in compatibility symbol ----------------------------------------------- */
/* Exact signature of the reference entry point (src/bpmax.c:129).  Computes on the GPU
 * and writes the caller's 6-level F pointer tree
 *   FTable[i1][j1][(i2+R)/N_tile][(j2+R)/N_tile][(i2+R)%N_tile][(j2+R)%N_tile]
 * for every i1<=j1, i2<=j2 (src/bpmax-wrapper.c:44).  The weight table is read through
 * the linked e_intra_score()/e_inter_score() plug point when those symbols are present in
 * the process, else the reference table is used.  Invalid parameters print the
 * reference's message and exit(-1), as the reference does (src/bpmax.c:131-133); MR, NR
 * and P are validated but the GPU schedule does not use them. */
void bpmax(long M, long N, long N_sec, long N_tile, long R, long MR, long NR, long P,
           int *seq1, int *seq2, float ******FTable);

/* ---- file formats -------------------------------------------------------------------- */
/* Reads the first FASTA-like record of `path` (header lines start with '>', letters are
 * mapped by bpmax_base_code).  Writes at most `cap` codes to out and returns the record
 * length, or a negative error.  Unlike the reference reader, a length different from the
 * expected one is reported to the caller instead of being replaced by random bases. */
int bpmax_read_sequence_file(const char *path, int *out, int cap);

/* Writes "FTable(i1,j1,i2,j2)=%0.2f\n" lines for every defined cell of a packed table,
 * in the reference's CHECKING order (src/bpmax-wrapper.c:318-342).  with_prompt=0 writes
 * only the values (the NO_PROMPT form). */
int bpmax_write_checking_dump(const char *path, const float *F_packed, int M, int N,
                              int with_prompt);

/* ---- isolated R0 max-plus kernel ------------------------------------------------------ */
/* For `batch` independent upper-triangular n x n matrix triples held on the DEVICE in the
 * engine's tiled layout (see bpmax_tri_tiled_floats), computes
 *   C[i][j] = max(C[i][j], max_{i<=k<=j} A[i][k] + B[k][j])  for i<=j.
 * A, B, C are device pointers, each holding batch*bpmax_tri_tiled_floats(n) floats. */
size_t bpmax_tri_tiled_floats(int n);
int bpmax_r0_tri_batch_device(bpmax_ctx *ctx, int n, int batch, const float *A_dev,
                              const float *B_dev, float *C_dev);
/* FP16 copies of `batch` tiled triangles: dst_dev holds batch*bpmax_tri_tiled_floats(n)
 * 16-bit values.  *exact = 1 when every value is -inf or an integer of magnitude <= 1024;
 * every sum of two such values is then exact in FP16, and
 * bpmax_r0_tri_batch_device_half on the copies gives C bit for bit as
 * bpmax_r0_tri_batch_device on the originals.  *exact = 0 otherwise (the copies are then
 * rounded toward zero and must not be used for exact results). */
int bpmax_tri_to_half(bpmax_ctx *ctx, int n, int batch, const float *src_dev,
                      unsigned short *dst_dev, int *exact);
/* bpmax_r0_tri_batch_device on FP16 operands (IEEE half bit patterns in the same tiled
 * layout; C stays FP32 and is combined in FP32), the reference's R0 max-plus micro-kernel
 * operation (micro_kernels/maxplus/max_plus.cpp) on packed halves.  Exact when every sum of
 * an A and a B value is an FP16 number, as bpmax_tri_to_half's *exact = 1 guarantees. */
int bpmax_r0_tri_batch_device_half(bpmax_ctx *ctx, int n, int batch,
                                   const unsigned short *A_dev, const unsigned short *B_dev,
                                   float *C_dev);
/* Host <-> tiled-device conversion helpers for those triples (dense row-major n x n on
 * the host; entries below the diagonal are ignored on upload, -inf on download). */
int bpmax_tri_upload(bpmax_ctx *ctx, int n, int batch, const float *host_dense, float *dev_tiled);
int bpmax_tri_download(bpmax_ctx *ctx, int n, int batch, const float *dev_tiled, float *host_dense);
/* Fills `batch` tiled triangles on the device with synthetic operands: cell (m, r, c),
 * r <= c < n, gets lo + splitmix64(seed ^ (m<<42 | r<<21 | c)) % (hi-lo+1); every other
 * cell -inf.  0 <= lo <= hi < 2^22, n < 2^21. */
int bpmax_tri_fill_random(bpmax_ctx *ctx, int n, int batch, unsigned long long seed, int lo, int hi,
                          float *dev_tiled);
void *bpmax_device_alloc(bpmax_ctx *ctx, size_t bytes);
int bpmax_device_free(bpmax_ctx *ctx, void *p);

/* ---- timing / introspection --------------------------------------------------------- */
/* When enabled, every kernel launch is bracketed by HIP events on the context stream and
 * per-kernel-class totals are accumulated (adds host overhead; off by default). */
int bpmax_ctx_set_timing(bpmax_ctx *ctx, int enable);
/* Like bpmax_ctx_set_timing, for the kernel classes whose bit (1 << class) is set in
 * `mask` only (0 disables timing). */
int bpmax_ctx_set_timing_mask(bpmax_ctx *ctx, unsigned mask);
/* Kernel classes for bpmax_ctx_kernel_stats. */
typedef enum {
    BPMAX_K_FILL = 0,
    BPMAX_K_S1,
    BPMAX_K_OUTER,      /* R0 + init + P1 + R3 + R4 (one launch per outer diagonal) */
    BPMAX_K_INNER_EXT,  /* R1/R2 cross-tile products (and S2's)                       */
    BPMAX_K_INNER_INTRA,/* intra-tile wavefront (R1/R2/P2 inside a tile; S2's)       */
    BPMAX_K_CONVERT,
    BPMAX_K_R0TRI,      /* isolated R0 kernel                                          */
    BPMAX_K_TASKS,      /* dependency-tile scheduler: every phase of a solve in one launch */
    BPMAX_K_NCLASSES
} bpmax_kernel_class;
/* Returns launches, total milliseconds and total algorithmic cell updates for a class
 * since the last reset. */
int bpmax_ctx_kernel_stats(bpmax_ctx *ctx, int kclass, long long *launches, double *ms,
                           double *updates);
int bpmax_ctx_reset_stats(bpmax_ctx *ctx);
/* Last BPMAX_SCHED_TASKS run: tasks run, and the polls its workgroups spent waiting for
 * dependencies (per task the largest count of any lane, summed; a poll is one device-scope
 * load plus a short sleep). */
int bpmax_ctx_task_stats(bpmax_ctx *ctx, long long *tasks, double *wait_polls);
/* Schedules:
 *   BPMAX_SCHED_PIPE (default): one launch per phase and block diagonal on two streams --
 *     R0 work of outer diagonal d1 on one, the inner closure of d1 on the other, joined by
 *     events at the tile diagonals each launch reads -- so the closure of d1 overlaps the R0
 *     work of d1 and d1+1.  Without timing the whole launch DAG is captured once per
 *     problem shape into a HIP graph and replayed (bpmax_ctx_set_graphs).
 *   BPMAX_SCHED_LAUNCH: the same launches on one stream, in reference fork/join order
 *     (per-phase timing without overlap).
 *   BPMAX_SCHED_TASKS: a whole solve as one launch of the dependency-tile scheduler; every
 *     tile task waits on per-tile ready counters of the tiles it reads.
 * chunk = K-steps per task of the task scheduler (0 keeps the current value).  The
 * BPMAX_SCHEDULE=pipe|launch|tasks environment variable sets the default. */
#define BPMAX_SCHED_TASKS 0
#define BPMAX_SCHED_LAUNCH 1
#define BPMAX_SCHED_PIPE 2
int bpmax_ctx_set_schedule(bpmax_ctx *ctx, int schedule, int chunk);
/* ---- autotuning ------------------------------------------------------------------------ */
/* Launch-schedule settings of one problem class.  A class is (M rounded up to a power of
 * two, N rounded up to a power of two, single pair or batch); a solve whose class has
 * settings uses them instead of the context's schedule.  No counterpart in the reference
 * (its tile and register-tile sizes are command-line parameters). */
typedef struct {
    int schedule;    /* BPMAX_SCHED_PIPE or BPMAX_SCHED_LAUNCH                          */
    int outer_split; /* R0 launches per outer diagonal (pipelined; 0: one per storage
                        block diagonal)                                                 */
    int outer_items; /* work items an R0 launch aims for, >= 64 (split-K granularity)   */
    int graphs;      /* HIP-graph replay of the pipelined schedule                      */
    int intra_wide;  /* in-tile closure launches of at most this many tiles use 1024-lane
                        workgroups (one cell per lane: latency), larger ones 256-lane
                        workgroups (one 2x2 block of cells per lane, one step per block
                        diagonal; 512-lane, two cells per lane, with
                        BPMAX_INTRA_BLOCKED=0).  With closure_ahead set, F's in-tile
                        launches are the merged 256-lane ones and this threshold only
                        selects the form of S2's                                          */
    int outer_kgroup; /* K values (storage tiles of the inner sum) per R0 work item; 0:
                        all of an output tile's.  1 keeps the items of one triangle in
                        step, so each source tile is read from HBM once per k1 and then
                        served from L2 to all its consumers (default 1)                  */
    int closure_ahead; /* 1 (default): the inner closure of F runs as one launch per
                        logical block diagonal DL, its in-tile workgroups (blocked form)
                        also running their tiles' two cross-tile steps that read DL-1, and
                        the other cross-tile steps of DL+1 riding in the same launch (half
                        as many dependent launches); 0: a cross-tile and an in-tile launch
                        per diagonal                                                     */
    int tile;        /* storage tile edge: 0 (default) = 32 for N <= 64, else 64; 32 = 32 for
                        every pair (less padding of N to whole tiles, a 32 x 32 R0 product)   */
} bpmax_tuning;
/* Time every candidate setting on `npairs` synthetic M x N pairs (median of `reps` runs
 * after a warm-up each), keep the fastest for the class and return it. */
int bpmax_tune_shape(bpmax_ctx *ctx, int M, int N, int npairs, int reps, bpmax_tuning *best,
                     double *best_ms);
/* The same on the caller's own pairs (class: their largest M and N, single or batch). */
int bpmax_tune_batch(bpmax_ctx *ctx, int npairs, const int *const *seq1s, const int *Ms,
                     const int *const *seq2s, const int *Ns, int reps, bpmax_tuning *best,
                     double *best_ms);
/* Set (t = NULL: clear) or read the settings of the class of (M, N, npairs);
 * get returns BPMAX_ENOTFOUND for an untuned class. */
int bpmax_ctx_set_tuning(bpmax_ctx *ctx, int M, int N, int npairs, const bpmax_tuning *t);
int bpmax_ctx_get_tuning(bpmax_ctx *ctx, int M, int N, int npairs, bpmax_tuning *t);
/* The settings for classes without an entry (schedule as bpmax_ctx_set_schedule). */
int bpmax_ctx_set_default_tuning(bpmax_ctx *ctx, const bpmax_tuning *t);
/* Enable or disable (default) HIP-graph capture and replay for BPMAX_SCHED_PIPE; without
 * timing only.  BPMAX_GRAPHS=1 in the environment sets the default. */
int bpmax_ctx_set_graphs(bpmax_ctx *ctx, int enable);
/* Enable (default) or disable FP16 operands for R0.  When every pair of a solve (or batch)
 * has 64 x 64 storage tiles and a score bound that FP16 represents exactly -- the weights
 * are multiples of 2^-p and bpmax_score_bound (which includes wmax * ((M+N)/2 + 1)) times
 * 2^p is at most 2^11 -- the inner
 * closure also writes each finished F tile as FP16 into a mirror (half the bytes of F,
 * extra device memory) and the R0 launches multiply those on packed halves (v_pk_add_f16 +
 * v_pk_maximum3_f16: 4 cell updates per 3 VALU instructions).  Every value is then exact,
 * so F and the score are bit-identical either way; other solves and the no-allocation
 * solve use FP32 operands.  The serving loop chooses per pair with the same bound (the
 * setting of the context at bpmax_server_start).  BPMAX_R0_HALF=0 in the environment sets
 * the default.  No counterpart in the reference (its R0 is FP32 AVX2). */
int bpmax_ctx_set_r0_half(bpmax_ctx *ctx, int enable);
/* An upper bound on every score (every F cell and every partial sum) of the pair seq1 x seq2
 * (codes 0..4, seq2 in its natural order) under the given tables (NULL: as
 * bpmax_engine_solve), the least of: wmax * ((M+N)/2 + 1); and, since a score is the weight
 * of a matching of the pair's bases with W(x,z) the larger weight of codes x, z in either
 * table and order, the least sum over bases of y[code] over y >= 0 with y[x] + y[z] >= W(x,z)
 * (the fractional matching optimum, by the vertices of that polyhedron).  R0 runs on FP16
 * operands for a pair when this bound times 2^p is at most 2^11 (weights multiples of
 * 2^-p).  Host only, no device call.  -1 for null or out-of-range input or a negative or
 * non-finite weight.  No counterpart in the reference. */
double bpmax_score_bound(const float *intra, const float *inter, const int *seq1, int M, const int *seq2, int N);
/* 1 if the last solve of the context ran its R0 launches on FP16 operands, else 0. */
int bpmax_ctx_r0_half_used(const bpmax_ctx *ctx);
/* The hipStream_t the context launches on (as void*). */
void *bpmax_ctx_stream(bpmax_ctx *ctx);
int bpmax_ctx_synchronize(bpmax_ctx *ctx);
/* Host-only check of the launch schedule that pairs of these shapes get under settings t:
 * builds the launch list and verifies, under the runner's ordering rules (stream order plus
 * the pipelined schedule's event waits), that every work item starts after all the work it
 * reads.  skip >= 0 drops that launch's wait first (to test the check itself).
 * *checked: dependencies verified.  BPMAX_EINVAL with a message in msg on a violation.
 * No counterpart in the reference (its OpenMP schedule is implicit in its loops). */
int bpmax_schedule_check(int npairs, const int *Ms, const int *Ns, const bpmax_tuning *t, int skip,
                         long long *checked, char *msg, int msglen);
/* Host-only check of the serving loop's plan (no GPU needed): replays `nreq` requests of
 * shapes Ms[k] x Ns[k], request k arriving once the server has planned arrive[k] ticks,
 * through the server's own admission and tick planner with the given budget, resident-pair
 * limit, ticks in flight and launch settings t, then runs the resulting stream operations
 * through a model of two FIFO streams with events and verifies that every work item starts
 * after all the work it reads, every event wait names an event recorded before it (and not
 * re-recorded since), resident extents never overlap and every operation on a pair finishes
 * before its extent is released.  skip >= 0 drops that event wait first (to test the check).
 * *checked: dependencies verified.  BPMAX_EINVAL with a message on a violation.  No
 * counterpart in the reference. */
int bpmax_server_check(int nreq, const int *Ms, const int *Ns, const int *arrive, size_t memory_budget,
                       int max_active, int depth, const bpmax_tuning *t, int skip, long long *checked,
                       char *msg, int msglen);
/* Cell updates of the full recurrence for an MxN pair:
 *   C(M+1,3)C(N+1,3) + 2C(M+1,2)C(N+1,3) + 2C(M+1,3)C(N+1,2)  (R0 + R1/R2 + R3/R4). */
double bpmax_cell_updates(int M, int N);

#ifdef __cplusplus
}
#endif
#endif /* BPMAX_H */
