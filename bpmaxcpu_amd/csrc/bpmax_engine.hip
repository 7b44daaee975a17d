// bpmax_engine.hip -- host side of the BPMax engine: context, device arena, schedule
// construction and the C ABI of include/bpmax.h.
//
// The schedule replaces the reference orchestrator bpmax() (src/bpmax.c:129-415, paths
// relative to TPDS/source/release/MPT_RT_V3/bpmax_complete_otf_ftable/):
//   S1 (one workgroup per pair); S2 by inner block diagonals; then for every outer
//   diagonal d1 = j1-i1: one k_outer launch over all triangles of the diagonal (R0, init,
//   P1, R3, R4 with split-K work items) and, per inner block diagonal D, one k_inner_ext
//   launch (cross-tile R1/R2) and one k_inner_intra launch (in-tile wavefront).
// Launches cover every pair of a batch at once, so batching is a property of the
// schedule, not of the kernels.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "bpmax.h"
#include "bpmax_engine_impl.h"
#include "bpmax_internal.h"
#include "bpmax_launch.h"

using namespace bpx;

namespace bpx {

const char *kClassNames[BPMAX_K_NCLASSES] = {"fill", "s1", "outer", "inner_ext", "inner_intra", "convert", "r0tri", "tasks"};

int choose_tile(int N) { return N <= 64 ? 32 : 64; }
// storage tile of an N-long strand under a tuning's `tile` setting (0: choose_tile)
int pick_tile(int N, int tile) { return tile == 32 ? 32 : choose_tile(N); }

unsigned long long splitmix64_host(unsigned long long x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

size_t pair_device_floats(int M, int N, int T)
{
    const long long nb = (N + T - 1) / T;
    const long long tstride = tri_ll(nb) * T * T;
    auto al = [](long long x) { return (x + 63) / 64 * 64; };
    return (size_t)(al(tri_ll(M) * tstride) + al(tstride) + al((long long)M * M));
}

// Exactness check of a weight table: every weight finite, >= 0 and a multiple of 2^-p for
// some p <= 12, with (largest weight) * (max pairs) * 2^p < 2^24, so every partial score
// is an exactly representable float and max-plus evaluation order cannot change a bit.
bool weights_exact(const float *intra, const float *inter, int M, int N, std::string *why)
{
    float wmax = 0.f;
    int p = -1;
    for (int cand = 0; cand <= 12 && p < 0; ++cand) {
        bool ok = true;
        for (int t = 0; t < 50 && ok; ++t) {
            const float x = t < 25 ? intra[t] : inter[t - 25];
            if (!std::isfinite(x) || x < 0.f) {
                if (why) *why = "weights must be finite and >= 0";
                return false;
            }
            const double s = std::ldexp((double)x, cand);
            ok = (s == std::floor(s));
        }
        if (ok) p = cand;
    }
    if (p < 0) {
        if (why) *why = "weights must be multiples of 2^-12";
        return false;
    }
    for (int t = 0; t < 25; ++t) wmax = std::max(wmax, std::max(intra[t], inter[t]));
    const double bound = (double)wmax * (((double)M + N) / 2.0 + 1.0) * std::ldexp(1.0, p);
    if (bound >= 16777216.0) {
        if (why) *why = "scores could exceed the exact float range for these weights and lengths";
        return false;
    }
    return true;
}

// The same bound against FP16 (11 significant bits): every score of an M x N pair is an
// integer multiple of 2^-p no larger than wmax * ((M+N)/2 + 1), so with that bound times 2^p
// at most 2^11 every score, and every R0 sum (never above the final value of its cell), is an
// exact FP16 number.  Weights that pass weights_exact only.
//
// With the sequences (s1, s2 non-null) and no negative weight, a tighter bound: each base
// pairs at most once and every term of the recurrence adds one pair's weight or sums scores
// over disjoint ranges, so a score is the weight of a matching of the pair's bases, with
// W(x,y) = the larger weight of codes x and y in either table and order.  Any y >= 0 with
// y[x] + y[z] >= W(x,z) for all codes x, z (x = z included) bounds every matching by
// sum over bases of y[code]; y = best/2 (best[x] = max_z W(x,z)) is one such, and lowering
// one y[x] at a time to the least value the others allow keeps the property.  Tried in turn,
// each only when the one before fails: best/2; the least sum after lowering in every order
// of the codes; the least sum over every vertex of that polyhedron, which is the minimum
// (equal to the largest fractional matching; for the default table 1.0 per base for
// uniform ACGU, M + N up to 1024, against 1.5 per base by length).  A vertex solves 5 of the
// 20 constraints as equalities; its basis is a forest of one-cycle components (odd cycles,
// loops, or a tight y >= 0), so it is a multiple of half the weights' grid, 2^-(p+1): each
// solution is rounded to that grid and kept only if it satisfies every constraint exactly.
static double vertex_bound(const double (&W)[5][5], const double (&cnt)[5], double grid, double upper)
{
    struct Row { int x, z; double rhs; }; // y[x] + y[z] >= rhs (z = x: 2 y[x]; z < 0: y[x] >= 0)
    Row rows[20];
    int nr = 0;
    for (int x = 0; x < 5; ++x)
        for (int z = x; z < 5; ++z) rows[nr++] = Row{x, z, W[x][z]};
    for (int x = 0; x < 5; ++x) rows[nr++] = Row{x, -1, 0.0};
    double best = upper;
    int pick[5] = {0, 1, 2, 3, 4};
    for (;;) {
        double A[5][6] = {};
        for (int r = 0; r < 5; ++r) {
            const Row &q = rows[pick[r]];
            A[r][q.x] += 1.0;
            if (q.z >= 0) A[r][q.z] += 1.0;
            A[r][5] = q.rhs;
        }
        bool ok = true;
        for (int c = 0; c < 5 && ok; ++c) { // Gaussian elimination, partial pivoting
            int pv = c;
            for (int r = c + 1; r < 5; ++r)
                if (std::fabs(A[r][c]) > std::fabs(A[pv][c])) pv = r;
            if (std::fabs(A[pv][c]) < 1e-9) {
                ok = false;
                break;
            }
            for (int k = 0; k < 6; ++k) std::swap(A[c][k], A[pv][k]);
            for (int r = 0; r < 5; ++r) {
                if (r == c) continue;
                const double f = A[r][c] / A[c][c];
                for (int k = c; k < 6; ++k) A[r][k] -= f * A[c][k];
            }
        }
        if (ok) {
            double y[5];
            for (int x = 0; x < 5; ++x) y[x] = std::round(A[x][5] / A[x][x] / grid) * grid;
            for (int r = 0; r < 20 && ok; ++r) {
                const Row &q = rows[r];
                ok = (q.z < 0 ? y[q.x] : (q.z == q.x ? 2.0 * y[q.x] : y[q.x] + y[q.z])) >= q.rhs;
            }
            if (ok) {
                double t = 0.0;
                for (int x = 0; x < 5; ++x) t += cnt[x] * y[x];
                best = std::min(best, t);
            }
        }
        int i = 4; // next 5-subset of the 20 rows, in lexicographic order
        while (i >= 0 && pick[i] == 15 + i) --i;
        if (i < 0) break;
        ++pick[i];
        for (int k = i + 1; k < 5; ++k) pick[k] = pick[k - 1] + 1;
    }
    return best;
}

// The smallest of the bounds above on every score of an M x N pair (in score units), trying
// the next one only while the bound is above `stop`.  p: the weights' grid exponent.
static double score_bound(const float *intra, const float *inter, int M, int N, const int *s1, const int *s2, int p,
                          double stop)
{
    float wmax = 0.f, wmin = 0.f;
    for (int t = 0; t < 25; ++t) {
        wmax = std::max(wmax, std::max(intra[t], inter[t]));
        wmin = std::min(wmin, std::min(intra[t], inter[t]));
    }
    double bound = (double)wmax * (((double)M + N) / 2.0 + 1.0);
    if (bound <= stop || !s1 || !s2 || wmin < 0.f) return bound;
    double W[5][5], y0[5], cnt[5] = {0, 0, 0, 0, 0};
    for (int x = 0; x < 5; ++x) {
        y0[x] = 0.0;
        for (int z = 0; z < 5; ++z) {
            W[x][z] = std::max({(double)intra[5 * x + z], (double)intra[5 * z + x], (double)inter[5 * x + z],
                                (double)inter[5 * z + x]});
            y0[x] = std::max(y0[x], W[x][z] / 2.0);
        }
    }
    for (int i = 0; i < M; ++i) cnt[s1[i]] += 1.0;
    for (int i = 0; i < N; ++i) cnt[s2[i]] += 1.0;
    auto total = [&](const double *y) {
        double t = 0.0;
        for (int x = 0; x < 5; ++x) t += cnt[x] * y[x];
        return t;
    };
    bound = std::min(bound, total(y0));
    if (bound <= stop) return bound;
    int order[5] = {0, 1, 2, 3, 4};
    do {
        double y[5];
        std::copy(y0, y0 + 5, y);
        for (int pass = 0; pass < 2; ++pass)
            for (int t = 0; t < 5; ++t) {
                const int x = order[t];
                double v = W[x][x] / 2.0;
                for (int z = 0; z < 5; ++z)
                    if (z != x) v = std::max(v, W[x][z] - y[z]);
                y[x] = std::max(0.0, v);
            }
        bound = std::min(bound, total(y));
    } while (std::next_permutation(order, order + 5));
    if (bound <= stop) return bound;
    return vertex_bound(W, cnt, std::ldexp(0.5, -p), bound);
}

static int grid_exponent(const float *intra, const float *inter)
{
    for (int cand = 0; cand <= 12; ++cand) {
        bool ok = true;
        for (int t = 0; t < 50 && ok; ++t) {
            const double s = std::ldexp((double)(t < 25 ? intra[t] : inter[t - 25]), cand);
            ok = (s == std::floor(s));
        }
        if (ok) return cand;
    }
    return -1;
}

bool weights_half_exact(const float *intra, const float *inter, int M, int N, const int *s1, const int *s2)
{
    const int p = grid_exponent(intra, inter);
    if (p < 0) return false;
    const double stop = std::ldexp(2048.0, -p);
    return score_bound(intra, inter, M, N, s1, s2, p, stop) <= stop;
}

} // namespace bpx

extern "C" double bpmax_score_bound(const float *intra, const float *inter, const int *seq1, int M, const int *seq2,
                                   int N)
{
    if (!seq1 || !seq2 || M < 1 || N < 1) return -1.0;
    for (int i = 0; i < M; ++i)
        if (seq1[i] < 0 || seq1[i] > 4) return -1.0;
    for (int i = 0; i < N; ++i)
        if (seq2[i] < 0 || seq2[i] > 4) return -1.0;
    float dflt_intra[25], dflt_inter[25], wt[50];
    bpmax_engine_score_table(dflt_intra, dflt_inter);
    memcpy(wt, intra ? intra : dflt_intra, 25 * sizeof(float));
    memcpy(wt + 25, inter ? inter : (intra ? intra : dflt_inter), 25 * sizeof(float));
    for (int t = 0; t < 50; ++t)
        if (!std::isfinite(wt[t]) || wt[t] < 0.f) return -1.0;
    const int p = grid_exponent(wt, wt + 25);
    return bpx::score_bound(wt, wt + 25, M, N, seq1, seq2, p < 0 ? 0 : p, -1.0);
}

// ----------------------------------------------------------------------------- ctx

int bpx_grow(bpmax_ctx *ctx, DevBuf &b, size_t bytes)
{
    if (b.cap >= bytes) return BPMAX_OK;
    if (b.p) {
        hipError_t e = hipFree(b.p);
        b.p = nullptr;
        b.cap = 0;
        if (e != hipSuccess) {
            ctx->err = std::string("hipFree: ") + hipGetErrorString(e);
            return BPMAX_EDEVICE;
        }
    }
    const size_t want = std::max<size_t>(bytes, 256);
    if (hipMalloc(&b.p, want) != hipSuccess) {
        (void)hipGetLastError();
        b.p = nullptr;
        ctx->err = "hipMalloc of " + std::to_string(want) + " bytes failed";
        return BPMAX_ENOMEM;
    }
    b.cap = want;
    return BPMAX_OK;
}

static void free_buf(DevBuf &b)
{
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

extern "C" int bpmax_ctx_create(bpmax_ctx **out, int device)
{
    if (!out) return BPMAX_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return BPMAX_EDEVICE;
    }
    if (device < 0 || device >= n) return BPMAX_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return BPMAX_EDEVICE;
    bpmax_ctx *c = new bpmax_ctx();
    c->device = device;
    if (const char *e = getenv("BPMAX_SCHEDULE")) {
        if (strcmp(e, "launch") == 0) c->schedule = BPMAX_SCHED_LAUNCH;
        else if (strcmp(e, "tasks") == 0) c->schedule = BPMAX_SCHED_TASKS;
        else c->schedule = BPMAX_SCHED_PIPE;
    }
    if (const char *e = getenv("BPMAX_TASK_CHUNK")) c->task_chunk = std::max(1, atoi(e));
    if (const char *e = getenv("BPMAX_GRAPHS")) c->graphs = atoi(e) != 0;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return BPMAX_EDEVICE;
    }
    if (hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return BPMAX_EDEVICE;
    }
    if (const char *e = getenv("BPMAX_OUTER_SPLIT")) c->outer_split = std::max(0, atoi(e));
    if (const char *e = getenv("BPMAX_OUTER_KGROUP")) c->outer_kgroup = std::max(0, atoi(e));
    if (const char *e = getenv("BPMAX_R0_AHEAD")) c->r0_ahead = atoi(e) != 0;
    if (const char *e = getenv("BPMAX_CLOSURE_AHEAD")) c->closure_ahead = atoi(e) != 0;
    if (const char *e = getenv("BPMAX_INTRA_BLOCKED")) c->intra_blocked = atoi(e) != 0;
    if (const char *e = getenv("BPMAX_R0_HALF")) c->r0_half = atoi(e) != 0;
    *out = c;
    return BPMAX_OK;
}

extern "C" int bpmax_ctx_destroy(bpmax_ctx *ctx)
{
    if (!ctx) return BPMAX_EINVAL;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    bpx_server_shutdown(ctx);
    free_buf(ctx->arena);
    free_buf(ctx->codes);
    free_buf(ctx->meta);
    free_buf(ctx->out);
    free_buf(ctx->scores);
    free_buf(ctx->r0meta);
    free_buf(ctx->r0half);
    free_buf(ctx->flags);
    (void)hipStreamSynchronize(ctx->stream2);
    if (ctx->graph) (void)hipGraphExecDestroy(ctx->graph);
    for (auto &e : ctx->ev_pool) (void)hipEventDestroy(e);
    for (auto &e : ctx->sync_ev) (void)hipEventDestroy(e);
    for (void *p : ctx->user_allocs) (void)hipFree(p);
    for (int b = 0; b < 2; ++b) {
        if (ctx->pin[b]) (void)hipHostFree(ctx->pin[b]);
        if (ctx->pin_ev[b]) (void)hipEventDestroy(ctx->pin_ev[b]);
    }
    (void)hipStreamDestroy(ctx->stream2);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return BPMAX_OK;
}

extern "C" const char *bpmax_last_error(const bpmax_ctx *ctx)
{
    return ctx ? ctx->err.c_str() : "null context";
}

extern "C" void *bpmax_ctx_stream(bpmax_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

extern "C" int bpmax_ctx_synchronize(bpmax_ctx *ctx)
{
    if (!ctx) return BPMAX_EINVAL;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return BPMAX_OK;
}

extern "C" int bpmax_ctx_set_timing(bpmax_ctx *ctx, int enable)
{
    if (!ctx) return BPMAX_EINVAL;
    ctx->timing = enable != 0;
    ctx->timing_mask = ~0u;
    return BPMAX_OK;
}

extern "C" int bpmax_ctx_set_timing_mask(bpmax_ctx *ctx, unsigned mask)
{
    if (!ctx) return BPMAX_EINVAL;
    ctx->timing = mask != 0;
    ctx->timing_mask = mask;
    return BPMAX_OK;
}

extern "C" int bpmax_ctx_set_schedule(bpmax_ctx *ctx, int schedule, int chunk)
{
    if (!ctx || chunk < 0) return BPMAX_EINVAL;
    if (schedule != BPMAX_SCHED_TASKS && schedule != BPMAX_SCHED_LAUNCH && schedule != BPMAX_SCHED_PIPE)
        return BPMAX_EINVAL;
    ctx->schedule = schedule;
    if (chunk > 0) ctx->task_chunk = chunk;
    return BPMAX_OK;
}

extern "C" int bpmax_ctx_set_graphs(bpmax_ctx *ctx, int enable)
{
    if (!ctx) return BPMAX_EINVAL;
    ctx->graphs = enable != 0;
    return BPMAX_OK;
}

extern "C" int bpmax_ctx_set_r0_half(bpmax_ctx *ctx, int enable)
{
    if (!ctx) return BPMAX_EINVAL;
    ctx->r0_half = enable != 0;
    return BPMAX_OK;
}

extern "C" int bpmax_ctx_r0_half_used(const bpmax_ctx *ctx) { return ctx && ctx->run_half ? 1 : 0; }


// ------------------------------------------------------------------------- autotuning
namespace bpx {
static int ceil_log2(int x)
{
    int k = 0;
    while ((1 << k) < x) ++k;
    return k;
}
// problem class: M and N rounded up to powers of two, single pair or batch
long long tuning_class(int M, int N, int npairs)
{
    return ((long long)ceil_log2(M) << 16) | ((long long)ceil_log2(N) << 8) | (npairs > 1 ? 1 : 0);
}
static bool tuning_valid(const bpmax_tuning *t)
{
    return t && (t->schedule == BPMAX_SCHED_PIPE || t->schedule == BPMAX_SCHED_LAUNCH) && t->outer_split >= 0 &&
           t->outer_items >= 64 && t->intra_wide >= 0 && t->outer_kgroup >= 0 && t->outer_kgroup <= 255 &&
           (t->closure_ahead == 0 || t->closure_ahead == 1) && (t->tile == 0 || t->tile == 32);
}
} // namespace bpx

extern "C" int bpmax_ctx_set_tuning(bpmax_ctx *ctx, int M, int N, int npairs, const bpmax_tuning *t)
{
    if (!ctx || M < 1 || N < 1 || npairs < 1) return BPMAX_EINVAL;
    const long long cls = tuning_class(M, N, npairs);
    if (!t) {
        ctx->tuning.erase(cls);
        return BPMAX_OK;
    }
    if (!tuning_valid(t)) return BPMAX_EINVAL;
    ctx->tuning[cls] = Tuning{t->schedule,   t->outer_split,  t->outer_items,   t->graphs ? 1 : 0,
                              t->intra_wide, t->outer_kgroup, t->closure_ahead, t->tile};
    return BPMAX_OK;
}

extern "C" int bpmax_ctx_set_default_tuning(bpmax_ctx *ctx, const bpmax_tuning *t)
{
    if (!ctx || !tuning_valid(t)) return BPMAX_EINVAL;
    ctx->schedule = t->schedule;
    ctx->outer_split = t->outer_split;
    ctx->outer_items = t->outer_items;
    ctx->graphs = t->graphs != 0;
    ctx->intra_wide = t->intra_wide;
    ctx->outer_kgroup = t->outer_kgroup;
    ctx->closure_ahead = t->closure_ahead;
    ctx->tile = t->tile;
    return BPMAX_OK;
}

extern "C" int bpmax_ctx_get_tuning(bpmax_ctx *ctx, int M, int N, int npairs, bpmax_tuning *t)
{
    if (!ctx || !t || M < 1 || N < 1 || npairs < 1) return BPMAX_EINVAL;
    auto it = ctx->tuning.find(tuning_class(M, N, npairs));
    if (it == ctx->tuning.end()) return BPMAX_ENOTFOUND;
    *t = bpmax_tuning{it->second.schedule,   it->second.outer_split,  it->second.outer_items,  it->second.graphs,
                      it->second.intra_wide, it->second.outer_kgroup, it->second.closure_ahead, it->second.tile};
    return BPMAX_OK;
}

// Times the candidate launch schedules on the given pairs and keeps the fastest for their
// class (largest M, largest N, single or batch): one stream or two-stream pipeline, R0
// launches per outer diagonal, R0 work items per launch; then, each under the best so far,
// the in-tile width threshold, the K group, closure ahead and graph replay.  Batches, whose
// runs cost more, try a shorter first list.
static int tune_pairs(bpmax_ctx *ctx, int npairs, const int *const *p1, const int *Ms, const int *const *p2,
                      const int *Ns, int reps, bpmax_tuning *best_out, double *best_ms)
{
    if (!ctx || npairs < 1 || !p1 || !p2 || !Ms || !Ns || reps < 1) return BPMAX_EINVAL;
    int mM = 0, mN = 0;
    for (int p = 0; p < npairs; ++p) {
        mM = std::max(mM, Ms[p]);
        mN = std::max(mN, Ns[p]);
    }
    std::vector<float> scores(npairs);
    auto run = [&]() -> int {
        if (npairs == 1) return bpmax_engine_solve(ctx, p1[0], Ms[0], p2[0], Ns[0], nullptr, nullptr, nullptr,
                                                   BPMAX_LAYOUT_NONE, scores.data());
        return bpmax_engine_solve_batch(ctx, npairs, p1, Ms, p2, Ns, nullptr, nullptr, scores.data());
    };
    // two rounds: the launch schedule at the default in-tile width threshold, then the
    // threshold (largest in-tile launch that uses the 1024-lane form) under the best schedule
    const int w0 = ctx->intra_wide, k0 = ctx->outer_kgroup, a0 = ctx->closure_ahead;
    std::vector<Tuning> cands;
    if (npairs == 1) {
        for (int items : {1024, 2048, 4096}) {
            cands.push_back(Tuning{BPMAX_SCHED_LAUNCH, 1, items, 0, w0, k0, a0});
            for (int split : {1, 2, 4, 0}) cands.push_back(Tuning{BPMAX_SCHED_PIPE, split, items, 0, w0, k0, a0});
        }
        cands.push_back(Tuning{BPMAX_SCHED_PIPE, 1, 4096, 1, w0, k0, a0});
    } else {
        cands.push_back(Tuning{BPMAX_SCHED_LAUNCH, 1, 2048, 0, w0, k0, a0});
        for (int items : {2048, 4096})
            for (int split : {1, 2, 4}) cands.push_back(Tuning{BPMAX_SCHED_PIPE, split, items, 0, w0, k0, a0});
    }
    const long long cls = tuning_class(mM, mN, npairs);
    const bool had = ctx->tuning.count(cls) != 0;
    const Tuning saved = had ? ctx->tuning[cls] : Tuning{};
    Tuning best{};
    double best_t = 1e300;
    auto trial = [&](const Tuning &c) -> int {
        ctx->tuning[cls] = c;
        int rc = run(); // builds and caches the schedule
        std::vector<double> ts;
        for (int r = 0; r < reps && rc == BPMAX_OK; ++r) {
            const auto t0 = std::chrono::steady_clock::now();
            rc = run();
            ts.push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
        }
        if (rc != BPMAX_OK) return rc;
        std::sort(ts.begin(), ts.end());
        const double med = ts[ts.size() / 2];
        if (med < best_t) {
            best_t = med;
            best = c;
        }
        return BPMAX_OK;
    };
    int rc = BPMAX_OK;
    for (const Tuning &c : cands)
        if ((rc = trial(c)) != BPMAX_OK) break;
    const Tuning sched_best = best;
    for (int wide : {0, 64, 256, 1024, 1 << 30}) {
        if (rc != BPMAX_OK || wide == w0) continue;
        Tuning c = sched_best;
        c.intra_wide = wide;
        rc = trial(c);
    }
    const Tuning wide_best = best;
    for (int kg : {0, 1, 2}) {
        if (rc != BPMAX_OK || kg == k0) continue;
        Tuning c = wide_best;
        c.outer_kgroup = kg;
        rc = trial(c);
    }
    const Tuning kg_best = best;
    for (int ahead : {0, 1}) {
        if (rc != BPMAX_OK || ahead == a0) continue;
        Tuning c = kg_best;
        c.closure_ahead = ahead;
        rc = trial(c);
    }
    // storage tile 32 instead of 64 (N > 64): less padding of N to whole tiles, against the
    // 64-wide product's lower LDS traffic per update
    const Tuning ahead0 = best;
    if (rc == BPMAX_OK && mN > 64) {
        Tuning c = ahead0;
        c.tile = c.tile == 32 ? 0 : 32;
        rc = trial(c);
    }
    // graph replay of the pipelined schedule: it saves most on small problems, whose solves
    // are chains of short launches (16x100: 1.79 -> 1.48 ms)
    const Tuning ahead_best = best;
    if (rc == BPMAX_OK && ahead_best.schedule == BPMAX_SCHED_PIPE) {
        Tuning c = ahead_best;
        c.graphs = !c.graphs;
        rc = trial(c);
    }
    if (rc != BPMAX_OK) {
        if (had) ctx->tuning[cls] = saved;
        else ctx->tuning.erase(cls);
        return rc;
    }
    ctx->tuning[cls] = best;
    if (best_out)
        *best_out = bpmax_tuning{best.schedule,   best.outer_split,  best.outer_items,  best.graphs,
                                 best.intra_wide, best.outer_kgroup, best.closure_ahead, best.tile};
    if (best_ms) *best_ms = best_t;
    return BPMAX_OK;
}

extern "C" int bpmax_tune_batch(bpmax_ctx *ctx, int npairs, const int *const *seq1s, const int *Ms,
                                const int *const *seq2s, const int *Ns, int reps, bpmax_tuning *best,
                                double *best_ms)
{
    return tune_pairs(ctx, npairs, seq1s, Ms, seq2s, Ns, reps, best, best_ms);
}

// The same on `npairs` synthetic M x N pairs (uniform ACGU, fixed seed).
extern "C" int bpmax_tune_shape(bpmax_ctx *ctx, int M, int N, int npairs, int reps, bpmax_tuning *best_out,
                                double *best_ms)
{
    if (!ctx || M < 1 || N < 1 || npairs < 1 || reps < 1) return BPMAX_EINVAL;
    std::vector<std::vector<int>> s1(npairs, std::vector<int>(M)), s2(npairs, std::vector<int>(N));
    unsigned long long x = 0x5EEDull;
    for (int p = 0; p < npairs; ++p) {
        for (int &c : s1[p]) c = (int)((x = splitmix64_host(x)) & 3);
        for (int &c : s2[p]) c = (int)((x = splitmix64_host(x)) & 3);
    }
    std::vector<const int *> p1(npairs), p2(npairs);
    std::vector<int> Ms(npairs, M), Ns(npairs, N);
    for (int p = 0; p < npairs; ++p) {
        p1[p] = s1[p].data();
        p2[p] = s2[p].data();
    }
    return tune_pairs(ctx, npairs, p1.data(), Ms.data(), p2.data(), Ns.data(), reps, best_out, best_ms);
}

extern "C" int bpmax_ctx_task_stats(bpmax_ctx *ctx, long long *tasks, double *wait_polls)
{
    if (!ctx) return BPMAX_EINVAL;
    if (tasks) *tasks = ctx->task_count;
    if (wait_polls) *wait_polls = (double)ctx->task_polls;
    return BPMAX_OK;
}

extern "C" int bpmax_ctx_reset_stats(bpmax_ctx *ctx)
{
    if (!ctx) return BPMAX_EINVAL;
    for (auto &s : ctx->stats) s = KStat{};
    return BPMAX_OK;
}

extern "C" int bpmax_ctx_kernel_stats(bpmax_ctx *ctx, int kclass, long long *launches, double *ms,
                                      double *updates)
{
    if (!ctx || kclass < 0 || kclass >= BPMAX_K_NCLASSES) return BPMAX_EINVAL;
    if (launches) *launches = ctx->stats[kclass].launches;
    if (ms) *ms = ctx->stats[kclass].ms;
    if (updates) *updates = ctx->stats[kclass].updates;
    return BPMAX_OK;
}

extern "C" double bpmax_cell_updates(int M, int N)
{
    auto c2 = [](double n) { return n * (n - 1) / 2.0; };
    auto c3 = [](double n) { return n * (n - 1) * (n - 2) / 6.0; };
    const double m = M + 1.0, n = N + 1.0;
    return c3(m) * c3(n) + 2.0 * c2(m) * c3(n) + 2.0 * c3(m) * c2(n);
}

extern "C" unsigned char bpmax_base_code(char c)
{
    switch (c) {
    case 'A': case 'a': case '0': return 0;
    case 'C': case 'c': case '1': return 1;
    case 'G': case 'g': case '2': return 2;
    case 'T': case 't': case 'U': case 'u': case '3': return 3;
    default: return 4;
    }
}

extern "C" void bpmax_engine_score_table(float intra[25], float inter[25])
{
    float t[25];
    for (int i = 0; i < 25; ++i) t[i] = 0.f;
    const int A = 0, C = 1, G = 2, U = 3;
    t[5 * A + U] = t[5 * U + A] = 1.0f;
    t[5 * C + G] = t[5 * G + C] = 3.0f;
    t[5 * G + U] = t[5 * U + G] = 1.5f;
    if (intra) memcpy(intra, t, sizeof t);
    if (inter) memcpy(inter, t, sizeof t);
}

extern "C" size_t bpmax_layout_bytes(int M, int N, bpmax_layout layout)
{
    if (M < 1 || N < 1) return 0;
    switch (layout) {
    case BPMAX_LAYOUT_PACKED_TRI_TRI: return (size_t)tri_ll(M) * (size_t)tri_ll(N) * sizeof(float);
    case BPMAX_LAYOUT_DENSE_4D: return (size_t)M * M * (size_t)N * N * sizeof(float);
    default: return 0;
    }
}

extern "C" size_t bpmax_engine_pair_device_bytes(int M, int N)
{
    if (M < 1 || N < 1) return 0;
    return pair_device_floats(M, N, choose_tile(N)) * sizeof(float);
}

// ----------------------------------------------------------------------- schedule
namespace bpx {

static void chunked(std::vector<WorkItem> &v, WorkItem base, int steps, int ch, bool epilogue_first)
{
    if (steps <= 0) {
        base.step_begin = base.step_end = 0;
        base.flags = epilogue_first ? WI_EPILOGUE : 0;
        v.push_back(base);
        return;
    }
    const int n = (steps + ch - 1) / ch;
    for (int c = 0; c < n; ++c) {
        WorkItem w = base;
        w.step_begin = c * ch;
        w.step_end = std::min(steps, (c + 1) * ch);
        w.flags = (n > 1 ? WI_ATOMIC : 0) | ((c == 0 && epilogue_first) ? WI_EPILOGUE : 0);
        v.push_back(w);
    }
}


// Items of `ch` steps over [sb, se) of one tile (cross-tile steps in the closure launches
// that carry them ahead), all max-combining.
static void chunked_range(std::vector<WorkItem> &v, WorkItem base, int sb, int se, int ch)
{
    for (int s = sb; s < se; s += ch) {
        WorkItem w = base;
        w.step_begin = s;
        w.step_end = std::min(se, s + ch);
        v.push_back(w);
    }
}

// R0 items from `first` on: workgroup b runs item xcd_work_index(b, n), so XCD x (= b % 8)
// runs the contiguous index range x*q + min(x, r) .. in order (q = n / 8, r = n % 8).  The
// last k items of every range are split into two halves of their steps, so that each XCD's
// share of the launch ends on short items (a shorter tail: 64x512 with 2048 items per
// launch 61.2 -> 58.4 ms).  Every range grows by k, which keeps the ranges of the new count
// aligned with them.  Both halves max-combine; the epilogue stays with the first.
static void split_tail(std::vector<WorkItem> &v, size_t first)
{
    const long long n = (long long)(v.size() - first);
    if (n < 64) return;
    const long long q = n / 8, r = n % 8;
    long long k = std::min<long long>(q / 4, 128);
    auto begin_of = [&](long long x) { return (long long)first + x * q + std::min(x, r); };
    auto size_of = [&](long long x) { return q + (x < r ? 1 : 0); };
    for (long long x = 0; x < 8; ++x) { // only items of >= 2 steps can be split
        long long run = 0;
        for (long long i = size_of(x) - 1; i >= 0 && run < k; --i) {
            const WorkItem &w = v[begin_of(x) + i];
            if (w.step_end - w.step_begin < 2) break;
            ++run;
        }
        k = std::min(k, run);
    }
    if (k <= 0) return;
    std::vector<WorkItem> out;
    out.reserve((size_t)(n + 8 * k));
    for (long long x = 0; x < 8; ++x)
        for (long long i = 0; i < size_of(x); ++i) {
            const WorkItem w = v[begin_of(x) + i];
            if (i < size_of(x) - k) {
                out.push_back(w);
                continue;
            }
            const int mid = w.step_begin + (w.step_end - w.step_begin) / 2;
            WorkItem a = w, b = w;
            a.step_end = mid;
            b.step_begin = mid;
            a.flags |= WI_ATOMIC;
            b.flags = (b.flags | WI_ATOMIC) & ~WI_EPILOGUE;
            out.push_back(a);
            out.push_back(b);
        }
    v.resize(first);
    v.insert(v.end(), out.begin(), out.end());
}

static int pick_chunk(long long total_steps, long long tiles)
{
    // aim for ~8 work items per CU (256 CUs) when the work allows it
    long long ch = total_steps / 2048;
    if (ch < 2) ch = 2;
    if (ch > 32) ch = 32;
    (void)tiles;
    return (int)ch;
}

// Launch lists are built from two pieces, both over the pairs of one storage-tile group g
// (pairs sharing T):
//   emit_s2:   S2 by logical block diagonal (cross-tile launch, then closure launch) for the
//              pairs with sel[p] set;
//   emit_tick: one step of the outer wavefront -- every selected pair p advances its outer
//              diagonal d1s[p] (pairs may sit on different diagonals): R0 launches over
//              `split` ranges of storage block diagonals (all storage tiles (I,I+Ds) of every
//              triangle (i1,i1+d1s[p]), Ds in the range), then the inner closure by logical
//              block diagonal DL (cross-tile launch, closure launch) -- or, with
//              closure_ahead, one launch per DL (k_inner_closure_blk): the in-tile items of
//              DL, whose workgroups first run their tiles' cross-tile steps D-1 and D (the two
//              that read diagonal DL-1), and cross-tile items with the other steps of the
//              tiles on DL+1 (they read diagonals below DL only).
// nbmax / nblmax: storage / logical tiles per side the tick's ranges and levels cover (the
// largest of the group); R0 range r ends at storage diagonal out_end of its last Ds.
// Every range / DL position gets a Launch entry, possibly empty, so that a runner can
// attach its events to each.  An R0 launch's `diag` is the last storage diagonal of its
// range; launches carry the tick index in `d1`.
void emit_s2(const std::vector<PairDesc> &P, const std::vector<char> &sel, int g, int T, int nblmax, Schedule &S)
{
    auto emit = [&](int kind, size_t first, int diag) {
        S.launches.push_back(Launch{kind, T, true, (int)first, (int)(S.items.size() - first), 0, g, -1, diag});
    };
    for (int D = 0; D < nblmax; ++D) {
        if (D > 0) {
            size_t f = S.items.size();
            long long tot = 0;
            for (size_t p = 0; p < P.size(); ++p)
                if (sel[p] && P[p].T == T && P[p].nblr > D) tot += (long long)(P[p].nblr - D) * D;
            const int ch = pick_chunk(tot, 0);
            for (size_t p = 0; p < P.size(); ++p) {
                if (!sel[p] || P[p].T != T || P[p].nblr <= D) continue;
                for (int I = 0; I + D < P[p].nblr; ++I)
                    chunked(S.items, WorkItem{(int)p, 0, 0, I, I + D, 0, 0, 0}, D, ch, false);
            }
            emit(BPMAX_K_INNER_EXT, f, D);
        }
        size_t f = S.items.size();
        for (size_t p = 0; p < P.size(); ++p) {
            if (!sel[p] || P[p].T != T || P[p].nblr <= D) continue;
            for (int I = 0; I + D < P[p].nblr; ++I) S.items.push_back(WorkItem{(int)p, 0, 0, I, I + D, 0, 0, 0});
        }
        emit(BPMAX_K_INNER_INTRA, f, D);
    }
}

// Which k1 terms of a triangle on outer diagonal d1 an R0 launch covers.
enum R0Part {
    R0_ALL,  // every k1 in [i1, j1), with the epilogue
    R0_BULK, // k1 in [i1+1, j1-2]: operands on outer diagonals <= d1-2
    R0_EDGE, // k1 = i1 and k1 = j1-1 (operands on diagonal d1-1), with the epilogue
};

// R0 launches of one step of the outer wavefront (see emit_tick) covering the k1 terms of
// `part` (R0_EDGE: both terms of a tile in one item of two steps per K), in `outer_split`
// ranges of storage diagonals, whose ends out_end receives.
static void emit_r0(const std::vector<PairDesc> &P, const std::vector<int> &d1s, int g, int T, int nbmax, int tick,
                    Schedule &S, std::vector<int> &out_end, int outer_split, int outer_items, int outer_kgroup,
                    R0Part part)
{
    auto active = [&](size_t p) { return P[p].T == T && d1s[p] >= 0 && d1s[p] < P[p].M; };
    auto nk1 = [&](int d1) { return part == R0_ALL ? d1 : (part == R0_BULK ? std::max(0, d1 - 2) : std::min(d1, 2)); };
    // R0 launches: storage diagonals in `split` contiguous ranges
    const int split = outer_split <= 0 ? nbmax : std::min(outer_split, nbmax);
    std::vector<int> range_end(split);
    out_end.assign(nbmax, 0);
    for (int r = 0, Ds = 0; r < split; ++r) {
        range_end[r] = (int)((long long)nbmax * (r + 1) / split) - 1;
        for (; Ds <= range_end[r]; ++Ds) out_end[Ds] = range_end[r];
    }
    for (int r = 0; r < split; ++r) {
        const int Da = r == 0 ? 0 : range_end[r - 1] + 1, Db = range_end[r];
        size_t f = S.items.size();
        // K values per item: a tile's K range split into groups of kg, so that every item of
        // a (triangle, k1 chunk) takes about one product step per k1 and the items that run
        // together stay aligned on k1 (each source tile read from L2 by all of its consumers
        // at about the same time)
        auto kg_of = [&](int d1) { return d1 == 0 || nbmax > 255 ? 0 : outer_kgroup; };
        auto groups = [&](int kg, int nK) { return kg > 0 ? (nK + kg - 1) / kg : 1; };
        long long steps = 0;
        double upd = 0;
        for (size_t p = 0; p < P.size(); ++p) {
            if (!active(p)) continue;
            const int d1 = d1s[p], kg = kg_of(d1);
            long long tiles = 0;
            for (int Ds = Da; Ds <= Db && Ds < P[p].nb; ++Ds)
                tiles += (long long)(P[p].M - d1) * (P[p].nb - Ds) * groups(kg, Ds + 1);
            steps += (long long)nk1(d1) * tiles;
            if (r == 0) {
                const double n = P[p].N;
                const double c3n = (n + 1) * n * (n - 1) / 6.0, c2n = (n + 1) * n / 2.0;
                upd += (double)(P[p].M - d1) * nk1(d1) * (c3n + 2.0 * c2n);
            }
        }
        // Items cover whole k1 values (ck of them, every K of each), ordered triangle,
        // k1-chunk, tile: workgroups running together then read the same source tiles (row I
        // of F(i1,k1), column J of F(k1+1,j1)) and share them in L2.
        const long long want = std::max(64, outer_items);
        const int ck = (int)std::max<long long>(1, (steps + want - 1) / want);
        for (size_t p = 0; p < P.size(); ++p) {
            if (!active(p)) continue;
            const int d1 = d1s[p], kg = kg_of(d1), n1 = nk1(d1);
            const int k1off = part == R0_BULK ? 1 : 0; // first k1 of the part, minus i1
            // k1 chunks; R0_ALL on d1 = 0 and R0_EDGE: one (the epilogue, or both edge terms)
            const int nck = part == R0_EDGE ? 1 : (n1 == 0 ? (part == R0_ALL ? 1 : 0) : (n1 + ck - 1) / ck);
            for (int i1 = 0; i1 + d1 < P[p].M; ++i1)
                for (int c = 0; c < nck; ++c)
                    for (int Ds = Da; Ds <= Db && Ds < P[p].nb; ++Ds)
                        for (int I = 0; I + Ds < P[p].nb; ++I)
                            for (int gi = 0, ng = groups(kg, Ds + 1); gi < ng; ++gi) {
                                const int koff = kg > 0 ? gi * kg : 0;
                                const int nK = kg > 0 ? std::min(kg, Ds + 1 - koff) : Ds + 1;
                                int sb, se, fl;
                                if (part == R0_EDGE) {
                                    // d1 = 0: the epilogue alone; d1 = 1: k1 = i1; else both edges
                                    sb = 0;
                                    se = std::min(d1, 2) * nK;
                                    fl = (d1 >= 3 || ng > 1 ? WI_ATOMIC : 0) | (gi == 0 ? WI_EPILOGUE : 0) |
                                         (d1 >= 2 ? WI_EDGE : 0);
                                } else {
                                    sb = (k1off + c * ck) * nK;
                                    se = (k1off + std::min(n1, (c + 1) * ck)) * nK;
                                    fl = part == R0_BULK ? WI_ATOMIC
                                                         : (nck > 1 || ng > 1 ? WI_ATOMIC : 0) |
                                                               (c == 0 && gi == 0 ? WI_EPILOGUE : 0);
                                }
                                if (kg > 0) fl |= (koff << WI_KOFF_SHIFT) | (nK << WI_KCNT_SHIFT);
                                S.items.push_back(WorkItem{(int)p, i1, i1 + d1, I, I + Ds, sb, se, fl});
                            }
        }
        split_tail(S.items, f);
        S.launches.push_back(Launch{BPMAX_K_OUTER, T, false, (int)f, (int)(S.items.size() - f), upd, g, tick, Db});
        S.launches.back().bulk = part == R0_BULK;
    }
}

// Inner-closure launches of one step of the outer wavefront (see emit_tick).
static void emit_closure(const std::vector<PairDesc> &P, const std::vector<int> &d1s, int g, int T, int nblmax,
                         int tick, Schedule &S, bool closure_ahead)
{
    auto active = [&](size_t p) { return P[p].T == T && d1s[p] >= 0 && d1s[p] < P[p].M; };
    auto emit = [&](int kind, size_t first, double upd, int diag) {
        S.launches.push_back(Launch{kind, T, false, (int)first, (int)(S.items.size() - first), upd, g, tick, diag});
    };
    // inner closure by logical block diagonal
    for (int D = 0; D < nblmax; ++D) {
        if (D > 0 && !closure_ahead) {
            size_t f = S.items.size();
            long long tot = 0;
            for (size_t p = 0; p < P.size(); ++p)
                if (active(p) && P[p].nblr > D) tot += (long long)(P[p].M - d1s[p]) * (P[p].nblr - D) * 2 * D;
            const int ch = pick_chunk(tot, 0);
            for (size_t p = 0; p < P.size(); ++p) {
                if (!active(p) || P[p].nblr <= D) continue;
                const int d1 = d1s[p];
                for (int i1 = 0; i1 + d1 < P[p].M; ++i1)
                    for (int I = 0; I + D < P[p].nblr; ++I)
                        chunked(S.items, WorkItem{(int)p, i1, i1 + d1, I, I + D, 0, 0, 0}, 2 * D, ch, false);
            }
            emit(BPMAX_K_INNER_EXT, f, 0, D);
        }
        size_t f = S.items.size();
        const int tin = closure_ahead ? TASK_INTRA << WI_TYPE_SHIFT : 0;
        // closure ahead: steps [0, D) and [D+2, 2D+2) of the tiles on D+1, in items longer
        // than the cross-tile launches' (these share launches with the in-tile work; 64x512,
        // 24x512 and 48x384 ran faster with 8..32 steps than with 4 or with the cross-tile
        // launches' chunks, and with FP16 cross-tile steps 64x512 ran 1 % faster again with
        // at least 16: 39.3 ms against 39.7 with 8; 24x512, 48x384 and 16x256 within 0.5 %)
        const int D1 = D + 1;
        long long tot = 0;
        for (size_t p = 0; p < P.size() && closure_ahead; ++p)
            if (active(p) && P[p].nblr > D1) tot += (long long)(P[p].M - d1s[p]) * (P[p].nblr - D1) * 2 * D;
        const int ch = (int)std::min<long long>(32, std::max<long long>(16, tot / 1024));
        const int tex = (TASK_EXT << WI_TYPE_SHIFT) | WI_ATOMIC;
        // by triangle, each triangle's in-tile items then its cross-tile items, so that every
        // XCD's contiguous share of the launch mixes the two kinds
        for (size_t p = 0; p < P.size(); ++p) {
            if (!active(p) || P[p].nblr <= D) continue;
            const int d1 = d1s[p];
            for (int i1 = 0; i1 + d1 < P[p].M; ++i1) {
                for (int I = 0; I + D < P[p].nblr; ++I)
                    S.items.push_back(WorkItem{(int)p, i1, i1 + d1, I, I + D, 0, 0, tin});
                if (!closure_ahead || D == 0) continue;
                for (int I = 0; I + D1 < P[p].nblr; ++I) {
                    const WorkItem base{(int)p, i1, i1 + d1, I, I + D1, 0, 0, tex};
                    chunked_range(S.items, base, 0, D, ch);
                    chunked_range(S.items, base, D + 2, 2 * D1, ch);
                }
            }
        }
        emit(BPMAX_K_INNER_INTRA, f, 0, D);
        S.launches.back().ahead = closure_ahead;
    }
}


void emit_tick(const std::vector<PairDesc> &P, const std::vector<int> &d1s, int g, int T, int nbmax, int nblmax,
               int tick, Schedule &S, std::vector<int> &out_end, int outer_split, int outer_items,
               int outer_kgroup, bool closure_ahead)
{
    emit_r0(P, d1s, g, T, nbmax, tick, S, out_end, outer_split, outer_items, outer_kgroup, R0_ALL);
    emit_closure(P, d1s, g, T, nblmax, tick, S, closure_ahead);
}

// The launch list of a whole solve of a set of pairs (descs already laid out), one group per
// storage tile size T: S2 of every pair, then one tick per outer diagonal d1 with every
// pair on d1 (pairs with M <= d1 are done).
void build_schedule(const std::vector<PairDesc> &P, Schedule &S, int outer_split, int outer_items, int outer_kgroup,
                    bool closure_ahead, bool r0_ahead)
{
    S.items.clear();
    S.launches.clear();
    S.groups.clear();
    for (auto &p : P) {
        int g = -1;
        for (size_t k = 0; k < S.groups.size(); ++k)
            if (S.groups[k].T == p.T) g = (int)k;
        if (g < 0) {
            S.groups.push_back(GroupInfo{p.T, p.nbl / p.nb, 0, 0, 0, {}});
            g = (int)S.groups.size() - 1;
        }
        GroupInfo &G = S.groups[g];
        G.maxM = std::max(G.maxM, p.M);
        G.maxnb = std::max(G.maxnb, p.nb);
        G.maxnbl = std::max(G.maxnbl, p.nblr);
    }
    const std::vector<char> all(P.size(), 1);
    std::vector<int> d1s(P.size());
    for (int g = 0; g < (int)S.groups.size(); ++g) {
        GroupInfo &G = S.groups[g];
        emit_s2(P, all, g, G.T, G.maxnbl, S);
        if (!r0_ahead) {
            for (int d1 = 0; d1 < G.maxM; ++d1) {
                std::fill(d1s.begin(), d1s.end(), d1);
                emit_tick(P, d1s, g, G.T, G.maxnb, G.maxnbl, d1, S, G.out_end, outer_split, outer_items,
                          outer_kgroup, closure_ahead);
            }
            continue;
        }
        // Step d1: the edge R0 of d1 (it reads diagonal d1-1), the bulk R0 of d1+1 (it reads
        // diagonals <= d1-1), then the closure of d1.  Both R0 parts become ready when the
        // closure of d1-1 ends; the edge part, which the closure of d1 waits for, goes first
        // on the R0 stream, and the bulk of d1+1 then runs beside the closure of d1.  Every
        // launch a later one waits for is emitted before it.
        for (int d1 = 0; d1 < G.maxM; ++d1) {
            std::fill(d1s.begin(), d1s.end(), d1);
            emit_r0(P, d1s, g, G.T, G.maxnb, d1, S, G.out_end, outer_split, outer_items, outer_kgroup, R0_EDGE);
            if (d1 + 1 < G.maxM) {
                std::fill(d1s.begin(), d1s.end(), d1 + 1);
                emit_r0(P, d1s, g, G.T, G.maxnb, d1 + 1, S, G.out_end, outer_split, outer_items, outer_kgroup,
                        R0_BULK);
            }
            std::fill(d1s.begin(), d1s.end(), d1);
            emit_closure(P, d1s, g, G.T, G.maxnbl, d1, S, closure_ahead);
        }
    }
}

// Task list for the dependency-tile scheduler (k_tasks), in a topological order:
// S1 of every pair; S2 logical tiles by block diagonal (cross-tile chunks before
// closures); then per outer diagonal d1, walking the logical block diagonals DL of the
// inner closure: before DL, the R0 chunks of every storage tile whose first logical
// sub-tile lies on DL; then the cross-tile chunks and closures of DL.  Every dependence
// of a task (see wait_deps in bpmax_kernels.hip) precedes it in this order.  flags
// receives each pair's initial counters (pending = contributions still to land).
void build_tasks(const std::vector<PairDesc> &P, int T, int chunk, std::vector<WorkItem> &tasks,
                 std::vector<int> &flags)
{
    tasks.clear();
    int maxM = 0, maxnbl = 0;
    for (auto &p : P)
        if (p.T == T) {
            maxM = std::max(maxM, p.M);
            maxnbl = std::max(maxnbl, p.nbl);
        }
    auto add = [&](int type, int p, int i1, int j1, int I, int J, int sb, int se, int extra) {
        tasks.push_back(WorkItem{p, i1, j1, I, J, sb, se, (type << WI_TYPE_SHIFT) | extra});
    };
    auto sctr = [&](const PairDesc &d, int X, int Y) -> int * {
        return flags.data() + d.flag + 2 + 2 * utri_ll(X, Y, d.nbl);
    };
    auto fctr = [&](const PairDesc &d, int i1, int j1, int X, int Y) -> int * {
        return flags.data() + d.flag + 2 + 2 * tri_ll(d.nbl) +
               2 * (utri_ll(i1, j1, d.M) * tri_ll(d.nbl) + utri_ll(X, Y, d.nbl));
    };
    for (size_t p = 0; p < P.size(); ++p)
        if (P[p].T == T) add(TASK_S1, (int)p, 0, 0, 0, 0, 0, 0, 0);
    for (int D = 0; D < maxnbl; ++D) {
        for (size_t p = 0; p < P.size(); ++p) {
            if (P[p].T != T || P[p].nbl <= D || D == 0) continue;
            for (int I = 0; I + D < P[p].nbl; ++I)
                for (int sb = 0; sb < D; sb += chunk) {
                    add(TASK_S2EXT, (int)p, 0, 0, I, I + D, sb, std::min(D, sb + chunk), WI_ATOMIC);
                    sctr(P[p], I, I + D)[1] += 1;
                }
        }
        for (size_t p = 0; p < P.size(); ++p) {
            if (P[p].T != T || P[p].nbl <= D) continue;
            for (int I = 0; I + D < P[p].nbl; ++I) add(TASK_S2INTRA, (int)p, 0, 0, I, I + D, 0, 0, 0);
        }
    }
    for (int d1 = 0; d1 < maxM; ++d1)
        for (int DL = 0; DL < maxnbl; ++DL) {
            // R0 chunks of storage tiles (I, I+Ds) whose lowest logical diagonal is DL
            for (size_t p = 0; p < P.size(); ++p) {
                const PairDesc &d = P[p];
                if (d.T != T || d.M <= d1) continue;
                const int rs = d.nbl / d.nb;
                for (int Ds = 0; Ds < d.nb; ++Ds) {
                    if (std::max(0, rs * Ds - (rs - 1)) != DL) continue;
                    for (int i1 = 0; i1 + d1 < d.M; ++i1)
                        for (int I = 0; I + Ds < d.nb; ++I) {
                            const int J = I + Ds, steps = d1 * (Ds + 1);
                            int n = 0;
                            if (steps == 0) {
                                add(TASK_OUTER, (int)p, i1, i1 + d1, I, J, 0, 0, WI_ATOMIC | WI_EPILOGUE);
                                n = 1;
                            } else {
                                for (int sb = 0; sb < steps; sb += chunk, ++n)
                                    add(TASK_OUTER, (int)p, i1, i1 + d1, I, J, sb, std::min(steps, sb + chunk),
                                        WI_ATOMIC | (sb == 0 ? WI_EPILOGUE : 0));
                            }
                            for (int x = I * rs; x < I * rs + rs; ++x)
                                for (int y = J * rs; y < J * rs + rs; ++y)
                                    if (x <= y) fctr(d, i1, i1 + d1, x, y)[1] += n;
                        }
                }
            }
            if (DL > 0)
                for (size_t p = 0; p < P.size(); ++p) {
                    const PairDesc &d = P[p];
                    if (d.T != T || d.M <= d1 || d.nbl <= DL) continue;
                    for (int i1 = 0; i1 + d1 < d.M; ++i1)
                        for (int I = 0; I + DL < d.nbl; ++I)
                            for (int sb = 0; sb < 2 * DL; sb += chunk) {
                                add(TASK_EXT, (int)p, i1, i1 + d1, I, I + DL, sb, std::min(2 * DL, sb + chunk), WI_ATOMIC);
                                fctr(d, i1, i1 + d1, I, I + DL)[1] += 1;
                            }
                }
            for (size_t p = 0; p < P.size(); ++p) {
                const PairDesc &d = P[p];
                if (d.T != T || d.M <= d1 || d.nbl <= DL) continue;
                for (int i1 = 0; i1 + d1 < d.M; ++i1)
                    for (int I = 0; I + DL < d.nbl; ++I) add(TASK_INTRA, (int)p, i1, i1 + d1, I, I + DL, 0, 0, 0);
            }
        }
}

} // namespace bpx

// Event-bracketed launch helper.
static int timed(bpmax_ctx *ctx, hipStream_t st, int kclass, double updates, const std::function<hipError_t()> &fn)
{
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const bool on = ctx->timing && ((ctx->timing_mask >> kclass) & 1u);
    if (on) {
        while (ctx->ev_pool.size() < ctx->ev_used + 2) {
            hipEvent_t e;
            HIPCHK(ctx, hipEventCreate(&e));
            ctx->ev_pool.push_back(e);
        }
        e0 = ctx->ev_pool[ctx->ev_used++];
        e1 = ctx->ev_pool[ctx->ev_used++];
        HIPCHK(ctx, hipEventRecord(e0, st));
    }
    hipError_t e = fn();
    if (e != hipSuccess) {
        ctx->err = std::string("kernel launch (") + kClassNames[kclass] + "): " + hipGetErrorString(e);
        return BPMAX_EDEVICE;
    }
    if (on) {
        HIPCHK(ctx, hipEventRecord(e1, st));
        ctx->pending.push_back(PendingTiming{kclass, updates, e0, e1});
    }
    return BPMAX_OK;
}

int bpx_collect_timing(bpmax_ctx *ctx)
{
    if (!ctx->timing || ctx->pending.empty()) {
        ctx->ev_used = 0;
        ctx->pending.clear();
        return BPMAX_OK;
    }
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    for (auto &p : ctx->pending) {
        float ms = 0.f;
        HIPCHK(ctx, hipEventElapsedTime(&ms, p.e0, p.e1));
        ctx->stats[p.kclass].launches += 1;
        ctx->stats[p.kclass].ms += ms;
        ctx->stats[p.kclass].updates += p.updates;
    }
    ctx->pending.clear();
    ctx->ev_used = 0;
    return BPMAX_OK;
}

// One phase launch of the schedule on stream st (nothing for an empty position).  A launch
// of more than kMaxItems work items (one workgroup each) goes out in pieces: the items of
// one launch are independent, and a grid stays far below the 2^32 threads a dispatch holds.
constexpr int kMaxItems = 1 << 22;
// half: R0 launches read the FP16 mirrors (every pair of the run has one, ctx->run_half).
static int run_launch(bpmax_ctx *ctx, hipStream_t st, const Launch &L, const PairDesc *dP, const WorkItem *dI,
                      float *arena, const int *dcodes, const float *dw, bool half)
{
    if (L.count == 0) return BPMAX_OK;
    if (L.count > kMaxItems) {
        for (int f = 0; f < L.count; f += kMaxItems) {
            Launch part = L;
            part.first = L.first + f;
            part.count = std::min(kMaxItems, L.count - f);
            part.updates = f == 0 ? L.updates : 0;
            const int rc = run_launch(ctx, st, part, dP, dI, arena, dcodes, dw, half);
            if (rc != BPMAX_OK) return rc;
        }
        return BPMAX_OK;
    }
    const WorkItem *items = dI + L.first;
    switch (L.kind) {
    case BPMAX_K_OUTER:
        return timed(ctx, st, L.kind, L.updates,
                     [&] { return launch_outer(st, L.T, dP, items, L.count, arena, dcodes, dw, half); });
    case BPMAX_K_INNER_EXT:
        return timed(ctx, st, L.kind, L.updates,
                     [&] { return launch_inner_ext(st, L.T, L.single_strand, dP, items, L.count, arena); });
    default:
        if (L.ahead)
            return timed(ctx, st, L.kind, L.updates,
                         [&] { return launch_inner_closure(st, L.T, dP, items, L.count, arena, dcodes, dw); });
        return timed(ctx, st, L.kind, L.updates, [&] {
            return launch_inner_intra(st, L.T, L.single_strand, dP, items, L.count, arena, dcodes, dw,
                                      L.count <= ctx->eff_intra_wide, ctx->intra_blocked != 0);
        });
    }
}

int bpx_run_launch(bpmax_ctx *ctx, hipStream_t st, const Launch &L, const PairDesc *dP, const WorkItem *dI,
                   float *arena, const int *dcodes, const float *dw)
{
    if (L.kind != BPMAX_K_OUTER || L.half_items <= 0) return run_launch(ctx, st, L, dP, dI, arena, dcodes, dw, false);
    // items of pairs with FP16 tables first, on FP16 operands; the rest on FP32 ones
    Launch h = L, f = L;
    h.count = L.half_items;
    h.updates = L.updates * L.half_items / L.count;
    f.first = L.first + L.half_items;
    f.count = L.count - L.half_items;
    f.updates = L.updates - h.updates;
    const int rc = run_launch(ctx, st, h, dP, dI, arena, dcodes, dw, true);
    return rc != BPMAX_OK ? rc : run_launch(ctx, st, f, dP, dI, arena, dcodes, dw, false);
}

// Dependency events of the pipelined schedule: fill, S2 done, end, then per T group one
// per (d1, storage diagonal) of R0 work and one per (d1, logical diagonal) of closure work.
static int ensure_sync_events(bpmax_ctx *ctx, const Schedule &S, std::vector<size_t> &out_base,
                              std::vector<size_t> &in_base)
{
    size_t n = 3; // fill, S2, end
    out_base.assign(S.groups.size(), 0);
    in_base.assign(S.groups.size(), 0);
    for (size_t g = 0; g < S.groups.size(); ++g) {
        const GroupInfo &G = S.groups[g];
        out_base[g] = n;
        n += (size_t)G.maxM * G.maxnb;
        in_base[g] = n;
        n += (size_t)G.maxM * G.maxnbl;
    }
    while (ctx->sync_ev.size() < n) {
        hipEvent_t e;
        HIPCHK(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
        ctx->sync_ev.push_back(e);
    }
    return BPMAX_OK;
}

// Last storage diagonal whose R0 a closure launch needs: the tiles it writes lie on logical
// diagonal L.diag (and L.diag+1 when it carries cross-tile steps ahead); logical diagonal
// DL lies in storage diagonals floor(DL/rs) .. ceil(DL/rs).
static int closure_r0_diag(const GroupInfo &G, const Launch &L)
{
    return std::min((L.diag + (L.ahead ? 1 : 0) + G.rs - 1) / G.rs, G.maxnb - 1);
}

// Pipelined schedule on two streams: sO runs the fill, S1 and the R0 launches, sI runs S2
// and every inner-closure launch.  R0 of (d1, storage diagonal Ds) reads tiles of the
// triangles on diagonals < d1 up to storage diagonal Ds, i.e. logical diagonal
// rs*Ds + rs-1 -- a whole R0 launch waits for the closure of d1-1 to reach that logical
// diagonal (for d1 = 0: for S2).  The closure of (d1, logical diagonal DL) needs R0 done on
// every storage tile that holds a logical tile of diagonal DL, i.e. storage diagonal
// ceil(DL/rs).  The closure of d1 thus overlaps the R0 launches of d1 (later storage
// diagonals) and of d1+1 (earlier ones).
// With the R0 of d1 split (build_schedule's r0_ahead), its bulk launches (k1 terms whose
// operands lie on diagonals <= d1-2) wait for the closure of d1-2 instead, and run beside
// the closure of d1-1; its edge launches (k1 = i1 and k1 = j1-1) wait for the closure of
// d1-1 like whole R0 launches.  The closure of d1 waits for the edge launches, which follow
// the bulk ones of d1 on sO.
static int enqueue_pipe(bpmax_ctx *ctx, int npairs, const Schedule &S, const PairDesc *dP, const WorkItem *dI,
                        float *arena, long long fl, const int *dcodes, const float *dw)
{
    hipStream_t sO = ctx->stream, sI = ctx->stream2;
    std::vector<size_t> ob, ib;
    int rc;
    if ((rc = ensure_sync_events(ctx, S, ob, ib)) != BPMAX_OK) return rc;
    hipEvent_t ev_fill = ctx->sync_ev[0], ev_s2 = ctx->sync_ev[1], ev_end = ctx->sync_ev[2];
    auto ev_out = [&](int g, int d1, int D) { return ctx->sync_ev[ob[g] + (size_t)d1 * S.groups[g].maxnb + D]; };
    auto ev_in = [&](int g, int d1, int D) { return ctx->sync_ev[ib[g] + (size_t)d1 * S.groups[g].maxnbl + D]; };
    auto in_need = [&](const Launch &L) {
        const GroupInfo &G = S.groups[L.group];
        return std::min(G.rs * L.diag + G.rs - 1, G.maxnbl - 1);
    };
    auto lag = [](const Launch &L) { return L.bulk ? 2 : 1; };
    auto on_r0_stream = [](const Launch &L) { return L.kind == BPMAX_K_OUTER; };
    // logical diagonals some R0 launch waits for: only those record an event (a record
    // between two closure launches costs a few microseconds on the critical path)
    std::vector<std::vector<char>> waited(S.groups.size());
    for (size_t g = 0; g < S.groups.size(); ++g) waited[g].assign(S.groups[g].maxnbl, 0);
    for (const Launch &L : S.launches)
        if (L.d1 >= 0 && on_r0_stream(L) && L.d1 - lag(L) >= 0) waited[L.group][in_need(L)] = 1;

    if ((rc = timed(ctx, sO, BPMAX_K_FILL, 0, [&] { return launch_fill(sO, arena, fl, -INFINITY); })) != BPMAX_OK)
        return rc;
    HIPCHK(ctx, hipEventRecord(ev_fill, sO));
    HIPCHK(ctx, hipStreamWaitEvent(sI, ev_fill, 0));
    // S1 is read by the R0 launches only: built on their stream, beside the S2 chain; so is
    // the FP16 mirrors' -inf fill, which the first closure (after the first R0 launch) needs
    if ((rc = timed(ctx, sO, BPMAX_K_S1, 0, [&] { return launch_s1(sO, dP, npairs, arena, dcodes, dw); })) !=
        BPMAX_OK)
        return rc;
    if (ctx->fill_half > 0 &&
        (rc = timed(ctx, sO, BPMAX_K_FILL, 0,
                    [&] { return launch_fill2(sO, arena + fl, 0, 0.f, ctx->fill_half, 0xFC00FC00u); })) != BPMAX_OK)
        return rc;
    for (const Launch &L : S.launches)
        if (L.d1 < 0 && (rc = run_launch(ctx, sI, L, dP, dI, arena, dcodes, dw, ctx->run_half)) != BPMAX_OK) return rc;
    // the FP16 cross-tile operands of the closure: shifted copies of the finished S2
    if (ctx->run_half && (rc = timed(ctx, sI, BPMAX_K_FILL, 0, [&] { return launch_s2_shift(sI, dP, npairs, arena); })) !=
                             BPMAX_OK)
        return rc;
    HIPCHK(ctx, hipEventRecord(ev_s2, sI));

    int o_g = -1, o_src = -2, o_wait = -1; // waits issued on sO for the closure of (group, o_src)
    int i_g = -1, i_d1 = -1, i_wait = -1;  // waits issued on sI for the R0 of (group, i_d1)
    for (const Launch &L : S.launches) {
        if (L.d1 < 0) continue;
        const GroupInfo &G = S.groups[L.group];
        if (on_r0_stream(L)) {
            const int src = L.d1 - lag(L);
            if (L.group != o_g || std::max(src, -1) != o_src) o_g = L.group, o_src = std::max(src, -1), o_wait = -1;
            if (src < 0) {
                if (o_wait < 0) HIPCHK(ctx, hipStreamWaitEvent(sO, ev_s2, 0));
                o_wait = G.maxnbl;
            } else {
                const int need = in_need(L);
                if (need > o_wait) {
                    HIPCHK(ctx, hipStreamWaitEvent(sO, ev_in(L.group, src, need), 0));
                    o_wait = need;
                }
            }
            if ((rc = run_launch(ctx, sO, L, dP, dI, arena, dcodes, dw, ctx->run_half)) != BPMAX_OK) return rc;
            HIPCHK(ctx, hipEventRecord(ev_out(L.group, L.d1, L.diag), sO));
        } else {
            if (L.group != i_g || L.d1 != i_d1) i_g = L.group, i_d1 = L.d1, i_wait = -1;
            const int need = G.out_end[closure_r0_diag(G, L)];
            if (need > i_wait) {
                HIPCHK(ctx, hipStreamWaitEvent(sI, ev_out(L.group, L.d1, need), 0));
                i_wait = need;
            }
            if ((rc = run_launch(ctx, sI, L, dP, dI, arena, dcodes, dw, ctx->run_half)) != BPMAX_OK) return rc;
            if (L.kind == BPMAX_K_INNER_INTRA && waited[L.group][L.diag])
                HIPCHK(ctx, hipEventRecord(ev_in(L.group, L.d1, L.diag), sI));
        }
    }
    HIPCHK(ctx, hipEventRecord(ev_end, sI));
    HIPCHK(ctx, hipStreamWaitEvent(sO, ev_end, 0));
    return BPMAX_OK;
}

// Capture the pipelined schedule into a graph (kept until the shapes or buffers change)
// and launch it on ctx->stream.
static int launch_pipe_graph(bpmax_ctx *ctx, const std::vector<long long> &key, int npairs, const Schedule &S,
                             const PairDesc *dP, const WorkItem *dI, float *arena, long long fl, const int *dcodes,
                             const float *dw)
{
    std::vector<long long> gk = key;
    for (const DevBuf *b : {&ctx->arena, &ctx->meta, &ctx->codes}) {
        gk.push_back((long long)(uintptr_t)b->p);
        gk.push_back((long long)b->cap);
    }
    gk.push_back((long long)(uintptr_t)arena);
    gk.push_back(ctx->eff_intra_wide); // kernel choices that vary between runs of one shape
    gk.push_back(ctx->run_half);
    gk.push_back(ctx->fill_half);
    gk.push_back(fl);
    if (!ctx->graph || gk != ctx->graph_key) {
        if (ctx->graph) HIPCHK(ctx, hipGraphExecDestroy(ctx->graph));
        ctx->graph = nullptr;
        ctx->graph_key.clear();
        std::vector<size_t> ob, ib;
        int rc;
        if ((rc = ensure_sync_events(ctx, S, ob, ib)) != BPMAX_OK) return rc;
        HIPCHK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        rc = enqueue_pipe(ctx, npairs, S, dP, dI, arena, fl, dcodes, dw);
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(ctx->stream, &g);
        if (rc != BPMAX_OK || ce != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            if (rc == BPMAX_OK) ctx->err = std::string("graph capture: ") + hipGetErrorString(ce);
            return BPMAX_EDEVICE;
        }
        const hipError_t ie = hipGraphInstantiate(&ctx->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess) {
            ctx->graph = nullptr;
            ctx->err = std::string("graph instantiate: ") + hipGetErrorString(ie);
            return BPMAX_EDEVICE;
        }
        ctx->graph_key = gk;
    }
    HIPCHK(ctx, hipGraphLaunch(ctx->graph, ctx->stream));
    return BPMAX_OK;
}

// ------------------------------------------------------------------ schedule check
// Host only.  Builds the launch list the pipelined (or one-stream) runner executes for the
// given pair shapes and settings and checks, under the runner's ordering rules, that every
// work item starts after all the work it reads has finished:
//   R0 of F(i1,j1) storage tile (I,J): the closure of F(i1,k1) and F(k1+1,j1), k1 in
//     [i1,j1), over the logical tiles of storage rows/columns I..J;
//   in-tile closure of a logical tile: every R0 item of its storage tile, and the closure
//     of the tiles left of it in its row and below it in its column; cross-tile before
//     in-tile;
//   cross-tile step s of tile (I,J), D = J-I: R0 of its storage tile and the closure of the
//     one tile it reads, (I,I+s) for s < D, else (I+1+s-D,J);
//   every cross-tile step of every tile run once (by a cross-tile item, or, with
//     closure_ahead, steps D-1 and D by the tile's in-tile item);
//   R0 of the first outer diagonal: all of S2 (of its tile-size group).
// The pipelined runner's rules, as enqueue_pipe applies them: stream order on the R0
// stream and on the closure stream; an R0 range of d1 waits for the closure of d1-1 up to
// logical diagonal rs*Ds+rs-1 of its last storage diagonal Ds (d1 = 0: for S2), and a
// closure level DL of d1 waits for the R0 range of d1 holding storage diagonal
// ceil(DL/rs).  skip >= 0 drops that launch's wait (a planted error for the check's test).
extern "C" int bpmax_schedule_check(int npairs, const int *Ms, const int *Ns, const bpmax_tuning *t, int skip,
                                    long long *checked, char *msg, int msglen)
{
    auto fail = [&](const std::string &m) {
        if (msg && msglen > 0) snprintf(msg, (size_t)msglen, "%s", m.c_str());
        return BPMAX_EINVAL;
    };
    if (npairs < 1 || !Ms || !Ns || !tuning_valid(t)) return fail("bad arguments");
    std::vector<PairDesc> P(npairs);
    for (int p = 0; p < npairs; ++p) {
        if (Ms[p] < 1 || Ns[p] < 1) return fail("bad shape");
        PairDesc d{};
        d.M = Ms[p];
        d.N = Ns[p];
        d.T = pick_tile(d.N, t->tile);
        d.nb = (d.N + d.T - 1) / d.T;
        d.nbl = d.nb * (d.T / inner_tile(d.T));
        d.nblr = (d.N + inner_tile(d.T) - 1) / inner_tile(d.T);
        P[p] = d;
    }
    const bool one = t->schedule == BPMAX_SCHED_LAUNCH;
    Schedule S;
    build_schedule(P, S, one ? 1 : t->outer_split, t->outer_items, t->outer_kgroup, t->closure_ahead != 0,
                   !one && t->schedule == BPMAX_SCHED_PIPE);
    const int n = (int)S.launches.size();
    // R0 launches on the R0 stream (sO); everything else runs on the closure stream
    auto on_o = [&](int i) { return !one && S.launches[i].kind == BPMAX_K_OUTER; };
    // the launch each launch waits for (-1: none), as enqueue_pipe derives it
    std::vector<int> wait(n, -1);
    if (!one) {
        std::map<std::tuple<int, int, int>, int> intra_of, outer_of; // (group, d1, diag) -> launch
        std::vector<int> s2_end(S.groups.size(), -1);
        for (int i = 0; i < n; ++i) {
            const Launch &L = S.launches[i];
            if (L.d1 < 0) s2_end[L.group] = i;
            else if (on_o(i)) outer_of[{L.group, L.d1, L.diag}] = i;
            else if (L.kind == BPMAX_K_INNER_INTRA) intra_of[{L.group, L.d1, L.diag}] = i;
        }
        int o_g = -1, o_src = -2, o_wait = -1, i_g = -1, i_d1 = -1, i_wait = -1;
        for (int i = 0; i < n; ++i) {
            const Launch &L = S.launches[i];
            if (L.d1 < 0) continue;
            const GroupInfo &G = S.groups[L.group];
            if (on_o(i)) {
                const int src = L.d1 - (L.bulk ? 2 : 1);
                if (L.group != o_g || std::max(src, -1) != o_src) o_g = L.group, o_src = std::max(src, -1), o_wait = -1;
                if (src < 0) {
                    if (o_wait < 0) wait[i] = s2_end[L.group];
                    o_wait = G.maxnbl;
                } else {
                    const int need = std::min(G.rs * L.diag + G.rs - 1, G.maxnbl - 1);
                    if (need > o_wait) {
                        wait[i] = intra_of.at({L.group, src, need});
                        o_wait = need;
                    }
                }
            } else if (L.kind != BPMAX_K_OUTER) {
                if (L.group != i_g || L.d1 != i_d1) i_g = L.group, i_d1 = L.d1, i_wait = -1;
                const int need = G.out_end[closure_r0_diag(G, L)];
                if (need > i_wait) {
                    wait[i] = outer_of.at({L.group, L.d1, need});
                    i_wait = need;
                }
            }
        }
        if (skip >= 0 && skip < n) wait[skip] = -1;
    }
    // happens-before frontier: every R0 launch <= hbO[i] and every closure launch <=
    // hbI[i] has finished when launch i starts (one stream: everything before i)
    std::vector<int> hbO(n, -1), hbI(n, -1);
    int prevO = -1, prevI = -1;
    for (int i = 0; i < n; ++i) {
        const bool outer = on_o(i);
        int o = -1, in = -1;
        auto absorb = [&](int j) {
            if (j < 0) return;
            o = std::max(o, hbO[j]);
            in = std::max(in, hbI[j]);
            if (on_o(j)) o = std::max(o, j);
            else in = std::max(in, j);
        };
        absorb(outer ? prevO : prevI);
        if (wait[i] >= i) return fail("a launch waits for a later one");
        absorb(wait[i]);
        hbO[i] = o;
        hbI[i] = in;
        (outer ? prevO : prevI) = i;
    }
    auto before = [&](int a, int b) { return on_o(a) ? a <= hbO[b] : a <= hbI[b]; };
    auto r0_before = [&](const std::pair<int, int> &r, int b) {
        return (r.first < 0 || before(r.first, b)) && (r.second < 0 || before(r.second, b));
    };
    using Key = std::tuple<int, int, int, int, int>;
    std::map<Key, int> closed, ext_at;
    // last R0 launch writing each storage tile, per stream (R0 stream, closure stream): the
    // earlier ones on a stream are ordered before it
    std::map<Key, std::pair<int, int>> r0_last;
    std::vector<int> s2_last(S.groups.size(), -1);
    for (int i = 0; i < n; ++i) {
        const Launch &L = S.launches[i];
        if (L.d1 < 0) s2_last[L.group] = std::max(s2_last[L.group], i);
        for (int k = L.first; k < L.first + L.count; ++k) {
            const WorkItem &w = S.items[k];
            const Key key{w.pair, w.i1, w.j1, w.I, w.J};
            if (L.kind == BPMAX_K_OUTER) {
                auto &r = r0_last.emplace(key, std::make_pair(-1, -1)).first->second;
                (on_o(i) ? r.first : r.second) = i;
            }
            else if (L.d1 >= 0 && L.kind == BPMAX_K_INNER_INTRA && wi_type(w.flags) != TASK_EXT) {
                if (closed.count(key)) return fail("a tile is closed twice");
                closed[key] = i;
            } else if (L.d1 >= 0) ext_at[key] = std::max(ext_at.count(key) ? ext_at[key] : -1, i);
        }
    }
    for (int p = 0; p < npairs; ++p)
        for (int i1 = 0; i1 < P[p].M; ++i1)
            for (int j1 = i1; j1 < P[p].M; ++j1)
                for (int X = 0; X < P[p].nblr; ++X)
                    for (int Y = X; Y < P[p].nblr; ++Y)
                        if (!closed.count(Key{p, i1, j1, X, Y})) return fail("a tile is never closed");
    // cross-tile step coverage of F tiles: each of the 2D steps exactly once
    std::map<Key, std::vector<char>> steps_run;
    auto run_step = [&](const Key &k, int D, int st) -> bool {
        auto &v = steps_run[k];
        if (v.empty()) v.assign(2 * D, 0);
        if (st < 0 || st >= 2 * D || v[st]) return false;
        v[st] = 1;
        return true;
    };
    for (int i = 0; i < n; ++i) {
        const Launch &L = S.launches[i];
        if (L.d1 < 0 || L.kind == BPMAX_K_OUTER) continue;
        for (int k = L.first; k < L.first + L.count; ++k) {
            const WorkItem &w = S.items[k];
            const Key key{w.pair, w.i1, w.j1, w.I, w.J};
            const int D = w.J - w.I;
            bool ok = true;
            if (L.kind == BPMAX_K_INNER_EXT || wi_type(w.flags) == TASK_EXT)
                for (int st = w.step_begin; st < w.step_end; ++st) ok = ok && run_step(key, D, st);
            else if (L.ahead && D > 0)
                ok = run_step(key, D, D - 1) && run_step(key, D, D);
            if (!ok) return fail("a cross-tile step is run twice or out of range");
        }
    }
    for (int p = 0; p < npairs; ++p)
        for (int i1 = 0; i1 < P[p].M; ++i1)
            for (int j1 = i1; j1 < P[p].M; ++j1)
                for (int X = 0; X < P[p].nblr; ++X)
                    for (int Y = X + 1; Y < P[p].nblr; ++Y) {
                        const auto it = steps_run.find(Key{p, i1, j1, X, Y});
                        if (it == steps_run.end() || std::count(it->second.begin(), it->second.end(), 1) != 2 * (Y - X))
                            return fail("a cross-tile step is never run");
                    }
    long long nchk = 0;
    std::map<Key, std::vector<char>> r0_cov; // (k1, K) steps run per R0 output tile
    for (int i = 0; i < n; ++i) {
        const Launch &L = S.launches[i];
        for (int k = L.first; k < L.first + L.count; ++k) {
            const WorkItem &w = S.items[k];
            const PairDesc &p = P[w.pair];
            const int rs = p.nbl / p.nb;
            char where[160];
            snprintf(where, sizeof where, " (launch %d, pair %d, triangle (%d,%d), tile (%d,%d))", i, w.pair, w.i1, w.j1,
                     w.I, w.J);
            if (L.kind == BPMAX_K_OUTER) {
                if (w.j1 == w.i1 && !before(s2_last[L.group], i)) return fail(std::string("R0 before S2") + where);
                if ((w.flags & WI_EPILOGUE) && w.j1 - w.i1 >= 4) { // P1 reads F(i1+1, j1-1)
                    const Key src{w.pair, w.i1 + 1, w.j1 - 1, w.I * rs, std::min((w.J + 1) * rs, p.nblr) - 1};
                    if (!before(closed.at(src), i)) return fail(std::string("R0 epilogue before its P1 term") + where);
                }
                // the item's k1 values; closure goes by logical diagonal, and the closure order
                // is checked below, so the last-closed tile of the square read suffices
                const int kcnt = (w.flags >> WI_KCNT_SHIFT) & 255, nK = kcnt ? kcnt : w.J - w.I + 1;
                std::vector<int> k1s;
                const int Kb = w.I + ((w.flags >> WI_KOFF_SHIFT) & 255);
                auto &cov = r0_cov[Key{w.pair, w.i1, w.j1, w.I, w.J}];
                if (cov.empty()) cov.assign((size_t)(w.j1 - w.i1) * (w.J - w.I + 1), 0);
                for (int st = w.step_begin; st < w.step_end; ++st) {
                    const int k1 = outer_k1(w, st, nK), K = Kb + st % nK;
                    if (k1 < w.i1 || k1 >= w.j1 || K > w.J) return fail(std::string("R0 step outside the triangle") + where);
                    char &c = cov[(size_t)(k1 - w.i1) * (w.J - w.I + 1) + (K - w.I)];
                    if (c) return fail(std::string("an R0 step is run twice") + where);
                    c = 1;
                    if (k1s.empty() || k1s.back() != k1) k1s.push_back(k1);
                }
                for (int k1 : k1s)
                    for (int side = 0; side < 2; ++side) {
                        const int a = side ? k1 + 1 : w.i1, b = side ? w.j1 : k1;
                        const Key src{w.pair, a, b, w.I * rs, std::min((w.J + 1) * rs, p.nblr) - 1};
                        if (!before(closed.at(src), i)) return fail(std::string("R0 before the closure it reads") + where);
                        ++nchk;
                    }
            } else if (L.d1 >= 0 && (L.kind == BPMAX_K_INNER_EXT || wi_type(w.flags) == TASK_EXT)) {
                // cross-tile steps: step s reads X tile (I, I+s) (R1, s < D) or
                // (I+1+s-D, J) (R2); it writes the tile, after its R0
                const auto r0 = r0_last.find(Key{w.pair, w.i1, w.j1, w.I / rs, w.J / rs});
                if (r0 == r0_last.end() || !r0_before(r0->second, i)) return fail(std::string("closure before R0") + where);
                const int D = w.J - w.I;
                for (int st = w.step_begin; st < w.step_end; ++st) {
                    const Key src = st < D ? Key{w.pair, w.i1, w.j1, w.I, w.I + st}
                                           : Key{w.pair, w.i1, w.j1, w.I + 1 + st - D, w.J};
                    if (!before(closed.at(src), i)) return fail(std::string("cross-tile step before its operand") + where);
                    ++nchk;
                }
            } else if (L.d1 >= 0) {
                const auto r0 = r0_last.find(Key{w.pair, w.i1, w.j1, w.I / rs, w.J / rs});
                if (r0 == r0_last.end() || !r0_before(r0->second, i)) return fail(std::string("closure before R0") + where);
                for (int Y = w.I; Y < w.J; ++Y)
                    if (!before(closed.at(Key{w.pair, w.i1, w.j1, w.I, Y}), i))
                        return fail(std::string("closure before the tile left of it") + where);
                for (int X = w.I + 1; X <= w.J; ++X)
                    if (!before(closed.at(Key{w.pair, w.i1, w.j1, X, w.J}), i))
                        return fail(std::string("closure before the tile below it") + where);
                if (L.kind == BPMAX_K_INNER_INTRA) {
                    const auto e = ext_at.find(Key{w.pair, w.i1, w.j1, w.I, w.J});
                    if (e != ext_at.end() && !before(e->second, i))
                        return fail(std::string("in-tile before cross-tile") + where);
                }
                nchk += 1 + (w.J - w.I) * 2;
            }
        }
    }
    for (int p = 0; p < npairs; ++p)
        for (int i1 = 0; i1 < P[p].M; ++i1)
            for (int j1 = i1 + 1; j1 < P[p].M; ++j1)
                for (int I = 0; I < P[p].nb; ++I)
                    for (int J = I; J < P[p].nb; ++J) {
                        const auto c = r0_cov.find(Key{p, i1, j1, I, J});
                        if (c == r0_cov.end() || std::count(c->second.begin(), c->second.end(), 0))
                            return fail("an R0 step of a tile is never run");
                    }
    if (checked) *checked = nchk;
    return BPMAX_OK;
}

static int run_pairs(bpmax_ctx *ctx, int npairs, const int *const *seq1s, const int *Ms, const int *const *seq2s,
                     const int *Ns, const float *intra_in, const float *inter_in, bool sync);

// Lay out pairs in the arena, upload codes/descriptors/weights/items, run the schedule.
// On success the F tables of all pairs are resident in ctx->arena and ctx->pairs_host
// describes them (ctx->result_valid).
int bpx_run_pairs(bpmax_ctx *ctx, int npairs, const int *const *seq1s, const int *Ms, const int *const *seq2s,
                  const int *Ns, const float *intra_in, const float *inter_in, bool sync)
{
    ctx->result_valid = false;
    const int rc = run_pairs(ctx, npairs, seq1s, Ms, seq2s, Ns, intra_in, inter_in, sync);
    ctx->result_valid = rc == BPMAX_OK;
    return rc;
}

static int run_pairs(bpmax_ctx *ctx, int npairs, const int *const *seq1s, const int *Ms, const int *const *seq2s,
                     const int *Ns, const float *intra_in, const float *inter_in, bool sync)
{
    if (npairs <= 0) return BPMAX_EINVAL;
    float wt[50];
    float dflt_intra[25], dflt_inter[25];
    bpmax_engine_score_table(dflt_intra, dflt_inter);
    memcpy(wt, intra_in ? intra_in : dflt_intra, 25 * sizeof(float));
    memcpy(wt + 25, inter_in ? inter_in : (intra_in ? intra_in : dflt_inter), 25 * sizeof(float));

    // Launch-schedule settings: the tuned ones of this problem class, else the context's.
    Tuning eff{ctx->schedule,   ctx->outer_split,  ctx->outer_items,  ctx->graphs ? 1 : 0,
               ctx->intra_wide, ctx->outer_kgroup, ctx->closure_ahead, ctx->tile};
    if (ctx->schedule != BPMAX_SCHED_TASKS) {
        int mM = 0, mN = 0;
        for (int p = 0; p < npairs; ++p) {
            mM = std::max(mM, Ms[p]);
            mN = std::max(mN, Ns[p]);
        }
        auto tu = ctx->tuning.find(tuning_class(mM, mN, npairs));
        if (tu != ctx->tuning.end()) eff = tu->second;
    }

    std::vector<PairDesc> &P = ctx->pairs_host;
    P.assign(npairs, PairDesc{});
    long long fl = 0, il = 0;
    auto al = [](long long x) { return (x + 63) / 64 * 64; };
    for (int p = 0; p < npairs; ++p) {
        const int M = Ms[p], N = Ns[p];
        if (M < 1 || N < 1 || !seq1s[p] || !seq2s[p]) {
            ctx->err = "pair " + std::to_string(p) + ": lengths must be >= 1 and sequences non-null";
            return BPMAX_EINVAL;
        }
        for (int i = 0; i < M; ++i)
            if (seq1s[p][i] < 0 || seq1s[p][i] > 4) {
                ctx->err = "seq1 codes must be in 0..4";
                return BPMAX_EINVAL;
            }
        for (int i = 0; i < N; ++i)
            if (seq2s[p][i] < 0 || seq2s[p][i] > 4) {
                ctx->err = "seq2 codes must be in 0..4";
                return BPMAX_EINVAL;
            }
        std::string why;
        if (!weights_exact(wt, wt + 25, M, N, &why)) {
            ctx->err = why;
            return BPMAX_EWEIGHTS;
        }
        PairDesc d{};
        d.M = M;
        d.N = N;
        d.T = pick_tile(N, eff.tile);
        d.nb = (N + d.T - 1) / d.T;
        d.nbl = d.nb * (d.T / inner_tile(d.T));
        d.nblr = (d.N + inner_tile(d.T) - 1) / inner_tile(d.T);
        d.tstride = tri_ll(d.nb) * d.T * d.T;
        d.F = fl;
        fl += al(tri_ll(M) * d.tstride);
        d.S2 = fl;
        fl += al(d.tstride);
        d.S1 = fl;
        fl += al((long long)M * M);
        d.a = il;
        il += al(M);
        d.b = il;
        il += al((long long)d.nb * d.T);
        P[p] = d;
    }
    // FP16 mirrors of F after all FP32 tables, when every pair is on 64 x 64 tiles and exact in
    // FP16 (the R0 launches then read them; the no-allocation mode keeps its scratch for FP32)
    bool half = ctx->r0_half && !ctx->carve && eff.schedule != BPMAX_SCHED_TASKS && ctx->schedule != BPMAX_SCHED_TASKS;
    for (int p = 0; p < npairs && half; ++p)
        half = P[p].T == 64 && weights_half_exact(wt, wt + 25, P[p].M, P[p].N, seq1s[p], seq2s[p]);
    long long fh = 0;
    if (half)
        for (auto &d : P) {
            d.Fh = fl + fh;
            fh += al((tri_ll(d.M) * d.tstride + 1) / 2);
        }
    ctx->run_half = half;
    ctx->fill_half = fh; // the F mirrors start at -inf; the S2 copies below are written whole
    if (half)
        for (auto &d : P) {
            d.S2r = fl + fh;
            fh += al((d.tstride + 1) / 2);
            d.S2c = fl + fh;
            fh += al((d.tstride + 1) / 2);
        }
    int rc;
    char *codes_dev = nullptr, *meta_dev = nullptr;
    size_t meta_room = 0;
    auto al256 = [](size_t x) { return (x + 255) / 256 * 256; };
    if (ctx->carve) { // no-allocation mode: arena | codes | meta carved from the caller's region
        const size_t ab = al256((size_t)(fl + fh) * sizeof(float)), cb = al256((size_t)il * sizeof(int));
        if (ab + cb + 256 > ctx->carve_bytes) {
            ctx->err = "the pair needs " + std::to_string(ab + cb) + " bytes of device scratch for its tables, more "
                       "than the " + std::to_string(ctx->carve_bytes) + " available";
            return BPMAX_ENOMEM;
        }
        ctx->arena_ext = ctx->carve;
        ctx->arena_ext_bytes = ab;
        codes_dev = ctx->carve + ab;
        meta_dev = codes_dev + cb;
        meta_room = ctx->carve_bytes - ab - cb;
    } else {
        if (ctx->arena_ext) {
            if ((size_t)(fl + fh) * sizeof(float) > ctx->arena_ext_bytes) {
                ctx->err = "pairs need " + std::to_string((size_t)(fl + fh) * sizeof(float)) +
                           " arena bytes, the slice holds " + std::to_string(ctx->arena_ext_bytes);
                return BPMAX_ENOMEM;
            }
        } else if ((rc = bpx_grow(ctx, ctx->arena, (size_t)(fl + fh) * sizeof(float))) != BPMAX_OK) {
            return rc;
        }
        if ((rc = bpx_grow(ctx, ctx->codes, (size_t)il * sizeof(int))) != BPMAX_OK) return rc;
        codes_dev = (char *)ctx->codes.p;
    }

    std::vector<int> &codes = ctx->codes_host;
    codes.assign((size_t)il, 4);
    for (int p = 0; p < npairs; ++p) {
        memcpy(codes.data() + P[p].a, seq1s[p], sizeof(int) * Ms[p]);
        for (int x = 0; x < Ns[p]; ++x) codes[P[p].b + x] = seq2s[p][Ns[p] - 1 - x];
    }
    // dependency counters of the task scheduler
    long long nflags = 0;
    for (auto &d : P) {
        d.flag = nflags;
        nflags += pair_flag_ints(d.M, d.nbl);
    }
    std::vector<int> Ts;
    for (auto &d : P)
        if (std::find(Ts.begin(), Ts.end(), d.T) == Ts.end()) Ts.push_back(d.T);
    std::sort(Ts.begin(), Ts.end());

    // no-allocation mode: no graph capture, no task-scheduler counters
    if (ctx->carve) {
        if (eff.schedule == BPMAX_SCHED_TASKS) eff.schedule = BPMAX_SCHED_PIPE;
        eff.graphs = 0;
    }
    const int sched = eff.schedule;
    ctx->eff_intra_wide = eff.intra_wide;

    // The schedule is rebuilt only when the shapes or the schedule settings change.
    std::vector<long long> key = {sched,           ctx->task_chunk,   eff.outer_split, eff.outer_items,
                                  eff.outer_kgroup, eff.closure_ahead, eff.tile,        npairs,
                                  ctx->r0_ahead};
    // (intra_wide changes kernel choices, not the schedule: the graph key carries it)
    for (auto &d : P) {
        key.push_back(d.M);
        key.push_back(d.N);
    }
    Schedule &S = ctx->sched;
    if (key != ctx->sched_key) {
        ctx->sched_key.clear();
        ctx->items_on_device = nullptr;
        if (sched == BPMAX_SCHED_TASKS) {
            S.items.clear();
            S.launches.clear();
            S.groups.clear();
            ctx->flags_host.assign((size_t)nflags, 0);
            std::vector<WorkItem> tl;
            for (int T : Ts) {
                build_tasks(P, T, ctx->task_chunk, tl, ctx->flags_host);
                S.launches.push_back(
                    Launch{BPMAX_K_TASKS, T, false, (int)S.items.size(), (int)tl.size(), 0, 0, -1, 0});
                S.items.insert(S.items.end(), tl.begin(), tl.end());
            }
        } else {
            // one stream: a single R0 launch per outer diagonal (nothing to overlap with)
            build_schedule(P, S, sched == BPMAX_SCHED_LAUNCH ? 1 : eff.outer_split, eff.outer_items,
                           eff.outer_kgroup, eff.closure_ahead != 0, sched == BPMAX_SCHED_PIPE && ctx->r0_ahead);
        }
        ctx->sched_key = key;
    }

    // meta buffer: weights | pair descs | items (items uploaded once per schedule)
    const size_t wbytes = 256;
    const size_t pbytes = (P.size() * sizeof(PairDesc) + 255) / 256 * 256;
    const size_t ibytes = S.items.size() * sizeof(WorkItem);
    bool send_items = true;
    if (ctx->carve) {
        if (wbytes + pbytes + ibytes > meta_room) {
            ctx->err = "the schedule needs " + std::to_string(wbytes + pbytes + ibytes) + " bytes of device scratch "
                       "beyond the tables, " + std::to_string(meta_room) + " are left";
            return BPMAX_ENOMEM;
        }
        ctx->items_on_device = nullptr; // the shared scratch keeps nothing between calls
    } else {
        const size_t meta_cap = ctx->meta.cap;
        if ((rc = bpx_grow(ctx, ctx->meta, wbytes + pbytes + ibytes)) != BPMAX_OK) return rc;
        if (ctx->meta.cap != meta_cap) ctx->items_on_device = nullptr;
        meta_dev = (char *)ctx->meta.p;
        send_items = ctx->items_on_device != ctx->meta.p;
    }
    ctx->meta_dev = meta_dev;
    ctx->staging.resize(wbytes + pbytes + (send_items ? ibytes : 0));
    memcpy(ctx->staging.data(), wt, sizeof wt);
    memcpy(ctx->staging.data() + wbytes, P.data(), P.size() * sizeof(PairDesc));
    if (send_items) memcpy(ctx->staging.data() + wbytes + pbytes, S.items.data(), ibytes);
    char *meta = meta_dev;
    const float *dw = (const float *)meta;
    const PairDesc *dP = (const PairDesc *)(meta + wbytes);
    const WorkItem *dI = (const WorkItem *)(meta + wbytes + pbytes);
    float *arena = bpx_arena(ctx);
    const int *dcodes = (const int *)codes_dev;

    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipMemcpyAsync(meta_dev, ctx->staging.data(), ctx->staging.size(), hipMemcpyHostToDevice,
                               ctx->stream));
    if (!ctx->carve) ctx->items_on_device = ctx->meta.p;
    HIPCHK(ctx, hipMemcpyAsync(codes_dev, codes.data(), codes.size() * sizeof(int), hipMemcpyHostToDevice,
                               ctx->stream));
    hipStream_t st = ctx->stream;

    if (sched == BPMAX_SCHED_TASKS) {
        if ((rc = timed(ctx, st, BPMAX_K_FILL, 0, [&] { return launch_fill(st, arena, fl, -INFINITY); })) !=
            BPMAX_OK)
            return rc;
        // [0..7] per T group: queue heads; [8] error word; then the counters
        const size_t hdr = 64;
        if ((rc = bpx_grow(ctx, ctx->flags, (hdr + ctx->flags_host.size()) * sizeof(int))) != BPMAX_OK) return rc;
        ctx->flags_staging.assign(hdr, 0);
        ctx->flags_staging.insert(ctx->flags_staging.end(), ctx->flags_host.begin(), ctx->flags_host.end());
        HIPCHK(ctx, hipMemcpyAsync(ctx->flags.p, ctx->flags_staging.data(), ctx->flags_staging.size() * sizeof(int),
                                   hipMemcpyHostToDevice, ctx->stream));
        int *dflags = (int *)ctx->flags.p;
        double upd = 0;
        for (auto &d : P) upd += bpmax_cell_updates(d.M, d.N);
        for (size_t g = 0; g < S.launches.size(); ++g) {
            const Launch &L = S.launches[g];
            rc = timed(ctx, st, BPMAX_K_TASKS, g == 0 ? upd : 0, [&] {
                return launch_tasks(st, L.T, dP, dI + L.first, L.count, dflags + g, dflags + hdr, dflags + 8, arena,
                                    dcodes, dw);
            });
            if (rc != BPMAX_OK) return rc;
        }
        int ew[6] = {0, 0, 0, 0, 0, 0}; // error word, pad, 64-bit poll count, tasks run, pad
        HIPCHK(ctx, hipMemcpyAsync(ew, dflags + 8, sizeof ew, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        const int err = ew[0];
        unsigned long long polls = 0;
        memcpy(&polls, ew + 2, sizeof polls);
        ctx->task_polls = polls;
        ctx->task_count = ew[4];
        if (err) {
            ctx->err = "task scheduler: a dependency wait timed out (kernel abandoned its tasks)";
            return BPMAX_EDEVICE;
        }
        return bpx_collect_timing(ctx);
    }

    if (sched == BPMAX_SCHED_PIPE) {
        if (eff.graphs && !ctx->timing && !ctx->carve)
            rc = launch_pipe_graph(ctx, key, npairs, S, dP, dI, arena, fl, dcodes, dw);
        else
            rc = enqueue_pipe(ctx, npairs, S, dP, dI, arena, fl, dcodes, dw);
        if (rc != BPMAX_OK) return rc;
    } else {
        if ((rc = timed(ctx, st, BPMAX_K_FILL, 0,
                        [&] { return launch_fill2(st, arena, fl, -INFINITY, ctx->fill_half, 0xFC00FC00u); })) !=
            BPMAX_OK)
            return rc;
        if ((rc = timed(ctx, st, BPMAX_K_S1, 0, [&] { return launch_s1(st, dP, npairs, arena, dcodes, dw); })) !=
            BPMAX_OK)
            return rc;
        bool shifted = !ctx->run_half;
        for (const Launch &L : S.launches) {
            if (!shifted && L.d1 >= 0) { // S2 is done: its shifted FP16 copies for the closure
                if ((rc = timed(ctx, st, BPMAX_K_FILL, 0, [&] { return launch_s2_shift(st, dP, npairs, arena); })) !=
                    BPMAX_OK)
                    return rc;
                shifted = true;
            }
            if ((rc = run_launch(ctx, st, L, dP, dI, arena, dcodes, dw, ctx->run_half)) != BPMAX_OK) return rc;
        }
    }
    // the host vectors (codes, staging) are context members, alive until the next run
    if (!sync) return BPMAX_OK;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return bpx_collect_timing(ctx);
}

int bpx_gather_scores(bpmax_ctx *ctx, float *scores, bool sync)
{
    const int n = (int)ctx->pairs_host.size();
    int rc;
    float *dscores = ctx->carve_scores;
    if (!dscores) {
        if ((rc = bpx_grow(ctx, ctx->scores, n * sizeof(float))) != BPMAX_OK) return rc;
        dscores = (float *)ctx->scores.p;
    }
    const size_t wbytes = 256;
    const PairDesc *dP = (const PairDesc *)(ctx->meta_dev + wbytes);
    HIPCHK(ctx, launch_scores(ctx->stream, dP, n, bpx_arena(ctx), dscores));
    HIPCHK(ctx, hipMemcpyAsync(scores, dscores, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    if (sync) HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return BPMAX_OK;
}

float *bpx_arena(bpmax_ctx *ctx) { return ctx->arena_ext ? (float *)ctx->arena_ext : (float *)ctx->arena.p; }

// Arena bytes the pairs need (the layout of bpx_run_pairs).
size_t bpx_arena_bytes(int npairs, const int *Ms, const int *Ns)
{
    auto al = [](long long x) { return (x + 63) / 64 * 64; };
    long long fl = 0;
    for (int p = 0; p < npairs; ++p) {
        const int T = choose_tile(Ns[p]), nb = (Ns[p] + T - 1) / T;
        const long long ts = tri_ll(nb) * T * T;
        fl += al(tri_ll(Ms[p]) * ts) + al(ts) + al((long long)Ms[p] * Ms[p]);
    }
    return (size_t)fl * sizeof(float);
}

extern "C" int bpmax_engine_solve(bpmax_ctx *ctx, const int *seq1, int M, const int *seq2, int N,
                                  const float *intra, const float *inter, float *F_out, bpmax_layout layout,
                                  float *score_out)
{
    if (!ctx) return BPMAX_EINVAL;
    if (layout != BPMAX_LAYOUT_NONE && !F_out) {
        ctx->err = "F_out is NULL for a layout that writes F";
        return BPMAX_EINVAL;
    }
    if (layout != BPMAX_LAYOUT_NONE && layout != BPMAX_LAYOUT_PACKED_TRI_TRI && layout != BPMAX_LAYOUT_DENSE_4D) {
        ctx->err = "unknown layout";
        return BPMAX_EINVAL;
    }
    int rc = bpx_run_pairs(ctx, 1, &seq1, &M, &seq2, &N, intra, inter, true);
    if (rc != BPMAX_OK) return rc;
    float score = 0.f;
    if ((rc = bpx_gather_scores(ctx, &score, true)) != BPMAX_OK) return rc;
    if (score_out) *score_out = score;
    if (layout == BPMAX_LAYOUT_NONE) return BPMAX_OK;
    return bpx_copy_out(ctx, bpx_arena(ctx), ctx->pairs_host[0], layout, F_out);
}

extern "C" int bpmax_engine_fetch(bpmax_ctx *ctx, int pair, bpmax_layout layout, float *F_out)
{
    if (!ctx) return BPMAX_EINVAL;
    if (!F_out || (layout != BPMAX_LAYOUT_PACKED_TRI_TRI && layout != BPMAX_LAYOUT_DENSE_4D)) {
        ctx->err = "fetch needs F_out and a packed or dense layout";
        return BPMAX_EINVAL;
    }
    if (!ctx->result_valid || ctx->arena_ext || pair < 0 || pair >= (int)ctx->pairs_host.size()) {
        ctx->err = "no resident result for pair " + std::to_string(pair);
        return BPMAX_EINVAL;
    }
    return bpx_copy_out(ctx, bpx_arena(ctx), ctx->pairs_host[pair], layout, F_out);
}

int bpx_copy_out(bpmax_ctx *ctx, const float *arena, const PairDesc &P, bpmax_layout layout, float *F_out)
{
    int rc;
    const size_t bytes = bpmax_layout_bytes(P.M, P.N, layout);
    if ((rc = bpx_grow(ctx, ctx->out, bytes)) != BPMAX_OK) return rc;
    float *dout = (float *)ctx->out.p;
    if (layout == BPMAX_LAYOUT_PACKED_TRI_TRI) {
        rc = timed(ctx, ctx->stream, BPMAX_K_CONVERT, 0, [&] { return launch_to_packed(ctx->stream, arena, P, dout); });
    } else {
        rc = timed(ctx, ctx->stream, BPMAX_K_CONVERT, 0, [&] {
            hipError_t e = launch_fill(ctx->stream, dout, (long long)(bytes / sizeof(float)), 0.0f);
            if (e != hipSuccess) return e;
            return launch_to_dense(ctx->stream, arena, P, dout);
        });
    }
    if (rc != BPMAX_OK) return rc;
    if ((rc = bpx_d2h(ctx, F_out, dout, bytes)) != BPMAX_OK) return rc;
    return bpx_collect_timing(ctx);
}

// memcpy on up to nth threads (large host copies out of pinned staging)
static void par_memcpy(void *dst, const void *src, size_t n, unsigned nth)
{
    if (n < (size_t(8) << 20) || nth <= 1) {
        memcpy(dst, src, n);
        return;
    }
    const size_t part = (n + nth - 1) / nth;
    std::vector<std::thread> th;
    for (unsigned t = 1; t < nth && t * part < n; ++t) {
        const size_t a = t * part, b = std::min(n, a + part);
        th.emplace_back([=] { memcpy((char *)dst + a, (const char *)src + a, b - a); });
    }
    memcpy(dst, src, std::min(n, part));
    for (auto &x : th) x.join();
}

// Device-to-host copy into pageable memory: in chunks through two pinned buffers, the DMA of
// chunk c+1 running while host threads copy chunk c out (a copy straight into pageable
// memory is staged by the runtime at a fraction of the link rate).  Synchronous.
int bpx_d2h(bpmax_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    constexpr size_t kChunk = size_t(64) << 20;
    if (bytes <= kChunk) {
        HIPCHK(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        return BPMAX_OK;
    }
    for (int b = 0; b < 2; ++b)
        if (!ctx->pin[b]) {
            if (hipHostMalloc(&ctx->pin[b], kChunk) != hipSuccess) {
                (void)hipGetLastError();
                ctx->pin[b] = nullptr;
                ctx->err = "hipHostMalloc of the copy staging failed";
                return BPMAX_ENOMEM;
            }
            HIPCHK(ctx, hipEventCreateWithFlags(&ctx->pin_ev[b], hipEventDisableTiming));
        }
    const size_t n = (bytes + kChunk - 1) / kChunk;
    auto len = [&](size_t c) { return std::min(kChunk, bytes - c * kChunk); };
    auto issue = [&](size_t c) -> hipError_t {
        const hipError_t e = hipMemcpyAsync(ctx->pin[c % 2], (const char *)src + c * kChunk, len(c),
                                            hipMemcpyDeviceToHost, ctx->stream);
        return e != hipSuccess ? e : hipEventRecord(ctx->pin_ev[c % 2], ctx->stream);
    };
    const unsigned nth = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
    HIPCHK(ctx, issue(0));
    for (size_t c = 0; c < n; ++c) {
        if (c + 1 < n) HIPCHK(ctx, issue(c + 1)); // the buffer it fills was emptied at c-1
        HIPCHK(ctx, hipEventSynchronize(ctx->pin_ev[c % 2]));
        par_memcpy((char *)dst + c * kChunk, ctx->pin[c % 2], len(c), nth);
    }
    return BPMAX_OK;
}

extern "C" int bpmax_engine_solve_batch(bpmax_ctx *ctx, int npairs, const int *const *seq1s, const int *Ms,
                                        const int *const *seq2s, const int *Ns, const float *intra,
                                        const float *inter, float *scores)
{
    if (!ctx || npairs <= 0 || !seq1s || !Ms || !seq2s || !Ns || !scores) return BPMAX_EINVAL;
    int rc = bpx_run_pairs(ctx, npairs, seq1s, Ms, seq2s, Ns, intra, inter, true);
    if (rc != BPMAX_OK) return rc;
    return bpx_gather_scores(ctx, scores, true);
}

// ------------------------------------------------------------------ isolated R0 kernel
extern "C" size_t bpmax_tri_tiled_floats(int n)
{
    if (n < 1) return 0;
    const int T = 64, nb = (n + T - 1) / T;
    return (size_t)tri_ll(nb) * T * T;
}

extern "C" void *bpmax_device_alloc(bpmax_ctx *ctx, size_t bytes)
{
    if (!ctx) return nullptr;
    void *p = nullptr;
    if (hipSetDevice(ctx->device) != hipSuccess || hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        ctx->err = "hipMalloc of " + std::to_string(bytes) + " bytes failed";
        return nullptr;
    }
    ctx->user_allocs.push_back(p);
    return p;
}

extern "C" int bpmax_device_free(bpmax_ctx *ctx, void *p)
{
    if (!ctx || !p) return BPMAX_EINVAL;
    auto it = std::find(ctx->user_allocs.begin(), ctx->user_allocs.end(), p);
    if (it == ctx->user_allocs.end()) return BPMAX_EINVAL;
    ctx->user_allocs.erase(it);
    HIPCHK(ctx, hipFree(p));
    return BPMAX_OK;
}

// Tile list of the isolated R0 (heavy tiles, long K ranges, first so the tail of the grid is
// short) into ctx->r0meta, then one timed launch.
static int r0_tri_run(bpmax_ctx *ctx, int n, int batch,
                      const std::function<hipError_t(int nb, const int2 *tiles, int ntiles)> &launch)
{
    if (batch > 65535) {
        ctx->err = "batch must be <= 65535";
        return BPMAX_EINVAL;
    }
    const int T = 64, nb = (n + T - 1) / T;
    std::vector<int2> tiles;
    for (int D = nb - 1; D >= 0; --D)
        for (int I = 0; I + D < nb; ++I) tiles.push_back(make_int2(I, I + D));
    int rc;
    if ((rc = bpx_grow(ctx, ctx->r0meta, tiles.size() * sizeof(int2))) != BPMAX_OK) return rc;
    HIPCHK(ctx, hipMemcpyAsync(ctx->r0meta.p, tiles.data(), tiles.size() * sizeof(int2), hipMemcpyHostToDevice,
                               ctx->stream));
    const double nn = n;
    const double upd = (double)batch * (nn + 2) * (nn + 1) * nn / 6.0;
    rc = timed(ctx, ctx->stream, BPMAX_K_R0TRI, upd,
               [&] { return launch(nb, (const int2 *)ctx->r0meta.p, (int)tiles.size()); });
    if (rc != BPMAX_OK) return rc;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return bpx_collect_timing(ctx);
}

extern "C" int bpmax_r0_tri_batch_device(bpmax_ctx *ctx, int n, int batch, const float *A_dev, const float *B_dev,
                                         float *C_dev)
{
    if (!ctx || n < 1 || batch < 1 || !A_dev || !B_dev || !C_dev) return BPMAX_EINVAL;
    return r0_tri_run(ctx, n, batch, [&](int nb, const int2 *tiles, int ntiles) {
        return launch_r0tri(ctx->stream, 64, nb, tiles, ntiles, batch, A_dev, B_dev, C_dev);
    });
}

extern "C" int bpmax_r0_tri_batch_device_half(bpmax_ctx *ctx, int n, int batch, const unsigned short *A_dev,
                                              const unsigned short *B_dev, float *C_dev)
{
    if (!ctx || n < 1 || batch < 1 || !A_dev || !B_dev || !C_dev) return BPMAX_EINVAL;
    return r0_tri_run(ctx, n, batch, [&](int nb, const int2 *tiles, int ntiles) {
        return launch_r0tri_h(ctx->stream, nb, tiles, ntiles, batch, A_dev, B_dev, C_dev);
    });
}

extern "C" int bpmax_tri_to_half(bpmax_ctx *ctx, int n, int batch, const float *src_dev, unsigned short *dst_dev,
                                 int *exact)
{
    if (!ctx || n < 1 || batch < 1 || !src_dev || !dst_dev || !exact) return BPMAX_EINVAL;
    const long long nb = (n + 63) / 64, elems = (long long)batch * tri_ll((int)nb) * 64 * 64;
    int rc;
    if ((rc = bpx_grow(ctx, ctx->r0half, 256)) != BPMAX_OK) return rc;
    int *flag = (int *)ctx->r0half.p, inexact = 1;
    HIPCHK(ctx, hipMemsetAsync(flag, 0, sizeof(int), ctx->stream));
    rc = timed(ctx, ctx->stream, BPMAX_K_CONVERT, 0,
               [&] { return launch_tri_to_half(ctx->stream, elems, src_dev, dst_dev, flag); });
    if (rc != BPMAX_OK) return rc;
    HIPCHK(ctx, hipMemcpyAsync(&inexact, flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    *exact = inexact == 0;
    return bpx_collect_timing(ctx);
}

extern "C" int bpmax_tri_upload(bpmax_ctx *ctx, int n, int batch, const float *host_dense, float *dev_tiled)
{
    if (!ctx || n < 1 || batch < 1 || !host_dense || !dev_tiled) return BPMAX_EINVAL;
    const int T = 64, nb = (n + T - 1) / T;
    const size_t bytes = (size_t)batch * n * n * sizeof(float);
    int rc;
    if ((rc = bpx_grow(ctx, ctx->out, bytes)) != BPMAX_OK) return rc;
    HIPCHK(ctx, hipMemcpyAsync(ctx->out.p, host_dense, bytes, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, launch_dense_to_tiled(ctx->stream, n, nb, T, batch, (const float *)ctx->out.p, dev_tiled));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return BPMAX_OK;
}

extern "C" int bpmax_tri_fill_random(bpmax_ctx *ctx, int n, int batch, unsigned long long seed, int lo, int hi,
                                     float *dev_tiled)
{
    if (!ctx || n < 1 || n >= (1 << 21) || batch < 1 || batch > 65535 || !dev_tiled || lo < 0 || hi < lo ||
        hi >= (1 << 22))
        return BPMAX_EINVAL;
    const int T = 64, nb = (n + T - 1) / T;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, launch_tri_random(ctx->stream, n, nb, T, batch, seed, lo, hi, dev_tiled));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return BPMAX_OK;
}

extern "C" int bpmax_tri_download(bpmax_ctx *ctx, int n, int batch, const float *dev_tiled, float *host_dense)
{
    if (!ctx || n < 1 || batch < 1 || !host_dense || !dev_tiled) return BPMAX_EINVAL;
    const int T = 64, nb = (n + T - 1) / T;
    const size_t bytes = (size_t)batch * n * n * sizeof(float);
    int rc;
    if ((rc = bpx_grow(ctx, ctx->out, bytes)) != BPMAX_OK) return rc;
    HIPCHK(ctx, launch_tiled_to_dense(ctx->stream, n, nb, T, batch, dev_tiled, (float *)ctx->out.p));
    HIPCHK(ctx, hipMemcpyAsync(host_dense, ctx->out.p, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return BPMAX_OK;
}

// ------------------------------------------------------------ no-allocation entry point
// Device scratch of bpmax_engine_solve_noalloc: one array per GPU, loaded with the library's
// code object and held until the process ends.  In host code `scratch` is only a stand-in
// for the symbol: its device address on each GPU comes from hipGetSymbolAddress (once per
// device, with that device current) and only that address is passed on.
static __device__ __align__(256) unsigned char scratch[BPMAX_NOALLOC_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static std::map<int, char *> g_scratch_ptr; // device -> address of `scratch` there

static int scratch_on_device(bpmax_ctx *ctx, char **out)
{
    std::lock_guard<std::mutex> g(g_scratch_mu);
    auto it = g_scratch_ptr.find(ctx->device);
    if (it != g_scratch_ptr.end()) {
        *out = it->second;
        return BPMAX_OK;
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        ctx->err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return BPMAX_EDEVICE;
    }
    g_scratch_ptr[ctx->device] = (char *)p;
    *out = (char *)p;
    return BPMAX_OK;
}

// Layout of the scratch for one call: the run's arena | codes | meta from the front (carved
// in bpx_run_pairs), then the scores (256 B) and the packed F (when asked for) at the end.
extern "C" int bpmax_engine_solve_noalloc(bpmax_ctx *ctx, const int *seq1, int M, const int *seq2, int N,
                                          const float *intra, const float *inter, float *F_out,
                                          bpmax_layout layout, float *score_out)
{
    if (!ctx) return BPMAX_EINVAL;
    if (layout != BPMAX_LAYOUT_NONE && layout != BPMAX_LAYOUT_PACKED_TRI_TRI) {
        ctx->err = "the no-allocation solve writes BPMAX_LAYOUT_NONE or BPMAX_LAYOUT_PACKED_TRI_TRI";
        return BPMAX_EINVAL;
    }
    if ((layout != BPMAX_LAYOUT_NONE && !F_out) || M < 1 || N < 1) {
        ctx->err = "lengths must be >= 1 and F_out set for a layout that writes F";
        return BPMAX_EINVAL;
    }
    const size_t kBytes = BPMAX_NOALLOC_SCRATCH_BYTES;
    const size_t out_bytes =
        layout == BPMAX_LAYOUT_NONE ? 0 : (bpmax_layout_bytes(M, N, layout) + 255) / 256 * 256;
    if (out_bytes + 512 > kBytes) {
        ctx->err = "the packed F table of " + std::to_string(M) + "x" + std::to_string(N) + " needs " +
                   std::to_string(out_bytes) + " bytes, more than the no-allocation scratch holds";
        return BPMAX_ENOMEM;
    }
    char *base = nullptr;
    int rc = scratch_on_device(ctx, &base);
    if (rc != BPMAX_OK) return rc;
    char *tail = base + kBytes - out_bytes - 256;
    ctx->carve = base;
    ctx->carve_bytes = kBytes - out_bytes - 256;
    ctx->carve_scores = (float *)tail;
    struct Restore { // the context's own buffers again after the call, whatever its outcome
        bpmax_ctx *c;
        ~Restore()
        {
            c->carve = nullptr;
            c->carve_bytes = 0;
            c->carve_scores = nullptr;
            c->arena_ext = nullptr;
            c->arena_ext_bytes = 0;
            c->result_valid = false; // the shared scratch is not a resident result
            c->items_on_device = nullptr;
            c->meta_dev = nullptr;
        }
    } restore{ctx};
    rc = bpx_run_pairs(ctx, 1, &seq1, &M, &seq2, &N, intra, inter, true);
    if (rc != BPMAX_OK) return rc;
    float score = 0.f;
    if ((rc = bpx_gather_scores(ctx, &score, true)) != BPMAX_OK) return rc;
    if (score_out) *score_out = score;
    if (layout == BPMAX_LAYOUT_NONE) return BPMAX_OK;
    float *dout = (float *)(tail + 256);
    HIPCHK(ctx, launch_to_packed(ctx->stream, bpx_arena(ctx), ctx->pairs_host[0], dout));
    HIPCHK(ctx, hipMemcpyAsync(F_out, dout, bpmax_layout_bytes(M, N, layout), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return BPMAX_OK;
}
