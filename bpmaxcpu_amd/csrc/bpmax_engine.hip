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
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "bpmax.h"
#include "bpmax_engine_impl.h"
#include "bpmax_internal.h"
#include "bpmax_launch.h"

using namespace bpx;

namespace bpx {

const char *kClassNames[BPMAX_K_NCLASSES] = {"fill", "s1", "outer", "inner_ext", "inner_intra", "convert", "r0tri"};

int choose_tile(int N) { return N <= 64 ? 32 : 64; }

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

} // namespace bpx

// ----------------------------------------------------------------------------- ctx
#define HIPCHK(ctx, expr)                                                                   \
    do {                                                                                    \
        hipError_t _e = (expr);                                                             \
        if (_e != hipSuccess) {                                                             \
            (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(_e);                 \
            return BPMAX_EDEVICE;                                                           \
        }                                                                                   \
    } while (0)

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
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return BPMAX_EDEVICE;
    }
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
    for (auto &e : ctx->ev_pool) (void)hipEventDestroy(e);
    for (void *p : ctx->user_allocs) (void)hipFree(p);
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

static int pick_chunk(long long total_steps, long long tiles)
{
    // aim for ~8 work items per CU (256 CUs) when the work allows it
    long long ch = total_steps / 2048;
    if (ch < 2) ch = 2;
    if (ch > 32) ch = 32;
    (void)tiles;
    return (int)ch;
}

// Build the launch list for a set of pairs (descs already laid out).
void build_schedule(const std::vector<PairDesc> &P, Schedule &S)
{
    S.items.clear();
    S.launches.clear();
    int maxM = 0, maxnb = 0;
    std::vector<int> Ts;
    for (auto &p : P) {
        maxM = std::max(maxM, p.M);
        maxnb = std::max(maxnb, p.nb);
        if (std::find(Ts.begin(), Ts.end(), p.T) == Ts.end()) Ts.push_back(p.T);
    }
    auto emit = [&](int kind, int T, bool ss, size_t first, double upd) {
        const int n = (int)(S.items.size() - first);
        if (n > 0) S.launches.push_back(Launch{kind, T, ss, (int)first, n, upd});
    };
    // S2 single-strand table, by inner block diagonal
    for (int D = 0; D < maxnb; ++D)
        for (int T : Ts) {
            if (D > 0) {
                size_t f = S.items.size();
                long long tot = 0;
                for (size_t p = 0; p < P.size(); ++p)
                    if (P[p].T == T && P[p].nb > D) tot += (long long)(P[p].nb - D) * D;
                const int ch = pick_chunk(tot, 0);
                for (size_t p = 0; p < P.size(); ++p) {
                    if (P[p].T != T || P[p].nb <= D) continue;
                    for (int I = 0; I + D < P[p].nb; ++I)
                        chunked(S.items, WorkItem{(int)p, 0, 0, I, I + D, 0, 0, 0}, D, ch, false);
                }
                emit(BPMAX_K_INNER_EXT, T, true, f, 0);
            }
            size_t f = S.items.size();
            for (size_t p = 0; p < P.size(); ++p) {
                if (P[p].T != T || P[p].nb <= D) continue;
                for (int I = 0; I + D < P[p].nb; ++I) S.items.push_back(WorkItem{(int)p, 0, 0, I, I + D, 0, 0, 0});
            }
            emit(BPMAX_K_INNER_INTRA, T, true, f, 0);
        }
    // F table, by outer diagonal
    for (int d1 = 0; d1 < maxM; ++d1) {
        for (int T : Ts) {
            size_t f = S.items.size();
            long long tot = 0;
            double upd = 0;
            for (size_t p = 0; p < P.size(); ++p) {
                if (P[p].T != T || P[p].M <= d1) continue;
                const long long nb = P[p].nb;
                long long per_tri = 0;
                for (long long D = 0; D < nb; ++D) per_tri += (nb - D) * (D + 1);
                tot += (long long)(P[p].M - d1) * d1 * per_tri;
                const double n = P[p].N;
                const double c3n = (n + 1) * n * (n - 1) / 6.0, c2n = (n + 1) * n / 2.0;
                upd += (double)(P[p].M - d1) * d1 * (c3n + 2.0 * c2n);
            }
            const int ch = pick_chunk(tot, 0);
            for (size_t p = 0; p < P.size(); ++p) {
                if (P[p].T != T || P[p].M <= d1) continue;
                for (int i1 = 0; i1 + d1 < P[p].M; ++i1)
                    for (int I = 0; I < P[p].nb; ++I)
                        for (int J = I; J < P[p].nb; ++J)
                            chunked(S.items, WorkItem{(int)p, i1, i1 + d1, I, J, 0, 0, 0}, d1 * (J - I + 1), ch, true);
            }
            emit(BPMAX_K_OUTER, T, false, f, upd);
        }
        for (int D = 0; D < maxnb; ++D)
            for (int T : Ts) {
                if (D > 0) {
                    size_t f = S.items.size();
                    long long tot = 0;
                    for (size_t p = 0; p < P.size(); ++p)
                        if (P[p].T == T && P[p].M > d1 && P[p].nb > D)
                            tot += (long long)(P[p].M - d1) * (P[p].nb - D) * 2 * D;
                    const int ch = pick_chunk(tot, 0);
                    for (size_t p = 0; p < P.size(); ++p) {
                        if (P[p].T != T || P[p].M <= d1 || P[p].nb <= D) continue;
                        for (int i1 = 0; i1 + d1 < P[p].M; ++i1)
                            for (int I = 0; I + D < P[p].nb; ++I)
                                chunked(S.items, WorkItem{(int)p, i1, i1 + d1, I, I + D, 0, 0, 0}, 2 * D, ch, false);
                    }
                    emit(BPMAX_K_INNER_EXT, T, false, f, 0);
                }
                size_t f = S.items.size();
                for (size_t p = 0; p < P.size(); ++p) {
                    if (P[p].T != T || P[p].M <= d1 || P[p].nb <= D) continue;
                    for (int i1 = 0; i1 + d1 < P[p].M; ++i1)
                        for (int I = 0; I + D < P[p].nb; ++I)
                            S.items.push_back(WorkItem{(int)p, i1, i1 + d1, I, I + D, 0, 0, 0});
                }
                emit(BPMAX_K_INNER_INTRA, T, false, f, 0);
            }
    }
}

} // namespace bpx

// Event-bracketed launch helper.
static int timed(bpmax_ctx *ctx, int kclass, double updates, const std::function<hipError_t()> &fn)
{
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ctx->timing) {
        while (ctx->ev_pool.size() < ctx->ev_used + 2) {
            hipEvent_t e;
            HIPCHK(ctx, hipEventCreate(&e));
            ctx->ev_pool.push_back(e);
        }
        e0 = ctx->ev_pool[ctx->ev_used++];
        e1 = ctx->ev_pool[ctx->ev_used++];
        HIPCHK(ctx, hipEventRecord(e0, ctx->stream));
    }
    hipError_t e = fn();
    if (e != hipSuccess) {
        ctx->err = std::string("kernel launch (") + kClassNames[kclass] + "): " + hipGetErrorString(e);
        return BPMAX_EDEVICE;
    }
    if (ctx->timing) {
        HIPCHK(ctx, hipEventRecord(e1, ctx->stream));
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

// Lay out pairs in the arena, upload codes/descriptors/weights/items, run the schedule.
// On success the F tables of all pairs are resident in ctx->arena and ctx->pairs_host
// describes them.
int bpx_run_pairs(bpmax_ctx *ctx, int npairs, const int *const *seq1s, const int *Ms, const int *const *seq2s,
                  const int *Ns, const float *intra_in, const float *inter_in)
{
    if (npairs <= 0) return BPMAX_EINVAL;
    float wt[50];
    float dflt_intra[25], dflt_inter[25];
    bpmax_engine_score_table(dflt_intra, dflt_inter);
    memcpy(wt, intra_in ? intra_in : dflt_intra, 25 * sizeof(float));
    memcpy(wt + 25, inter_in ? inter_in : (intra_in ? intra_in : dflt_inter), 25 * sizeof(float));

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
        d.T = choose_tile(N);
        d.nb = (N + d.T - 1) / d.T;
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
    int rc;
    if ((rc = bpx_grow(ctx, ctx->arena, (size_t)fl * sizeof(float))) != BPMAX_OK) return rc;
    if ((rc = bpx_grow(ctx, ctx->codes, (size_t)il * sizeof(int))) != BPMAX_OK) return rc;

    std::vector<int> codes((size_t)il, 4);
    for (int p = 0; p < npairs; ++p) {
        memcpy(codes.data() + P[p].a, seq1s[p], sizeof(int) * Ms[p]);
        for (int x = 0; x < Ns[p]; ++x) codes[P[p].b + x] = seq2s[p][Ns[p] - 1 - x];
    }
    Schedule &S = ctx->sched;
    build_schedule(P, S);

    // meta buffer: weights | pair descs | items
    const size_t wbytes = 256;
    const size_t pbytes = (P.size() * sizeof(PairDesc) + 255) / 256 * 256;
    const size_t ibytes = S.items.size() * sizeof(WorkItem);
    if ((rc = bpx_grow(ctx, ctx->meta, wbytes + pbytes + ibytes)) != BPMAX_OK) return rc;
    ctx->staging.resize(wbytes + pbytes + ibytes);
    memcpy(ctx->staging.data(), wt, sizeof wt);
    memcpy(ctx->staging.data() + wbytes, P.data(), P.size() * sizeof(PairDesc));
    memcpy(ctx->staging.data() + wbytes + pbytes, S.items.data(), ibytes);
    char *meta = (char *)ctx->meta.p;
    const float *dw = (const float *)meta;
    const PairDesc *dP = (const PairDesc *)(meta + wbytes);
    const WorkItem *dI = (const WorkItem *)(meta + wbytes + pbytes);
    float *arena = (float *)ctx->arena.p;
    const int *dcodes = (const int *)ctx->codes.p;

    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipMemcpyAsync(ctx->meta.p, ctx->staging.data(), ctx->staging.size(), hipMemcpyHostToDevice,
                               ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(ctx->codes.p, codes.data(), codes.size() * sizeof(int), hipMemcpyHostToDevice,
                               ctx->stream));
    hipStream_t st = ctx->stream;
    if ((rc = timed(ctx, BPMAX_K_FILL, 0, [&] { return launch_fill(st, arena, fl, -INFINITY); })) != BPMAX_OK)
        return rc;
    if ((rc = timed(ctx, BPMAX_K_S1, 0, [&] { return launch_s1(st, dP, npairs, arena, dcodes, dw); })) != BPMAX_OK)
        return rc;
    for (const Launch &L : S.launches) {
        const WorkItem *items = dI + L.first;
        switch (L.kind) {
        case BPMAX_K_OUTER:
            rc = timed(ctx, L.kind, L.updates,
                       [&] { return launch_outer(st, L.T, dP, items, L.count, arena, dcodes, dw); });
            break;
        case BPMAX_K_INNER_EXT:
            rc = timed(ctx, L.kind, L.updates,
                       [&] { return launch_inner_ext(st, L.T, L.single_strand, dP, items, L.count, arena); });
            break;
        default:
            rc = timed(ctx, L.kind, L.updates, [&] {
                return launch_inner_intra(st, L.T, L.single_strand, dP, items, L.count, arena, dcodes, dw);
            });
            break;
        }
        if (rc != BPMAX_OK) return rc;
    }
    // the host vectors (codes, staging) must outlive the async copies
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return bpx_collect_timing(ctx);
}

int bpx_gather_scores(bpmax_ctx *ctx, float *scores)
{
    const int n = (int)ctx->pairs_host.size();
    int rc;
    if ((rc = bpx_grow(ctx, ctx->scores, n * sizeof(float))) != BPMAX_OK) return rc;
    const size_t wbytes = 256;
    const PairDesc *dP = (const PairDesc *)((char *)ctx->meta.p + wbytes);
    HIPCHK(ctx, launch_scores(ctx->stream, dP, n, (const float *)ctx->arena.p, (float *)ctx->scores.p));
    HIPCHK(ctx, hipMemcpyAsync(scores, ctx->scores.p, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return BPMAX_OK;
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
    int rc = bpx_run_pairs(ctx, 1, &seq1, &M, &seq2, &N, intra, inter);
    if (rc != BPMAX_OK) return rc;
    float score = 0.f;
    if ((rc = bpx_gather_scores(ctx, &score)) != BPMAX_OK) return rc;
    if (score_out) *score_out = score;
    if (layout == BPMAX_LAYOUT_NONE) return BPMAX_OK;
    const PairDesc &P = ctx->pairs_host[0];
    const size_t bytes = bpmax_layout_bytes(M, N, layout);
    if ((rc = bpx_grow(ctx, ctx->out, bytes)) != BPMAX_OK) return rc;
    float *dout = (float *)ctx->out.p;
    const float *arena = (const float *)ctx->arena.p;
    if (layout == BPMAX_LAYOUT_PACKED_TRI_TRI) {
        rc = timed(ctx, BPMAX_K_CONVERT, 0, [&] { return launch_to_packed(ctx->stream, arena, P, dout); });
    } else {
        rc = timed(ctx, BPMAX_K_CONVERT, 0, [&] {
            hipError_t e = launch_fill(ctx->stream, dout, (long long)(bytes / sizeof(float)), 0.0f);
            if (e != hipSuccess) return e;
            return launch_to_dense(ctx->stream, arena, P, dout);
        });
    }
    if (rc != BPMAX_OK) return rc;
    HIPCHK(ctx, hipMemcpyAsync(F_out, dout, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return bpx_collect_timing(ctx);
}

extern "C" int bpmax_engine_solve_batch(bpmax_ctx *ctx, int npairs, const int *const *seq1s, const int *Ms,
                                        const int *const *seq2s, const int *Ns, const float *intra,
                                        const float *inter, float *scores)
{
    if (!ctx || npairs <= 0 || !seq1s || !Ms || !seq2s || !Ns || !scores) return BPMAX_EINVAL;
    int rc = bpx_run_pairs(ctx, npairs, seq1s, Ms, seq2s, Ns, intra, inter);
    if (rc != BPMAX_OK) return rc;
    return bpx_gather_scores(ctx, scores);
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

extern "C" int bpmax_r0_tri_batch_device(bpmax_ctx *ctx, int n, int batch, const float *A_dev, const float *B_dev,
                                         float *C_dev)
{
    if (!ctx || n < 1 || batch < 1 || !A_dev || !B_dev || !C_dev) return BPMAX_EINVAL;
    if (batch > 65535) {
        ctx->err = "batch must be <= 65535";
        return BPMAX_EINVAL;
    }
    const int T = 64, nb = (n + T - 1) / T;
    std::vector<int2> tiles;
    // heavy tiles (long K ranges) first so the tail of the grid is short
    for (int D = nb - 1; D >= 0; --D)
        for (int I = 0; I + D < nb; ++I) tiles.push_back(make_int2(I, I + D));
    int rc;
    if ((rc = bpx_grow(ctx, ctx->r0meta, tiles.size() * sizeof(int2))) != BPMAX_OK) return rc;
    HIPCHK(ctx, hipMemcpyAsync(ctx->r0meta.p, tiles.data(), tiles.size() * sizeof(int2), hipMemcpyHostToDevice,
                               ctx->stream));
    const double nn = n;
    const double upd = (double)batch * (nn + 2) * (nn + 1) * nn / 6.0;
    rc = timed(ctx, BPMAX_K_R0TRI, upd, [&] {
        return launch_r0tri(ctx->stream, T, nb, (const int2 *)ctx->r0meta.p, (int)tiles.size(), batch, A_dev, B_dev,
                            C_dev);
    });
    if (rc != BPMAX_OK) return rc;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
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
