// bpmax_engine_impl.h -- private definition of bpmax_ctx and engine helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <functional>
#include <string>
#include <vector>

#include "bpmax.h"
#include "bpmax_internal.h"

namespace bpx {

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

struct Launch {
    int kind;            // bpmax_kernel_class
    int T;
    bool single_strand;  // S2 phases
    int first, count;    // work-item range
    double updates;      // algorithmic cell updates (outer phases)
};

struct Schedule {
    std::vector<WorkItem> items;
    std::vector<Launch> launches;
};

struct KStat {
    long long launches = 0;
    double ms = 0.0;
    double updates = 0.0;
};

struct PendingTiming {
    int kclass;
    double updates;
    hipEvent_t e0, e1;
};

struct Server; // continuous-batching server (bpmax_server.cpp)

int choose_tile(int N);
size_t pair_device_floats(int M, int N, int T);
bool weights_exact(const float *intra, const float *inter, int M, int N, std::string *why);
void build_schedule(const std::vector<PairDesc> &P, Schedule &S);

} // namespace bpx

struct bpmax_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    bpx::DevBuf arena, codes, meta, out, scores, r0meta;
    std::vector<bpx::PairDesc> pairs_host;
    bpx::Schedule sched;
    std::vector<char> staging;
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<bpx::PendingTiming> pending;
    bpx::KStat stats[BPMAX_K_NCLASSES];
    std::vector<void *> user_allocs;
    bpx::Server *server = nullptr;
};

int bpx_grow(bpmax_ctx *ctx, bpx::DevBuf &b, size_t bytes);
int bpx_run_pairs(bpmax_ctx *ctx, int npairs, const int *const *seq1s, const int *Ms, const int *const *seq2s,
                  const int *Ns, const float *intra, const float *inter);
int bpx_gather_scores(bpmax_ctx *ctx, float *scores);
int bpx_collect_timing(bpmax_ctx *ctx);
void bpx_server_shutdown(bpmax_ctx *ctx);
