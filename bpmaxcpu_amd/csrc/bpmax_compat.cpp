// bpmax_compat.cpp -- the reference's C entry point and file formats on top of the engine.
//
// Paths are relative to TPDS/source/release/MPT_RT_V3/bpmax_complete_otf_ftable/ of the
// reference.
#include <dlfcn.h>
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <memory>
#include <mutex>
#include <thread>
#include <string>
#include <vector>

#include "bpmax.h"

namespace {

// The weight plug point: the reference resolves e_intra_score / e_inter_score at link
// time (src/external_functions.h:33-35).  A process that links the reference's
// external_functions unit (compiled as C++) or a C replacement exports them; otherwise
// the reference's default table is used.
typedef float (*score_fn)(int, int);

score_fn find_score(const char *cxx_name, const char *c_name)
{
    void *p = dlsym(RTLD_DEFAULT, cxx_name);
    if (!p) p = dlsym(RTLD_DEFAULT, c_name);
    return (score_fn)p;
}

std::mutex g_mu;
bpmax_ctx *g_ctx = nullptr; // lazily created context behind the compatibility symbol

} // namespace

// Exact signature of src/bpmax.c:129.
extern "C" void bpmax(long M, long N, long N_sec, long N_tile, long R, long MR, long NR, long P, int *seq1,
                      int *seq2, float ******FTable)
{
    // same parameter check and failure behaviour as src/bpmax.c:131-133
    if (!((M >= 1 && N >= 8 && N_sec >= 2 && N_tile >= 4 && R >= 0 && N_tile >= R + 1 && MR >= 1 && NR >= 1 &&
           P >= 1 && 0 >= P - 128))) {
        printf("The value of parameters are not valid.\n");
        exit(-1);
    }
    float intra[25], inter[25];
    bpmax_engine_score_table(intra, inter);
    score_fn fi = find_score("_Z13e_intra_scoreii", "e_intra_score");
    score_fn fe = find_score("_Z13e_inter_scoreii", "e_inter_score");
    for (int x = 0; x < 5; ++x)
        for (int y = 0; y < 5; ++y) {
            if (fi) intra[5 * x + y] = fi(x, y);
            if (fe) inter[5 * x + y] = fe(x, y);
        }

    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ctx) {
        const char *dev = getenv("BPMAX_DEVICE");
        int rc = bpmax_ctx_create(&g_ctx, dev ? atoi(dev) : 0);
        if (rc != BPMAX_OK) {
            printf("bpmax: no usable GPU (error %d)\n", rc);
            exit(-1);
        }
    }
    // left uninitialised: the engine writes every element (zeroing 1 GB would cost ~0.1 s)
    std::unique_ptr<float[]> packed(new float[(size_t)(M * (M + 1) / 2) * (size_t)(N * (N + 1) / 2)]);
    float score = 0.f;
    int rc = bpmax_engine_solve(g_ctx, seq1, (int)M, seq2, (int)N, intra, inter, packed.get(),
                                BPMAX_LAYOUT_PACKED_TRI_TRI, &score);
    if (rc != BPMAX_OK) {
        printf("bpmax: engine error %d: %s\n", rc, bpmax_last_error(g_ctx));
        exit(-1);
    }
    // scatter into the caller's 6-level tree, F(i1,j1,i2,j2) =
    // FTable[i1][j1][(i2+R)/N_tile][(j2+R)/N_tile][(i2+R)%N_tile][(j2+R)%N_tile]
    // (src/bpmax-wrapper.c:44)
    const size_t TN = (size_t)(N * (N + 1) / 2);
    std::vector<std::pair<long, long>> tris; // (i1, j1) in packed order: balanced work units
    for (long i1 = 0; i1 < M; ++i1)
        for (long j1 = i1; j1 < M; ++j1) tris.emplace_back(i1, j1);
    auto scatter = [&](long u) {
        const long i1 = tris[u].first, j1 = tris[u].second;
        {
            const float *tri = packed.get() + (size_t)u * TN;
            for (long i2 = 0; i2 < N; ++i2) {
                const float *row = tri + (size_t)(i2 * N - i2 * (i2 - 1) / 2) - i2;
                const long ri = i2 + R;
                for (long j2 = i2; j2 < N;) {
                    const long cj = j2 + R;
                    const long j3 = cj % N_tile;
                    const long run = std::min<long>(N_tile - j3, N - j2);
                    memcpy(FTable[i1][j1][ri / N_tile][cj / N_tile][ri % N_tile] + j3, row + j2, sizeof(float) * run);
                    j2 += run;
                }
            }
        }
    };
    const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::atomic<long> next{0};
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nth; ++t)
        pool.emplace_back([&] {
            for (long u; (u = next.fetch_add(1)) < (long)tris.size();) scatter(u);
        });
    for (auto &th : pool) th.join();
}

// The reference compiles its driver (src/bpmax-wrapper.c) as C++ (src/Makefile: CC=icpc,
// -std=c++11), so a driver built that way calls the C++-mangled name of the prototype at
// src/bpmax-wrapper.c:51.  Export that name too, so either build links unchanged.
extern "C" void bpmax_cxx_linkage(long M, long N, long N_sec, long N_tile, long R, long MR, long NR, long P,
                                  int *seq1, int *seq2, float ******FTable) __asm__("_Z5bpmaxllllllllPiS_PPPPPPf");
extern "C" void bpmax_cxx_linkage(long M, long N, long N_sec, long N_tile, long R, long MR, long NR, long P,
                                  int *seq1, int *seq2, float ******FTable)
{
    bpmax(M, N, N_sec, N_tile, R, MR, NR, P, seq1, seq2, FTable);
}

extern "C" int bpmax_read_sequence_file(const char *path, int *out, int cap)
{
    // FASTA-like record: '>' header lines skipped, letters/digits/'&' kept, ';' ends the
    // record, the next '>' at a line start ends it too (src/external_functions.c:338-377).
    if (!path || (cap > 0 && !out)) return BPMAX_EINVAL;
    FILE *f = fopen(path, "rt");
    if (!f) return BPMAX_EINVAL;
    int n = 0, c, last = '\n', state = 0;
    bool any = false;
    while ((c = fgetc(f)) != EOF) {
        if (last == '\n' && c == '>') {
            if (any) break;
            state = 1;
        } else if (c == '\n') {
            state = 0;
        } else if (state == 0) {
            const bool alnum = (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || (c >= '0' && c <= '9') || c == '&';
            if (alnum) {
                if (n < cap) out[n] = bpmax_base_code((char)c);
                ++n;
                any = true;
            } else if (c == ';') {
                break;
            }
        }
        last = c;
    }
    fclose(f);
    return any ? n : BPMAX_EINVAL;
}

extern "C" int bpmax_write_checking_dump(const char *path, const float *F, int M, int N, int with_prompt)
{
    if (!path || !F || M < 1 || N < 1) return BPMAX_EINVAL;
    FILE *f = fopen(path, "w");
    if (!f) return BPMAX_EINVAL;
    size_t idx = 0;
    for (int i1 = 0; i1 < M; ++i1)
        for (int j1 = i1; j1 < M; ++j1)
            for (int i2 = 0; i2 < N; ++i2)
                for (int j2 = i2; j2 < N; ++j2, ++idx) {
                    if (with_prompt) fprintf(f, "FTable(%d,%d,%d,%d)=", i1, j1, i2, j2);
                    fprintf(f, "%0.2f\n", F[idx]);
                }
    fclose(f);
    return BPMAX_OK;
}
