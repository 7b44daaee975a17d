/*
 * bpmax_oracle.c -- CPU restatement of the BPMax recurrence.  TEST INFRASTRUCTURE ONLY.
 *
 * This file is the checker for the HIP engine: only tests/, the smoke entry point and
 * bench.py's CPU-baseline leg may load it.  The product library never links it and has
 * no CPU fallback.
 *
 * Semantics follow the Alpha specification of the reference,
 *   TPDS/source/release/MPT_RT_V3/bpmax_complete_otf_ftable/compilation_scripts/bpmax_verify.ab:16-71
 * and its memoized evaluator
 *   TPDS/source/release/MPT_RT_V3/bpmax_complete_otf_ftable/src/bpmax_verify.c:518-721,
 * evaluated here as a plain wavefront (outer diagonal d1 = j1-i1, then inner diagonal
 * d2 = j2-i2) instead of a recursion.  Every term is a single two-operand float sum and
 * the reduction is max, so evaluation order does not change a single bit.
 *
 * Conventions (same as the reference):
 *   - a = seq1 codes, b[x] = seq2[N-1-x]  (strand 2 is indexed reversed),
 *     codes 0..4 = A,C,G,U/T,N  (src/external_functions.c:74-85);
 *   - intra[5*x+y] is e_intra_score(x,y), inter[5*x+y] is e_inter_score(x,y)
 *     (src/external_functions.c:21-68);
 *   - only cells with i1<=j1 and i2<=j2 are defined.
 *
 * Output layout "packed tri-tri": triangle (i1,j1) in row-major upper order over M, each
 * triangle holding its tri(N) cells (i2, j2>=i2) in row-major upper order -- the order of
 * the reference's CHECKING dump (src/bpmax-wrapper.c:318-342).
 *
 * Parity pins: tests/test_oracle.py checks this file against fixtures written by the
 * reference binary (oracle/_ref, built from the reference sources by oracle/Makefile),
 * against brute-force Nussinov enumeration for S1/S2, and against invariants.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NEG_INF (-INFINITY)

static inline float fmax2(float x, float y) { return x > y ? x : y; }
static inline int64_t tri64(int64_t n) { return n * (n + 1) / 2; }
/* row-major upper-triangle index of (i,j), i<=j<n */
static inline int64_t utri(int64_t i, int64_t j, int64_t n) { return i * n - i * (i - 1) / 2 + (j - i); }

int64_t oracle_packed_cells(int M, int N) { return tri64(M) * tri64(N); }

/*
 * Single-strand table (Nussinov with minimum hairpin of 3):
 *   S[i][j] = 0                                                   if j-i <= 3
 *           = max(S[i+1][j-1] + intra(x_i,x_j), max_{i<=k<j} S[i][k] + S[k+1][j])  otherwise
 * bpmax_verify.ab:17-26 (S1 over seq1, S2 over reversed seq2).  S is dense n*n row-major;
 * entries below the diagonal are set to -inf.
 */
void oracle_single_strand(int n, const int *x, const float *intra, float *S)
{
    for (int64_t i = 0; i < (int64_t)n * n; ++i) S[i] = NEG_INF;
    for (int d = 0; d < n; ++d) {
        for (int i = 0; i + d < n; ++i) {
            int j = i + d;
            float v;
            if (d <= 3) {
                v = 0.0f;
            } else {
                v = S[(int64_t)(i + 1) * n + (j - 1)] + intra[5 * x[i] + x[j]];
                for (int k = i; k < j; ++k)
                    v = fmax2(v, S[(int64_t)i * n + k] + S[(int64_t)(k + 1) * n + j]);
            }
            S[(int64_t)i * n + j] = v;
        }
    }
}

/*
 * Full F table.  F(i1,j1,i2,j2), bpmax_verify.ab:45-71:
 *   base  i1==j1 && i2==j2 : inter(a[i1], b[i2])
 *   else max of
 *     P1 = F(i1+1,j1-1,i2,j2) + intra(a[i1],a[j1])      if j1-i1 > 3, else the literal 0
 *     P2 = F(i1,j1,i2+1,j2-1) + intra(b[i2],b[j2])      if j2-i2 > 3, else the literal 0
 *     S1(i1,j1) + S2(i2,j2)
 *     R0 = max_{i1<=k1<j1, i2<=k2<j2} F(i1,k1,i2,k2) + F(k1+1,j1,k2+1,j2)   (i1<j1 && i2<j2)
 *     R1 = max_{i2<=k2<j2} F(i1,j1,i2,k2) + S2(k2+1,j2)                      (i2<j2)
 *     R2 = max_{i2<=k2<j2} S2(i2,k2) + F(i1,j1,k2+1,j2)                      (i2<j2)
 *     R3 = max_{i1<=k1<j1} S1(i1,k1) + F(k1+1,j1,i2,j2)                      (i1<j1)
 *     R4 = max_{i1<=k1<j1} F(i1,k1,i2,j2) + S1(k1+1,j1)                      (i1<j1)
 *   (reductions over empty ranges contribute the literal 0 of the spec's case arms).
 *
 * Evaluation: per triangle (i1,j1) in outer-diagonal order, first the "outer" terms that
 * only read earlier triangles (S1+S2, P1, R0, R3, R4) for every (i2,j2), then the inner
 * wavefront over d2 adding P2, R1, R2.  Triangles on one outer diagonal are independent.
 *
 * F must hold oracle_packed_cells(M,N) floats.  Returns 0, or -1 on bad arguments /
 * allocation failure.
 */
/* Planted-mistake switches for the checker's own tests (0 = the faithful recurrence):
 * each bit removes or corrupts one part so tests can show the fixtures catch it. */
enum {
    MUT_NO_P1 = 1, MUT_NO_P2 = 2, MUT_NO_R0 = 4, MUT_R0_NOSHIFT = 8, MUT_NO_R1 = 16,
    MUT_NO_R2 = 32, MUT_NO_R3 = 64, MUT_NO_R4 = 128, MUT_INTER_TRANSPOSED = 256,
    MUT_SEQ2_UNREVERSED = 512, MUT_HAIRPIN_2 = 1024
};

int oracle_bpmax_mut(int M, int N, const int *seq1, const int *seq2, const float *intra,
                     const float *inter, float *F, float *score, int mut);

int oracle_bpmax(int M, int N, const int *seq1, const int *seq2, const float *intra,
                 const float *inter, float *F, float *score)
{
    return oracle_bpmax_mut(M, N, seq1, seq2, intra, inter, F, score, 0);
}

int oracle_bpmax_mut(int M, int N, const int *seq1, const int *seq2, const float *intra,
                     const float *inter, float *F, float *score, int mut)
{
    if (M < 1 || N < 1 || !seq1 || !seq2 || !intra || !inter || !F) return -1;
    for (int i = 0; i < M; ++i) if (seq1[i] < 0 || seq1[i] > 4) return -1;
    for (int i = 0; i < N; ++i) if (seq2[i] < 0 || seq2[i] > 4) return -1;

    const int64_t TN = tri64(N);
    int *b = (int *)malloc(sizeof(int) * N);
    float *S1 = (float *)malloc(sizeof(float) * (size_t)M * M);
    float *S2 = (float *)malloc(sizeof(float) * (size_t)N * N);
    float *S2T = (float *)malloc(sizeof(float) * (size_t)N * N); /* S2T[j][k] = S2[k][j] */
    if (!b || !S1 || !S2 || !S2T) { free(b); free(S1); free(S2); free(S2T); return -1; }
    for (int x = 0; x < N; ++x) b[x] = (mut & MUT_SEQ2_UNREVERSED) ? seq2[x] : seq2[N - 1 - x];
    oracle_single_strand(M, seq1, intra, S1);
    oracle_single_strand(N, b, intra, S2);
    for (int k = 0; k < N; ++k)
        for (int j = 0; j < N; ++j) S2T[(int64_t)j * N + k] = S2[(int64_t)k * N + j];

#define TRI(i1, j1) (F + utri((i1), (j1), M) * TN)

    for (int d1 = 0; d1 < M; ++d1) {
#pragma omp parallel for schedule(dynamic, 1)
        for (int i1 = 0; i1 < M - d1; ++i1) {
            const int j1 = i1 + d1;
            float *C = TRI(i1, j1);
            /* ---- outer terms ---- */
            const float s1 = S1[(int64_t)i1 * M + j1];
            for (int i2 = 0; i2 < N; ++i2) {
                float *row = C + utri(i2, i2, N);
                const float *s2row = S2 + (int64_t)i2 * N;
                for (int j2 = i2; j2 < N; ++j2) row[j2 - i2] = s1 + s2row[j2];
            }
            if (d1 == 0) {
                for (int i2 = 0; i2 < N; ++i2)
                    C[utri(i2, i2, N)] = (mut & MUT_INTER_TRANSPOSED) ? inter[5 * b[i2] + seq1[i1]]
                                                                      : inter[5 * seq1[i1] + b[i2]];
            } else {
                /* the spec's literal 0 arms (always dominated by S1+S2 >= 0) */
                for (int64_t c = 0; c < TN; ++c) C[c] = fmax2(C[c], 0.0f);
            }
            if (d1 > 3 && !(mut & MUT_NO_P1)) { /* P1 */
                const float w = intra[5 * seq1[i1] + seq1[j1]];
                const float *P = TRI(i1 + 1, j1 - 1);
                for (int64_t c = 0; c < TN; ++c) C[c] = fmax2(C[c], P[c] + w);
            }
            for (int k1 = i1; k1 < j1; ++k1) {
                const float *A = TRI(i1, k1);       /* F(i1,k1,.,.)   */
                const float *B = TRI(k1 + 1, j1);   /* F(k1+1,j1,.,.) */
                const float s1a = S1[(int64_t)i1 * M + k1];        /* S1(i1,k1)   */
                const float s1b = S1[(int64_t)(k1 + 1) * M + j1];  /* S1(k1+1,j1) */
                /* R3, R4 (elementwise over the inner triangle) */
                for (int64_t c = 0; c < TN; ++c) {
                    if (!(mut & MUT_NO_R3)) C[c] = fmax2(C[c], s1a + B[c]);
                    if (!(mut & MUT_NO_R4)) C[c] = fmax2(C[c], A[c] + s1b);
                }
                if (mut & MUT_NO_R0) continue;
                const int sh = (mut & MUT_R0_NOSHIFT) ? 0 : 1;
                /* R0: C[i2][j2] max= A[i2][k2] + B[k2+1][j2] for i2<=k2<j2 */
                for (int i2 = 0; i2 < N; ++i2) {
                    float *crow = C + utri(i2, i2, N) - i2;          /* crow[j2] */
                    const float *arow = A + utri(i2, i2, N) - i2;    /* arow[k2] */
                    for (int k2 = i2; k2 < N - 1; ++k2) {
                        const float a = arow[k2];
                        const float *brow = B + utri(k2 + sh, k2 + sh, N) - (k2 + sh); /* brow[j2] */
                        for (int j2 = k2 + 1; j2 < N; ++j2) crow[j2] = fmax2(crow[j2], a + brow[j2]);
                    }
                }
            }
            /* ---- inner wavefront: P2, R1, R2 ---- */
            for (int d2 = 1; d2 < N; ++d2) {
                for (int i2 = 0; i2 + d2 < N; ++i2) {
                    const int j2 = i2 + d2;
                    float v = C[utri(i2, j2, N)];
                    if (d2 > ((mut & MUT_HAIRPIN_2) ? 2 : 3) && !(mut & MUT_NO_P2))
                        v = fmax2(v, C[utri(i2 + 1, j2 - 1, N)] + intra[5 * b[i2] + b[j2]]);
                    else
                        v = fmax2(v, 0.0f);
                    const float *frow = C + utri(i2, i2, N) - i2;   /* frow[k2] = F(i2,k2) */
                    const float *s2col = S2T + (int64_t)j2 * N;     /* s2col[k] = S2(k,j2) */
                    const float *s2row = S2 + (int64_t)i2 * N;      /* s2row[k] = S2(i2,k) */
                    for (int k2 = i2; k2 < j2; ++k2) {
                        if (!(mut & MUT_NO_R1)) v = fmax2(v, frow[k2] + s2col[k2 + 1]);          /* R1 */
                        if (!(mut & MUT_NO_R2)) v = fmax2(v, s2row[k2] + C[utri(k2 + 1, j2, N)]); /* R2 */
                    }
                    C[utri(i2, j2, N)] = v;
                }
            }
        }
    }
    if (score) *score = TRI(0, M - 1)[utri(0, N - 1, N)];
#undef TRI
    free(b); free(S1); free(S2); free(S2T);
    return 0;
}

/*
 * Isolated R0 kernel on one upper-triangular n*n pair (dense row-major storage):
 *   C[i][j] = max(C[i][j], max_{i<=k<=j} A[i][k] + B[k][j])   for i<=j,
 * the "R0 max-plus kernel isolated" workload; lower-triangle entries are left untouched.
 */
void oracle_maxplus_tri(int n, const float *A, const float *B, float *C)
{
    for (int i = 0; i < n; ++i)
        for (int k = i; k < n; ++k) {
            const float a = A[(int64_t)i * n + k];
            const float *brow = B + (int64_t)k * n;
            float *crow = C + (int64_t)i * n;
            for (int j = k; j < n; ++j) crow[j] = fmax2(crow[j], a + brow[j]);
        }
}

int oracle_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
