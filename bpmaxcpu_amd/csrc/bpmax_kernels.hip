// bpmax_kernels.hip -- hand-written gfx950 kernels of the BPMax engine.
//
// Max-plus ("tropical") products are not multiply-accumulates, so nothing here uses MFMA:
// the hot loop is VALU v_add_f32 + v_max3_f32 on per-lane 4x4 register tiles fed from
// LDS with 128-bit reads (1.5 VALU instructions per cell update; packed FP32 ops are
// disabled at compile time).  Scores are exact multiples of 2^-8 far below 2^24, so every
// sum is exact and every reduction order gives the same bits.
//
// Phases (reference functions they replace, paths relative to
// TPDS/source/release/MPT_RT_V3/bpmax_complete_otf_ftable/src/):
//   k_outer       R0 double max-plus + init + P1 + R3 + R4 for one outer diagonal
//                 (bpmax_outer_reductions.c:123-180, matrix_max_plus_section.c:153-257,
//                  reg_3_24.c:9-84, bpmax_ftable_init.c:116-215,
//                  bpmax_outer_south_west.c:115-160, bpmax_r3_section.c:114-139,
//                  bpmax_r4_section.c:114-139)
//   k_inner_ext   cross-tile parts of R1/R2 (F) and of S2's split term
//                 (bpmax_inner_reductions.c:140-178, bpmax_single_strand_s2_tile.c:124-262)
//   k_inner_intra in-tile dependence wavefront with P2 (bpmax_inner_reductions_diagonal_tile.c,
//                 bpmax_inner_reductions_finalize.c, bpmax_single_strand_diagonal_tile.c,
//                 bpmax_single_strand_finalize.c)
//   k_s1          strand-1 single-strand table (bpmax_single_strand_s1.c:115-166)
#include <hip/hip_runtime.h>
#include <cstdint>
#include "bpmax_internal.h"
#include "bpmax_launch.h"

namespace bpx {

#define NEG_INF (-__builtin_inff())

template <int T>
struct Tc {
    static constexpr int R = 4;        // register tile edge: each lane owns R x R cells
    static constexpr int TX = T / R;   // lanes per tile row
    static constexpr int NT = TX * TX; // lanes per product workgroup (256 for T=64)
    static constexpr int LD = T + 4;   // LDS row stride in floats (keeps 16-byte rows)
};

__device__ __forceinline__ const float *tile_at(const float *tri, int I, int J, int nb, int T)
{
    return tri + utri_ll(I, J, nb) * (long long)(T * T);
}
__device__ __forceinline__ float *tile_at(float *tri, int I, int J, int nb, int T)
{
    return tri + utri_ll(I, J, nb) * (long long)(T * T);
}

// As[k][r] = X[I*T + r][K*T + k + cshift]  (cshift 0 or -1); -inf where the source tile
// lies below the diagonal.  With lastcol_only, every column but k = T-1 reads -inf.
template <int T>
__device__ __forceinline__ void load_A(float *As, const float *X, int nb, int I, int K, int cshift,
                                       bool lastcol_only)
{
    constexpr int NT = Tc<T>::NT, LD = Tc<T>::LD;
    for (int e = threadIdx.x; e < T * T; e += NT) {
        const int r = e / T, k = e % T;
        const int y = K * T + k + cshift;
        float v = NEG_INF;
        if (y >= 0 && !(lastcol_only && k != T - 1)) {
            const int Y = y / T;
            if (Y >= I && Y < nb) v = tile_at(X, I, Y, nb, T)[r * T + (y % T)];
        }
        As[k * LD + r] = v;
    }
}

// Bs[k][c] = X[K*T + k + rshift][J*T + c]  (rshift 0 or +1); -inf rows below the diagonal.
template <int T>
__device__ __forceinline__ void load_B(float *Bs, const float *X, int nb, int K, int J, int rshift)
{
    constexpr int NT = Tc<T>::NT, LD = Tc<T>::LD;
    constexpr int Q = T / 4;
    for (int e = threadIdx.x; e < T * Q; e += NT) {
        const int k = e / Q, c4 = (e % Q) * 4;
        const int x = K * T + k + rshift;
        const int Xt = x / T;
        float4 v = make_float4(NEG_INF, NEG_INF, NEG_INF, NEG_INF);
        if (Xt <= J && Xt < nb) v = *(const float4 *)(tile_at(X, Xt, J, nb, T) + (x % T) * T + c4);
        *(float4 *)(Bs + k * LD + c4) = v;
    }
}

// acc[i][j] = max(acc[i][j], max_k As[k][ty*4+i] + Bs[k][tx*4+j]) over the whole tile depth.
template <int T>
__device__ __forceinline__ void product(float (&acc)[4][4], const float *As, const float *Bs, int ty,
                                        int tx)
{
    constexpr int LD = Tc<T>::LD;
#pragma unroll 4
    for (int k = 0; k < T; k += 2) {
        const float4 a0 = *(const float4 *)(As + k * LD + ty * 4);
        const float4 a1 = *(const float4 *)(As + (k + 1) * LD + ty * 4);
        const float4 b0 = *(const float4 *)(Bs + k * LD + tx * 4);
        const float4 b1 = *(const float4 *)(Bs + (k + 1) * LD + tx * 4);
        const float av0[4] = {a0.x, a0.y, a0.z, a0.w}, av1[4] = {a1.x, a1.y, a1.z, a1.w};
        const float bv0[4] = {b0.x, b0.y, b0.z, b0.w}, bv1[4] = {b1.x, b1.y, b1.z, b1.w};
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acc[i][j] = fmaxf(acc[i][j], fmaxf(av0[i] + bv0[j], av1[i] + bv1[j]));
    }
}

// Write the lane's 4x4 cells of tile `dst`; atomic = max-combine with what is there
// (valid because every defined cell's final value is >= 0, where float order equals the
// signed-integer order of the bit patterns).
template <int T>
__device__ __forceinline__ void store_acc(float *dst, const float (&acc)[4][4], int ty, int tx, bool atomic)
{
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float *row = dst + (ty * 4 + i) * T + tx * 4;
        if (atomic) {
#pragma unroll
            for (int j = 0; j < 4; ++j) atomicMax((int *)(row + j), __float_as_int(acc[i][j]));
        } else {
            *(float4 *)row = make_float4(acc[i][0], acc[i][1], acc[i][2], acc[i][3]);
        }
    }
}

// ------------------------------------------------------------------------------ outer
// One work item = one output tile (I,J) of triangle (i1,j1) on the current outer diagonal
// and a range of K-steps; step s <-> (k1 = i1 + s/(J-I+1), K = I + s%(J-I+1)):
//   C[I][J] max= F(i1,k1)[I][K] (x) F(k1+1,j1)[K*T+1 .. K*T+T][J]        (R0, k2+1 shift)
// Items flagged WI_EPILOGUE add S1(i1,j1)+S2 (or the inter base case on d1 = 0), P1 and
// the R3/R4 streams of the same tile.
template <int T>
__global__ __launch_bounds__(Tc<T>::NT) void k_outer(const PairDesc *__restrict__ pairs,
                                                     const WorkItem *__restrict__ items,
                                                     float *__restrict__ arena,
                                                     const int *__restrict__ codes,
                                                     const float *__restrict__ w)
{
    constexpr int TX = Tc<T>::TX, LD = Tc<T>::LD;
    __shared__ __attribute__((aligned(16))) float As[T * LD];
    __shared__ __attribute__((aligned(16))) float Bs[T * LD];
    const WorkItem it = items[blockIdx.x];
    const PairDesc P = pairs[it.pair];
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    const int i1 = it.i1, j1 = it.j1, I = it.I, J = it.J, nb = P.nb, M = P.M;
    float acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = NEG_INF;

    const int nK = J - I + 1;
    for (int s = it.step_begin; s < it.step_end; ++s) {
        const int k1 = i1 + s / nK, K = I + s % nK;
        const float *A = arena + P.F + utri_ll(i1, k1, M) * P.tstride;
        const float *B = arena + P.F + utri_ll(k1 + 1, j1, M) * P.tstride;
        __syncthreads();
        load_A<T>(As, A, nb, I, K, 0, false);
        load_B<T>(Bs, B, nb, K, J, 1);
        __syncthreads();
        product<T>(acc, As, Bs, ty, tx);
    }

    const long long toff = utri_ll(I, J, nb) * (long long)(T * T);
    if (it.flags & WI_EPILOGUE) {
        const float *S1 = arena + P.S1;
        const float *S2t = arena + P.S2 + toff;
        const int *a = codes + P.a, *b = codes + P.b;
        const float s1 = S1[i1 * M + j1];
        const int d1 = j1 - i1;
        float v[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) v[i][j] = s1 + S2t[(ty * 4 + i) * T + tx * 4 + j];
        if (d1 == 0) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int r = I * T + ty * 4 + i, c = J * T + tx * 4 + j;
                    if (r == c) v[i][j] = w[25 + 5 * a[i1] + b[r]]; // inter(a[i1], b[i2])
                }
        }
        if (d1 >= 4) {
            const float w1 = w[5 * a[i1] + a[j1]];
            const float *Pt = arena + P.F + utri_ll(i1 + 1, j1 - 1, M) * P.tstride + toff;
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) v[i][j] = fmaxf(v[i][j], Pt[(ty * 4 + i) * T + tx * 4 + j] + w1);
        }
        for (int k1 = i1; k1 < j1; ++k1) {
            const float s1a = S1[i1 * M + k1], s1b = S1[(k1 + 1) * M + j1];
            const float *Bt = arena + P.F + utri_ll(k1 + 1, j1, M) * P.tstride + toff;
            const float *At = arena + P.F + utri_ll(i1, k1, M) * P.tstride + toff;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float4 bb = *(const float4 *)(Bt + (ty * 4 + i) * T + tx * 4);
                const float4 aa = *(const float4 *)(At + (ty * 4 + i) * T + tx * 4);
                v[i][0] = fmaxf(v[i][0], fmaxf(s1a + bb.x, aa.x + s1b));
                v[i][1] = fmaxf(v[i][1], fmaxf(s1a + bb.y, aa.y + s1b));
                v[i][2] = fmaxf(v[i][2], fmaxf(s1a + bb.z, aa.z + s1b));
                v[i][3] = fmaxf(v[i][3], fmaxf(s1a + bb.w, aa.w + s1b));
            }
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = fmaxf(acc[i][j], v[i][j]);
    }
    float *C = arena + P.F + utri_ll(i1, j1, M) * P.tstride + toff;
    store_acc<T>(C, acc, ty, tx, (it.flags & WI_ATOMIC) != 0);
}

// ------------------------------------------------------------------------ inner ext
// Cross-tile part of the inner closure for tile (I,J), D = J-I >= 1, of X = F(i1,j1)
// (MODE_S = false) or X = S2 (MODE_S = true), S = S2.  Steps:
//   F: s <  D : R1  X[I][K] (x) S[K*T+1 ..][J]          K = I+s      (k2 in tile K)
//      s >= D : R2  S[I][K'*T-1 ..] (x) X[K'][J]        K' = I+1+s-D (k2+1 in tile K')
//   S: s <  D : S[I][K] (x) S[K*T+1 ..][J], K = I+s; for K = I only column T-1 of S[I][I]
//               (the remaining columns of tile I pair with cells of (I,J) itself and are
//               handled by the in-tile wavefront).
template <int T, bool MODE_S>
__global__ __launch_bounds__(Tc<T>::NT) void k_inner_ext(const PairDesc *__restrict__ pairs,
                                                         const WorkItem *__restrict__ items,
                                                         float *__restrict__ arena)
{
    constexpr int TX = Tc<T>::TX, LD = Tc<T>::LD;
    __shared__ __attribute__((aligned(16))) float As[T * LD];
    __shared__ __attribute__((aligned(16))) float Bs[T * LD];
    const WorkItem it = items[blockIdx.x];
    const PairDesc P = pairs[it.pair];
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    const int I = it.I, J = it.J, nb = P.nb, D = J - I;
    float *S = arena + P.S2;
    float *X = MODE_S ? S : arena + P.F + utri_ll(it.i1, it.j1, P.M) * P.tstride;
    float acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = NEG_INF;
    for (int s = it.step_begin; s < it.step_end; ++s) {
        __syncthreads();
        if (MODE_S) {
            const int K = I + s;
            load_A<T>(As, S, nb, I, K, 0, K == I);
            load_B<T>(Bs, S, nb, K, J, 1);
        } else if (s < D) {
            const int K = I + s;
            load_A<T>(As, X, nb, I, K, 0, false);
            load_B<T>(Bs, S, nb, K, J, 1);
        } else {
            const int K2 = I + 1 + (s - D);
            load_A<T>(As, S, nb, I, K2, -1, false);
            load_B<T>(Bs, X, nb, K2, J, 0);
        }
        __syncthreads();
        product<T>(acc, As, Bs, ty, tx);
    }
    store_acc<T>(tile_at(X, I, J, nb, T), acc, ty, tx, true);
}

// ---------------------------------------------------------------------- inner intra
// In-tile closure of tile (I,J) of X (F(i1,j1) or S2): processes the cells by
// anti-diagonal step s = b + (T-1-a) (local row a, column b); every dependence of a cell
// (same row to the left, same column below, P2 at (a+1,b-1)) lies on an earlier step.
// For each cell, G lanes split the in-tile terms
//   R1: X[a][b'] + SR[b'+1][b]   b' in [diag ? a : 0, b-1]   SR = S2 tile (J,J)
//   R2: SL[a][a'-1] + X[a'][b]   a' in [a+1, diag ? b : T-1] SL = S2 tile (I,I)
// and combine by lane shuffles; then P2 = X[a+1][b-1] + intra(b[r], b[c]) when c-r >= 4.
// For S2 on a diagonal tile SL = SR = X itself (both factors of S2's split term are in
// the tile); S2 additionally starts every upper cell at 0 (its short-interval value).
template <int T, bool MODE_S>
__global__ __launch_bounds__(256) void k_inner_intra(const PairDesc *__restrict__ pairs,
                                                     const WorkItem *__restrict__ items,
                                                     float *__restrict__ arena,
                                                     const int *__restrict__ codes,
                                                     const float *__restrict__ w)
{
    constexpr int NTI = 256, G = NTI / T, L = T + 1;
    __shared__ float Xs[T * L], Sa[T * L], Sb[T * L];
    __shared__ float below[T], left[T];
    __shared__ float corner;
    __shared__ int cr[T], cc[T];
    const WorkItem it = items[blockIdx.x];
    const PairDesc P = pairs[it.pair];
    const int I = it.I, J = it.J, nb = P.nb;
    const bool diag = (I == J);
    float *S = arena + P.S2;
    float *X = MODE_S ? S : arena + P.F + utri_ll(it.i1, it.j1, P.M) * P.tstride;
    const int tid = threadIdx.x;
    const float *Xt = tile_at(X, I, J, nb, T);
    for (int e = tid; e < T * T; e += NTI) {
        const int a = e / T, b = e % T;
        float v = Xt[e];
        if (MODE_S && I * T + a <= J * T + b) v = fmaxf(v, 0.0f);
        Xs[a * L + b] = v;
    }
    const float *SL = Sa, *SR = Sb;
    if (diag) {
        if (MODE_S) {
            SL = Xs;
            SR = Xs;
        } else {
            const float *St = tile_at(S, I, I, nb, T);
            for (int e = tid; e < T * T; e += NTI) Sa[(e / T) * L + e % T] = St[e];
            SR = Sa;
        }
    } else {
        const float *S1t = tile_at(S, I, I, nb, T), *S2t = tile_at(S, J, J, nb, T);
        for (int e = tid; e < T * T; e += NTI) {
            Sa[(e / T) * L + e % T] = S1t[e];
            Sb[(e / T) * L + e % T] = S2t[e];
        }
        const float *Bt = tile_at(X, I + 1, J, nb, T);
        const float *Lt = tile_at(X, I, J - 1, nb, T);
        for (int e = tid; e < T; e += NTI) {
            below[e] = Bt[e];
            left[e] = Lt[e * T + T - 1];
        }
        if (tid == 0) corner = (J - I >= 2) ? tile_at(X, I + 1, J - 1, nb, T)[T - 1] : NEG_INF;
    }
    const int *bc = codes + P.b;
    for (int e = tid; e < T; e += NTI) {
        cr[e] = bc[I * T + e];
        cc[e] = bc[J * T + e];
    }
    __syncthreads();

    const int q = tid / G, g = tid % G;
    for (int s = diag ? T : 0; s <= 2 * T - 2; ++s) {
        const int alo = s < T - 1 ? T - 1 - s : 0;
        const int ahi = s > T - 1 ? 2 * T - 2 - s : T - 1;
        const int a = alo + q;
        const bool live = a <= ahi;
        const int b = a + s - (T - 1);
        float v = NEG_INF;
        if (live) {
            const int lo1 = diag ? a : 0;
            for (int bp = lo1 + g; bp < b; bp += G) v = fmaxf(v, Xs[a * L + bp] + SR[(bp + 1) * L + b]);
            const int hi2 = diag ? b : T - 1;
            for (int ap = a + 1 + g; ap <= hi2; ap += G) v = fmaxf(v, SL[a * L + ap - 1] + Xs[ap * L + b]);
        }
#pragma unroll
        for (int off = 1; off < G; off <<= 1) v = fmaxf(v, __shfl_xor(v, off));
        if (live && g == 0) {
            v = fmaxf(v, Xs[a * L + b]);
            const int r = I * T + a, c = J * T + b;
            if (c - r >= 4) {
                float p;
                if (a + 1 < T && b >= 1) p = Xs[(a + 1) * L + b - 1];
                else if (b >= 1) p = below[b - 1];
                else if (a + 1 < T) p = left[a + 1];
                else p = corner;
                v = fmaxf(v, p + w[5 * cr[a] + cc[b]]);
            }
            Xs[a * L + b] = v;
        }
        __syncthreads();
    }
    float *Xo = tile_at(X, I, J, nb, T);
    for (int e = tid; e < T * T; e += NTI) Xo[e] = Xs[(e / T) * L + e % T];
}

// ------------------------------------------------------------------------------ S1
// Strand-1 single-strand table, dense M x M, one workgroup per pair, diagonal by diagonal
// (bpmax_verify.ab:17-21).
__global__ __launch_bounds__(256) void k_s1(const PairDesc *__restrict__ pairs, float *__restrict__ arena,
                                            const int *__restrict__ codes, const float *__restrict__ w)
{
    const PairDesc P = pairs[blockIdx.x];
    float *S1 = arena + P.S1;
    const int *a = codes + P.a;
    const int M = P.M;
    for (int e = threadIdx.x; e < M * M; e += blockDim.x) S1[e] = NEG_INF;
    __syncthreads();
    for (int d = 0; d < M; ++d) {
        for (int i = threadIdx.x; i + d < M; i += blockDim.x) {
            const int j = i + d;
            float v = 0.0f;
            if (d > 3) {
                v = S1[(i + 1) * M + j - 1] + w[5 * a[i] + a[j]];
                for (int k = i; k < j; ++k) v = fmaxf(v, S1[i * M + k] + S1[(k + 1) * M + j]);
            }
            S1[i * M + j] = v;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------- utilities
__global__ void k_fill(float *__restrict__ p, long long n, float v)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// F (tiled) -> packed tri-tri rows: block (i2, outer triangle t).
__global__ void k_to_packed(const float *__restrict__ arena, PairDesc P, float *__restrict__ out)
{
    const int i2 = blockIdx.x;
    const long long t = blockIdx.y;
    const int N = P.N, T = P.T, nb = P.nb;
    const float *tri = arena + P.F + t * P.tstride;
    float *dst = out + t * tri_ll(N) + utri_ll(i2, i2, N) - i2;
    const int I = i2 / T, ri = i2 % T;
    for (int j2 = i2 + threadIdx.x; j2 < N; j2 += blockDim.x)
        dst[j2] = tile_at(tri, I, j2 / T, nb, T)[ri * T + j2 % T];
}

// F (tiled) -> dense [M][M][N][N]: block (i2, i1*M + j1); only i1<=j1, i2<=j2 written.
__global__ void k_to_dense(const float *__restrict__ arena, PairDesc P, float *__restrict__ out)
{
    const int i2 = blockIdx.x;
    const int i1 = blockIdx.y / P.M, j1 = blockIdx.y % P.M;
    if (j1 < i1) return;
    const int N = P.N, T = P.T, nb = P.nb;
    const float *tri = arena + P.F + utri_ll(i1, j1, P.M) * P.tstride;
    float *dst = out + (((long long)i1 * P.M + j1) * N + i2) * N;
    const int I = i2 / T, ri = i2 % T;
    for (int j2 = i2 + threadIdx.x; j2 < N; j2 += blockDim.x)
        dst[j2] = tile_at(tri, I, j2 / T, nb, T)[ri * T + j2 % T];
}

// Gather F(0,M-1,0,N-1) of each pair into scores[p].
__global__ void k_scores(const PairDesc *__restrict__ pairs, int npairs, const float *__restrict__ arena,
                         float *__restrict__ scores)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs) return;
    const PairDesc P = pairs[p];
    const int c = P.N - 1;
    const float *tri = arena + P.F + utri_ll(0, P.M - 1, P.M) * P.tstride;
    scores[p] = tile_at(tri, 0, c / P.T, P.nb, P.T)[c % P.T];
}

// ------------------------------------------------------------------- isolated R0
// C[i][j] max= max_{i<=k<=j} A[i][k] + B[k][j] on a batch of tiled upper triangles;
// block = (output tile index within the triangle, matrix).  tiles holds the (I,J) list.
template <int T>
__global__ __launch_bounds__(Tc<T>::NT) void k_r0tri(int nb, const int2 *__restrict__ tiles,
                                                     const float *__restrict__ A,
                                                     const float *__restrict__ B, float *__restrict__ C)
{
    constexpr int TX = Tc<T>::TX, LD = Tc<T>::LD;
    __shared__ __attribute__((aligned(16))) float As[T * LD];
    __shared__ __attribute__((aligned(16))) float Bs[T * LD];
    const int2 IJ = tiles[blockIdx.x];
    const long long tstride = tri_ll(nb) * T * T;
    const long long m = blockIdx.y;
    const float *Am = A + m * tstride, *Bm = B + m * tstride;
    float *Cm = C + m * tstride;
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    float acc[4][4];
    const float *Ct = tile_at(Cm, IJ.x, IJ.y, nb, T);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float4 c = *(const float4 *)(Ct + (ty * 4 + i) * T + tx * 4);
        acc[i][0] = c.x;
        acc[i][1] = c.y;
        acc[i][2] = c.z;
        acc[i][3] = c.w;
    }
    for (int K = IJ.x; K <= IJ.y; ++K) {
        __syncthreads();
        load_A<T>(As, Am, nb, IJ.x, K, 0, false);
        load_B<T>(Bs, Bm, nb, K, IJ.y, 0);
        __syncthreads();
        product<T>(acc, As, Bs, ty, tx);
    }
    store_acc<T>(tile_at(Cm, IJ.x, IJ.y, nb, T), acc, ty, tx, false);
}

// Dense row-major n x n batch <-> tiled triangles (lower cells -> -inf on both sides).
__global__ void k_dense_to_tiled(int n, int nb, int T, const float *__restrict__ src, float *__restrict__ dst)
{
    const long long m = blockIdx.y;
    const long long tstride = tri_ll(nb) * T * T;
    const int r = blockIdx.x;
    for (int c = threadIdx.x; c < nb * T; c += blockDim.x) {
        if (c / T < r / T) continue;
        float v = NEG_INF;
        if (r < n && c < n && r <= c) v = src[m * n * n + (long long)r * n + c];
        tile_at(dst + m * tstride, r / T, c / T, nb, T)[(r % T) * T + c % T] = v;
    }
}
__global__ void k_tiled_to_dense(int n, int nb, int T, const float *__restrict__ src, float *__restrict__ dst)
{
    const long long m = blockIdx.y;
    const long long tstride = tri_ll(nb) * T * T;
    const int r = blockIdx.x;
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        float v = NEG_INF;
        if (r <= c) v = tile_at(src + m * tstride, r / T, c / T, nb, T)[(r % T) * T + c % T];
        dst[m * n * n + (long long)r * n + c] = v;
    }
}

// ---------------------------------------------------------------- launch wrappers
template <int T>
static hipError_t outer_T(hipStream_t st, const PairDesc *pairs, const WorkItem *items, int n, float *arena,
                          const int *codes, const float *w)
{
    hipLaunchKernelGGL(k_outer<T>, dim3(n), dim3(Tc<T>::NT), 0, st, pairs, items, arena, codes, w);
    return hipGetLastError();
}
hipError_t launch_outer(hipStream_t st, int T, const PairDesc *pairs, const WorkItem *items, int n,
                        float *arena, const int *codes, const float *w)
{
    if (n <= 0) return hipSuccess;
    switch (T) {
    case 32: return outer_T<32>(st, pairs, items, n, arena, codes, w);
    case 64: return outer_T<64>(st, pairs, items, n, arena, codes, w);
    default: return hipErrorInvalidValue;
    }
}

template <int T, bool S>
static hipError_t ext_T(hipStream_t st, const PairDesc *pairs, const WorkItem *items, int n, float *arena)
{
    hipLaunchKernelGGL((k_inner_ext<T, S>), dim3(n), dim3(Tc<T>::NT), 0, st, pairs, items, arena);
    return hipGetLastError();
}
hipError_t launch_inner_ext(hipStream_t st, int T, bool single_strand, const PairDesc *pairs,
                            const WorkItem *items, int n, float *arena)
{
    if (n <= 0) return hipSuccess;
    if (T == 32) return single_strand ? ext_T<32, true>(st, pairs, items, n, arena) : ext_T<32, false>(st, pairs, items, n, arena);
    if (T == 64) return single_strand ? ext_T<64, true>(st, pairs, items, n, arena) : ext_T<64, false>(st, pairs, items, n, arena);
    return hipErrorInvalidValue;
}

template <int T, bool S>
static hipError_t intra_T(hipStream_t st, const PairDesc *pairs, const WorkItem *items, int n, float *arena,
                          const int *codes, const float *w)
{
    hipLaunchKernelGGL((k_inner_intra<T, S>), dim3(n), dim3(256), 0, st, pairs, items, arena, codes, w);
    return hipGetLastError();
}
hipError_t launch_inner_intra(hipStream_t st, int T, bool single_strand, const PairDesc *pairs,
                              const WorkItem *items, int n, float *arena, const int *codes, const float *w)
{
    if (n <= 0) return hipSuccess;
    if (T == 32) return single_strand ? intra_T<32, true>(st, pairs, items, n, arena, codes, w) : intra_T<32, false>(st, pairs, items, n, arena, codes, w);
    if (T == 64) return single_strand ? intra_T<64, true>(st, pairs, items, n, arena, codes, w) : intra_T<64, false>(st, pairs, items, n, arena, codes, w);
    return hipErrorInvalidValue;
}

hipError_t launch_s1(hipStream_t st, const PairDesc *pairs, int npairs, float *arena, const int *codes,
                     const float *w)
{
    if (npairs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_s1, dim3(npairs), dim3(256), 0, st, pairs, arena, codes, w);
    return hipGetLastError();
}

hipError_t launch_fill(hipStream_t st, float *p, long long n, float v)
{
    if (n <= 0) return hipSuccess;
    long long blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t launch_to_packed(hipStream_t st, const float *arena, const PairDesc &P, float *out)
{
    hipLaunchKernelGGL(k_to_packed, dim3(P.N, (unsigned)tri_ll(P.M)), dim3(256), 0, st, arena, P, out);
    return hipGetLastError();
}
hipError_t launch_to_dense(hipStream_t st, const float *arena, const PairDesc &P, float *out)
{
    hipLaunchKernelGGL(k_to_dense, dim3(P.N, (unsigned)(P.M * P.M)), dim3(256), 0, st, arena, P, out);
    return hipGetLastError();
}
hipError_t launch_scores(hipStream_t st, const PairDesc *pairs, int npairs, const float *arena, float *scores)
{
    if (npairs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_scores, dim3((npairs + 63) / 64), dim3(64), 0, st, pairs, npairs, arena, scores);
    return hipGetLastError();
}

hipError_t launch_r0tri(hipStream_t st, int T, int nb, const int2 *tiles, int ntiles, int batch, const float *A,
                        const float *B, float *C)
{
    if (ntiles <= 0 || batch <= 0) return hipSuccess;
    if (T == 64) {
        hipLaunchKernelGGL(k_r0tri<64>, dim3(ntiles, batch), dim3(Tc<64>::NT), 0, st, nb, tiles, A, B, C);
    } else if (T == 32) {
        hipLaunchKernelGGL(k_r0tri<32>, dim3(ntiles, batch), dim3(Tc<32>::NT), 0, st, nb, tiles, A, B, C);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
hipError_t launch_dense_to_tiled(hipStream_t st, int n, int nb, int T, int batch, const float *src, float *dst)
{
    hipLaunchKernelGGL(k_dense_to_tiled, dim3(nb * T, batch), dim3(256), 0, st, n, nb, T, src, dst);
    return hipGetLastError();
}
hipError_t launch_tiled_to_dense(hipStream_t st, int n, int nb, int T, int batch, const float *src, float *dst)
{
    hipLaunchKernelGGL(k_tiled_to_dense, dim3(n, batch), dim3(256), 0, st, n, nb, T, src, dst);
    return hipGetLastError();
}

} // namespace bpx
