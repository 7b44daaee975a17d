// bpmax_internal.h -- structures shared by the host engine and the HIP kernels.
//
// Device layout of one sequence pair ("resident pair"):
//   * F: tri(M) outer triangles (i1,j1) in row-major upper order over M; each triangle
//     holds tri(nb) square T x T tiles (I,J), I<=J, in row-major upper order over nb =
//     ceil(N/T); each tile is row-major.  Inner index x = i2 or j2 lives in tile x/T at
//     offset x%T; indices N..nb*T-1 are padding.  Cells below the diagonal inside the
//     diagonal tiles hold -inf, which makes every tile product mask itself.
//   * S2: one more triangle in the same tiled layout (strand-2 single-strand table, over
//     reversed seq2).
//   * S1: dense M x M (strand-1 single-strand table).
//   * codes: seq1[M] and b[nb*T] = reversed seq2, padded with code 4 (N).
#pragma once
#include <cstdint>

namespace bpx {

constexpr int kMaxTile = 64;

struct PairDesc {
    int M, N;          // strand lengths
    int nb;            // tiles per inner side
    int T;             // tile edge (all pairs of one launch share it)
    long long F;       // float offset of triangle (0,0)
    long long tstride; // floats per outer triangle = tri(nb)*T*T
    long long S2;      // float offset of the tiled S2 triangle
    long long S1;      // float offset of dense S1
    long long a;       // int offset of seq1 codes
    long long b;       // int offset of reversed, padded seq2 codes
};

// One output tile of one phase.  Steps index the K-dimension of that tile's max-plus
// reduction (see the kernels for the meaning per phase).
struct WorkItem {
    int pair;
    int i1, j1;        // outer triangle (unused by the S2 phases)
    int I, J;          // inner tile
    int step_begin, step_end;
    int flags;
};

enum : int {
    WI_ATOMIC = 1,     // several items write this tile: combine with atomic max
    WI_EPILOGUE = 2,   // this item also adds init/P1/R3/R4 (outer phase)
};

__host__ __device__ inline long long tri_ll(long long n) { return n * (n + 1) / 2; }
__host__ __device__ inline long long utri_ll(long long i, long long j, long long n)
{
    return i * n - i * (i - 1) / 2 + (j - i);
}

} // namespace bpx
