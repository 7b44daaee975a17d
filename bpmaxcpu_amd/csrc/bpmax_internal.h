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
//   * Fh (optional): F again in FP16, the R0 kernel's operand copy when every score of the
//     pair is an integer multiple of 2^-p below 2^(11-p) (exact in FP16); cells that the
//     closure never writes hold FP16 -inf.
#pragma once
#include <cstdint>

namespace bpx {

constexpr int kMaxTile = 64;

struct PairDesc {
    int M, N;          // strand lengths
    int nb;            // storage tiles per inner side
    int T;             // storage tile edge (all pairs of one launch share it)
    int nbl;           // logical tiles per inner side for the inner phases (edge min(T, 32))
    int nblr;          // logical tiles per side that hold real cells, ceil(N / edge): tiles
                       // (X, Y) with Y >= nblr lie wholly in the padding, which no real cell
                       // reads, and are never closed
    long long F;       // float offset of triangle (0,0)
    long long tstride; // floats per outer triangle = tri(nb)*T*T
    long long S2;      // float offset of the tiled S2 triangle
    long long S1;      // float offset of dense S1
    long long Fh = -1; // float offset of the FP16 mirror of F (same tiled layout, one half
                       // per cell; written when the inner closure finalizes a tile), or -1
    long long S2r = -1; // float offsets of FP16 copies of S2 shifted by one row (S2r[x][y] =
    long long S2c = -1; // S2[x+1][y]) and by one column (S2c[x][y] = S2[x][y-1]), the cross-tile
                        // operands of R1 and R2 (same tiled layout; -inf outside S2), or -1
    long long a;       // int offset of seq1 codes
    long long b;       // int offset of reversed, padded seq2 codes
    long long flag;    // int offset of the pair's dependency counters (task scheduler):
                       //   [0] S1 done, [1] S2 tiles done, then (done, pending) per S2 tile,
                       //   then (done, pending) per F tile (triangle-major, tile-minor)
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
    WI_EDGE = 4,       // outer items: the steps of k1 = i1, then of k1 = j1-1 (the two terms
                       // that read the triangles one outer diagonal below)
    WI_TYPE_SHIFT = 4, // task type (scheduler), 4 bits
    WI_KOFF_SHIFT = 8, // outer items: first K of the item's K range minus I, 8 bits
    WI_KCNT_SHIFT = 16, // outer items: number of K values of the range (0: all of I..J)
};
__host__ __device__ inline int wi_type(int flags) { return (flags >> WI_TYPE_SHIFT) & 15; }
// k1 of step s of an outer item whose steps run nK K values per k1
__host__ __device__ inline int outer_k1(const WorkItem &it, int s, int nK)
{
    const int q = s / nK;
    return (it.flags & WI_EDGE) ? (q ? it.j1 - 1 : it.i1) : it.i1 + q;
}

enum : int { TASK_S1 = 1, TASK_S2EXT, TASK_S2INTRA, TASK_OUTER, TASK_EXT, TASK_INTRA };

// counters of one pair: S1 done, S2 done-count, (done, pending) per logical S2 tile and
// per logical F tile; nbl = logical tiles per inner side
__host__ __device__ inline long long pair_flag_ints(long long M, long long nbl)
{
    return 2 + 2 * (nbl * (nbl + 1) / 2) + 2 * (M * (M + 1) / 2) * (nbl * (nbl + 1) / 2);
}

// edge of the logical tiles the inner phases run on, for storage tile T
__host__ __device__ constexpr int inner_tile(int T) { return T < 32 ? T : 32; }

__host__ __device__ inline long long tri_ll(long long n) { return n * (n + 1) / 2; }
__host__ __device__ inline long long utri_ll(long long i, long long j, long long n)
{
    return i * n - i * (i - 1) / 2 + (j - i);
}

} // namespace bpx
