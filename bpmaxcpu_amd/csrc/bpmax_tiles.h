// bpmax_tiles.h -- device-side tile machinery shared by the BPMax kernels.
//
// Storage: every inner triangle (an outer triangle's N x N upper triangle, or S2) is kept
// as square T x T storage tiles, row-major, in row-major upper-tile order (see
// bpmax_internal.h).  Phases address it through "logical" LT x LT tiles (LT divides T):
// logical tile (X,Y) is the LT x LT block at storage tile (X/RS, Y/RS), offset
// ((X%RS)*LT, (Y%RS)*LT), row stride T, RS = T/LT.  R0 runs on LT = T (64 x 64, 4 x 4
// register tiles); the inner closure (cross-tile R1/R2, in-tile wavefront, S2) runs on
// LT = 32, which halves the wavefront's dependent steps and quarters its per-step work.
#pragma once
#include <hip/hip_runtime.h>
#include "bpmax_internal.h"

namespace bpx {

#define NEG_INF (-__builtin_inff())

constexpr int kThreads = 256; // every phase body runs on 256 lanes (4 waves)

// Workgroups go to the 8 XCDs round-robin in linear dispatch order and every XCD has its
// own L2.  Work index of linear block b out of n: XCD x = b % 8 gets the contiguous range
// starting at x*q + min(x, r) (q = n/8, r = n%8), so neighbouring work items -- the tiles
// of one triangle, which read the same source tiles -- share one L2.  A bijection on [0,n).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_work_index(int b, int n)
{
    const int x = b % kXcds, l = b / kXcds, q = n / kXcds, r = n % kXcds;
    return x * q + (x < r ? x : r) + l;
}

// Product tiles: each of the 4 waves owns one QS x QS quadrant of the LT x LT output tile
// (wave w: rows (w/2)*QS.., columns (w%2)*QS..), each lane an R x R block of it.  On the
// diagonal steps of a triangular product a quadrant's k-range shrinks (see quad_k_range).
template <int LT>
struct Tc {
    static constexpr int R = LT / 16;   // register tile edge: each lane owns R x R cells
    static constexpr int QS = LT / 2;   // quadrant edge
    static constexpr int QL = QS / R;   // lanes per quadrant row
    static constexpr int LD = LT + 4;   // LDS row stride of product operands (16-byte rows)
    static constexpr int PROD_FLOATS = 2 * LT * LD;
};

// The staging and product helpers below run on 256 lanes; `tid` is the lane's index among
// them (threadIdx.x, or its index within a 256-lane half of a wider workgroup).

// First row r0 and column c0 of this lane's R x R block.
template <int LT>
__device__ __forceinline__ void lane_block(int &r0, int &c0, int tid = threadIdx.x)
{
    constexpr int R = Tc<LT>::R, QS = Tc<LT>::QS, QL = Tc<LT>::QL;
    const int w = tid / 64, l = tid % 64;
    r0 = (w >> 1) * QS + (l / QL) * R;
    c0 = (w & 1) * QS + (l % QL) * R;
}

// Useful k-range [kb, ke) of this wave's quadrant in one product step:
//   adiag: A is a diagonal tile (A[r][k] = -inf for k < r): rows >= (w/2)*QS need k >= that;
//   bdiag: B is a diagonal tile (B[k][c] = -inf for k > c): columns < (w%2+1)*QS need k below;
//   lastcol: only column LT-1 of A is finite.
// Multiples of QS; wave-uniform.
template <int LT>
__device__ __forceinline__ void quad_k_range(bool adiag, bool bdiag, bool lastcol, int &kb, int &ke,
                                             int tid = threadIdx.x)
{
    constexpr int QS = Tc<LT>::QS;
    const int w = __builtin_amdgcn_readfirstlane(tid / 64);
    kb = lastcol ? LT - QS : (adiag ? (w >> 1) * QS : 0);
    ke = bdiag ? ((w & 1) + 1) * QS : LT;
}

// Balanced form of quad_k_range for the product steps of a 256-lane workgroup.  On a diagonal
// step the quadrants' ranges differ (K = I: waves 0,1 2 QS, waves 2,3 QS; K = J: waves 1,3
// 2 QS, waves 0,2 QS; both: QS, 2 QS, 0, QS), and the step's barrier makes every wave wait
// for the longest.  Here a lighter wave also runs a slice [hb, he) of a heavier wave's range
// on that wave's cells: wave 2 helps wave 1, wave 3 wave 0, wave 0 wave 3 (a fixed target
// per wave, so one extra accumulator serves all steps of an item), and no wave runs more than
// 1.5 QS per step (QS when K = I = J).  Own range [kb, ke); ranges are multiples of QS/2.
__device__ __forceinline__ int balance_target(int w) { return w == 0 ? 3 : (w == 2 ? 1 : (w == 3 ? 0 : -1)); }
template <int LT>
__device__ __forceinline__ void balanced_k_range(bool adiag, bool bdiag, int w, int &kb, int &ke, int &hb, int &he)
{
    constexpr int QS = Tc<LT>::QS, H = QS / 2;
    kb = 0, ke = LT, hb = 0, he = 0;
    if (adiag && bdiag) {
        kb = w == 0 || w == 2 ? 0 : QS;
        ke = w == 0 ? QS : (w == 2 ? 0 : LT);
        if (w == 2) he = QS;
    } else if (adiag) {
        kb = w < 2 ? H : QS;
        if (w >= 2) he = H;
    } else if (bdiag) {
        ke = w == 0 || w == 2 ? QS : LT - H;
        if (w == 0 || w == 2) hb = LT - H, he = LT;
    }
}

// End of a balanced product: each helper wave hands its slice (acc2, on the cells of wave
// balance_target(w)) to that wave through LDS (R*R*64 floats per target wave, 4 regions), and
// the target max-merges it.  `x` must be free LDS; every lane of the workgroup calls this.
template <int LT>
__device__ __forceinline__ void balance_merge(float (&acc)[Tc<LT>::R][Tc<LT>::R],
                                              const float (&acc2)[Tc<LT>::R][Tc<LT>::R], float *x, int w)
{
    constexpr int R = Tc<LT>::R, RG = R * R * 64;
    const int l = threadIdx.x % 64, hw = balance_target(w);
    __syncthreads();
    if (hw >= 0) {
#pragma unroll
        for (int i = 0; i < R; ++i)
#pragma unroll
            for (int j = 0; j < R; ++j) x[hw * RG + i * R * 64 + l * R + j] = acc2[i][j];
    }
    __syncthreads();
    if (w != 2) { // waves 0, 1 and 3 are targets
#pragma unroll
        for (int i = 0; i < R; ++i)
#pragma unroll
            for (int j = 0; j < R; ++j) acc[i][j] = fmaxf(acc[i][j], x[w * RG + i * R * 64 + l * R + j]);
    }
}

// Storage tile (I,J) (runtime T), for the layout conversions.
__device__ __forceinline__ const float *tile_at(const float *tri, int I, int J, int nb, int T)
{
    return tri + utri_ll(I, J, nb) * (long long)(T * T);
}
__device__ __forceinline__ float *tile_at(float *tri, int I, int J, int nb, int T)
{
    return tri + utri_ll(I, J, nb) * (long long)(T * T);
}

__device__ __forceinline__ float *tri_at(float *arena, const PairDesc &P, int i1, int j1)
{
    return arena + P.F + utri_ll(i1, j1, P.M) * P.tstride;
}

// Logical LT x LT tile (X,Y) of a triangle with nb storage tiles per side (X <= Y).
template <int T, int LT>
__device__ __forceinline__ float *ltile(const float *tri, int X, int Y, int nb)
{
    constexpr int RS = T / LT;
    return (float *)tri + utri_ll(X / RS, Y / RS, nb) * (long long)(T * T) + (X % RS) * LT * T + (Y % RS) * LT;
}

// Tile staging in two phases so that global-memory latency overlaps compute: fetch_*
// issues every global load of a tile into registers (unrolled, branch-free: an absent
// source tile is read from a valid address and replaced by -inf), commit_* writes the
// registers to LDS.  PER = float4 loads per lane per tile.
template <int LT>
struct Stage {
    static constexpr int Q = LT / 4, PER = LT * Q / kThreads;
};

// A operand: As[k][r] = X[I*LT + r][K*LT + k + cshift] (cshift 0 or -1), stored k-major
// for the product's vector row reads.  -inf where the source tile lies below the diagonal;
// with lastcol_only every column but k = LT-1 is -inf.  Lane mapping: within a wave, 16
// consecutive rows x 4 consecutive float4 column chunks, which makes the transposed LDS
// stores conflict-free and keeps 64-byte row runs in each global load.
template <int LT>
struct FetchA {
    float4 v[Stage<LT>::PER];
    float prev;          // X[I*LT + r][K*LT - 1] for lane r < LT (cshift = -1)
    bool have, have_prev;
};

template <int LT>
__device__ __forceinline__ void a_coords(int p, int &r, int &c4, int tid)
{
    const int idx = tid + p * kThreads;
    const int B = idx / 64, l = idx % 64;
    r = 16 * (B % (LT / 16)) + l % 16;
    c4 = 4 * (4 * (B / (LT / 16)) + l / 16);
}

template <int T, int LT>
__device__ __forceinline__ void fetch_A(FetchA<LT> &f, const float *X, int nb, int I, int K, int cshift,
                                        int tid = threadIdx.x)
{
    const int nbl = nb * (T / LT);
    f.have = K >= I && K < nbl;
    const float *t = ltile<T, LT>(X, I, f.have ? K : I, nb);
#pragma unroll
    for (int p = 0; p < Stage<LT>::PER; ++p) {
        int r, c4;
        a_coords<LT>(p, r, c4, tid);
        f.v[p] = *(const float4 *)(t + r * T + c4);
    }
    f.have_prev = cshift < 0 && K - 1 >= I;
    f.prev = NEG_INF;
    if (cshift < 0) {
        const float *tp = ltile<T, LT>(X, I, f.have_prev ? K - 1 : I, nb);
        f.prev = tp[(tid % LT) * T + LT - 1];
    }
}

template <int LT>
__device__ __forceinline__ void commit_A(float *As, const FetchA<LT> &f, int cshift, bool lastcol_only,
                                         int tid = threadIdx.x)
{
    constexpr int LD = Tc<LT>::LD;
#pragma unroll
    for (int p = 0; p < Stage<LT>::PER; ++p) {
        int r, c4;
        a_coords<LT>(p, r, c4, tid);
        const float e[4] = {f.v[p].x, f.v[p].y, f.v[p].z, f.v[p].w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int k = c4 + q - cshift;
            float val = f.have ? e[q] : NEG_INF;
            if (lastcol_only && k != LT - 1) val = NEG_INF;
            if (k < LT) As[k * LD + r] = val;
        }
    }
    if (cshift < 0 && tid < LT) As[tid] = (f.have_prev && !lastcol_only) ? f.prev : NEG_INF;
}

__device__ __forceinline__ float sel(bool ok, float v) { return ok ? v : NEG_INF; }
__device__ __forceinline__ float4 pick(bool first, const float4 &a, const float4 &b)
{
    return make_float4(first ? a.x : b.x, first ? a.y : b.y, first ? a.z : b.z, first ? a.w : b.w);
}

// B operand: Bs[k][c] = X[K*LT + k + rshift][J*LT + c] (rshift 0 or +1); rows of tiles
// below the diagonal (or past the last tile) are -inf.
template <int LT>
struct FetchB {
    float4 v[Stage<LT>::PER];
    unsigned have; // bit p: row of chunk p exists
};

template <int T, int LT>
__device__ __forceinline__ void fetch_B(FetchB<LT> &f, const float *X, int nb, int K, int J, int rshift,
                                        int tid = threadIdx.x)
{
    constexpr int Q = Stage<LT>::Q;
    const int nbl = nb * (T / LT);
    f.have = 0;
#pragma unroll
    for (int p = 0; p < Stage<LT>::PER; ++p) {
        const int idx = tid + p * kThreads;
        const int k = idx / Q, c4 = (idx % Q) * 4;
        const int x = K * LT + k + rshift;
        const int Xt = x / LT;
        const bool ok = Xt <= J && Xt < nbl;
        f.have |= ok ? (1u << p) : 0u;
        const float *t = ltile<T, LT>(X, ok ? Xt : J, J, nb);
        f.v[p] = *(const float4 *)(t + (ok ? x % LT : 0) * T + c4);
    }
}

template <int LT>
__device__ __forceinline__ void commit_B(float *Bs, const FetchB<LT> &f, int tid = threadIdx.x)
{
    constexpr int LD = Tc<LT>::LD, Q = Stage<LT>::Q;
#pragma unroll
    for (int p = 0; p < Stage<LT>::PER; ++p) {
        const int idx = tid + p * kThreads;
        const int k = idx / Q, c4 = (idx % Q) * 4;
        // select per component: a select of whole float4 values is lowered through scratch
        const bool ok = (f.have >> p) & 1u;
        *(float4 *)(Bs + k * LD + c4) = make_float4(sel(ok, f.v[p].x), sel(ok, f.v[p].y), sel(ok, f.v[p].z),
                                                    sel(ok, f.v[p].w));
    }
}

// v_max3_f32 as one instruction.  Written out because the compiler regroups chains of
// maxima into v_max_f32 + v_max3_f32 pairs, 1.75 instead of 1.5 VALU per cell update.
__device__ __forceinline__ float max3(float a, float b, float c)
{
    float r;
    asm("v_max3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// acc[i][j] = max(acc[i][j], max_{kb<=k<ke} As[k][r0+i] + Bs[k][c0+j]) (kb, ke multiples
// of KU, by default QS): two k per round, v_add_f32 x2 + v_max3_f32 per pair of cell updates.
template <int LT, int KU = Tc<LT>::QS>
__device__ __forceinline__ void product(float (&acc)[Tc<LT>::R][Tc<LT>::R], const float *As, const float *Bs,
                                        int r0, int c0, int kb, int ke)
{
    constexpr int LD = Tc<LT>::LD, R = Tc<LT>::R;
    for (int k0 = kb; k0 < ke; k0 += KU) {
#pragma unroll
        for (int k = k0; k < k0 + KU; k += 2) {
            float av0[R], av1[R], bv0[R], bv1[R];
#pragma unroll
            for (int i = 0; i < R; ++i) {
                av0[i] = As[k * LD + r0 + i];
                av1[i] = As[(k + 1) * LD + r0 + i];
                bv0[i] = Bs[k * LD + c0 + i];
                bv1[i] = Bs[(k + 1) * LD + c0 + i];
            }
#pragma unroll
            for (int i = 0; i < R; ++i)
#pragma unroll
                for (int j = 0; j < R; ++j)
                    acc[i][j] = max3(acc[i][j], av0[i] + bv0[j], av1[i] + bv1[j]);
        }
    }
}

// Max-combine the lane's R x R cells into tile `dst` (row stride T) -- every defined cell's
// final value is >= 0, where float order equals the signed-integer order of the bit
// patterns -- or store them.
template <int T, int LT>
__device__ __forceinline__ void store_acc(float *dst, const float (&acc)[Tc<LT>::R][Tc<LT>::R], int r0, int c0,
                                          bool atomic)
{
    constexpr int R = Tc<LT>::R;
#pragma unroll
    for (int i = 0; i < R; ++i) {
        float *row = dst + (r0 + i) * T + c0;
#pragma unroll
        for (int j = 0; j < R; ++j) {
            if (atomic) atomicMax((int *)(row + j), __float_as_int(acc[i][j]));
            else row[j] = acc[i][j];
        }
    }
}

// Max over aligned groups of G consecutive lanes: quad permutes on the DPP path (no LDS)
// for the first two levels, a lane shuffle beyond.
template <int G>
__device__ __forceinline__ float lane_group_max(float v)
{
    if (G >= 2) v = fmaxf(v, __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false)));
    if (G >= 4) v = fmaxf(v, __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, false)));
    // row_half_mirror: lane i <-> 7-i inside each half-row, pairs the two quads of 8 lanes
    if (G >= 8) v = fmaxf(v, __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x141, 0xF, 0xF, false)));
#pragma unroll
    for (int off = 8; off < G; off <<= 1) v = fmaxf(v, __shfl_xor(v, off));
    return v;
}

// LDS layout of the in-tile closure on an LT x LT tile, in this order: a guard of L
// floats, Sa (LT rows of stride L = LT+4: the S2 tile (I,I)), Sb (LT+1 rows: S2 (J,J) and a
// pad row), a guard row of LX = 2*LT+4 floats, Xs (LT+1 rows of stride LX: the tile and a
// pad row; columns LT.. of every row are pads), then the boundary row/column/corner, the
// weight table and the codes.  Everything before BELOW starts at -inf (pads and guards)
// and the tiles are stored over it.  About 19 KB at LT = 32, so that a 512-lane closure
// workgroup fits on a CU beside four R0 workgroups.
template <int LT>
struct Intra {
    static constexpr int L = LT + 4, LX = 2 * LT + 4;
    static constexpr int SA = L, SB = SA + LT * L, XS = SB + (LT + 1) * L + LX;
    static constexpr int BELOW = XS + (LT + 1) * LX, FILL_END = BELOW;
    static constexpr int WL = BELOW + 2 * LT + 4, CODES = WL + 28;
    static constexpr int FLOATS = CODES + 2 * LT;
};

// acc = max(acc, x, y) in place, one v_max3_f32.
__device__ __forceinline__ void max3_into(float &acc, float x, float y)
{
    asm("v_max3_f32 %0, %0, %1, %2" : "+v"(acc) : "v"(x), "v"(y));
}

constexpr int cmax(int a, int b) { return a > b ? a : b; }
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// ------------------------------------------------------- direct-to-LDS staging (64 x 64 R0)
// The 64 x 64 R0 product steps stage their operands with global_load_lds_dwordx4: one
// wave-wide load writes 16 bytes per lane straight into LDS at a wave-uniform base + 16 *
// lane, so tiles reach LDS without passing through VGPRs or a commit pass.  A 64-deep K
// step runs as two 32-deep halves on two LDS stages: the loads of half h+1 fly while half h
// computes, one barrier per half.
//
// Stage layout (floats):
//   A   64 rows x 8 float4: A[r][32*kh + 4*k4 .. +3] at float4 slot r*8 + (k4 ^ ((r >> 2) & 7))
//       (row-major 32-wide halves; k4 swizzled per 4-row group, so that the eight 4-row
//       groups a wave reads at one k4 fall in eight different bank groups)
//   B   33 rows x 64: rows 32*kh .. 32*kh+32 of the unshifted B operand, row-major (the
//       33rd row serves the k2+1 shift of R0's last k of the half)
//   S1  4 floats: scalar terms that travel with the stage
struct DlStage {
    static constexpr int A = 0, B = 64 * 32, S1 = B + 33 * 64, FLOATS = S1 + 4;
};
static_assert(DlStage::FLOATS % 4 == 0, "stages stay 16-byte aligned");

// A row of -inf in device memory: the source of staged rows that lie outside the triangle.
static __device__ __attribute__((aligned(256))) const unsigned kNegInfRow[64] = {
#define BPX_NI4 0xFF800000u, 0xFF800000u, 0xFF800000u, 0xFF800000u
#define BPX_NI16 BPX_NI4, BPX_NI4, BPX_NI4, BPX_NI4
    BPX_NI16, BPX_NI16, BPX_NI16, BPX_NI16
#undef BPX_NI16
#undef BPX_NI4
};

__device__ __forceinline__ unsigned lds_offset(const void *p)
{
    // the low 32 bits of a generic address in the LDS aperture are the LDS byte offset
    return __builtin_amdgcn_readfirstlane((unsigned)(size_t)p);
}

// 16 (g2lds16) or 4 (g2lds4) bytes per active lane from global address g (per lane) into LDS
// at lds + lane * size (lds wave-uniform).  Written as asm: with the compiler's builtin it
// orders every later LDS read of the wave behind these loads (it cannot tell the two stages
// apart), which serialises the prefetch with the product.  These loads have no register
// result; a stage is published by vm_wait_all() and a barrier.
// M0 is reserved to the compiler, which ignores it in a clobber list: these asm blocks are
// correct only while no compiler-generated code of the file uses M0 (tests/test_abi.py
// compiles the kernels to assembly and checks that every use of M0 is one of these pairs).
__device__ __forceinline__ void g2lds16(const void *g, const void *lds)
{
    asm volatile("s_mov_b32 m0, %1\n\tglobal_load_lds_dwordx4 %0, off" : : "v"(g), "s"(lds_offset(lds)) : "memory");
}
__device__ __forceinline__ void g2lds4(const void *g, const void *lds)
{
    asm volatile("s_mov_b32 m0, %1\n\tglobal_load_lds_dword %0, off" : : "v"(g), "s"(lds_offset(lds)) : "memory");
}
__device__ __forceinline__ void vm_wait_all() { asm volatile("s_waitcnt vmcnt(0)" : : : "memory"); }

// Stage the A half [32*kh, 32*kh+32) of the 64 x 64 tile At (row stride 64): wave w loads
// rows 16w .. 16w+15, lane l row 16w + 8c + l/8 at swizzled slot l%8.
__device__ __forceinline__ void dl_stage_A(float *stage, const float *At, int kh, int w, int l)
{
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int r = 16 * w + 8 * c + l / 8, k4 = (l % 8) ^ ((r >> 2) & 7);
        g2lds16(At + r * 64 + 32 * kh + 4 * k4, stage + DlStage::A + (2 * w + c) * 256);
    }
}

// Stage rows 32*kh + q, q in [8w, 8w+8), of the 64 x 64 tile Bt: lane l row 8w + 4c + l/16,
// columns 4*(l%16) ..
__device__ __forceinline__ void dl_stage_B32(float *stage, const float *Bt, int kh, int w, int l)
{
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int q = 8 * w + 4 * c + l / 16;
        g2lds16(Bt + (32 * kh + q) * 64 + 4 * (l % 16), stage + DlStage::B + (8 * w + 4 * c) * 64);
    }
}

// acc[i][j] max= A[r0+i][k] + B[k+RS][c0+j] (stage rows) over the 16 k of the stage's half
// from kq (0 or 16): per 4 k, four A float4 (the lane's rows) and four B float4 (its
// columns) for 64 updates, v_add_f32 x 2 + v_max3_f32 per pair.
template <int RS>
__device__ __forceinline__ void dl_prod16(float (&acc)[4][4], const float *stage, int r0, int c0, int kq)
{
    const float4 *A = (const float4 *)(stage + DlStage::A) + r0 * 8;
    const float *B = stage + DlStage::B + (kq + RS) * 64 + c0;
    const int sw = ((r0 >> 2) & 7) ^ (kq >> 2);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        float4 a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = A[i * 8 + (sw ^ u)];
#pragma unroll
        for (int v = 0; v < 4; ++v) b[v] = *(const float4 *)(B + (4 * u + v) * 64);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float av[4] = {a[i].x, a[i].y, a[i].z, a[i].w};
            const float b0[4] = {b[0].x, b[0].y, b[0].z, b[0].w}, b1[4] = {b[1].x, b[1].y, b[1].z, b[1].w};
            const float b2[4] = {b[2].x, b[2].y, b[2].z, b[2].w}, b3[4] = {b[3].x, b[3].y, b[3].z, b[3].w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                acc[i][j] = max3(acc[i][j], av[0] + b0[j], av[1] + b1[j]);
                acc[i][j] = max3(acc[i][j], av[2] + b2[j], av[3] + b3[j]);
            }
        }
    }
}

// dl_prod16 over [lo, hi) (multiples of 16) clipped to the stage's half [32*kh, 32*kh+32).
template <int RS>
__device__ __forceinline__ void dl_product(float (&acc)[4][4], const float *stage, int r0, int c0, int kh, int lo,
                                           int hi)
{
    lo = lo > 32 * kh ? lo : 32 * kh;
    hi = hi < 32 * kh + 32 ? hi : 32 * kh + 32;
    for (int k = lo; k < hi; k += 16) dl_prod16<RS>(acc, stage, r0, c0, k - 32 * kh);
}

// ------------------------------------------------- FP16 operands (64 x 64 R0, exact scores)
// When every score of a pair is an integer multiple of 2^-p below 2^(11-p), FP16 holds all
// of them exactly, every R0 sum (bounded by the cell's final value) is exact, and max-plus
// runs on packed halves: per two k and two columns, v_pk_add_f16 x 2 (the A value broadcast
// to both halves by op_sel) + v_pk_maximum3_f16 -- 4 cell updates in 3 VALU instructions
// against 2 in 3 for FP32, with half the LDS and HBM bytes per operand.  The operands come
// from the FP16 mirror of F (PairDesc::Fh), same tiled layout, 64 halves per tile row.
//
// Stage layout (bytes), one whole 64-deep K step (FP16 halves the bytes of a step, so two
// such stages take the LDS of two 32-deep FP32 ones: one barrier and one load wait per step):
//   A  4 blocks (block w: rows 16w .. 16w+15, staged by wave w) of 16 rows x 8 chunks of 16 B
//      (8 consecutive k); block w starts at w*ABLK (2112: each block 64 B further into the
//      256 B bank window), chunk c of block row rr at slot 8*rr + (c ^ ((rr >> 2) & 3)), so the
//      eight 4-row groups one wave reads for a chunk fall in eight different bank groups;
//   B  65 rows x 128 B: rows 0 .. 63 of the unshifted B tile and, as row 64, row 0 of the
//      next tile down (R0's k2+1 shift), row-major;
//   S1 16 B: the scalar R3/R4 terms that travel with the stage (FP32).
typedef unsigned short h16bits;
struct HStage {
    static constexpr int ABLK = 2112, A = 0, B = 4 * ABLK, S1 = B + 65 * 128, BYTES = S1 + 16;
};
static_assert(HStage::BYTES % 16 == 0, "stages stay 16-byte aligned");

// A row of FP16 -inf in device memory (staged rows that lie outside the triangle).
static __device__ __attribute__((aligned(256))) const unsigned kNegInfRowH[32] = {
#define BPX_HI4 0xFC00FC00u, 0xFC00FC00u, 0xFC00FC00u, 0xFC00FC00u
    BPX_HI4, BPX_HI4, BPX_HI4, BPX_HI4, BPX_HI4, BPX_HI4, BPX_HI4, BPX_HI4
#undef BPX_HI4
};

// (a.lo + b.lo, a.lo + b.hi) and (a.hi + b.lo, a.hi + b.hi): A's k (resp. k+1) value broadcast
// against two B columns.
__device__ __forceinline__ unsigned pk_add_lo(unsigned a, unsigned b)
{
    unsigned r;
    asm("v_pk_add_f16 %0, %1, %2 op_sel_hi:[0,1]" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ unsigned pk_add_hi(unsigned a, unsigned b)
{
    unsigned r;
    asm("v_pk_add_f16 %0, %1, %2 op_sel:[1,0] op_sel_hi:[1,1]" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ unsigned pk_add(unsigned a, unsigned b)
{
    unsigned r;
    asm("v_pk_add_f16 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
// acc = max(acc, x, y) on both halves, one v_pk_maximum3_f16 (no NaN ever occurs here)
__device__ __forceinline__ void pk_max3_into(unsigned &acc, unsigned x, unsigned y)
{
    asm("v_pk_maximum3_f16 %0, %0, %1, %2" : "+v"(acc) : "v"(x), "v"(y));
}
__device__ __forceinline__ unsigned pk_pack(float lo, float hi)
{
    return __builtin_bit_cast(unsigned, __builtin_amdgcn_cvt_pkrtz(lo, hi));
}
__device__ __forceinline__ float pk_lo(unsigned v) { return (float)__builtin_bit_cast(_Float16, (h16bits)(v & 0xFFFFu)); }
__device__ __forceinline__ float pk_hi(unsigned v) { return (float)__builtin_bit_cast(_Float16, (h16bits)(v >> 16)); }

// Stage the 64 x 64 FP16 tile At as the A operand: wave w stages rows 16w .. 16w+15, two
// loads of 8 rows each; lane l of load q takes row 16w + 8q + l/8 at slot 64q + l (its chunk is
// l%8 XOR the row's swizzle).
__device__ __forceinline__ void h_stage_A(char *st, const h16bits *At, int w, int l)
{
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int rr = 8 * q + (l >> 3), c = (l & 7) ^ ((rr >> 2) & 3);
        g2lds16(At + (16 * w + rr) * 64 + 8 * c, st + HStage::A + w * HStage::ABLK + 1024 * q);
    }
}
// Stage rows 16w .. 16w+15 of the FP16 tile Bt (B rows 0 .. 63): lane l of load q row
// 16w + 8q + l/8, chunk l%8.
__device__ __forceinline__ void h_stage_B(char *st, const h16bits *Bt, int w, int l)
{
#pragma unroll
    for (int q = 0; q < 2; ++q)
        g2lds16(Bt + (16 * w + 8 * q + (l >> 3)) * 64 + 8 * (l & 7), st + HStage::B + (2 * w + q) * 1024);
}

// Per-lane A addressing of rows r0 .. r0+3 (r0 a multiple of 4, so the four rows lie in one
// 16-row block, 128 B apart, and share one swizzle): row r0's base and 16 x its swizzle.
// Chunk c of row r0+i is at base + 128 i + ((16 c) ^ sw16).
struct HRows {
    int base, sw16;
};
__device__ __forceinline__ HRows h_rows(int r0)
{
    const int rr = r0 & 15;
    return HRows{HStage::A + (r0 >> 4) * HStage::ABLK + 128 * rr, 16 * ((rr >> 2) & 3)};
}

// acc[i][p] (columns c0+2p, c0+2p+1 packed) max= A[r0+i][k] + B[k+RS][c] over the 8 k from
// kq (a multiple of 8): four 16 B A reads, eight 8 B B reads, 96 VALU for 128 cell updates.
// The four sums of a row's two accumulators are formed before either maximum, so that no
// maximum reads a sum of the instruction just before it (that would cost a wait state).
template <int RS>
__device__ __forceinline__ void h_prod8(unsigned (&acc)[4][2], const char *st, const HRows &rows, int c0, int kq)
{
    const char *ar = st + rows.base + ((2 * kq) ^ rows.sw16); // chunk kq / 8
    uint4 a[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) a[i] = *(const uint4 *)(ar + 128 * i);
    uint2 b[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) b[j] = *(const uint2 *)(st + HStage::B + (kq + RS + j) * 128 + 2 * c0);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const unsigned av = m == 0 ? a[i].x : m == 1 ? a[i].y : m == 2 ? a[i].z : a[i].w;
            const unsigned t0 = pk_add_lo(av, b[2 * m].x), t1 = pk_add_hi(av, b[2 * m + 1].x);
            const unsigned t2 = pk_add_lo(av, b[2 * m].y), t3 = pk_add_hi(av, b[2 * m + 1].y);
            pk_max3_into(acc[i][0], t0, t1);
            pk_max3_into(acc[i][1], t2, t3);
        }
    }
}

// h_prod8 over [lo, hi) (multiples of 8).
template <int RS>
__device__ __forceinline__ void h_product(unsigned (&acc)[4][2], const char *st, const HRows &rows, int c0, int lo,
                                          int hi)
{
    for (int k = lo; k < hi; k += 8) h_prod8<RS>(acc, st, rows, c0, k);
}

// ------------------------------------------ FP16 cross-tile steps (32 x 32 logical tiles)
// One step of R1/R2 across logical tiles with the S2 operand taken from its shifted FP16 copy
// (PairDesc::S2r / S2c), so that both operands are plain logical tiles.  Stage of one step
// (bytes): A two blocks of 16 rows x 4 chunks of 16 B (8 k), block b at b*ABLK (1088: the
// second block 64 B further into the 256 B bank window), chunk c of block row rr at slot
// 4*rr + (c ^ ((rr >> 2) & 3)); B 32 rows x 64 B.  Waves 0 and 1 stage A's blocks, waves 2
// and 3 B's halves, one 16 B load per lane.
struct XStage {
    static constexpr int ABLK = 1088, A = 0, B = 2 * ABLK, BYTES = B + 32 * 64;
};

__device__ __forceinline__ void x_stage(char *st, const h16bits *At, const h16bits *Bt, int w, int l)
{
    if (w < 2) {
        const int rr = l >> 2, c = (l & 3) ^ ((rr >> 2) & 3);
        g2lds16(At + (16 * w + rr) * 64 + 8 * c, st + XStage::A + w * XStage::ABLK);
    } else {
        g2lds16(Bt + (16 * (w - 2) + (l >> 2)) * 64 + 8 * (l & 3), st + XStage::B + (w - 2) * 1024);
    }
}

// acc[i][p] (columns c0+2p, c0+2p+1) max= A[r0+i][k] + B[k][c] over the 8 k of chunk kc.
__device__ __forceinline__ void x_prod8(unsigned (&acc)[4][2], const char *st, int r0, int c0, int kc)
{
    uint4 a[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int r = r0 + i, rr = r & 15;
        a[i] = *(const uint4 *)(st + XStage::A + (r >> 4) * XStage::ABLK + 64 * rr + 16 * (kc ^ ((rr >> 2) & 3)));
    }
    uint2 b[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) b[j] = *(const uint2 *)(st + XStage::B + (8 * kc + j) * 64 + 2 * c0);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const unsigned av = m == 0 ? a[i].x : m == 1 ? a[i].y : m == 2 ? a[i].z : a[i].w;
            const unsigned t0 = pk_add_lo(av, b[2 * m].x), t1 = pk_add_hi(av, b[2 * m + 1].x);
            const unsigned t2 = pk_add_lo(av, b[2 * m].y), t3 = pk_add_hi(av, b[2 * m + 1].y);
            pk_max3_into(acc[i][0], t0, t1); // (sums first: see h_prod8)
            pk_max3_into(acc[i][1], t2, t3);
        }
    }
}
// at most one (two) of the wave's vector-memory operations outstanding: with loads completing
// in order, every load but the wave's last one (two) has landed
__device__ __forceinline__ void vm_wait_le1() { asm volatile("s_waitcnt vmcnt(1)" : : : "memory"); }
__device__ __forceinline__ void vm_wait_le2() { asm volatile("s_waitcnt vmcnt(2)" : : : "memory"); }

} // namespace bpx
