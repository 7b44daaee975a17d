// bpmax_kernels.hip -- hand-written gfx950 kernels of the BPMax engine.
//
// Max-plus ("tropical") products are not multiply-accumulates, so nothing here uses MFMA:
// the hot loop is VALU v_add_f32 + v_max3_f32 on per-lane register tiles fed from LDS
// with vector reads (1.5 VALU instructions per cell update; packed FP32 ops are disabled
// at compile time).  Scores are exact multiples of 2^-p far below 2^24, so every sum is
// exact and every reduction order gives the same bits.
//
// Phase bodies (reference functions they replace, paths relative to
// TPDS/source/release/MPT_RT_V3/bpmax_complete_otf_ftable/src/):
//   outer_body  R0 double max-plus + init + P1 + R3 + R4 of one 64 x 64 storage tile
//               (bpmax_outer_reductions.c:123-180, matrix_max_plus_section.c:153-257,
//                reg_3_24.c:9-84, bpmax_ftable_init.c:116-215,
//                bpmax_outer_south_west.c:115-160, bpmax_r3_section.c:114-139,
//                bpmax_r4_section.c:114-139)
//   ext_body    cross-tile parts of R1/R2 (F) and of S2's split term on 32 x 32 logical
//               tiles (bpmax_inner_reductions.c:140-178, bpmax_single_strand_s2_tile.c:124-262)
//   intra_body  in-tile dependence wavefront with P2 on 32 x 32 logical tiles
//               (bpmax_inner_reductions_diagonal_tile.c, bpmax_inner_reductions_finalize.c,
//                bpmax_single_strand_diagonal_tile.c, bpmax_single_strand_finalize.c)
//   s1_body     strand-1 single-strand table (bpmax_single_strand_s1.c:115-166)
// They run either as one launch per phase and diagonal (k_outer / k_inner_ext /
// k_inner_intra) or as tasks of the dependency-tile scheduler k_tasks, which replaces the
// reference's per-diagonal fork/join (src/bpmax.c:283-324) by per-tile ready counters.
#include <hip/hip_runtime.h>
#include <cstdint>
#include "bpmax_internal.h"
#include "bpmax_launch.h"
#include "bpmax_tiles.h"

namespace bpx {

// logical tile edge of the inner phases for storage tile T
template <int T>
struct Inner {
    static constexpr int LT = T < 32 ? T : 32;
    static constexpr int RS = T / LT;
};

// ------------------------------------------------------------------------------ outer
// Output storage tile (I,J) of triangle (i1,j1) over K-steps [step_begin, step_end) of the
// item's K range Kb .. Kb+nK-1 (all of I..J by default):
// step s <-> (k1 = i1 + s/nK, K = Kb + s%nK):
//   C[I][J] max= F(i1,k1)[I][K] (x) F(k1+1,j1)[K*T+1 .. K*T+T][J]        (R0, k2+1 shift)
// The same steps carry R3 and R4, whose operands are tiles the R0 steps already stage:
//   K = J: A = F(i1,k1)[I][J]   -> R4  F(i1,k1) + S1(k1+1,j1)
//   K = I: B = F(k1+1,j1)[I][J] shifted down one row; with row I*T staged above it
//          (Bs[-1]) the unshifted tile is Bs[r-1]  -> R3  S1(i1,k1) + F(k1+1,j1)
// so a split of the steps over work items splits R3/R4 with them.  With WI_EPILOGUE also
// S1(i1,j1)+S2 (the inter base case on the diagonal when d1 = 0) and P1.
// T = 64 stages operands directly into two LDS stages (DlStage); T = 32 through registers
// (As, B's extra row, Bs).
template <int T>
struct OuterLds {
    static constexpr int FLOATS = T == 64 ? 2 * DlStage::FLOATS : Tc<T>::PROD_FLOATS + Tc<T>::LD;
};

// Epilogue of an outer item (WI_EPILOGUE): S1(i1,j1) + S2, the inter base case on the
// diagonal when d1 = 0, and P1 = F(i1+1,j1-1) + intra weight when d1 >= 4.
template <int T>
__device__ __forceinline__ void outer_epilogue(const PairDesc &P, int i1, int j1, long long toff, const float *arena,
                                               const int *__restrict__ codes, const float *__restrict__ w,
                                               float (&acc)[Tc<T>::R][Tc<T>::R], int r0, int c0, int I, int J)
{
    constexpr int R = Tc<T>::R;
    const float *S2t = arena + P.S2 + toff;
    const int *a = codes + P.a, *b = codes + P.b;
    const float s1 = arena[P.S1 + i1 * P.M + j1];
    const int d1 = j1 - i1;
    const float w1 = d1 >= 4 ? w[5 * a[i1] + a[j1]] : 0.f;
    const float *Pt = d1 >= 4 ? tri_at((float *)arena, P, i1 + 1, j1 - 1) + toff : S2t;
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const int lr = r0 + i, lc = c0 + j;
            float v = s1 + S2t[lr * T + lc];
            if (d1 == 0 && I * T + lr == J * T + lc) v = w[25 + 5 * a[i1] + b[I * T + lr]]; // inter base
            if (d1 >= 4) v = fmaxf(v, Pt[lr * T + lc] + w1);
            acc[i][j] = fmaxf(acc[i][j], v);
        }
}

// Max-combine a 64 x 64 tile held as 4 x 4 lane blocks into dst (row stride 64) by whole
// rows: the blocks go through LDS (`x`, 64 x 68 floats, free on entry; every lane of the
// workgroup calls this) and each wave instruction then covers 64 consecutive floats of one
// row.  Straight from the lane blocks one instruction touches 8 rows x 8 float4, and the
// memory side pays for every touched line (64x512: 54.4 -> 51.1 ms per solve).
__device__ __forceinline__ void store_rows_atomic(float *dst, const float (&acc)[4][4], int r0, int c0, float *x,
                                                  int wv, int l)
{
    constexpr int LX = 68;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 4; ++i)
        *(float4 *)(x + (r0 + i) * LX + c0) = make_float4(acc[i][0], acc[i][1], acc[i][2], acc[i][3]);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int r = 16 * wv + q;
        atomicMax((int *)(dst + r * 64 + l), __float_as_int(x[r * LX + l]));
    }
}

// The item's own and helper k ranges (see outer_body_rs): own [mb, me) and, for a wave that
// helps, wave hw's cells over [hb, he); helped: this wave receives a helper's slice.
struct OuterRanges {
    int mb = 0, me = 64, hw = -1, hb = 0, he = 0;
    bool helped = false, balance = false;
};
template <int T, bool BAL>
__device__ __forceinline__ OuterRanges outer_ranges(const WorkItem &it, int nK, int Kb, int wv)
{
    constexpr int QS = Tc<T>::QS, H = QS / 2;
    const int qr = wv >> 1, qc = wv & 1, I = it.I, J = it.J;
    OuterRanges g;
    g.me = T;
    g.balance = BAL && nK == 1 && (Kb == I || Kb == J) && it.step_begin < it.step_end;
    if (!g.balance) return g;
    if (Kb == I && Kb == J) { // w0 [0,QS)  w1 [0,2QS) -> w1 [QS,2QS) + w2 [0,QS)  w3 [QS,2QS)
        g.mb = wv == 3 || wv == 1 ? QS : 0;
        g.me = wv == 0 ? QS : (wv == 2 ? 0 : T);
        if (wv == 2) g.hw = 1, g.hb = 0, g.he = QS;
        g.helped = wv == 1;
    } else if (Kb == I) { // rows of qr = 1 need k >= QS: qr 0 [H,2QS), qr 1 [QS,2QS) + [0,H) of (0,qc)
        g.mb = qr ? QS : H;
        if (qr) g.hw = qc, g.hb = 0, g.he = H;
        g.helped = qr == 0;
    } else { // columns of qc = 0 need k < QS: qc 1 [0,2QS-H), qc 0 [0,QS) + [2QS-H,2QS) of (qr,1)
        g.me = qc ? T - H : QS;
        if (!qc) g.hw = 2 * qr + 1, g.hb = T - H, g.he = T;
        g.helped = qc == 1;
    }
    return g;
}

// End of a balanced item: helpers hand their slices (acc2) to the helped waves through LDS.
template <int T>
__device__ __forceinline__ void outer_merge(const OuterRanges &g, float (&acc)[Tc<T>::R][Tc<T>::R],
                                            const float (&acc2)[Tc<T>::R][Tc<T>::R], float *smem, int wv)
{
    constexpr int R = Tc<T>::R, RG = R * R * 64;
    const int l = threadIdx.x % 64;
    __syncthreads();
    if (g.hw >= 0) {
#pragma unroll
        for (int i = 0; i < R; ++i)
#pragma unroll
            for (int j = 0; j < R; ++j) smem[g.hw * RG + i * R * 64 + l * R + j] = acc2[i][j];
    }
    __syncthreads();
    if (g.helped) {
#pragma unroll
        for (int i = 0; i < R; ++i)
#pragma unroll
            for (int j = 0; j < R; ++j) acc[i][j] = fmaxf(acc[i][j], smem[wv * RG + i * R * 64 + l * R + j]);
    }
}

// 64 x 64 storage tiles: operands staged straight into LDS (DlStage), two 32-deep halves per
// K step.  Half h = 2s + kh of step s: A = F(i1,k1) tile (I,K) columns 32kh.., B = rows
// K*64 + 32kh + q (q = 0..32) of F(k1+1,j1), tile column J (row 64 is row 0 of tile K+1, or
// -inf when K = J); R0 reads B row k+1 for k (the k2+1 shift).  R3 (K = I) reads the
// unshifted B rows of its half, R4 (K = J) the A columns of its half; their S1 terms are
// staged with the first half of the step.
template <bool BAL>
__device__ void outer_body_dl(const PairDesc &P, const WorkItem &it, float *__restrict__ arena,
                              const int *__restrict__ codes, const float *__restrict__ w, float *smem)
{
    constexpr int T = 64, R = 4, QS = Tc<T>::QS, H = QS / 2;
    int r0, c0;
    lane_block<T>(r0, c0);
    const int i1 = it.i1, j1 = it.j1, I = it.I, J = it.J, nb = P.nb, M = P.M;
    const long long toff = utri_ll(I, J, nb) * (long long)(T * T);
    const float *S1 = arena + P.S1;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / 64), l = threadIdx.x % 64;
    const int kcnt = (it.flags >> WI_KCNT_SHIFT) & 255;
    const int nK = kcnt ? kcnt : J - I + 1, Kb = I + ((it.flags >> WI_KOFF_SHIFT) & 255);
    float acc[R][R], acc2[R][R];
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
        for (int j = 0; j < R; ++j) acc[i][j] = acc2[i][j] = NEG_INF;
    const OuterRanges g = outer_ranges<T, BAL>(it, nK, Kb, wv);
    int hr0 = 0, hc0 = 0;
    if (g.hw >= 0) lane_block<T>(hr0, hc0, g.hw * 64 + l);
    // padding (N not a multiple of T): see outer_body_rs
    const int kreal = (P.N - (nb - 1) * T + H - 1) / H * H;
    auto quad_pad = [&](int x) {
        return (I == nb - 1 && I * T + (x >> 1) * QS >= P.N) || (J == nb - 1 && J * T + (x & 1) * QS >= P.N);
    };
    const bool own_pad = quad_pad(wv), help_pad = g.hw >= 0 && quad_pad(g.hw);
    // operand tiles of half h: as outer_body_h (one-K items keep their tile offsets and take
    // k1 without a division; triangle indices in 32 bits; P read once)
    auto tile_off = [&](int a, int b) { return utri_ll(a, b, nb) * (long long)(T * T); };
    auto tri32 = [&](int a, int b) { return a * M - (a * (a - 1) >> 1) + (b - a); };
    const long long tstride = P.tstride;
    const float *F = arena + P.F;
    const long long oIK = tile_off(I, Kb), oKJ = tile_off(Kb, J), oK1J = Kb + 1 <= J ? tile_off(Kb + 1, J) : 0;
    auto issue = [&](int h) {
        const int s = h >> 1, kh = h & 1;
        int k1, K;
        long long aoff = oIK, boff = oKJ, b1off = oK1J;
        if (nK == 1) {
            k1 = (it.flags & WI_EDGE) ? (s ? j1 - 1 : i1) : i1 + s;
            K = Kb;
        } else {
            k1 = outer_k1(it, s, nK);
            K = Kb + s % nK;
            aoff = tile_off(I, K);
            boff = tile_off(K, J);
            b1off = K + 1 <= J ? tile_off(K + 1, J) : 0;
        }
        float *st = smem + (h & 1) * DlStage::FLOATS;
        const float *Bf = F + (long long)tri32(k1 + 1, j1) * tstride;
        dl_stage_A(st, F + (long long)tri32(i1, k1) * tstride + aoff, kh, wv, l);
        dl_stage_B32(st, Bf + boff, kh, wv, l);
        if (wv == 0 && l < 16) { // B row 32 of the half: row 32 of tile K, or row 0 of tile K+1 (-inf past J)
            const float *src = kh == 0 ? Bf + boff + 32 * T : (K + 1 <= J ? Bf + b1off : (const float *)kNegInfRow);
            g2lds16(src + 4 * l, st + DlStage::B + 32 * 64);
        }
        if (kh == 0 && l == 0) {
            if (wv == 1 && K == I) g2lds4(S1 + i1 * M + k1, st + DlStage::S1);
            if (wv == 2 && K == J) g2lds4(S1 + (k1 + 1) * M + j1, st + DlStage::S1 + 1);
        }
    };
    const int h0 = 2 * it.step_begin, h1 = 2 * it.step_end;
    if (h0 < h1) issue(h0);
    float s1a = NEG_INF, s1b = NEG_INF;
    for (int h = h0; h < h1; ++h) {
        const int s = h >> 1, kh = h & 1, K = nK == 1 ? Kb : Kb + s % nK;
        const float *st = smem + (h & 1) * DlStage::FLOATS;
        vm_wait_all();
        __syncthreads();
        if (kh == 0) { // read before the next loads are issued into the other stage
            s1a = st[DlStage::S1];
            s1b = st[DlStage::S1 + 1];
        }
        if (h + 1 < h1) issue(h + 1); // in flight while this half computes
        if (K == J && kh == (wv & 1)) { // R4: A(r, c) = F(i1,k1)[I*T + r][J*T + c]
            const float4 *A = (const float4 *)(st + DlStage::A) + r0 * 8 + ((l % 8) ^ ((r0 >> 2) & 7));
#pragma unroll
            for (int i = 0; i < R; ++i) {
                const float4 v = A[i * 8];
                const float e[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int j = 0; j < R; ++j) acc[i][j] = fmaxf(acc[i][j], e[j] + s1b);
            }
        }
        if (K == I && kh == (wv >> 1)) { // R3: B row (r - 32kh) = F(k1+1,j1)[I*T + r][J*T + c]
            const float *B = st + DlStage::B + (r0 - 32 * kh) * 64 + c0;
#pragma unroll
            for (int i = 0; i < R; ++i) {
                const float4 v = *(const float4 *)(B + i * 64);
                const float e[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int j = 0; j < R; ++j) acc[i][j] = fmaxf(acc[i][j], s1a + e[j]);
            }
        }
        int kb = g.mb, ke = g.me;
        if (!g.balance) quad_k_range<T>(K == I, K == J, false, kb, ke);
        const int klim = K == nb - 1 ? kreal : T;
        if (!own_pad) dl_product<1>(acc, st, r0, c0, kh, kb, ke < klim ? ke : klim);
        if (g.hw >= 0 && !help_pad) dl_product<1>(acc2, st, hr0, hc0, kh, g.hb, g.he < klim ? g.he : klim);
    }
    if (g.balance) outer_merge<T>(g, acc, acc2, smem, wv);
    if (it.flags & WI_EPILOGUE) outer_epilogue<T>(P, i1, j1, toff, arena, codes, w, acc, r0, c0, I, J);
    float *dst = tri_at(arena, P, i1, j1) + toff;
    if (it.flags & WI_ATOMIC) store_rows_atomic(dst, acc, r0, c0, smem, wv, l);
    else store_acc<T, T>(dst, acc, r0, c0, false);
}

// Cell mapping of an FP16 R0 item.  With one K per item every step of the item is of one
// kind.  K = I only (A a diagonal tile: row r needs k >= r): wave w owns rows 16w .. 16w+15
// (lane l: rows 16w + 4(l/16), columns 4(l%16)) and needs k in [16w, 64).  K = J only (B the
// shifted diagonal tile: column c needs k < c): columns 16w .. 16w+15 (lane l: rows 4(l/4),
// columns 16w + 4(l%4)) and k in [0, 16w+16).  Those ranges, 64, 48, 32 and 16 k, are balanced
// to 40 per wave: the two longest hand 8-k-aligned slices to the two shortest, which run them on
// the longer waves' cells into their second accumulator.  Other items keep the quadrants of
// outer_ranges (48 k per wave at most on a K = I or K = J step).
enum : int { HMAP_QUAD, HMAP_ROWS, HMAP_COLS };
__device__ __forceinline__ void h_cells(int map, int w, int l, int &r0, int &c0)
{
    if (map == HMAP_ROWS) {
        r0 = 16 * w + 4 * (l >> 4);
        c0 = 4 * (l & 15);
    } else if (map == HMAP_COLS) {
        r0 = 4 * (l >> 2);
        c0 = 16 * w + 4 * (l & 3);
    } else {
        lane_block<64>(r0, c0, w * 64 + l);
    }
}
__device__ __forceinline__ OuterRanges h_band_ranges(int map, int w)
{
    OuterRanges g;
    g.balance = true;
    if (map == HMAP_ROWS) { // needs: w0 [0,64) w1 [16,64) w2 [32,64) w3 [48,64)
        constexpr int mb[4] = {0, 16, 32, 48}, me[4] = {40, 56, 64, 64};
        g.mb = mb[w];
        g.me = me[w];
        if (w == 2) g.hw = 1, g.hb = 56, g.he = 64;
        if (w == 3) g.hw = 0, g.hb = 40, g.he = 64;
        g.helped = w < 2;
    } else { // needs: w0 [0,16) w1 [0,32) w2 [0,48) w3 [0,64)
        constexpr int me[4] = {16, 32, 40, 40};
        g.mb = 0;
        g.me = me[w];
        if (w == 0) g.hw = 3, g.hb = 40, g.he = 64;
        if (w == 1) g.hw = 2, g.hb = 40, g.he = 48;
        g.helped = w >= 2;
    }
    return g;
}

// FP16 form of outer_body_dl (pairs whose scores FP16 holds exactly, PairDesc::Fh >= 0):
// the R0 operands are staged from the FP16 mirror of F, a whole 64-deep K step per stage
// (HStage, 16.8 KB: the LDS of one 32-deep FP32 half), so each step takes one barrier and one
// load wait; they are multiplied on packed halves (h_product: 4 cell updates per 3 VALU
// instructions).  R3/R4 read the same staged tiles.  The lane's 4 x 4 cells are held as 4 x 2
// packed pairs and turned back into FP32 (exactly) for the epilogue and the store, which are
// outer_body_dl's.
template <bool BAL>
__device__ void outer_body_h(const PairDesc &P, const WorkItem &it, float *__restrict__ arena,
                             const int *__restrict__ codes, const float *__restrict__ w, char *smem)
{
    constexpr int T = 64, R = 4, QS = Tc<T>::QS, H = QS / 2;
    const int i1 = it.i1, j1 = it.j1, I = it.I, J = it.J, nb = P.nb, M = P.M;
    const long long toff = utri_ll(I, J, nb) * (long long)(T * T);
    const float *S1 = arena + P.S1;
    const h16bits *Fh = (const h16bits *)(arena + P.Fh);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / 64), l = threadIdx.x % 64;
    const int kcnt = (it.flags >> WI_KCNT_SHIFT) & 255;
    const int nK = kcnt ? kcnt : J - I + 1, Kb = I + ((it.flags >> WI_KOFF_SHIFT) & 255);
    unsigned acc[R][2], acc2[R][2];
#pragma unroll
    for (int i = 0; i < R; ++i) acc[i][0] = acc[i][1] = acc2[i][0] = acc2[i][1] = 0xFC00FC00u;
    const bool one_k = BAL && nK == 1 && it.step_begin < it.step_end;
    const int map = !one_k || (Kb == I && Kb == J) ? HMAP_QUAD : (Kb == I ? HMAP_ROWS : (Kb == J ? HMAP_COLS : HMAP_QUAD));
    const OuterRanges g = map == HMAP_QUAD ? outer_ranges<T, BAL>(it, nK, Kb, wv) : h_band_ranges(map, wv);
    int r0, c0, hr0 = 0, hc0 = 0;
    h_cells(map, wv, l, r0, c0);
    if (g.hw >= 0) h_cells(map, g.hw, l, hr0, hc0);
    const HRows rows = h_rows(r0);
    // padding (N not a multiple of T): see outer_body_rs; a wave whose quadrant or band lies
    // wholly in padding rows or columns runs no product
    const int kreal = (P.N - (nb - 1) * T + H - 1) / H * H;
    auto quad_pad = [&](int x) {
        if (map == HMAP_ROWS) return I == nb - 1 && I * T + 16 * x >= P.N;
        if (map == HMAP_COLS) return J == nb - 1 && J * T + 16 * x >= P.N;
        return (I == nb - 1 && I * T + (x >> 1) * QS >= P.N) || (J == nb - 1 && J * T + (x & 1) * QS >= P.N);
    };
    const bool own_pad = quad_pad(wv), help_pad = g.hw >= 0 && quad_pad(g.hw);
    // operand tiles of step s.  Items of one K (the usual case) keep their tile offsets and
    // take k1 without a division; triangle indices in 32 bits (M * M < 2^31)
    auto tile_off = [&](int a, int b) { return utri_ll(a, b, nb) * (long long)(T * T); };
    auto tri32 = [&](int a, int b) { return a * M - (a * (a - 1) >> 1) + (b - a); };
    const long long tstride = P.tstride; // (P lives in memory: read once, not per step)
    const long long oIK = tile_off(I, Kb), oKJ = tile_off(Kb, J), oK1J = Kb + 1 <= J ? tile_off(Kb + 1, J) : 0;
    auto issue = [&](int s) {
        int k1, K;
        long long aoff = oIK, boff = oKJ, b1off = oK1J;
        if (nK == 1) {
            k1 = (it.flags & WI_EDGE) ? (s ? j1 - 1 : i1) : i1 + s;
            K = Kb;
        } else {
            k1 = outer_k1(it, s, nK);
            K = Kb + s % nK;
            aoff = tile_off(I, K);
            boff = tile_off(K, J);
            b1off = K + 1 <= J ? tile_off(K + 1, J) : 0;
        }
        char *st = smem + (s & 1) * HStage::BYTES;
        const h16bits *Bf = Fh + (long long)tri32(k1 + 1, j1) * tstride;
        h_stage_A(st, Fh + (long long)tri32(i1, k1) * tstride + aoff, wv, l);
        h_stage_B(st, Bf + boff, wv, l);
        if (wv == 0 && l < 8) { // B row 64: row 0 of tile K+1 (-inf past J)
            const h16bits *src = K + 1 <= J ? Bf + b1off : (const h16bits *)kNegInfRowH;
            g2lds16(src + 8 * l, st + HStage::B + 64 * 128);
        }
        if (l == 0) {
            if (wv == 1 && K == I) g2lds4(S1 + i1 * M + k1, st + HStage::S1);
            if (wv == 2 && K == J) g2lds4(S1 + (k1 + 1) * M + j1, st + HStage::S1 + 4);
        }
    };
    if (it.step_begin < it.step_end) issue(it.step_begin);
    for (int s = it.step_begin; s < it.step_end; ++s) {
        const int K = nK == 1 ? Kb : Kb + s % nK;
        const char *st = smem + (s & 1) * HStage::BYTES;
        vm_wait_all();
        __syncthreads();
        // read before the next loads are issued into the other stage
        const float sa = *(const float *)(st + HStage::S1), sb = *(const float *)(st + HStage::S1 + 4);
        if (s + 1 < it.step_end) issue(s + 1); // in flight while this step computes
        if (K == J) { // R4: A(r, c) = F(i1,k1)[I*T + r][J*T + c], columns c0 .. c0+3
            const unsigned s1b = pk_pack(sb, sb);
            const int c = c0 >> 3, off = 2 * (c0 & 7);
#pragma unroll
            for (int i = 0; i < R; ++i) {
                const uint2 v = *(const uint2 *)(st + rows.base + 128 * i + ((16 * c) ^ rows.sw16) + off);
                const unsigned t0 = pk_add(v.x, s1b), t1 = pk_add(v.y, s1b);
                pk_max3_into(acc[i][0], t0, t0);
                pk_max3_into(acc[i][1], t1, t1);
            }
        }
        if (K == I) { // R3: B row r = F(k1+1,j1)[I*T + r][J*T + c]
            const unsigned s1a = pk_pack(sa, sa);
#pragma unroll
            for (int i = 0; i < R; ++i) {
                const uint2 v = *(const uint2 *)(st + HStage::B + (r0 + i) * 128 + 2 * c0);
                const unsigned t0 = pk_add(s1a, v.x), t1 = pk_add(s1a, v.y);
                pk_max3_into(acc[i][0], t0, t0);
                pk_max3_into(acc[i][1], t1, t1);
            }
        }
        int kb = g.mb, ke = g.me;
        if (!g.balance) quad_k_range<T>(K == I, K == J, false, kb, ke);
        const int klim = K == nb - 1 ? kreal : T;
        if (!own_pad) h_product<1>(acc, st, rows, c0, kb, ke < klim ? ke : klim);
        if (g.hw >= 0 && !help_pad) h_product<1>(acc2, st, h_rows(hr0), hc0, g.hb, g.he < klim ? g.he : klim);
    }
    if (g.balance) { // helpers hand their slices to the helped waves (as outer_merge, packed)
        unsigned *x = (unsigned *)smem;
        __syncthreads();
        if (g.hw >= 0) {
#pragma unroll
            for (int i = 0; i < R; ++i) {
                x[g.hw * 512 + i * 128 + 2 * l] = acc2[i][0];
                x[g.hw * 512 + i * 128 + 2 * l + 1] = acc2[i][1];
            }
        }
        __syncthreads();
        if (g.helped) {
#pragma unroll
            for (int i = 0; i < R; ++i) {
                pk_max3_into(acc[i][0], x[wv * 512 + i * 128 + 2 * l], acc[i][0]);
                pk_max3_into(acc[i][1], x[wv * 512 + i * 128 + 2 * l + 1], acc[i][1]);
            }
        }
    }
    float acc32[R][R];
#pragma unroll
    for (int i = 0; i < R; ++i) {
        acc32[i][0] = pk_lo(acc[i][0]);
        acc32[i][1] = pk_hi(acc[i][0]);
        acc32[i][2] = pk_lo(acc[i][1]);
        acc32[i][3] = pk_hi(acc[i][1]);
    }
    if (it.flags & WI_EPILOGUE) outer_epilogue<T>(P, i1, j1, toff, arena, codes, w, acc32, r0, c0, I, J);
    float *dst = tri_at(arena, P, i1, j1) + toff;
    if (it.flags & WI_ATOMIC) store_rows_atomic(dst, acc32, r0, c0, (float *)smem, wv, l);
    else store_acc<T, T>(dst, acc32, r0, c0, false);
}

// BAL: balance the waves' k ranges on diagonal steps (below); off in the task kernel, whose
// register budget it would exceed.
template <int T, bool BAL>
__device__ void outer_body_rs(const PairDesc &P, const WorkItem &it, float *__restrict__ arena,
                              const int *__restrict__ codes, const float *__restrict__ w, float *smem)
{
    constexpr int LD = Tc<T>::LD, R = Tc<T>::R, Q = Stage<T>::Q;
    float *As = smem, *Bs = smem + (T + 1) * LD;
    int r0, c0;
    lane_block<T>(r0, c0);
    const int i1 = it.i1, j1 = it.j1, I = it.I, J = it.J, nb = P.nb, M = P.M;
    const long long toff = utri_ll(I, J, nb) * (long long)(T * T);
    const float *S1 = arena + P.S1;
    float acc[R][R];
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
        for (int j = 0; j < R; ++j) acc[i][j] = NEG_INF;

    // the item's K values: Kb .. Kb+nK-1 (all of I..J unless the flags name a sub-range)
    const int kcnt = (it.flags >> WI_KCNT_SHIFT) & 255;
    const int nK = kcnt ? kcnt : J - I + 1, Kb = I + ((it.flags >> WI_KOFF_SHIFT) & 255);
    FetchA<T> fa;
    FetchB<T> fb;
    float4 brow = make_float4(NEG_INF, NEG_INF, NEG_INF, NEG_INF); // row I*T of F(k1+1,j1), lanes < Q
    // S1 terms of the fetched step (R3: S1(i1,k1), R4: S1(k1+1,j1)), loaded with its tiles so
    // that the step's wait for them is the wait for the tiles, not one for the next prefetch
    float s1a_f = NEG_INF, s1b_f = NEG_INF;
    auto fetch = [&](int s) {
        const int k1 = outer_k1(it, s, nK), K = Kb + s % nK;
        const float *Bt = tri_at(arena, P, k1 + 1, j1);
        fetch_A<T, T>(fa, tri_at(arena, P, i1, k1), nb, I, K, 0);
        fetch_B<T, T>(fb, Bt, nb, K, J, 1);
        if (K == I && threadIdx.x < Q) brow = *(const float4 *)(Bt + toff + 4 * threadIdx.x);
        if (K == I) s1a_f = S1[i1 * M + k1];
        if (K == J) s1b_f = S1[(k1 + 1) * M + j1];
    };
    if (it.step_begin < it.step_end) fetch(it.step_begin);
    // Items of one K value: every step has the same diagonal type.  On diagonal steps the
    // quadrants' k ranges differ (quad_k_range), so a wave with a shorter range also runs a
    // slice of a longer one of another wave's quadrant into acc2, which that wave max-merges
    // at the end of the item: the slowest wave's k range per step drops from 2 QS to 1.5 QS
    // (K = I or K = J) or QS (K = I = J).  Max is exact, so the split changes no result.
    // The ranges are fixed per item here (the per-step balanced_k_range, which also serves
    // items of several K, timed 0.4 ms slower on a 64x512 solve of one-K items).
    constexpr int QS = Tc<T>::QS, H = QS / 2;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
    const OuterRanges g = outer_ranges<T, BAL>(it, nK, Kb, wv);
    const int mb = g.mb, me = g.me, hw = g.hw, hb = g.hb, he = g.he;
    const bool balance = g.balance;
    float acc2[R][R];
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
        for (int j = 0; j < R; ++j) acc2[i][j] = NEG_INF;
    int hr0 = 0, hc0 = 0;
    if (hw >= 0) lane_block<T>(hr0, hc0, hw * 64 + (int)(threadIdx.x % 64));
    // Padding (N not a multiple of T): inner indices N.. of the last storage tile hold no real
    // cell, and no real cell reads them -- a real cell (r, c) takes k2 in [r, c), so k2 <= N-2.
    // On steps with K = nb-1 the product stops at the last real k (rounded up to H), and a wave
    // whose quadrant lies wholly in padding rows (I = nb-1) or columns (J = nb-1) runs no
    // product at all; the padding cells keep values no real cell uses.
    const int kreal = (P.N - (nb - 1) * T + H - 1) / H * H; // <= T
    auto quad_pad = [&](int w) {
        return (I == nb - 1 && I * T + (w >> 1) * QS >= P.N) || (J == nb - 1 && J * T + (w & 1) * QS >= P.N);
    };
    const bool own_pad = quad_pad(wv), help_pad = hw >= 0 && quad_pad(hw);
    for (int s = it.step_begin; s < it.step_end; ++s) {
        const int K = Kb + s % nK;
        __syncthreads();
        commit_A<T>(As, fa, 0, false);
        commit_B<T>(Bs, fb);
        if (K == I && threadIdx.x < Q) *(float4 *)(Bs - LD + 4 * threadIdx.x) = brow;
        const float s1a = s1a_f, s1b = s1b_f;
        __syncthreads();
        if (s + 1 < it.step_end) fetch(s + 1); // in flight while this step computes
        if (K == J) { // R4: As[c][r] = F(i1,k1)[I*T + r][J*T + c]
#pragma unroll
            for (int i = 0; i < R; ++i)
#pragma unroll
                for (int j = 0; j < R; ++j) acc[i][j] = fmaxf(acc[i][j], As[(c0 + j) * LD + r0 + i] + s1b);
        }
        if (K == I) { // R3: Bs[r-1][c] = F(k1+1,j1)[I*T + r][J*T + c]
#pragma unroll
            for (int i = 0; i < R; ++i)
#pragma unroll
                for (int j = 0; j < R; ++j) acc[i][j] = fmaxf(acc[i][j], s1a + Bs[(r0 + i - 1) * LD + c0 + j]);
        }
        int kb = mb, ke = me;
        if (!balance) quad_k_range<T>(K == I, K == J, false, kb, ke);
        const int klim = K == nb - 1 ? kreal : T;
        if (!own_pad) product<T, H>(acc, As, Bs, r0, c0, kb, ke < klim ? ke : klim);
        if (hw >= 0 && !help_pad) product<T, H>(acc2, As, Bs, hr0, hc0, hb, he < klim ? he : klim);
    }
    if (balance) outer_merge<T>(g, acc, acc2, smem, wv); // through the (now free) operand LDS

    if (it.flags & WI_EPILOGUE) outer_epilogue<T>(P, i1, j1, toff, arena, codes, w, acc, r0, c0, I, J);
    store_acc<T, T>(tri_at(arena, P, i1, j1) + toff, acc, r0, c0, (it.flags & WI_ATOMIC) != 0);
}

template <int T, bool BAL = true>
__device__ __forceinline__ void outer_body(const PairDesc &P, const WorkItem &it, float *__restrict__ arena,
                                           const int *__restrict__ codes, const float *__restrict__ w, float *smem)
{
    if constexpr (T == 64) outer_body_dl<BAL>(P, it, arena, codes, w, smem);
    else outer_body_rs<T, BAL>(P, it, arena, codes, w, smem);
}

// ------------------------------------------------------------------------ inner ext
// Cross-tile part of the inner closure for logical tile (I,J), D = J-I >= 1, of
// X = F(i1,j1) (MODE_S = false) or X = S2 (MODE_S = true), S = S2.  Steps:
//   F: s <  D : R1  X[I][K] (x) S[K*LT+1 ..][J]          K = I+s      (k2 in tile K)
//      s >= D : R2  S[I][K'*LT-1 ..] (x) X[K'][J]        K' = I+1+s-D (k2+1 in tile K')
//   S: s <  D : S[I][K] (x) S[K*LT+1 ..][J], K = I+s; for K = I only column LT-1 of S[I][I]
//               (the remaining columns of tile I pair with cells of (I,J) itself and are
//               handled by the in-tile wavefront).
// Step s of F reads X tile (I,I+s) (R1) or (I+1+s-D, J) (R2): logical diagonal s or
// 2D-1-s, so only steps D-1 and D read diagonal D-1.
struct ExtStep {
    int aK, cshift;         // A: logical tile (I, aK) of X (a_x) or S, column shift cshift
    int bK, rshift;         // B: rows of logical tile row bK shifted by rshift, column tile J
    bool a_x, b_x;          // operand taken from X (else from S2)
    bool lastcol, adiag, bdiag;
};

template <bool MODE_S>
__device__ __forceinline__ ExtStep ext_step(int I, int J, int s)
{
    const int D = J - I;
    if (MODE_S) return ExtStep{I + s, 0, I + s, 1, false, false, s == 0, false, false};
    if (s < D) return ExtStep{I + s, 0, I + s, 1, true, false, false, s == 0, false}; // R1
    const int K2 = I + 1 + (s - D);                                                    // R2
    return ExtStep{K2, -1, K2, 0, false, true, false, false, K2 == J};
}

// Cross-tile products split k over the four waves: on every step, wave w takes k in
// [w*LT/4, (w+1)*LT/4) on the whole LT x LT tile, and each lane keeps a 4 x 4 block of cells
// (lane l: rows 4*(l / (LT/4)), columns 4*(l % (LT/4))).  A lane then reads 8 operand floats
// per k for 16 cell updates (a 2 x 2 block over all k would read 4 for 4), half the LDS bytes
// per update, for the same VALU work.  The four partial maxima meet once per item, in LDS.
template <int LT>
struct Xk {
    static_assert(LT == 32, "cross-tile k split: 64 lanes of 4 x 4 blocks per 32 x 32 tile");
    static constexpr int KQ = LT / 4;                 // k per wave per step
    static constexpr int LD = Tc<LT>::LD;             // row stride of the partials in LDS
    static constexpr int RED_FLOATS = 4 * LT * LD;    // one partial tile per wave
};

// acc[i][j] = max(acc[i][j], max_{kb<=k<ke} As[k][r0+i] + Bs[k][c0+j]), kb and ke even: two k
// per round, v_add_f32 x 2 + v_max3_f32 per pair of updates.
template <int LT>
__device__ __forceinline__ void product44(float (&acc)[4][4], const float *As, const float *Bs, int r0, int c0,
                                          int kb, int ke)
{
    constexpr int LD = Tc<LT>::LD;
#pragma unroll 1 // unrolled, the loads of later rounds are hoisted past 64 VGPRs
    for (int k = kb; k < ke; k += 2) {
        const float4 a0 = *(const float4 *)(As + k * LD + r0), a1 = *(const float4 *)(As + (k + 1) * LD + r0);
        const float4 b0 = *(const float4 *)(Bs + k * LD + c0), b1 = *(const float4 *)(Bs + (k + 1) * LD + c0);
        const float av0[4] = {a0.x, a0.y, a0.z, a0.w}, av1[4] = {a1.x, a1.y, a1.z, a1.w};
        const float bv0[4] = {b0.x, b0.y, b0.z, b0.w}, bv1[4] = {b1.x, b1.y, b1.z, b1.w};
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = max3(acc[i][j], av0[i] + bv0[j], av1[i] + bv1[j]);
    }
}

// Steps [sb, se) of the cross-tile part of logical tile (I,J) of X on 256 lanes.  On return
// `out` holds the maxima of the four cells (tid / (LT/4), 4*(tid % (LT/4)) + 0..3) of the
// tile (-inf with no steps); smem is free again.
template <int T, bool MODE_S>
__device__ void ext_accumulate(const PairDesc &P, float *X, int I, int J, int sb, int se, float4 &out, float *arena,
                               float *smem)
{
    constexpr int LT = Inner<T>::LT;
    constexpr int LD = Tc<LT>::LD, KQ = Xk<LT>::KQ, BPR = LT / 4; // 4 x 4 blocks per row
    float *As = smem, *Bs = smem + LT * LD;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / 64), l = threadIdx.x % 64;
    const int r0 = 4 * (l / BPR), c0 = 4 * (l % BPR);
    const int nb = P.nb;
    float *S = arena + P.S2;
    float acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = NEG_INF;
    FetchA<LT> fa;
    FetchB<LT> fb;
    ExtStep st{};
    auto fetch = [&](int s) {
        st = ext_step<MODE_S>(I, J, s);
        fetch_A<T, LT>(fa, st.a_x ? X : S, nb, I, st.aK, st.cshift);
        fetch_B<T, LT>(fb, st.b_x ? X : S, nb, st.bK, J, st.rshift);
    };
    if (sb < se) fetch(sb);
    for (int s = sb; s < se; ++s) {
        __syncthreads();
        commit_A<LT>(As, fa, st.cshift, st.lastcol);
        commit_B<LT>(Bs, fb);
        // with only column LT-1 of A finite, the pair of k values (LT-2, LT-1) is the step
        const int kb = st.lastcol ? (w == 3 ? LT - 2 : LT) : w * KQ, ke = st.lastcol ? LT : (w + 1) * KQ;
        __syncthreads();
        if (s + 1 < se) fetch(s + 1);
        product44<LT>(acc, As, Bs, r0, c0, kb, ke);
    }
    // the four partial tiles meet in LDS; lane tid takes four cells of one row
    float *red = smem;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 4; ++i)
        *(float4 *)(red + w * LT * LD + (r0 + i) * LD + c0) = make_float4(acc[i][0], acc[i][1], acc[i][2], acc[i][3]);
    __syncthreads();
    const int row = threadIdx.x / BPR, col = 4 * (threadIdx.x % BPR);
    float4 m = *(const float4 *)(red + row * LD + col);
#pragma unroll
    for (int q = 1; q < 4; ++q) {
        const float4 v = *(const float4 *)(red + q * LT * LD + row * LD + col);
        m = make_float4(fmaxf(m.x, v.x), fmaxf(m.y, v.y), fmaxf(m.z, v.z), fmaxf(m.w, v.w));
    }
    out = m;
    __syncthreads();
}

// FP16 form of ext_accumulate for F (pairs with FP16 mirrors, T = 64): the same steps, every
// operand a plain logical tile -- F tiles from the mirror, S2 from its row-shifted (R1's B) or
// column-shifted (R2's A) FP16 copy -- staged straight into LDS three steps ahead (four
// stages), one barrier per step; wave w takes the 8 k of chunk w of every step on the whole
// 32 x 32 tile (4 x 4 cells per lane, packed), and the four partial tiles meet in LDS at the
// end.  `out` as ext_accumulate's; smem (>= 4 XStage) is free again on return.
static_assert(4 * XStage::BYTES <= 4 * Xk<32>::RED_FLOATS && 4 * XStage::BYTES <= 4 * Intra<32>::FLOATS &&
                  4 * 512 * 4 <= 4 * XStage::BYTES,
              "four FP16 cross-tile stages (and the packed partial tiles) fit the closure kernels' LDS");
__device__ void ext_accumulate_h(const PairDesc &P, int i1, int j1, int I, int J, int sb, int se, float4 &out,
                                 float *arena, char *smem)
{
    constexpr int NS = 4;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / 64), l = threadIdx.x % 64;
    const int r0 = 4 * (l >> 3), c0 = 4 * (l & 7), nb = P.nb, D = J - I;
    const h16bits *Xh = (const h16bits *)(arena + P.Fh) + utri_ll(i1, j1, P.M) * P.tstride;
    const h16bits *Sr = (const h16bits *)(arena + P.S2r), *Sc = (const h16bits *)(arena + P.S2c);
    auto lt = [&](const h16bits *tri, int x, int y) {
        return tri + utri_ll(x >> 1, y >> 1, nb) * 4096 + (x & 1) * 32 * 64 + (y & 1) * 32;
    };
    unsigned acc[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[i][0] = acc[i][1] = 0xFC00FC00u;
    auto issue = [&](int s) {
        char *st = smem + (s % NS) * XStage::BYTES;
        if (s < D) { // R1: X(I, I+s) + S2 rows shifted by one (tile (I+s, J) of S2r)
            x_stage(st, lt(Xh, I, I + s), lt(Sr, I + s, J), w, l);
        } else { // R2: S2 columns shifted by one (tile (I, K2) of S2c) + X(K2, J)
            const int K2 = I + 1 + (s - D);
            x_stage(st, lt(Sc, I, K2), lt(Xh, K2, J), w, l);
        }
    };
    for (int q = 0; q < NS - 1 && sb + q < se; ++q) issue(sb + q);
    for (int s = sb; s < se; ++s) {
        // every wave issues one load per step: those of the later steps already in flight may stay
        const int later = se - 1 - s < NS - 2 ? se - 1 - s : NS - 2;
        if (later >= 2) vm_wait_le2();
        else if (later == 1) vm_wait_le1();
        else vm_wait_all();
        __syncthreads();
        if (s + NS - 1 < se) issue(s + NS - 1); // into the stage of step s-1
        x_prod8(acc, smem + (s % NS) * XStage::BYTES, r0, c0, w);
    }
    unsigned *red = (unsigned *)smem; // partial tiles: [wave][row][16 pairs]
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        red[w * 512 + (r0 + i) * 16 + c0 / 2] = acc[i][0];
        red[w * 512 + (r0 + i) * 16 + c0 / 2 + 1] = acc[i][1];
    }
    __syncthreads();
    const int row = threadIdx.x >> 3, col = 4 * (threadIdx.x & 7);
    unsigned m0 = red[row * 16 + col / 2], m1 = red[row * 16 + col / 2 + 1];
#pragma unroll
    for (int q = 1; q < 4; ++q) {
        pk_max3_into(m0, red[q * 512 + row * 16 + col / 2], m0);
        pk_max3_into(m1, red[q * 512 + row * 16 + col / 2 + 1], m1);
    }
    out = make_float4(pk_lo(m0), pk_hi(m0), pk_lo(m1), pk_hi(m1));
    __syncthreads();
}

// Always max-combined into the tile (other items of the same tile run concurrently).
template <int T, bool MODE_S>
__device__ void ext_body(const PairDesc &P, const WorkItem &it, float *__restrict__ arena, float *smem)
{
    constexpr int LT = Inner<T>::LT;
    float *X = MODE_S ? arena + P.S2 : tri_at(arena, P, it.i1, it.j1);
    float4 m;
    if (T == 64 && !MODE_S && P.S2r >= 0)
        ext_accumulate_h(P, it.i1, it.j1, it.I, it.J, it.step_begin, it.step_end, m, arena, (char *)smem);
    else
        ext_accumulate<T, MODE_S>(P, X, it.I, it.J, it.step_begin, it.step_end, m, arena, smem);
    // through LDS, so that each wave instruction max-combines whole row segments (LT
    // consecutive floats) instead of one float of every fourth column
    const int row = threadIdx.x / (LT / 4), col = 4 * (threadIdx.x % (LT / 4));
    *(float4 *)(smem + row * LT + col) = m;
    __syncthreads();
    float *dst = ltile<T, LT>(X, it.I, it.J, P.nb);
#pragma unroll
    for (int q = 0; q < LT * LT / kThreads; ++q) {
        const int e = q * kThreads + (int)threadIdx.x;
        atomicMax((int *)(dst + (e / LT) * T + e % LT), __float_as_int(smem[e]));
    }
}

// ---------------------------------------------------------------------- inner intra
// In-tile closure of logical tile (I,J) of X (F(i1,j1) or S2).  Cell (a,b) (local row a,
// column b) sits on anti-diagonal s = b + (LT-1-a); every in-tile term of it pairs two
// cells on earlier anti-diagonals:
//   R1: X[a][k] + SR[k+1][b]   k in [diag ? a : 0, b-1]    SR = S2 tile (J,J)
//   R2: SL[a][k-1] + X[k][b]   k in [a+1, diag ? b : LT-1] SL = S2 tile (I,I)
//   P2: X[a+1][b-1] + intra(b[r], b[c]) when c-r >= 4 (across the tile edge: the row
//       below / column left / corner, loaded up front)
// (for S2 on a diagonal tile SL = SR = X itself, and S2 starts every upper cell at 0).
//
// Push formulation: each lane owns NC cells (q + m*QN, b), m < NC, of column b = tid % LT
// (q = tid / LT) and keeps their running maxima in registers.  Step s finalizes
// anti-diagonal s (owners write max(acc, P2) to LDS), one barrier, then every lane folds
// in the terms whose newest factor just became final: for each owned cell the one new
// cell of its row, X[a][a+s-(LT-1)] (R1), and the one new cell of its column,
// X[b-s+(LT-1)][b] (R2), each with its S2 partner.  A partner that is not final yet still
// holds a valid lower bound (every stored value is the score of some structure), so the
// early sum never exceeds the optimum, and the exact pair is added when the later factor
// finalizes.  Terms whose indices leave the tile must add nothing:
//   * row factors of negative column: X rows have stride LX = 2*LT+4, so columns -LT..-1
//     of row a are pad columns of row a-1 (or of a guard row), all -inf;
//   * column factors of rows past LT-1: the row index is clamped to the -inf pad row LT;
//   * every other out-of-tile index (partners of -inf factors; factors of cells already
//     final, whose registers are no longer stored) only has to stay inside the LDS block,
//     which the guard row before Sa and the order Sa, Sb, Xs ensure.
// All LDS reads of a step are conflict-free: row factors are broadcasts, column factors
// and S2 partners run along consecutive columns.
//
// late (blocked form only): the tile's two cross-tile steps that read logical diagonal J-I-1
// (R1 with K = J-1, R2 with K' = I+1) are run here first, when the launch leaves them out of
// the cross-tile items (see k_inner_closure_blk).
//
// BLK (F, LT*LT/4 lanes): the blocked form below instead of the push wavefront.
template <int T, bool MODE_S, int NT = kThreads, bool BLK = false>
__device__ __forceinline__ void intra_body(const PairDesc &P, int i1, int j1, int I, int J, float *__restrict__ arena,
                           const int *__restrict__ codes, const float *__restrict__ w, float *smem,
                           bool late = false)
{
    constexpr int LT = Inner<T>::LT;
    using C = Intra<LT>;
    constexpr int L = C::L, LX = C::LX;
    float *Xs = smem + C::XS, *Sa = smem + C::SA, *Sb = smem + C::SB;
    float *below = smem + C::BELOW, *left = below + LT, *cornerp = left + LT;
    float *wl = smem + C::WL;
    int *cr = (int *)(smem + C::CODES), *cc = cr + LT;
    const int nb = P.nb;
    const bool diag = (I == J);
    float *S = arena + P.S2;
    float *X = MODE_S ? S : tri_at(arena, P, i1, j1);
    const int tid = threadIdx.x;

    // late cross-tile steps: staged in the LDS that the set-up below overwrites, results kept
    // in registers until the tile is in LDS.  On FP16 operands they run while the tile loads
    // below are in flight (nothing else writes tile (I,J) during the launch, so the values
    // loaded stay current); the FP32 form runs first, as its registers and the tile's do not
    // fit together.
    constexpr bool kLate = !MODE_S && BLK;
    float4 emax = make_float4(NEG_INF, NEG_INF, NEG_INF, NEG_INF);
    const bool late_steps = kLate && late && J > I, late_h = T == 64 && P.S2r >= 0;
    if constexpr (kLate) {
        if (late_steps && !late_h) ext_accumulate<T, false>(P, X, I, J, J - I - 1, J - I + 1, emax, arena, smem);
    }

    // tiles: every global load issued before any LDS store (float4, unrolled, branch-free;
    // lanes past the tile's LT*Q float4 chunks re-read chunk 0 and store nothing)
    constexpr int Q = LT / 4, CH = LT * Q, PER = (CH + NT - 1) / NT;
    const float *Xt = ltile<T, LT>(X, I, J, nb);
    const float *SLt = ltile<T, LT>(S, I, I, nb), *SRt = ltile<T, LT>(S, J, J, nb);
    float4 vx[PER], va[PER], vb[PER];
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int idx = (tid + p * NT) % CH, a = idx / Q, c4 = (idx % Q) * 4;
        vx[p] = *(const float4 *)(Xt + a * T + c4);
        va[p] = *(const float4 *)(SLt + a * T + c4);
        vb[p] = *(const float4 *)(SRt + a * T + c4);
    }
    float bl = NEG_INF, lf = NEG_INF, cn = NEG_INF;
    if (!diag && tid < LT) {
        bl = ltile<T, LT>(X, I + 1, J, nb)[tid];
        lf = ltile<T, LT>(X, I, J - 1, nb)[tid * T + LT - 1];
        if (tid == 0 && J - I >= 2) cn = ltile<T, LT>(X, I + 1, J - 1, nb)[LT - 1];
    }
    if constexpr (kLate) {
        if (late_steps && late_h) ext_accumulate_h(P, i1, j1, I, J, J - I - 1, J - I + 1, emax, arena, (char *)smem);
    }
    const int *bc = codes + P.b;
    int vcr = 0, vcc = 0;
    if (tid < LT) {
        vcr = bc[I * LT + tid];
        vcc = bc[J * LT + tid];
    }
    const float vw = tid < 25 ? w[tid] : 0.f;
    // -inf: the guards, pad columns LT.. of every row of Sa, Sb and Xs, pad rows LT
    // (everything before BELOW that is not a tile cell: no overlap with the tile stores)
    for (int e = tid; e < C::FILL_END; e += NT) {
        bool cell;
        if (e >= C::XS) cell = (e - C::XS) % LX < LT && (e - C::XS) / LX < LT;
        else if (e >= C::SB) cell = (e - C::SB) % L < LT && (e - C::SB) / L < LT;
        else cell = e >= C::SA && (e - C::SA) % L < LT && (e - C::SA) / L < LT;
        if (!cell) smem[e] = NEG_INF;
    }
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int idx = tid + p * NT, a = idx / Q, c4 = (idx % Q) * 4;
        if (CH % NT != 0 && idx >= CH) break;
        float4 x = vx[p];
        if (MODE_S) { // S2 starts every upper cell at 0, its short-interval value
            const int r = I * LT + a, c = J * LT + c4;
            if (r <= c) x.x = fmaxf(x.x, 0.f);
            if (r <= c + 1) x.y = fmaxf(x.y, 0.f);
            if (r <= c + 2) x.z = fmaxf(x.z, 0.f);
            if (r <= c + 3) x.w = fmaxf(x.w, 0.f);
        }
        *(float4 *)(Xs + a * LX + c4) = x;
        *(float4 *)(Sa + a * L + c4) = va[p];                    // S2 tile (I,I)
        *(float4 *)(Sb + a * L + c4) = pick(diag, va[p], vb[p]); // S2 tile (J,J)
    }
    if (tid < LT) {
        below[tid] = bl;
        left[tid] = lf;
        cr[tid] = vcr;
        cc[tid] = vcc;
    }
    if (tid == 0) cornerp[0] = cn;
    if (tid < 25) wl[tid] = vw;
    if constexpr (kLate) {
        if (late_steps) { // fold the late steps into the tile (each cell has one owner)
            __syncthreads();
            float *c = Xs + (tid / (LT / 4)) * LX + 4 * (tid % (LT / 4));
            c[0] = fmaxf(c[0], emax.x);
            c[1] = fmaxf(c[1], emax.y);
            c[2] = fmaxf(c[2], emax.z);
            c[3] = fmaxf(c[3], emax.w);
        }
    }
    // SL = S2 (I,I), SR = S2 (J,J) (both in Sb on a diagonal tile); S2's own diagonal tiles
    // take both factors from the tile being computed
    const bool fresh = MODE_S && diag;
    const float *SL = diag ? (MODE_S ? Xs : Sb) : Sa;
    const float *SR = fresh ? Xs : Sb;
    const int LSL = fresh ? LX : L, LSR = fresh ? LX : L; // their row strides
    __syncthreads();

    if constexpr (BLK) {
        // Blocked wavefront: lane (A, Bc) owns the 2 x 2 block of cells rows 2A..2A+1,
        // columns 2Bc..2Bc+1.  Its cells lie on anti-diagonals s-1, s, s, s+1, and every
        // in-tile factor of them lies in a block of a smaller block diagonal u = Bc - A or
        // in the block itself, so the tile closes in one step per block diagonal (LT-1
        // steps, LT/2 on a diagonal tile) instead of one per anti-diagonal:
        //   step u: fold in the terms whose factor cells were finalized at step u-1 -- the
        //     two of each cell's row in block (A, A+u-1) (R1) and the two of its column in
        //     block (Bc-u+1, Bc) (R2), each with its S2 partner; at u = Bc - A close the
        //     block in registers: BL, then TL and BR (each with the BL term), then TR.
        // Factor cells outside the tile read the -inf pads (rows: the pad row LT; columns
        // -LT..-1: the pad columns of the row above or the guard row); their partner
        // indices are clamped into the tile's rows and columns.  On a diagonal tile the
        // lower-left cell of a diagonal block is below the diagonal and stays -inf.
        static_assert(!MODE_S && NT * 4 == LT * LT, "one 2 x 2 block per lane");
        constexpr int HB = LT / 2;
        const int A = tid / HB, Bc = tid % HB, a0 = 2 * A, b0 = 2 * Bc, town = Bc - A;
        const bool active = !diag || A <= Bc;
        float acc[2][2], wv[2][2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int a = a0 + i, c = b0 + j;
                acc[i][j] = Xs[a * LX + c];
                const bool ok = (J * LT + c) - (I * LT + a) >= 4;
                wv[i][j] = ok ? wl[5 * cr[a] + cc[c]] : NEG_INF;
            }
        // P2 factors X[a+1][b-1] of TL, BL, BR (TR's is BL), across the tile edge when outside
        auto p2at = [&](int row, int col) -> const float * {
            return row < LT ? (col >= 0 ? Xs + row * LX + col : left + row) : (col >= 0 ? below + col : cornerp);
        };
        const float *pTL = p2at(a0 + 1, b0 - 1), *pBL = p2at(a0 + 2, b0 - 1), *pBR = p2at(a0 + 2, b0);
        const float sl_in = SL[a0 * L + a0], sr_in = SR[(b0 + 1) * L + b0 + 1]; // partners inside the block
        const int u0 = diag ? 0 : 1 - HB;
        const int oX = (int)(Xs - smem), oSR = (int)(SR - smem), oSL = (int)(SL - smem);
        for (int u = u0; u < HB; ++u) {
            // offsets (into smem) of step u's operands, set before its barrier so that the
            // reads after it issue at once
            const int k0 = 2 * (A + u - 1);  // row factors: columns k0, k0+1 (>= -LT)
            const int k1 = 2 * (Bc - u + 1); // column factors: rows k1, k1+1 (> a0+1)
            int o_xr = oX + a0 * LX + k0;
            int o_sr = oSR + clampi(k0 + 1, 0, LT - 1) * L + b0; // SR[k0+1][b0..]
            int o_xc0 = oX + min(k1, LT) * LX + b0, o_xc1 = oX + min(k1 + 1, LT) * LX + b0;
            int o_sl = oSL + a0 * L + min(k1 - 1, L - 2);
            asm volatile("" : "+v"(o_xr), "+v"(o_sr), "+v"(o_xc0), "+v"(o_xc1), "+v"(o_sl));
            __syncthreads(); // block diagonal u-1 is final
            if (!active || u > town) continue;
            // the P2 factors serve the closing step (u == town) only, and are final by then;
            // read with the step's other operands, so that a step makes one LDS round trip
            const float p2bl = *pBL, p2tl = *pTL, p2br = *pBR;
            if (u > u0) {
                const float2 xr0 = *(const float2 *)(smem + o_xr);
                const float2 xr1 = *(const float2 *)(smem + o_xr + LX);
                const float2 sr0 = *(const float2 *)(smem + o_sr);     // SR[k0+1][b0..]
                const float2 sr1 = *(const float2 *)(smem + o_sr + L); // SR[k0+2][b0..]
                const float2 xc0 = *(const float2 *)(smem + o_xc0);
                const float2 xc1 = *(const float2 *)(smem + o_xc1);
                const float sl00 = smem[o_sl], sl01 = smem[o_sl + 1];
                const float sl10 = smem[o_sl + L], sl11 = smem[o_sl + L + 1];
                max3_into(acc[0][0], xr0.x + sr0.x, xr0.y + sr1.x);
                max3_into(acc[0][1], xr0.x + sr0.y, xr0.y + sr1.y);
                max3_into(acc[1][0], xr1.x + sr0.x, xr1.y + sr1.x);
                max3_into(acc[1][1], xr1.x + sr0.y, xr1.y + sr1.y);
                max3_into(acc[0][0], sl00 + xc0.x, sl01 + xc1.x);
                max3_into(acc[0][1], sl00 + xc0.y, sl01 + xc1.y);
                max3_into(acc[1][0], sl10 + xc0.x, sl11 + xc1.x);
                max3_into(acc[1][1], sl10 + xc0.y, sl11 + xc1.y);
            }
            if (u == town) {
                const float xbl = fmaxf(acc[1][0], p2bl + wv[1][0]);
                const float xtl = fmaxf(max3(acc[0][0], sl_in + xbl, NEG_INF), p2tl + wv[0][0]);
                const float xbr = fmaxf(max3(acc[1][1], xbl + sr_in, NEG_INF), p2br + wv[1][1]);
                const float xtr = fmaxf(max3(acc[0][1], xtl + sr_in, sl_in + xbr), xbl + wv[0][1]);
                *(float2 *)(Xs + a0 * LX + b0) = make_float2(xtl, xtr);
                *(float2 *)(Xs + (a0 + 1) * LX + b0) = make_float2(xbl, xbr);
            }
        }
    } else {
        constexpr int NC = LT * LT / NT, QN = NT / LT; // cells per lane, lanes per column
        static_assert(NT % LT == 0 && NC * QN == LT, "cell mapping");
        const int b = tid % LT, q = tid / LT;
        // per cell: running maximum, P2 weight (-inf where P2 does not apply, so that
        // max(acc, p2 + w) needs no select), finalize step, P2 factor address
        float acc[NC], wv[NC];
        int sown[NC];
        const float *p2p[NC];
    #pragma unroll
        for (int m = 0; m < NC; ++m) {
            const int a = q + m * QN;
            acc[m] = Xs[a * LX + b];
            const bool ok = (J * LT + b) - (I * LT + a) >= 4;
            wv[m] = ok ? wl[5 * cr[a] + cc[b]] : NEG_INF; // P2 weight intra(b[I*LT+a], b[J*LT+b])
            sown[m] = b + (LT - 1) - a;
            const int row = a + 1, col = b - 1; // P2 factor X[a+1][b-1], or across the tile edge
            p2p[m] = row < LT ? (col >= 0 ? Xs + row * LX + col : left + row) : (col >= 0 ? below + col : cornerp);
        }
        const int s0 = diag ? LT - 1 : 0;
        // cells of anti-diagonal s0 start final
    #pragma unroll
        for (int m = 0; m < NC; ++m)
            if (s0 == sown[m]) Xs[(q + m * QN) * LX + b] = fmaxf(acc[m], *p2p[m] + wv[m]);
        // the P2 factor read at step s serves the finalize of step s+1: it belongs to the
        // lane's cell on anti-diagonal s+2 (if any) and lies on anti-diagonal s
        auto p2_next = [&](int s) -> const float * {
            const int d = sown[0] - (s + 2); // that cell is m = d / QN
            const float *p = cornerp;
    #pragma unroll
            for (int m = 0; m < NC; ++m) p = d == m * QN ? p2p[m] : p;
            return p;
        };
        float p2 = *p2_next(s0 - 1);
        if constexpr (!MODE_S) {
            // F tiles: S2 partners at constant stride L.  Per step one pointer per factor kind
            // moves by a constant; cells of one lane sit QN rows apart (immediate offsets).
            const float *pxr = Xs + q * LX + q - (LT - 1) + s0;           // X[q][q+s-(LT-1)]
            const float *psr = SR + (q + s0 - (LT - 2)) * L + b;          // SR[q+s-(LT-2)][b]
            const float *psl = SL + q * L + b - s0 + (LT - 2);            // SL[q][b-s+(LT-2)]
            for (int s = s0; s < 2 * LT - 2; ++s) {
                __syncthreads(); // anti-diagonal s is final
                const float xcol = Xs[min(b - s + (LT - 1), LT) * LX + b]; // past the last row: pad row LT
                float xr[NC], sr[NC], sl[NC];
    #pragma unroll
                for (int m = 0; m < NC; ++m) {
                    xr[m] = pxr[m * QN * (LX + 1)];
                    sr[m] = psr[m * QN * L];
                    sl[m] = psl[m * QN * L];
                }
                const float p2n = *p2_next(s);
    #pragma unroll
                for (int m = 0; m < NC; ++m) {
                    max3_into(acc[m], xr[m] + sr[m], sl[m] + xcol);
                    if (s + 1 == sown[m]) Xs[(q + m * QN) * LX + b] = fmaxf(acc[m], p2 + wv[m]);
                }
                p2 = p2n;
                pxr += 1;
                psr += L;
                psl -= 1;
            }
        } else {
            // S2 tiles: partners may be cells of the tile itself (fresh), read after the
            // barrier and with clamped indices
            for (int s = s0; s < 2 * LT - 2; ++s) {
                __syncthreads(); // anti-diagonal s is final
                const int ac = b - s + (LT - 1);
                const float xcol = Xs[min(ac, LT) * LX + b];
                const int ksl = ac - 1 < 0 ? 0 : ac - 1;
                float xr[NC], sr[NC], sl[NC];
    #pragma unroll
                for (int m = 0; m < NC; ++m) {
                    const int a = q + m * QN, kr = a + s - (LT - 1);
                    xr[m] = Xs[a * LX + kr];
                    sr[m] = SR[clampi(kr + 1, 0, LT) * LSR + b];
                    sl[m] = SL[a * LSL + ksl];
                }
                const float p2n = *p2_next(s);
    #pragma unroll
                for (int m = 0; m < NC; ++m) {
                    max3_into(acc[m], xr[m] + sr[m], sl[m] + xcol);
                    if (s + 1 == sown[m]) Xs[(q + m * QN) * LX + b] = fmaxf(acc[m], p2 + wv[m]);
                }
                p2 = p2n;
            }
        }
    }
    __syncthreads();
    float *Xo = ltile<T, LT>(X, I, J, nb);
    // the tile is final: also its FP16 copy, the R0 operand of pairs that have one
    h16bits *Xh = nullptr;
    if (!MODE_S && P.Fh >= 0)
        Xh = (h16bits *)(arena + P.Fh) + utri_ll(i1, j1, P.M) * P.tstride + (Xo - X);
    for (int idx = tid; idx < CH; idx += NT) {
        const int a = idx / Q, c4 = (idx % Q) * 4;
        const float4 v = *(const float4 *)(Xs + a * LX + c4);
        *(float4 *)(Xo + a * T + c4) = v;
        if (Xh) *(uint2 *)(Xh + a * T + c4) = make_uint2(pk_pack(v.x, v.y), pk_pack(v.z, v.w));
    }
}

// ------------------------------------------------------------------------------ S1
// Strand-1 single-strand table, dense M x M, one workgroup, diagonal by diagonal
// (compilation_scripts/bpmax_verify.ab:17-21).
__device__ void s1_body(const PairDesc &P, float *__restrict__ arena, const int *__restrict__ codes,
                        const float *__restrict__ w)
{
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

// ------------------------------------------------------- one launch per phase kernels
template <int T>
__global__ __launch_bounds__(kThreads) void k_outer(const PairDesc *__restrict__ pairs,
                                                    const WorkItem *__restrict__ items, float *__restrict__ arena,
                                                    const int *__restrict__ codes, const float *__restrict__ w)
{
    __shared__ __attribute__((aligned(16))) float smem[OuterLds<T>::FLOATS];
    const WorkItem it = items[xcd_work_index(blockIdx.x, gridDim.x)];
    outer_body<T>(pairs[it.pair], it, arena, codes, w, smem);
}

// FP16-operand form of k_outer<64> (every pair of the launch has an FP16 mirror of F).
// LDS: two 64-deep FP16 stages (also room for the 64 x 68 floats of the row-wise atomic
// combine).
__global__ __launch_bounds__(kThreads) void k_outer_h(const PairDesc *__restrict__ pairs,
                                                      const WorkItem *__restrict__ items, float *__restrict__ arena,
                                                      const int *__restrict__ codes, const float *__restrict__ w)
{
    constexpr int BYTES = cmax(2 * HStage::BYTES, 64 * 68 * 4);
    __shared__ __attribute__((aligned(16))) char smem[BYTES];
    const WorkItem it = items[xcd_work_index(blockIdx.x, gridDim.x)];
    outer_body_h<true>(pairs[it.pair], it, arena, codes, w, smem);
}

// The closure kernels (cross-tile and in-tile) are the critical path of the pipelined
// schedule and run beside its R0 launches.  They fit on a CU next to four R0 workgroups
// (at most 64 VGPRs per SIMD lane between them: 4 waves of up to 64 -- the cross-tile
// kernel, the blocked in-tile form -- or 8 of up to 32 -- the 512-lane in-tile forms; 9 /
// 19 KB of LDS) and raise their wave
// priority, so they start at once and issue ahead of the R0 waves instead of queueing
// behind them.
template <int T, bool MODE_S>
__global__ __launch_bounds__(kThreads, 8) void k_inner_ext(const PairDesc *__restrict__ pairs,
                                                        const WorkItem *__restrict__ items, float *__restrict__ arena)
{
    __shared__ __attribute__((aligned(16))) float smem[Xk<Inner<T>::LT>::RED_FLOATS];
    __builtin_amdgcn_s_setprio(3); // latency-critical chain: issue ahead of R0 waves
    const WorkItem it = items[xcd_work_index(blockIdx.x, gridDim.x)];
    ext_body<T, MODE_S>(pairs[it.pair], it, arena, smem);
}

// The in-tile closure in two widths: NT = 1024 lanes (one cell per lane, 4 waves per SIMD
// to interleave the short dependent chain of each anti-diagonal step: lowest latency, for
// launches of few tiles) or NT = 512 (two cells per lane: the per-step pointer and barrier
// work shared, half the waves per tile, for launches of many tiles).  S2 uses 1024 or 256.
template <int T, bool MODE_S, int NT>
__global__ __launch_bounds__(NT, MODE_S ? 1 : 8) void k_inner_intra(const PairDesc *__restrict__ pairs,
                                                    const WorkItem *__restrict__ items,
                                                    float *__restrict__ arena, const int *__restrict__ codes,
                                                    const float *__restrict__ w)
{
    __shared__ __attribute__((aligned(16))) float smem[Intra<Inner<T>::LT>::FLOATS];
    __builtin_amdgcn_s_setprio(3);
    const WorkItem it = items[xcd_work_index(blockIdx.x, gridDim.x)];
    intra_body<T, MODE_S, NT>(pairs[it.pair], it.i1, it.j1, it.I, it.J, arena, codes, w, smem);
}

// In-tile closure of F in the blocked form (one 2 x 2 block of cells per lane, 256 lanes).
template <int T>
__global__ __launch_bounds__(kThreads, 8) void k_inner_intra_blk(const PairDesc *__restrict__ pairs,
                                                              const WorkItem *__restrict__ items,
                                                              float *__restrict__ arena,
                                                              const int *__restrict__ codes,
                                                              const float *__restrict__ w)
{
    __shared__ __attribute__((aligned(16))) float smem[Intra<Inner<T>::LT>::FLOATS];
    __builtin_amdgcn_s_setprio(3);
    const WorkItem it = items[xcd_work_index(blockIdx.x, gridDim.x)];
    intra_body<T, false, kThreads, true>(pairs[it.pair], it.i1, it.j1, it.I, it.J, arena, codes, w, smem);
}

// The inner closure of F by logical block diagonal DL in one launch per diagonal instead of
// two (cross-tile, then in-tile): the launch for DL holds
//   * in-tile items (type TASK_INTRA) of the tiles on DL, each of which first runs the two
//     cross-tile steps of its tile that read diagonal DL-1 (intra_body, late), and
//   * cross-tile items (type TASK_EXT) with the other steps of the tiles on DL+1, which
//     read diagonals below DL only and so run beside the in-tile work of DL.
// 256 lanes: the blocked in-tile form and the cross-tile body; fits beside four R0
// workgroups like k_inner_ext.
template <int T>
__global__ __launch_bounds__(kThreads, 8) void k_inner_closure_blk(const PairDesc *__restrict__ pairs,
                                                                const WorkItem *__restrict__ items,
                                                                float *__restrict__ arena,
                                                                const int *__restrict__ codes,
                                                                const float *__restrict__ w)
{
    static_assert(Xk<Inner<T>::LT>::RED_FLOATS <= Intra<Inner<T>::LT>::FLOATS, "cross-tile LDS fits the in-tile LDS");
    __shared__ __attribute__((aligned(16))) float smem[Intra<Inner<T>::LT>::FLOATS];
    __builtin_amdgcn_s_setprio(3);
    const WorkItem it = items[xcd_work_index(blockIdx.x, gridDim.x)];
    const PairDesc &P = pairs[it.pair];
    if (wi_type(it.flags) == TASK_EXT) ext_body<T, false>(P, it, arena, smem);
    else intra_body<T, false, kThreads, true>(P, it.i1, it.j1, it.I, it.J, arena, codes, w, smem, true);
}


__global__ __launch_bounds__(kThreads) void k_s1(const PairDesc *__restrict__ pairs, float *__restrict__ arena,
                                                 const int *__restrict__ codes, const float *__restrict__ w)
{
    s1_body(pairs[blockIdx.x], arena, codes, w);
}

// ------------------------------------------------------- dependency-tile scheduler
// Every task of a solve (S1, S2 tiles, then per outer diagonal the R0 chunks of every
// storage tile and the cross-tile chunks and in-tile closures of every logical tile) is
// one entry of a list in a topological order.  A fixed grid of persistent workgroups takes
// tickets from an atomic head, one task at a time, so a task only ever waits on tasks
// holding smaller tickets, which are already running or finished: no deadlock, whatever the
// grid size, and every workgroup leaves once the tickets run out.  Each logical F / S2 tile has
// two counters in `flags`:
//   done    = 1 once its in-tile closure has stored the final values;
//   pending = number of max-combining contributions (R0 chunks of its storage tile,
//             cross-tile chunks) still to land; the closure starts when it reaches 0.
// Waits poll with relaxed agent-scope loads (thread 0 of the workgroup) and end with one
// acquire fence; producers release with a fence before the counter update.  A wait that
// exceeds kSpinLimit polls sets *err and abandons the task (and every later waiter gives
// up when it sees *err), so the kernel always drains.
constexpr int kSpinLimit = 1 << 22;

__device__ __forceinline__ int ld_relaxed(const int *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct Waiter {
    const int *err;
    int polls = 0;
    bool ok = true;
    __device__ bool until(const int *p, int want)
    {
        while (ok && ld_relaxed(p) != want) {
            __builtin_amdgcn_s_sleep(2);
            if (++polls > kSpinLimit || ((polls & 1023) == 0 && ld_relaxed(err) != 0)) ok = false;
        }
        return ok;
    }
};

// counters of logical tile (X,Y) (nbl logical tiles per side)
__device__ __forceinline__ int *f_ctr(int *flags, const PairDesc &P, int i1, int j1, int X, int Y)
{
    return flags + P.flag + 2 + 2 * tri_ll(P.nbl) + 2 * (utri_ll(i1, j1, P.M) * tri_ll(P.nbl) + utri_ll(X, Y, P.nbl));
}
__device__ __forceinline__ int *s_ctr(int *flags, const PairDesc &P, int X, int Y)
{
    return flags + P.flag + 2 + 2 * utri_ll(X, Y, P.nbl);
}

// Dependencies of a task, enumerated so that every lane of the workgroup can poll its share
// of them at once: dependency i is a counter and the value it must reach (want); entries
// that do not apply come back with p = nullptr.  The set is, per task type:
//   S2 cross-tile chunk: per step s (K = I+s) the tiles (I,K), (K,J) (K > I) and (K+1,J) done;
//   S2 in-tile: its pending contributions landed; for I < J the tiles (I,I), (J,J), (I+1,J),
//     (I,J-1) and, for J-I >= 2, (I+1,J-1) done;
//   R0 chunk: S1 done, every S2 tile done, then per step (k1, K) the logical sub-tiles of
//     storage tiles (I,K) of F(i1,k1) and (K,J), (K+1,J) of F(k1+1,j1); with the epilogue
//     also storage tile (I,J) of F(i1+1,j1-1) (P1, d1 >= 4) and of F(k1+1,j1), F(i1,k1) for
//     every k1 (R3/R4);
//   cross-tile chunk: every S2 tile done, then per step the one tile it reads;
//   in-tile: every S2 tile done, its pending contributions landed, and for I < J the tiles
//     (I+1,J), (I,J-1) and, for J-I >= 2, (I+1,J-1).
struct Dep {
    const int *p;
    int want;
};

// logical sub-tile q (0 .. rs*rs-1) of storage tile (SI,SJ) of triangle (i1,j1); the lower
// sub-tiles of a diagonal storage tile (x > y) do not exist
__device__ __forceinline__ Dep storage_dep(int *flags, const PairDesc &P, int rs, int i1, int j1, int SI, int SJ,
                                           int q)
{
    const int x = SI * rs + q / rs, y = SJ * rs + q % rs;
    if (x > y) return Dep{nullptr, 0};
    return Dep{f_ctr(flags, P, i1, j1, x, y), 1};
}

__device__ int dep_count(const PairDesc &P, const WorkItem &it)
{
    const int I = it.I, J = it.J, ns = it.step_end - it.step_begin, rs = P.nbl / P.nb;
    switch (wi_type(it.flags)) {
    case TASK_S2EXT: return 3 * ns;
    case TASK_S2INTRA: return 6;
    case TASK_OUTER: {
        int n = 2 + 3 * rs * rs * ns;
        if (it.flags & WI_EPILOGUE) n += rs * rs * (1 + 2 * (it.j1 - it.i1));
        return n;
    }
    case TASK_EXT: return 1 + ns;
    case TASK_INTRA: return 5;
    default: return 0;
    }
    (void)I;
    (void)J;
}

__device__ Dep dep_at(const PairDesc &P, const WorkItem &it, int *flags, int i)
{
    const int I = it.I, J = it.J, rs = P.nbl / P.nb;
    const Dep none{nullptr, 0};
    const Dep s2_all{flags + P.flag + 1, (int)tri_ll(P.nbl)};
    switch (wi_type(it.flags)) {
    case TASK_S2EXT: {
        const int s = it.step_begin + i / 3, K = I + s;
        switch (i % 3) {
        case 0: return Dep{s_ctr(flags, P, I, K), 1};
        case 1: return K > I ? Dep{s_ctr(flags, P, K, J), 1} : none;
        default: return Dep{s_ctr(flags, P, K + 1, J), 1};
        }
    }
    case TASK_S2INTRA:
        if (i == 0) return Dep{s_ctr(flags, P, I, J) + 1, 0};
        if (I == J) return none;
        switch (i) {
        case 1: return Dep{s_ctr(flags, P, I, I), 1};
        case 2: return Dep{s_ctr(flags, P, J, J), 1};
        case 3: return Dep{s_ctr(flags, P, I + 1, J), 1};
        case 4: return Dep{s_ctr(flags, P, I, J - 1), 1};
        default: return J - I >= 2 ? Dep{s_ctr(flags, P, I + 1, J - 1), 1} : none;
        }
    case TASK_OUTER: { // storage coordinates
        const int i1 = it.i1, j1 = it.j1, nK = J - I + 1, q2 = rs * rs;
        if (i == 0) return Dep{flags + P.flag, 1}; // S1
        if (i == 1) return s2_all;
        i -= 2;
        const int nst = 3 * q2 * (it.step_end - it.step_begin);
        if (i < nst) {
            const int s = it.step_begin + i / (3 * q2), r = i % (3 * q2), g = r / q2, q = r % q2;
            const int k1 = i1 + s / nK, K = I + s % nK;
            if (g == 0) return storage_dep(flags, P, rs, i1, k1, I, K, q);
            if (g == 1) return storage_dep(flags, P, rs, k1 + 1, j1, K, J, q);
            return K + 1 <= J ? storage_dep(flags, P, rs, k1 + 1, j1, K + 1, J, q) : none;
        }
        i -= nst; // epilogue
        const int g = i / q2, q = i % q2;
        if (g == 0) return j1 - i1 >= 4 ? storage_dep(flags, P, rs, i1 + 1, j1 - 1, I, J, q) : none;
        const int k1 = i1 + (g - 1) / 2;
        return (g - 1) % 2 == 0 ? storage_dep(flags, P, rs, k1 + 1, j1, I, J, q)
                                : storage_dep(flags, P, rs, i1, k1, I, J, q);
    }
    case TASK_EXT: { // logical coordinates
        if (i == 0) return s2_all;
        const int s = it.step_begin + i - 1, D = J - I;
        return s < D ? Dep{f_ctr(flags, P, it.i1, it.j1, I, I + s), 1}
                     : Dep{f_ctr(flags, P, it.i1, it.j1, I + 1 + (s - D), J), 1};
    }
    case TASK_INTRA:
        if (i == 0) return s2_all;
        if (i == 1) return Dep{f_ctr(flags, P, it.i1, it.j1, I, J) + 1, 0};
        if (I == J) return none;
        if (i == 2) return Dep{f_ctr(flags, P, it.i1, it.j1, I + 1, J), 1};
        if (i == 3) return Dep{f_ctr(flags, P, it.i1, it.j1, I, J - 1), 1};
        return J - I >= 2 ? Dep{f_ctr(flags, P, it.i1, it.j1, I + 1, J - 1), 1} : none;
    default: return none;
    }
}

// Every lane of the workgroup: wait for everything task `it` reads, the dependencies spread
// over the lanes (one round of polls instead of a chain of them on one thread).  Each lane's
// agent-scope acquire after its polls pairs with the producers' releases, and the barrier
// that follows carries that to every lane.  Returns false on a timeout or when another task
// gave up (*err set).
// *spins (shared) receives the largest number of unsatisfied polls of any lane: the task's
// time spent waiting, in polls (each a device-scope load plus s_sleep 2).
__device__ bool wait_deps(const PairDesc &P, const WorkItem &it, int *flags, const int *err, int *spins)
{
    const int n = dep_count(P, it);
    Waiter W{err};
    for (int i = threadIdx.x; i < n && W.ok; i += blockDim.x) {
        const Dep d = dep_at(P, it, flags, i);
        if (d.p) W.until(d.p, d.want);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (W.polls) atomicMax(spins, W.polls);
    const bool ok = __syncthreads_and(W.ok ? 1 : 0) != 0;
    if (!ok && threadIdx.x == 0) atomicOr((int *)err, 1);
    return ok;
}

// Thread 0 only, after every lane's stores were released.
__device__ void signal_done(const PairDesc &P, const WorkItem &it, int *flags)
{
    const int type = wi_type(it.flags);
    switch (type) {
    case TASK_S1:
        __hip_atomic_store(flags + P.flag, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        break;
    case TASK_S2EXT:
        __hip_atomic_fetch_add(s_ctr(flags, P, it.I, it.J) + 1, -1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        break;
    case TASK_S2INTRA:
        __hip_atomic_store(s_ctr(flags, P, it.I, it.J), 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_add(flags + P.flag + 1, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        break;
    case TASK_OUTER: { // every upper logical sub-tile of the storage tile
        const int rs = P.nbl / P.nb;
        for (int x = it.I * rs; x < it.I * rs + rs; ++x)
            for (int y = it.J * rs; y < it.J * rs + rs; ++y)
                if (x <= y)
                    __hip_atomic_fetch_add(f_ctr(flags, P, it.i1, it.j1, x, y) + 1, -1, __ATOMIC_RELEASE,
                                           __HIP_MEMORY_SCOPE_AGENT);
        break;
    }
    case TASK_EXT:
        __hip_atomic_fetch_add(f_ctr(flags, P, it.i1, it.j1, it.I, it.J) + 1, -1, __ATOMIC_RELEASE,
                               __HIP_MEMORY_SCOPE_AGENT);
        break;
    case TASK_INTRA:
        __hip_atomic_store(f_ctr(flags, P, it.i1, it.j1, it.I, it.J), 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        break;
    }
}

template <int T>
__global__ __launch_bounds__(kThreads) void k_tasks(const PairDesc *__restrict__ pairs,
                                                    const WorkItem *__restrict__ tasks, int ntasks, int *head,
                                                    int *flags, int *err, float *__restrict__ arena,
                                                    const int *__restrict__ codes, const float *__restrict__ w)
{
    constexpr int LT = Inner<T>::LT;
    constexpr int SM = cmax(OuterLds<T>::FLOATS, cmax(Xk<LT>::RED_FLOATS, Intra<LT>::FLOATS));
    __shared__ __attribute__((aligned(16))) float smem[SM];
    __shared__ int s_task, s_spins;
    // persistent: every workgroup takes tickets until the list is exhausted (or a task gave
    // up, *err), so the grid is bounded whatever the number of tasks.  err[2..3] count, as a
    // 64-bit sum, the polls workgroups spent waiting (the largest lane count per task);
    // err[4] counts tasks run.
    unsigned long long waited = 0;
    int ran = 0;
    for (;;) {
        __syncthreads(); // s_task and smem of the previous task are free
        if (threadIdx.x == 0) {
            s_task = ld_relaxed(err) != 0 ? ntasks : atomicAdd(head, 1);
            s_spins = 0;
        }
        __syncthreads();
        const int t = s_task;
        if (t >= ntasks) break;
        const WorkItem it = tasks[t];
        const PairDesc P = pairs[it.pair];
        if (!wait_deps(P, it, flags, err, &s_spins)) break;
        waited += (unsigned)s_spins;
        ++ran;
        switch (wi_type(it.flags)) {
        case TASK_S1: s1_body(P, arena, codes, w); break;
        case TASK_S2EXT: ext_body<T, true>(P, it, arena, smem); break;
        case TASK_S2INTRA: intra_body<T, true>(P, 0, 0, it.I, it.J, arena, codes, w, smem); break;
        case TASK_OUTER: outer_body<T, false>(P, it, arena, codes, w, smem); break;
        case TASK_EXT: ext_body<T, false>(P, it, arena, smem); break;
        case TASK_INTRA: intra_body<T, false>(P, it.i1, it.j1, it.I, it.J, arena, codes, w, smem); break;
        }
        __threadfence();
        __syncthreads();
        if (threadIdx.x == 0) signal_done(P, it, flags);
    }
    if (threadIdx.x == 0 && ran) {
        atomicAdd((unsigned long long *)(err + 2), waited);
        atomicAdd(err + 4, ran);
    }
}

// ------------------------------------------------------------------------- utilities
__global__ void k_fill(float *__restrict__ p, long long n, float v)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// p[0, n) = v and the n2 words after them = the bit pattern v2 (the FP16 mirrors: -inf in
// both halves), in one launch.
__global__ void k_fill2(float *__restrict__ p, long long n, float v, long long n2, unsigned v2)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n + n2; i += stride) {
        if (i < n) p[i] = v;
        else ((unsigned *)p)[i] = v2;
    }
}

// FP16 copies of S2 shifted by one row and by one column (PairDesc::S2r, S2c), the plain-tile
// cross-tile operands of the FP16 closure: block (x-chunk, pair), grid-stride over the cells
// of the pair's upper storage tiles.  S2 values outside the upper triangle read as -inf.
__global__ void k_s2_shift(const PairDesc *__restrict__ pairs, const float *__restrict__ arena,
                           float *__restrict__ arena_w)
{
    const PairDesc P = pairs[blockIdx.y];
    if (P.S2r < 0) return;
    const int T = P.T, nb = P.nb, TT = T * T, n = nb * T;
    const float *S = arena + P.S2;
    auto val = [&](int x, int y) -> float {
        if (y < 0 || x > y || x >= n || y >= n) return NEG_INF;
        return S[utri_ll(x / T, y / T, nb) * TT + (x % T) * T + y % T];
    };
    h16bits *Sr = (h16bits *)(arena_w + P.S2r), *Sc = (h16bits *)(arena_w + P.S2c);
    const long long cells = tri_ll(nb) * TT;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < cells; e += (long long)gridDim.x * blockDim.x) {
        const int t = (int)(e / TT), rc = (int)(e % TT), r = rc / T, c = rc % T;
        int I = 0;
        while (utri_ll(I + 1, I + 1, nb) <= t) ++I;
        const int J = I + (t - (int)utri_ll(I, I, nb)), x = I * T + r, y = J * T + c;
        Sr[e] = __builtin_bit_cast(h16bits, (_Float16)val(x + 1, y));
        Sc[e] = __builtin_bit_cast(h16bits, (_Float16)val(x, y - 1));
    }
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
// 64 x 64 tiles: operands staged straight into two LDS stages (DlStage), two 32-deep halves
// per K step, B unshifted.
__device__ __forceinline__ void r0tri_dl(int nb, const int2 *__restrict__ tiles, const float *__restrict__ A,
                                         const float *__restrict__ B, float *__restrict__ C)
{
    constexpr int T = 64, R = 4;
    __shared__ __attribute__((aligned(16))) float smem[2 * DlStage::FLOATS];
    const int wi = xcd_work_index(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
    const int2 IJ = tiles[wi % gridDim.x];
    const int I = IJ.x, J = IJ.y;
    const long long tstride = tri_ll(nb) * T * T;
    const long long m = wi / gridDim.x;
    const float *Am = A + m * tstride, *Bm = B + m * tstride;
    float *Cm = C + m * tstride;
    int r0, c0;
    lane_block<T>(r0, c0);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / 64), l = threadIdx.x % 64, hw = balance_target(wv);
    float acc[R][R], acc2[R][R];
    const float *Ct = tile_at(Cm, I, J, nb, T);
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
        for (int j = 0; j < R; ++j) {
            acc[i][j] = Ct[(r0 + i) * T + c0 + j];
            acc2[i][j] = NEG_INF;
        }
    int hr0 = 0, hc0 = 0;
    if (hw >= 0) lane_block<T>(hr0, hc0, hw * 64 + l);
    auto issue = [&](int h) {
        const int K = I + (h >> 1), kh = h & 1;
        float *st = smem + (h & 1) * DlStage::FLOATS;
        dl_stage_A(st, tile_at(Am, I, K, nb, T), kh, wv, l);
        dl_stage_B32(st, tile_at(Bm, K, J, nb, T), kh, wv, l);
    };
    const int nh = 2 * (J - I + 1);
    issue(0);
    // Wait for C here: the compiler counts only its own loads, so a wait it placed at acc's
    // first use inside the loop would also wait for the stage loads in flight there.
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
        for (int j = 0; j < R; ++j) asm volatile("" : "+v"(acc[i][j]));
    for (int h = 0; h < nh; ++h) {
        const int K = I + (h >> 1), kh = h & 1;
        vm_wait_all();
        __syncthreads();
        if (h + 1 < nh) issue(h + 1);
        const float *st = smem + (h & 1) * DlStage::FLOATS;
        int kb, ke, hb, he;
        balanced_k_range<T>(K == I, K == J, wv, kb, ke, hb, he);
        dl_product<0>(acc, st, r0, c0, kh, kb, ke);
        if (hw >= 0) dl_product<0>(acc2, st, hr0, hc0, kh, hb, he);
    }
    balance_merge<T>(acc, acc2, smem, wv);
    store_acc<T, T>(tile_at(Cm, I, J, nb, T), acc, r0, c0, false);
}

template <int T>
__global__ __launch_bounds__(kThreads) void k_r0tri(int nb, const int2 *__restrict__ tiles,
                                                    const float *__restrict__ A, const float *__restrict__ B,
                                                    float *__restrict__ C)
{
    if constexpr (T == 64) {
        r0tri_dl(nb, tiles, A, B, C);
        return;
    }
    constexpr int LD = Tc<T>::LD, R = Tc<T>::R;
    __shared__ __attribute__((aligned(16))) float As[T * LD];
    __shared__ __attribute__((aligned(16))) float Bs[T * LD];
    const int w = xcd_work_index(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
    const int2 IJ = tiles[w % gridDim.x];
    const long long tstride = tri_ll(nb) * T * T;
    const long long m = w / gridDim.x;
    const float *Am = A + m * tstride, *Bm = B + m * tstride;
    float *Cm = C + m * tstride;
    int r0, c0;
    lane_block<T>(r0, c0);
    float acc[R][R];
    const float *Ct = tile_at(Cm, IJ.x, IJ.y, nb, T);
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
        for (int j = 0; j < R; ++j) acc[i][j] = Ct[(r0 + i) * T + c0 + j];
    FetchA<T> fa;
    FetchB<T> fb;
    fetch_A<T, T>(fa, Am, nb, IJ.x, IJ.x, 0);
    fetch_B<T, T>(fb, Bm, nb, IJ.x, IJ.y, 0);
    // every tile's K range has both diagonal steps: balanced k ranges (balanced_k_range)
    constexpr int H = Tc<T>::QS / 2;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / 64), hw = balance_target(wv);
    float acc2[R][R];
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
        for (int j = 0; j < R; ++j) acc2[i][j] = NEG_INF;
    int hr0 = 0, hc0 = 0;
    if (hw >= 0) lane_block<T>(hr0, hc0, hw * 64 + (int)(threadIdx.x % 64));
    for (int K = IJ.x; K <= IJ.y; ++K) {
        __syncthreads();
        commit_A<T>(As, fa, 0, false);
        commit_B<T>(Bs, fb);
        __syncthreads();
        if (K + 1 <= IJ.y) {
            fetch_A<T, T>(fa, Am, nb, IJ.x, K + 1, 0);
            fetch_B<T, T>(fb, Bm, nb, K + 1, IJ.y, 0);
        }
        int kb, ke, hb, he;
        balanced_k_range<T>(K == IJ.x, K == IJ.y, wv, kb, ke, hb, he);
        product<T, H>(acc, As, Bs, r0, c0, kb, ke);
        if (hw >= 0) product<T, H>(acc2, As, Bs, hr0, hc0, hb, he);
    }
    balance_merge<T>(acc, acc2, As, wv);
    store_acc<T, T>(tile_at(Cm, IJ.x, IJ.y, nb, T), acc, r0, c0, false);
}

// FP16 copy of isolated R0 operands, four values per lane and step.  A value is exact for the
// FP16 product when it is -inf or an integer of magnitude <= 1024: every sum of two such values
// is then -inf or an integer of magnitude <= 2048, which FP16 holds exactly, and so is every
// maximum of them.  Any other value (a fraction, a larger magnitude, +inf, NaN) sets *inexact.
__global__ void k_tri_to_half(const float4 *__restrict__ src, long long n4, uint2 *__restrict__ dst,
                              int *__restrict__ inexact)
{
    bool bad = false;
    auto exact = [](float v) { return v == NEG_INF || (fabsf(v) <= 1024.0f && v == rintf(v)); };
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n4; q += (long long)gridDim.x * blockDim.x) {
        const float4 v = src[q];
        bad |= !(exact(v.x) && exact(v.y) && exact(v.z) && exact(v.w));
        dst[q] = make_uint2(pk_pack(v.x, v.y), pk_pack(v.z, v.w));
    }
    if (bad) atomicOr(inexact, 1);
}

// Isolated R0 on FP16 operands (copies made by k_tri_to_half, or any whose sums FP16 holds): r0tri_dl's work and tile order
// with a whole 64-deep K step per LDS stage (HStage, one barrier and one load wait per step)
// and products on packed halves (h_product, 4 cell updates per 3 VALU instructions).  The
// waves' k ranges on the diagonal steps are balanced_k_range's, the helpers' slices merged on
// packed halves at the end; C is max-combined in FP32 (its values need not be FP16-exact).
__global__ __launch_bounds__(kThreads) void k_r0tri_h(int nb, const int2 *__restrict__ tiles,
                                                      const h16bits *__restrict__ A, const h16bits *__restrict__ B,
                                                      float *__restrict__ C)
{
    constexpr int T = 64, R = 4;
    __shared__ __attribute__((aligned(16))) char smem[2 * HStage::BYTES];
    const int wi = xcd_work_index(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
    const int2 IJ = tiles[wi % gridDim.x];
    const int I = IJ.x, J = IJ.y;
    const long long tstride = tri_ll(nb) * T * T;
    const long long m = wi / gridDim.x;
    const h16bits *Am = A + m * tstride, *Bm = B + m * tstride;
    float *Ct = C + m * tstride + utri_ll(I, J, nb) * (long long)(T * T);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / 64), l = threadIdx.x % 64, hw = balance_target(wv);
    int r0, c0, hr0 = 0, hc0 = 0;
    lane_block<T>(r0, c0);
    if (hw >= 0) lane_block<T>(hr0, hc0, hw * 64 + l);
    const HRows rows = h_rows(r0), hrows = h_rows(hr0);
    unsigned acc[R][2], acc2[R][2];
#pragma unroll
    for (int i = 0; i < R; ++i) acc[i][0] = acc[i][1] = acc2[i][0] = acc2[i][1] = 0xFC00FC00u;
    float4 cv[R];
#pragma unroll
    for (int i = 0; i < R; ++i) cv[i] = *(const float4 *)(Ct + (r0 + i) * T + c0);
    auto tile = [&](const h16bits *X, int a, int b) { return X + utri_ll(a, b, nb) * (long long)(T * T); };
    auto issue = [&](int K) {
        char *st = smem + ((K - I) & 1) * HStage::BYTES;
        h_stage_A(st, tile(Am, I, K), wv, l);
        h_stage_B(st, tile(Bm, K, J), wv, l);
    };
    issue(I);
    for (int K = I; K <= J; ++K) {
        const char *st = smem + ((K - I) & 1) * HStage::BYTES;
        vm_wait_all();
        __syncthreads();
        if (K + 1 <= J) issue(K + 1); // into the stage every wave finished reading before this barrier
        int kb, ke, hb, he;
        balanced_k_range<T>(K == I, K == J, wv, kb, ke, hb, he);
        h_product<0>(acc, st, rows, c0, kb, ke);
        if (hw >= 0) h_product<0>(acc2, st, hrows, hc0, hb, he);
    }
    // helpers hand their slices to their target waves (waves 0, 1 and 3) through LDS
    unsigned *x = (unsigned *)smem;
    __syncthreads();
    if (hw >= 0) {
#pragma unroll
        for (int i = 0; i < R; ++i) {
            x[hw * 512 + i * 128 + 2 * l] = acc2[i][0];
            x[hw * 512 + i * 128 + 2 * l + 1] = acc2[i][1];
        }
    }
    __syncthreads();
    if (wv != 2) {
#pragma unroll
        for (int i = 0; i < R; ++i) {
            pk_max3_into(acc[i][0], x[wv * 512 + i * 128 + 2 * l], acc[i][0]);
            pk_max3_into(acc[i][1], x[wv * 512 + i * 128 + 2 * l + 1], acc[i][1]);
        }
    }
#pragma unroll
    for (int i = 0; i < R; ++i) {
        float4 c = cv[i];
        c.x = fmaxf(c.x, pk_lo(acc[i][0]));
        c.y = fmaxf(c.y, pk_hi(acc[i][0]));
        c.z = fmaxf(c.z, pk_lo(acc[i][1]));
        c.w = fmaxf(c.w, pk_hi(acc[i][1]));
        *(float4 *)(Ct + (r0 + i) * T + c0) = c;
    }
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
// Synthetic R0 operands: cell (m, r, c), r <= c < n, holds lo + h % (hi - lo + 1) with h the
// splitmix64 hash of seed ^ (m << 42 | r << 21 | c); every other cell (and padding) -inf.
__host__ __device__ inline unsigned long long splitmix64(unsigned long long x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
__global__ void k_tri_random(int n, int nb, int T, unsigned long long seed, int lo, int hi, float *__restrict__ dst)
{
    const long long m = blockIdx.y;
    const long long tstride = tri_ll(nb) * T * T;
    const int r = blockIdx.x;
    const unsigned span = (unsigned)(hi - lo + 1);
    for (int c = threadIdx.x; c < nb * T; c += blockDim.x) {
        if (c / T < r / T) continue;
        float v = NEG_INF;
        if (r < n && c < n && r <= c) {
            const unsigned long long key = ((unsigned long long)m << 42) | ((unsigned long long)r << 21) | (unsigned)c;
            v = (float)(lo + (int)(splitmix64(seed ^ key) % span));
        }
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
#define BPX_T_DISPATCH(T, CALL)                                                                         \
    switch (T) {                                                                                        \
    case 32: { constexpr int TT = 32; CALL; } break;                                                    \
    case 64: { constexpr int TT = 64; CALL; } break;                                                    \
    default: return hipErrorInvalidValue;                                                               \
    }

hipError_t launch_outer(hipStream_t st, int T, const PairDesc *pairs, const WorkItem *items, int n, float *arena,
                        const int *codes, const float *w, bool half)
{
    if (n <= 0) return hipSuccess;
    if (half && T == 64)
        hipLaunchKernelGGL(k_outer_h, dim3(n), dim3(kThreads), 0, st, pairs, items, arena, codes, w);
    else
        BPX_T_DISPATCH(T, hipLaunchKernelGGL(k_outer<TT>, dim3(n), dim3(kThreads), 0, st, pairs, items, arena, codes, w));
    return hipGetLastError();
}

hipError_t launch_inner_ext(hipStream_t st, int T, bool single_strand, const PairDesc *pairs,
                            const WorkItem *items, int n, float *arena)
{
    if (n <= 0) return hipSuccess;
    if (single_strand) {
        BPX_T_DISPATCH(T, hipLaunchKernelGGL((k_inner_ext<TT, true>), dim3(n), dim3(kThreads), 0, st, pairs, items, arena));
    } else {
        BPX_T_DISPATCH(T, hipLaunchKernelGGL((k_inner_ext<TT, false>), dim3(n), dim3(kThreads), 0, st, pairs, items, arena));
    }
    return hipGetLastError();
}

hipError_t launch_inner_intra(hipStream_t st, int T, bool single_strand, const PairDesc *pairs,
                              const WorkItem *items, int n, float *arena, const int *codes, const float *w, bool wide,
                              bool blocked)
{
    if (n <= 0) return hipSuccess;
    if (blocked && !single_strand && !wide) {
        BPX_T_DISPATCH(T, hipLaunchKernelGGL(k_inner_intra_blk<TT>, dim3(n), dim3(kThreads), 0, st, pairs, items,
                                             arena, codes, w));
        return hipGetLastError();
    }
#define BPX_INTRA(S, NT) \
    BPX_T_DISPATCH(T, hipLaunchKernelGGL((k_inner_intra<TT, S, NT>), dim3(n), dim3(NT), 0, st, pairs, items, arena, codes, w))
    if (single_strand) {
        if (wide) { BPX_INTRA(true, 1024); } else { BPX_INTRA(true, kThreads); }
    } else {
        if (wide) { BPX_INTRA(false, 1024); } else { BPX_INTRA(false, 2 * kThreads); }
    }
#undef BPX_INTRA
    return hipGetLastError();
}

hipError_t launch_inner_closure(hipStream_t st, int T, const PairDesc *pairs, const WorkItem *items, int n,
                                float *arena, const int *codes, const float *w)
{
    if (n <= 0) return hipSuccess;
    BPX_T_DISPATCH(T, hipLaunchKernelGGL(k_inner_closure_blk<TT>, dim3(n), dim3(kThreads), 0, st, pairs, items, arena,
                                         codes, w));
    return hipGetLastError();
}

hipError_t launch_s1(hipStream_t st, const PairDesc *pairs, int npairs, float *arena, const int *codes,
                     const float *w)
{
    if (npairs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_s1, dim3(npairs), dim3(kThreads), 0, st, pairs, arena, codes, w);
    return hipGetLastError();
}

hipError_t launch_tasks(hipStream_t st, int T, const PairDesc *pairs, const WorkItem *tasks, int ntasks, int *head,
                        int *flags, int *err, float *arena, const int *codes, const float *w)
{
    if (ntasks <= 0) return hipSuccess;
    // persistent workgroups: twice what fits on the device at once (4 per CU), enough to keep
    // every slot busy while the last tasks drain
    const int grid = ntasks < 2048 ? ntasks : 2048;
    BPX_T_DISPATCH(T, hipLaunchKernelGGL(k_tasks<TT>, dim3(grid), dim3(kThreads), 0, st, pairs, tasks, ntasks,
                                         head, flags, err, arena, codes, w));
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
hipError_t launch_s2_shift(hipStream_t st, const PairDesc *pairs, int npairs, float *arena)
{
    if (npairs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_s2_shift, dim3(64, npairs), dim3(256), 0, st, pairs, arena, arena);
    return hipGetLastError();
}
hipError_t launch_fill2(hipStream_t st, float *p, long long n, float v, long long n2, unsigned v2)
{
    if (n2 <= 0) return launch_fill(st, p, n, v);
    long long blocks = (n + n2 + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill2, dim3((unsigned)blocks), dim3(256), 0, st, p, n, v, n2, v2);
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
    BPX_T_DISPATCH(T, hipLaunchKernelGGL(k_r0tri<TT>, dim3(ntiles, batch), dim3(kThreads), 0, st, nb, tiles, A, B, C));
    return hipGetLastError();
}
hipError_t launch_tri_to_half(hipStream_t st, long long n, const float *src, unsigned short *dst, int *inexact)
{
    if (n <= 0 || n % 4) return hipErrorInvalidValue;
    const long long n4 = n / 4;
    long long blocks = (n4 + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL(k_tri_to_half, dim3((unsigned)blocks), dim3(256), 0, st, (const float4 *)src, n4, (uint2 *)dst,
                       inexact);
    return hipGetLastError();
}
hipError_t launch_r0tri_h(hipStream_t st, int nb, const int2 *tiles, int ntiles, int batch, const unsigned short *Ah,
                          const unsigned short *Bh, float *C)
{
    if (ntiles <= 0 || batch <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_r0tri_h, dim3(ntiles, batch), dim3(kThreads), 0, st, nb, tiles, Ah, Bh, C);
    return hipGetLastError();
}
hipError_t launch_dense_to_tiled(hipStream_t st, int n, int nb, int T, int batch, const float *src, float *dst)
{
    hipLaunchKernelGGL(k_dense_to_tiled, dim3(nb * T, batch), dim3(256), 0, st, n, nb, T, src, dst);
    return hipGetLastError();
}
hipError_t launch_tri_random(hipStream_t st, int n, int nb, int T, int batch, unsigned long long seed, int lo, int hi,
                             float *dst)
{
    hipLaunchKernelGGL(k_tri_random, dim3(nb * T, batch), dim3(256), 0, st, n, nb, T, seed, lo, hi, dst);
    return hipGetLastError();
}
hipError_t launch_tiled_to_dense(hipStream_t st, int n, int nb, int T, int batch, const float *src, float *dst)
{
    hipLaunchKernelGGL(k_tiled_to_dense, dim3(n, batch), dim3(256), 0, st, n, nb, T, src, dst);
    return hipGetLastError();
}

} // namespace bpx
