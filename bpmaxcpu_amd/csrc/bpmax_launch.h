// bpmax_launch.h -- host-side launch wrappers of the kernels in bpmax_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>
#include "bpmax_internal.h"

namespace bpx {

// half: every pair of the launch has an FP16 mirror of F (PairDesc::Fh); with T = 64 the
// R0 operands then come from the mirrors (k_outer_h)
hipError_t launch_outer(hipStream_t st, int T, const PairDesc *pairs, const WorkItem *items, int n,
                        float *arena, const int *codes, const float *w, bool half = false);
hipError_t launch_inner_ext(hipStream_t st, int T, bool single_strand, const PairDesc *pairs,
                            const WorkItem *items, int n, float *arena);
// F: wide = 1024-lane workgroups (one cell per lane), else 512 (two cells per lane) or, with
// blocked, 256 (one 2 x 2 block per lane, one step per block diagonal); S2: 1024 or 256
hipError_t launch_inner_intra(hipStream_t st, int T, bool single_strand, const PairDesc *pairs,
                              const WorkItem *items, int n, float *arena, const int *codes, const float *w,
                              bool wide, bool blocked);
// in-tile closure of one logical diagonal (blocked form) with the next diagonal's early
// cross-tile steps (items typed TASK_INTRA / TASK_EXT), 256-lane workgroups
hipError_t launch_inner_closure(hipStream_t st, int T, const PairDesc *pairs, const WorkItem *items, int n,
                                float *arena, const int *codes, const float *w);
hipError_t launch_tasks(hipStream_t st, int T, const PairDesc *pairs, const WorkItem *tasks, int ntasks, int *head,
                        int *flags, int *err, float *arena, const int *codes, const float *w);
hipError_t launch_s1(hipStream_t st, const PairDesc *pairs, int npairs, float *arena, const int *codes,
                     const float *w);
hipError_t launch_fill(hipStream_t st, float *p, long long n, float v);
// FP16 row- and column-shifted copies of S2 for the pairs that have them (PairDesc::S2r, S2c)
hipError_t launch_s2_shift(hipStream_t st, const PairDesc *pairs, int npairs, float *arena);
// p[0, n) = v, then n2 words of the bit pattern v2
hipError_t launch_fill2(hipStream_t st, float *p, long long n, float v, long long n2, unsigned v2);
hipError_t launch_to_packed(hipStream_t st, const float *arena, const PairDesc &P, float *out);
hipError_t launch_to_dense(hipStream_t st, const float *arena, const PairDesc &P, float *out);
hipError_t launch_scores(hipStream_t st, const PairDesc *pairs, int npairs, const float *arena, float *scores);
hipError_t launch_r0tri(hipStream_t st, int T, int nb, const int2 *tiles, int ntiles, int batch, const float *A,
                        const float *B, float *C);
// FP16 copy of n floats (n a multiple of 4); *inexact is set (never cleared) when a value is
// not -inf or an integer of magnitude <= 1024
hipError_t launch_tri_to_half(hipStream_t st, long long n, const float *src, unsigned short *dst, int *inexact);
// isolated R0 on 64 x 64 tiles from those FP16 copies
hipError_t launch_r0tri_h(hipStream_t st, int nb, const int2 *tiles, int ntiles, int batch, const unsigned short *Ah,
                          const unsigned short *Bh, float *C);
hipError_t launch_dense_to_tiled(hipStream_t st, int n, int nb, int T, int batch, const float *src, float *dst);
hipError_t launch_tiled_to_dense(hipStream_t st, int n, int nb, int T, int batch, const float *src, float *dst);
hipError_t launch_tri_random(hipStream_t st, int n, int nb, int T, int batch, unsigned long long seed, int lo, int hi,
                             float *dst);

} // namespace bpx
