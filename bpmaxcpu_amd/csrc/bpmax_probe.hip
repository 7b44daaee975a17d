// bpmax_probe.hip -- calibration kernels for hardware counters (libbpmax_probe.so).
// MEASUREMENT SUPPORT ONLY: not part of libbpmax_hip.so.
//
// k_valu_probe issues a known number of VALU instructions per wave: `iters` rounds of 16
// v_max3_f32 and 16 v_add_f32 (inline asm, so nothing is folded or regrouped), with the
// loop counter in scalar registers.  Reading SQ_INSTS_VALU / SQ_ACTIVE_INST_VALU /
// SQ_BUSY_CYCLES over two launches with different `iters` gives the counters' scale
// (instructions counted per wave instruction issued) and the fixed per-wave overhead, which
// turns the counters of the engine's kernels into wave-instruction counts.
#include <hip/hip_runtime.h>

#define MAX3_4                                                                    \
    "v_max3_f32 %0, %0, %1, %2\n v_max3_f32 %0, %0, %2, %1\n v_max3_f32 %0, %0, %1, %2\n" \
    " v_max3_f32 %0, %0, %2, %1\n"
#define ADD_4 "v_add_f32 %1, %1, %2\n v_add_f32 %2, %2, %1\n v_add_f32 %1, %1, %2\n v_add_f32 %2, %2, %1\n"

__global__ __launch_bounds__(256) void k_valu_probe(float *out, int iters)
{
    float a = (float)threadIdx.x, b = 1.0f, c = -1.0f;
    for (int i = 0; i < iters; ++i) {
        asm volatile(MAX3_4 MAX3_4 MAX3_4 MAX3_4 : "+v"(a) : "v"(b), "v"(c));
        asm volatile(ADD_4 ADD_4 ADD_4 ADD_4 : "+v"(a), "+v"(b), "+v"(c));
    }
    out[blockIdx.x * blockDim.x + threadIdx.x] = a + b + c;
}

// The same with packed FP16 instructions (the R0 product's kind): `iters` rounds of 16
// v_pk_maximum3_f16 and 16 v_pk_add_f16.  Packed 16-bit VALU instructions issue at half the
// rate of FP32 ones on this part, which this probe shows next to k_valu_probe.
#define PKMAX3_4                                                                              \
    "v_pk_maximum3_f16 %0, %0, %1, %2\n v_pk_maximum3_f16 %0, %0, %2, %1\n"                 \
    " v_pk_maximum3_f16 %0, %0, %1, %2\n v_pk_maximum3_f16 %0, %0, %2, %1\n"
#define PKADD_4                                                                               \
    "v_pk_add_f16 %1, %1, %2\n v_pk_add_f16 %2, %2, %1\n v_pk_add_f16 %1, %1, %2\n v_pk_add_f16 %2, %2, %1\n"

__global__ __launch_bounds__(256) void k_valu_probe_pk(float *out, int iters)
{
    unsigned a = threadIdx.x, b = 0x3c003c00u, c = 0xbc00bc00u;
    for (int i = 0; i < iters; ++i) {
        asm volatile(PKMAX3_4 PKMAX3_4 PKMAX3_4 PKMAX3_4 : "+v"(a) : "v"(b), "v"(c));
        asm volatile(PKADD_4 PKADD_4 PKADD_4 PKADD_4 : "+v"(a), "+v"(b), "+v"(c));
    }
    out[blockIdx.x * blockDim.x + threadIdx.x] = (float)(a ^ b ^ c);
}

// Launches k_valu_probe on `blocks` 256-lane workgroups (4 waves each) with `iters` rounds
// of 32 VALU instructions, waits, and returns 0 or the HIP error code.  `out` receives
// blocks*256 floats (device memory, allocated and freed here).
// packed: k_valu_probe_pk instead (same instruction counts)
static int probe(int blocks, int iters, bool packed, double *ms)
{
    if (blocks < 1 || iters < 0) return -1;
    float *out = nullptr;
    if (hipMalloc(&out, sizeof(float) * 256 * (size_t)blocks) != hipSuccess) return -2;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, nullptr);
    if (packed) hipLaunchKernelGGL(k_valu_probe_pk, dim3(blocks), dim3(256), 0, nullptr, out, iters);
    else hipLaunchKernelGGL(k_valu_probe, dim3(blocks), dim3(256), 0, nullptr, out, iters);
    (void)hipEventRecord(e1, nullptr);
    hipError_t e = hipEventSynchronize(e1);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (ms) *ms = t;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(out);
    return e == hipSuccess ? 0 : (int)e;
}

extern "C" int bpmax_probe_valu(int blocks, int iters, double *ms) { return probe(blocks, iters, false, ms); }
extern "C" int bpmax_probe_valu_pk(int blocks, int iters, double *ms) { return probe(blocks, iters, true, ms); }
