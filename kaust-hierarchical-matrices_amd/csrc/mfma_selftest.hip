// One MFMA per element type with the lane layout documented in mfma.h, so the
// layout assumed by the H-matvec kernels can be checked against a host GEMM.
#include "hmx_internal.h"
#include "mfma.h"

namespace hmx {

template <class T, class M = Mfma<T>>
__global__ void __launch_bounds__(64) k_mfma_selftest(const double* __restrict__ A, const double* __restrict__ B,
                                                     double* __restrict__ D) {
    constexpr int KL = M::KL, K = 4 * KL;
    const int l = threadIdx.x;
    T a[KL], b[KL];
    for (int e = 0; e < KL; ++e) {
        const int k = KL * (l / 16) + e;
        a[e] = (T)A[(l % 16) * K + k];
        b[e] = (T)B[k * 16 + (l % 16)];
    }
    typename M::accv acc = {0, 0, 0, 0};
    acc = M::op(M::make(a), M::make(b), acc);
    for (int v = 0; v < 4; ++v) D[M::out_row(l / 16, v) * 16 + (l % 16)] = (double)acc[v];
}

void mfma_selftest(Context& c, int type, const double* A, const double* B, double* D) {
    if (type == 0) k_mfma_selftest<double><<<1, 64, 0, c.stream>>>(A, B, D);
    else if (type == 1) k_mfma_selftest<float><<<1, 64, 0, c.stream>>>(A, B, D);
    else if (type == 2) k_mfma_selftest<_Float16><<<1, 64, 0, c.stream>>>(A, B, D);
    else if (type == 3) k_mfma_selftest<_Float16, Mfma32h><<<1, 64, 0, c.stream>>>(A, B, D);
    else throw Error(HMX_EINVAL, "mfma selftest: type must be 0 (f64), 1 (f32), 2 (f16) or 3 (f16, K = 32)");
    HMX_LAUNCHED();
    HMX_HIP(hipStreamSynchronize(c.stream));
}

// Streaming-read ceiling: every byte of a caller buffer read once per pass
// with 16-byte non-temporal loads, 4 in flight per lane, grid-stride.  The
// H-matvec kernels' bandwidth is quoted against the spec peak and against
// this measured ceiling.
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) k_stream_read(const u32x4_t* __restrict__ p, uint64_t n, uint32_t* __restrict__ sink) {
    uint64_t i = (uint64_t)blockIdx.x * 1024 + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * 1024;
    uint32_t acc = 0;
    for (; i + 768 < n; i += stride) {
        u32x4_t v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = __builtin_nontemporal_load(p + i + 256 * u);
#pragma unroll
        for (int u = 0; u < 4; ++u) acc ^= v[u].x ^ v[u].y ^ v[u].z ^ v[u].w;
    }
    for (; i < n; i += 256) acc ^= p[i].x;  // tail (n not a multiple of 1024)
    if (acc == 0x9e3779b9u) sink[0] = acc;   // keeps the loads; practically never taken
}

// FP64 arithmetic ceiling: every wave runs `iters` rounds of 8 independent
// accumulation chains, either v_mfma_f64_16x16x4f64 (unit 0; 2048 flops
// each) or v_fma_f64 (unit 1; 8 chains x 8 steps, 2 flops per lane each), on
// register operands only.  The matrix-core products and the samplers are
// quoted against the spec peak and against this measured rate.
typedef double f64x4_t __attribute__((ext_vector_type(4)));
template <int UNIT>
__global__ void __launch_bounds__(256) k_fp64_peak(int iters, double seed, double* __restrict__ sink) {
    const double x = seed + 1e-3 * (threadIdx.x & 63), y = 1.0 - 1e-9 * blockIdx.x;
    double s = 0.0;
    if constexpr (UNIT == 0) {
        f64x4_t acc[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[c] = f64x4_t{0.0, 0.0, 0.0, 0.0};
        for (int i = 0; i < iters; ++i)
#pragma unroll
            for (int c = 0; c < 8; ++c) acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, acc[c], 0, 0, 0);
#pragma unroll
        for (int c = 0; c < 8; ++c) s += acc[c][0] + acc[c][3];
    } else {
        double a[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) a[c] = x + c;
        for (int i = 0; i < iters; ++i)
#pragma unroll
            for (int u = 0; u < 8; ++u)
#pragma unroll
                for (int c = 0; c < 8; ++c) a[c] = fma(a[c], y, x);
#pragma unroll
        for (int c = 0; c < 8; ++c) s += a[c];
    }
    if (s == 0.123456789) sink[0] = s;  // keeps the arithmetic; practically never taken
}

void fp64_peak(Context& c, int unit, int iters, float* tflops) {
    HMX_CHECK(unit == 0 || unit == 1, HMX_EINVAL, "FP64 probe: unit 0 (matrix cores) or 1 (vector FMA)");
    HMX_CHECK(iters >= 1 && iters <= (1 << 20), HMX_EINVAL, "FP64 probe: 1 <= iters <= 2^20");
    DBuf<double> sink(1);
    hipDeviceProp_t prop;
    HMX_HIP(hipGetDeviceProperties(&prop, c.device));
    const unsigned blocks = (unsigned)prop.multiProcessorCount * 8;  // 8 waves per SIMD
    auto launch = [&](int it) {
        if (unit == 0) k_fp64_peak<0><<<blocks, 256, 0, c.stream>>>(it, 0.5, sink.p);
        else k_fp64_peak<1><<<blocks, 256, 0, c.stream>>>(it, 0.5, sink.p);
        HMX_LAUNCHED();
    };
    launch(std::max(1, iters / 10));  // warm-up
    hipEvent_t a, b;
    HMX_HIP(hipEventCreate(&a));
    HMX_HIP(hipEventCreate(&b));
    HMX_HIP(hipEventRecord(a, c.stream));
    launch(iters);
    HMX_HIP(hipEventRecord(b, c.stream));
    HMX_HIP(hipEventSynchronize(b));
    float t = 0;
    HMX_HIP(hipEventElapsedTime(&t, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    const double waves = (double)blocks * 4;
    const double flops = unit == 0 ? waves * iters * 8 * 2048.0 : waves * 64 * iters * 64 * 2.0;
    *tflops = (float)(flops / (t * 1e-3) / 1e12);
}

void stream_read(Context& c, const void* buf, uint64_t bytes, int iters, float* ms) {
    HMX_CHECK(bytes >= (1u << 20) && bytes % 16 == 0, HMX_EINVAL, "stream probe: buffer of >= 1 MiB, a multiple of 16 bytes");
    HMX_CHECK(iters >= 1, HMX_EINVAL, "stream probe: iters >= 1");
    DBuf<uint32_t> sink(1);
    hipEvent_t a, b;
    HMX_HIP(hipEventCreate(&a));
    HMX_HIP(hipEventCreate(&b));
    const uint64_t n = bytes / 16;
    k_stream_read<<<16384, 256, 0, c.stream>>>((const u32x4_t*)buf, n, sink.p);  // warm-up
    HMX_LAUNCHED();
    HMX_HIP(hipEventRecord(a, c.stream));
    for (int r = 0; r < iters; ++r) {
        k_stream_read<<<16384, 256, 0, c.stream>>>((const u32x4_t*)buf, n, sink.p);
        HMX_LAUNCHED();
    }
    HMX_HIP(hipEventRecord(b, c.stream));
    HMX_HIP(hipEventSynchronize(b));
    float t = 0;
    HMX_HIP(hipEventElapsedTime(&t, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms = t / iters;
}

}  // namespace hmx
