// One MFMA per element type with the lane layout documented in mfma.h, so the
// layout assumed by the H-matvec kernels can be checked against a host GEMM.
#include "hmx_internal.h"
#include "mfma.h"

namespace hmx {

template <class T>
__global__ void __launch_bounds__(64) k_mfma_selftest(const double* __restrict__ A, const double* __restrict__ B,
                                                     double* __restrict__ D) {
    using M = Mfma<T>;
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
    else throw Error(HMX_EINVAL, "mfma selftest: type must be 0 (f64), 1 (f32) or 2 (f16)");
    HMX_LAUNCHED();
    HMX_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace hmx
