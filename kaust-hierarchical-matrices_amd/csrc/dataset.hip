// Point clouds, random right-hand sides, the device scan used for rank
// offsets, and the sampled-row dense check.
#include "hmx_device.h"
#include "hmx_internal.h"

namespace hmx {

// generateDataset (src/generateDataset.cu:5-19): element e = j*n + i of the
// SoA array gets a 53-bit uniform from Philox(counter = e, key = seed).
__global__ void __launch_bounds__(256) k_points(uint64_t total, uint64_t seed, double* __restrict__ p) {
    uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    uint32_t c[4] = {(uint32_t)e, (uint32_t)(e >> 32), 0u, kTagPoints};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    p[e] = u53(c[0], c[1]);
}

// n x k column-major N(0,1) (replaces src/helperFunctions.cu:59-65)
__global__ void __launch_bounds__(256) k_rhs(uint32_t n, uint32_t k, uint64_t seed, double* __restrict__ x) {
    uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (uint64_t)n * k) return;
    uint32_t s = (uint32_t)(e / n), i = (uint32_t)(e % n);
    x[e] = gauss_at(seed, kTagRhs, 0, s, i);
}

void generate_dataset(Context& c, uint32_t n, uint32_t d, uint64_t seed, double* d_points) {
    uint64_t total = (uint64_t)n * d;
    if (!total) return;
    k_points<<<(unsigned)((total + 255) / 256), 256, 0, c.stream>>>(total, seed, d_points);
    HMX_LAUNCHED();
}

struct PointsJob {
    uint64_t base;  // first value of the cloud among all the launch's values
    uint64_t seed;
    double* out;
};

// many clouds in one launch; cloud k's values are exactly k_points' for its seed
__global__ void __launch_bounds__(256) k_points_batch(uint64_t total, const PointsJob* __restrict__ jobs, int nj) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    int lo = 0, hi = nj - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].base <= e) lo = mid;
        else hi = mid - 1;
    }
    const PointsJob J = jobs[lo];
    const uint64_t i = e - J.base;
    uint32_t c[4] = {(uint32_t)i, (uint32_t)(i >> 32), 0u, kTagPoints};
    philox4x32_10(c, (uint32_t)J.seed, (uint32_t)(J.seed >> 32));
    J.out[i] = u53(c[0], c[1]);
}

void generate_dataset_batch(Context& c, const std::vector<uint64_t>& counts, const std::vector<uint64_t>& seeds,
                            const std::vector<double*>& outs) {
    std::vector<PointsJob> jobs;
    uint64_t total = 0;
    for (size_t k = 0; k < counts.size(); ++k) {
        if (!counts[k]) continue;
        jobs.push_back(PointsJob{total, seeds[k], outs[k]});
        total += counts[k];
    }
    if (!total) return;
    DBuf<PointsJob> dj(jobs.size());
    HMX_HIP(hipMemcpyAsync(dj.p, jobs.data(), jobs.size() * sizeof(PointsJob), hipMemcpyHostToDevice, c.stream));
    k_points_batch<<<(unsigned)((total + 255) / 256), 256, 0, c.stream>>>(total, dj.p, (int)jobs.size());
    HMX_LAUNCHED();
    HMX_HIP(hipStreamSynchronize(c.stream));  // the job table goes out of scope
}

// k vectors (column-major, ld n) between original point order and the
// KD-tree order: position i of the KD order holds original point perm[i]
__global__ void __launch_bounds__(256) k_reorder(uint32_t n, uint32_t k, int toOriginal,
                                                 const int32_t* __restrict__ perm, const double* __restrict__ in,
                                                 double* __restrict__ out) {
    const uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (uint64_t)n * k) return;
    const uint64_t s = e / n, i = e % n;
    const uint64_t o = s * n + (uint64_t)perm[i];
    if (toOriginal) out[o] = in[e];
    else out[e] = in[o];
}

void reorder_vectors(Context& c, uint32_t n, uint32_t k, int toOriginal, const int32_t* d_perm, const double* d_in,
                     double* d_out) {
    const uint64_t total = (uint64_t)n * k;
    if (total == 0) return;
    k_reorder<<<(unsigned)((total + 255) / 256), 256, 0, c.stream>>>(n, k, toOriginal, d_perm, d_in, d_out);
    HMX_LAUNCHED();
}

void generate_random_vector(Context& c, uint32_t k, uint32_t n, uint64_t seed, double* d_out) {
    uint64_t total = (uint64_t)n * k;
    if (!total) return;
    k_rhs<<<(unsigned)((total + 255) / 256), 256, 0, c.stream>>>(n, k, seed, d_out);
    HMX_LAUNCHED();
}

// ---- exclusive scan int32 -> int64 ---------------------------------------
// Three passes: per-block sums, one-block scan of the sums, per-block scan.
constexpr int kScanBlock = 1024;  // elements per block (256 threads x 4)

__global__ void __launch_bounds__(256) k_scan_sums(const int32_t* __restrict__ in, size_t count,
                                                   int64_t* __restrict__ sums) {
    __shared__ int64_t red[256];
    size_t base = (size_t)blockIdx.x * kScanBlock;
    int64_t s = 0;
    for (int q = 0; q < 4; ++q) {
        size_t i = base + threadIdx.x + q * 256;
        if (i < count) s += in[i];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = red[0];
}

// in-place exclusive scan of nb block sums by one block; writes total to sums[nb]
__global__ void __launch_bounds__(256) k_scan_top(int64_t* __restrict__ sums, size_t nb) {
    __shared__ int64_t buf[256];
    __shared__ int64_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (size_t base = 0; base < nb; base += 256) {
        size_t i = base + threadIdx.x;
        int64_t v = i < nb ? sums[i] : 0;
        buf[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {  // Hillis-Steele inclusive scan
            int64_t t = threadIdx.x >= (unsigned)off ? buf[threadIdx.x - off] : 0;
            __syncthreads();
            buf[threadIdx.x] += t;
            __syncthreads();
        }
        int64_t incl = buf[threadIdx.x];
        if (i < nb) sums[i] = carry + incl - v;
        __syncthreads();
        if (threadIdx.x == 255) carry += incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[nb] = carry;
}

__global__ void __launch_bounds__(256) k_scan_block(const int32_t* __restrict__ in, size_t count,
                                                    const int64_t* __restrict__ sums, int64_t* __restrict__ out) {
    __shared__ int64_t buf[256];
    size_t base = (size_t)blockIdx.x * kScanBlock + (size_t)threadIdx.x * 4;
    int64_t v[4];
    int64_t loc = 0;
    for (int q = 0; q < 4; ++q) {
        v[q] = (base + q < count) ? in[base + q] : 0;
        loc += v[q];
    }
    buf[threadIdx.x] = loc;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        int64_t t = threadIdx.x >= (unsigned)off ? buf[threadIdx.x - off] : 0;
        __syncthreads();
        buf[threadIdx.x] += t;
        __syncthreads();
    }
    int64_t run = sums[blockIdx.x] + buf[threadIdx.x] - loc;
    for (int q = 0; q < 4; ++q) {
        if (base + q < count) out[base + q] = run;
        run += v[q];
    }
}

__global__ void k_scan_total(const int64_t* sums, size_t nb, int64_t* out, size_t count) { out[count] = sums[nb]; }

void exclusive_scan_i32_i64(Context& c, const int32_t* in, int64_t* out, size_t count) {
    size_t nb = (count + kScanBlock - 1) / kScanBlock;
    DBuf<int64_t> sums(nb + 1);
    if (nb) {
        k_scan_sums<<<(unsigned)nb, 256, 0, c.stream>>>(in, count, sums.p);
        HMX_LAUNCHED();
    }
    k_scan_top<<<1, 256, 0, c.stream>>>(sums.p, nb);
    HMX_LAUNCHED();
    if (nb) {
        k_scan_block<<<(unsigned)nb, 256, 0, c.stream>>>(in, count, sums.p, out);
        HMX_LAUNCHED();
    }
    k_scan_total<<<1, 1, 0, c.stream>>>(sums.p, nb, out, count);
    HMX_LAUNCHED();
    HMX_HIP(hipStreamSynchronize(c.stream));  // sums is freed on return
}

// ---- sampled-row dense product (EXPAND_MATRIX check on a row sample) ------
// One block per sampled row; threads stride over the n columns; block
// reduction in a fixed order, so the result is deterministic.
template <int K>
__global__ void __launch_bounds__(256) k_dense_rows(uint32_t n, uint32_t d, const double* __restrict__ p,
                                                    const int32_t* __restrict__ perm, double ell,
                                                    const int32_t* __restrict__ rows, uint32_t nr, uint32_t k,
                                                    const double* __restrict__ X, double* __restrict__ Y) {
    __shared__ double red[K][256];
    uint32_t r = blockIdx.x;
    int32_t pr = perm[rows[r]];
    double xr[3] = {0, 0, 0};
    for (uint32_t j = 0; j < d; ++j) xr[j] = p[(size_t)j * n + pr];
    for (uint32_t s0 = 0; s0 < k; s0 += K) {
        double acc[K];
#pragma unroll
        for (int q = 0; q < K; ++q) acc[q] = 0.0;
        for (uint32_t j = threadIdx.x; j < n; j += 256) {
            int32_t pj = perm[j];
            double xc[3] = {0, 0, 0};
            for (uint32_t t = 0; t < d; ++t) xc[t] = p[(size_t)t * n + pj];
            double kv = kernel_eval_d(xr, xc, (int)d, ell);
#pragma unroll
            for (int q = 0; q < K; ++q)
                if (s0 + q < k) acc[q] += kv * X[(size_t)(s0 + q) * n + j];
        }
#pragma unroll
        for (int q = 0; q < K; ++q) red[q][threadIdx.x] = acc[q];
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (threadIdx.x < (unsigned)w)
#pragma unroll
                for (int q = 0; q < K; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0)
            for (int q = 0; q < K && s0 + q < k; ++q) Y[(size_t)(s0 + q) * nr + r] = red[q][0];
        __syncthreads();
    }
}

void dense_rows(Context& c, uint32_t n, uint32_t d, const double* p, const int32_t* perm, double ell, uint32_t nr,
                const int32_t* rows, uint32_t k, const double* X, double* Y) {
    HMX_CHECK(d >= 1 && d <= 3, HMX_EINVAL, "dense rows: d must be 1..3");
    if (!nr || !k) return;
    k_dense_rows<8><<<nr, 256, 0, c.stream>>>(n, d, p, perm, ell, rows, nr, k, X, Y);
    HMX_LAUNCHED();
}

}  // namespace hmx
