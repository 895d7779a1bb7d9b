// KD-tree construction, FULL_TREE division (src/kDTreeConstruction.cu:14-231).
//
// Per level: (1) segmented min/max over every dimension -> bounding boxes and
// split dimension (first maximal span, CUB ArgMax semantics); (2) stable
// ascending sort of each segment by the split coordinate; (3) each segment is
// split at a + ceil((e-a)/2).  Offsets depend only on n, so they are computed
// on the host.  Sort: segments of <= 2048 points are sorted by one workgroup
// in LDS (bitonic network on the unique (key, position) pair, which equals a
// stable sort); larger segments use a segmented LSD radix sort (8 passes of 8
// bits, stable wave-level multisplit).
#include <algorithm>
#include <cfloat>

#include "hmx_device.h"
#include "hmx_internal.h"

namespace hmx {

constexpr int kBoxChunk = 4096;
constexpr int kSmallSeg = 2048;
constexpr int kRadixChunk = 2048;

// partial min/max of one chunk of one segment, all dims (d <= 3)
__global__ void __launch_bounds__(256) k_bbox_partial(uint32_t n, uint32_t d, const double* __restrict__ p,
                                                      const int32_t* __restrict__ perm,
                                                      const int2* __restrict__ chunks,  // {start, end}
                                                      double* __restrict__ pmin, double* __restrict__ pmax) {
    __shared__ double smin[3][256], smax[3][256];
    int2 ch = chunks[blockIdx.x];
    double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int i = ch.x + threadIdx.x; i < ch.y; i += 256) {
        int32_t pi = perm[i];
        for (uint32_t j = 0; j < d; ++j) {
            double v = p[(size_t)j * n + pi];
            mn[j] = v < mn[j] ? v : mn[j];
            mx[j] = v > mx[j] ? v : mx[j];
        }
    }
    for (uint32_t j = 0; j < d; ++j) { smin[j][threadIdx.x] = mn[j]; smax[j][threadIdx.x] = mx[j]; }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < (unsigned)w)
            for (uint32_t j = 0; j < d; ++j) {
                double a = smin[j][threadIdx.x + w], b = smax[j][threadIdx.x + w];
                if (a < smin[j][threadIdx.x]) smin[j][threadIdx.x] = a;
                if (b > smax[j][threadIdx.x]) smax[j][threadIdx.x] = b;
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (uint32_t j = 0; j < d; ++j) {
            pmin[(size_t)blockIdx.x * d + j] = smin[j][0];
            pmax[(size_t)blockIdx.x * d + j] = smax[j][0];
        }
}

// one thread per segment: fold the chunk partials, write the level's boxes
// and the split dimension (findSpan + ArgMax, src/kDTreeHelpers.cu:52-64)
__global__ void k_bbox_final(uint32_t nseg, uint32_t d, const int32_t* __restrict__ segChunk,
                             const double* __restrict__ pmin, const double* __restrict__ pmax,
                             double* __restrict__ bbmin, double* __restrict__ bbmax, int32_t* __restrict__ splitDim) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int c = segChunk[s]; c < segChunk[s + 1]; ++c)
        for (uint32_t j = 0; j < d; ++j) {
            double a = pmin[(size_t)c * d + j], b = pmax[(size_t)c * d + j];
            if (a < mn[j]) mn[j] = a;
            if (b > mx[j]) mx[j] = b;
        }
    int dim = 0;
    double best = mx[0] - mn[0];
    for (uint32_t j = 1; j < d; ++j) {
        double sp = mx[j] - mn[j];
        if (sp > best) { best = sp; dim = (int)j; }
    }
    for (uint32_t j = 0; j < d; ++j) { bbmin[(size_t)s * d + j] = mn[j]; bbmax[(size_t)s * d + j] = mx[j]; }
    splitDim[s] = dim;
}

// ---- small segments: bitonic sort in LDS ---------------------------------
__global__ void __launch_bounds__(256) k_sort_small(uint32_t n, const double* __restrict__ p,
                                                    int32_t* __restrict__ perm, const int32_t* __restrict__ segs,
                                                    const int32_t* __restrict__ offsets,
                                                    const int32_t* __restrict__ splitDim) {
    __shared__ uint64_t key[kSmallSeg];
    __shared__ int32_t pos[kSmallSeg];
    __shared__ int32_t val[kSmallSeg];
    int s = segs[blockIdx.x];
    int a = offsets[s], e = offsets[s + 1];
    int size = e - a;
    if (size <= 1) return;
    int P = 1;
    while (P < size) P <<= 1;
    const double* pd = p + (size_t)splitDim[s] * n;
    for (int i = threadIdx.x; i < P; i += 256) {
        if (i < size) {
            int32_t v = perm[a + i];
            val[i] = v;
            key[i] = dkey(pd[v]);
            pos[i] = i;
        } else {
            key[i] = ~0ull;
            pos[i] = P + i;
        }
    }
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < P; i += 256) {
                int ixj = i ^ j;
                if (ixj > i) {
                    uint64_t ka = key[i], kb = key[ixj];
                    int pa = pos[i], pb = pos[ixj];
                    bool gt = (ka > kb) || (ka == kb && pa > pb);
                    bool asc = (i & k) == 0;
                    if (gt == asc) {
                        key[i] = kb; key[ixj] = ka;
                        pos[i] = pb; pos[ixj] = pa;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < size; i += 256) perm[a + i] = val[pos[i]];
}

// ---- large segments: segmented LSD radix sort ------------------------------
__global__ void __launch_bounds__(256) k_make_keys(uint32_t n, const double* __restrict__ p,
                                                   const int32_t* __restrict__ perm, const int4* __restrict__ chunks,
                                                   const int32_t* __restrict__ splitDim, uint64_t* __restrict__ keys) {
    int4 ch = chunks[blockIdx.x];  // {start, end, segment, -}
    const double* pd = p + (size_t)splitDim[ch.z] * n;
    for (int i = ch.x + threadIdx.x; i < ch.y; i += 256) keys[i] = dkey(pd[perm[i]]);
}

__global__ void __launch_bounds__(256) k_radix_hist(const uint64_t* __restrict__ keys, const int4* __restrict__ chunks,
                                                    int shift, int32_t* __restrict__ hist) {
    __shared__ int32_t cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    int4 ch = chunks[blockIdx.x];
    for (int i = ch.x + threadIdx.x; i < ch.y; i += 256) atomicAdd(&cnt[(keys[i] >> shift) & 255], 1);
    __syncthreads();
    hist[(size_t)blockIdx.x * 256 + threadIdx.x] = cnt[threadIdx.x];
}

// one block per large segment; thread t owns digit t
__global__ void __launch_bounds__(256) k_radix_offsets(const int32_t* __restrict__ hist,
                                                       const int4* __restrict__ segInfo,  // {start, chunk0, chunk1, -}
                                                       int32_t* __restrict__ choff) {
    __shared__ int32_t buf[256];
    int4 si = segInfo[blockIdx.x];
    int d = threadIdx.x;
    int tot = 0;
    for (int c = si.y; c < si.z; ++c) tot += hist[(size_t)c * 256 + d];
    buf[d] = tot;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        int t = d >= off ? buf[d - off] : 0;
        __syncthreads();
        buf[d] += t;
        __syncthreads();
    }
    int run = si.x + buf[d] - tot;
    for (int c = si.y; c < si.z; ++c) {
        choff[(size_t)c * 256 + d] = run;
        run += hist[(size_t)c * 256 + d];
    }
}

// stable scatter: 4 waves, wave w owns elements [w*512, w*512+512) of the chunk
__global__ void __launch_bounds__(256) k_radix_scatter(const uint64_t* __restrict__ keys,
                                                       const int32_t* __restrict__ vin, const int4* __restrict__ chunks,
                                                       int shift, const int32_t* __restrict__ choff,
                                                       uint64_t* __restrict__ kout, int32_t* __restrict__ vout) {
    __shared__ int32_t cnt[4][256];
    __shared__ int32_t woff[4][256];
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < 4 * 256; i += 256) (&cnt[0][0])[i] = 0;
    __syncthreads();
    int4 ch = chunks[blockIdx.x];
    const int per = kRadixChunk / 4;
    int32_t rk[per / 64];
    int32_t dg[per / 64];
    uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
#pragma unroll
    for (int r = 0; r < per / 64; ++r) {
        int i = ch.x + w * per + r * 64 + lane;
        bool valid = i < ch.y;
        int d = valid ? (int)((keys[i] >> shift) & 255) : 0;
        uint64_t mask = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            bool on = (d >> bit) & 1;
            uint64_t bal = __ballot(valid && on);
            mask &= on ? bal : ~bal;
        }
        int before = valid ? cnt[w][d] : 0;
        int rl = __popcll(mask & lt);
        __builtin_amdgcn_wave_barrier();
        if (valid && rl == 0) cnt[w][d] = before + __popcll(mask);
        __builtin_amdgcn_wave_barrier();
        rk[r] = before + rl;
        dg[r] = valid ? d : -1;
    }
    __syncthreads();
    {
        int d = threadIdx.x;
        int run = 0;
        for (int q = 0; q < 4; ++q) {
            woff[q][d] = run;
            run += cnt[q][d];
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < per / 64; ++r) {
        if (dg[r] < 0) continue;
        int i = ch.x + w * per + r * 64 + lane;
        int dst = choff[(size_t)blockIdx.x * 256 + dg[r]] + woff[w][dg[r]] + rk[r];
        kout[dst] = keys[i];
        vout[dst] = vin[i];
    }
}

void build_kdtree(Context& c, KDTree& t, const double* d_points) {
    const uint32_t n = t.n, d = t.d, b = t.b;
    HMX_CHECK(d >= 1 && d <= 3, HMX_EINVAL, "dimensionOfInputPoints must be 1, 2 or 3");
    HMX_CHECK(b >= 1, HMX_EINVAL, "bucketSize must be >= 1");
    HMX_CHECK(n >= 1, HMX_EINVAL, "numberOfInputPoints must be >= 1");
    // FULL_TREE and DIVIDE_IN_HALF run the same construction in the reference
    // (the shared branches of src/kDTreeConstruction.cu); POWER_OF_TWO_ON_LEFT
    // cuts every level at multiples of a power-of-two size that starts at
    // upperPowerOfTwo(n) and halves while it exceeds b (fillOffsets,
    // src/kDTreeHelpers.cu:29-43)
    const bool p2l = t.divMethod == HMX_POWER_OF_TWO_ON_LEFT;
    uint32_t L, size = 0;
    if (p2l) {
        size = upper_pow2(n);
        L = 0;
        while (size > b) { size >>= 1; ++L; }
        t.numSegments = (n + size - 1) / size;
    } else {
        L = kd_num_levels(n, b);
        t.numSegments = 1u << L;
    }
    t.numLevels = L;
    const size_t nbox = ((size_t)2 << L) - 1;
    t.perm.alloc(n);
    t.offsets.alloc(t.numSegments + 1);
    t.bbmin.alloc(nbox * d);
    t.bbmax.alloc(nbox * d);
    DBuf<int32_t> splitDim(std::max<size_t>(1, t.numSegments));

    // identity permutation (initIndexMap, src/kDTreeHelpers.cu:12-26)
    std::vector<int32_t> h_id(n);
    for (uint32_t i = 0; i < n; ++i) h_id[i] = (int32_t)i;
    HMX_HIP(hipMemcpyAsync(t.perm.p, h_id.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, c.stream));

    DBuf<int32_t> permB(n);
    DBuf<uint64_t> keyA(n), keyB(n);
    std::vector<int32_t> off = {0, (int32_t)n};
    DBuf<int32_t> d_off(t.numSegments + 1);

    for (uint32_t level = 0; level < L; ++level) {
        const uint32_t lsize = p2l ? size << (L - level) : 0;
        if (p2l) {  // segments [s*lsize, min((s+1)*lsize, n))
            const uint32_t ns = (n + lsize - 1) / lsize;
            off.assign(ns + 1, 0);
            for (uint32_t q = 0; q < ns; ++q) off[q] = (int32_t)(q * lsize);
            off[ns] = (int32_t)n;
        }
        const uint32_t nseg = (uint32_t)off.size() - 1;
        // ---- bounding boxes
        std::vector<int2> bchunks;
        std::vector<int32_t> segChunk(nseg + 1);
        for (uint32_t s = 0; s < nseg; ++s) {
            segChunk[s] = (int32_t)bchunks.size();
            for (int32_t a = off[s]; a < off[s + 1]; a += kBoxChunk)
                bchunks.push_back(make_int2(a, std::min(a + kBoxChunk, off[s + 1])));
        }
        segChunk[nseg] = (int32_t)bchunks.size();
        DBuf<int2> d_bch(std::max<size_t>(1, bchunks.size()));
        DBuf<int32_t> d_sch(nseg + 1);
        DBuf<double> pmin(std::max<size_t>(1, bchunks.size() * d)), pmax(std::max<size_t>(1, bchunks.size() * d));
        HMX_HIP(hipMemcpyAsync(d_bch.p, bchunks.data(), bchunks.size() * sizeof(int2), hipMemcpyHostToDevice,
                               c.stream));
        HMX_HIP(hipMemcpyAsync(d_sch.p, segChunk.data(), (nseg + 1) * sizeof(int32_t), hipMemcpyHostToDevice,
                               c.stream));
        HMX_HIP(hipMemcpyAsync(d_off.p, off.data(), (nseg + 1) * sizeof(int32_t), hipMemcpyHostToDevice, c.stream));
        if (!bchunks.empty()) {
            k_bbox_partial<<<(unsigned)bchunks.size(), 256, 0, c.stream>>>(n, d, d_points, t.perm.p, d_bch.p, pmin.p,
                                                                            pmax.p);
            HMX_LAUNCHED();
        }
        size_t base = ((size_t)1 << level) - 1;
        k_bbox_final<<<(nseg + 255) / 256, 256, 0, c.stream>>>(nseg, d, d_sch.p, pmin.p, pmax.p,
                                                                t.bbmin.p + base * d, t.bbmax.p + base * d, splitDim.p);
        HMX_LAUNCHED();

        // ---- sort each segment by its split coordinate
        std::vector<int32_t> small;
        std::vector<int4> rchunks;
        std::vector<int4> rsegs;
        for (uint32_t s = 0; s < nseg; ++s) {
            int32_t a = off[s], e = off[s + 1];
            if (e - a <= 1) continue;
            if (e - a <= kSmallSeg) {
                small.push_back((int32_t)s);
            } else {
                int4 si = make_int4(a, (int)rchunks.size(), 0, 0);
                for (int32_t x = a; x < e; x += kRadixChunk)
                    rchunks.push_back(make_int4(x, std::min(x + kRadixChunk, e), (int)s, 0));
                si.z = (int)rchunks.size();
                rsegs.push_back(si);
            }
        }
        if (!rchunks.empty()) {
            DBuf<int4> d_rch(rchunks.size()), d_rseg(rsegs.size());
            DBuf<int32_t> hist(rchunks.size() * 256), choff(rchunks.size() * 256);
            HMX_HIP(hipMemcpyAsync(d_rch.p, rchunks.data(), rchunks.size() * sizeof(int4), hipMemcpyHostToDevice,
                                   c.stream));
            HMX_HIP(hipMemcpyAsync(d_rseg.p, rsegs.data(), rsegs.size() * sizeof(int4), hipMemcpyHostToDevice,
                                   c.stream));
            const unsigned nch = (unsigned)rchunks.size();
            k_make_keys<<<nch, 256, 0, c.stream>>>(n, d_points, t.perm.p, d_rch.p, splitDim.p, keyA.p);
            HMX_LAUNCHED();
            uint64_t* kin = keyA.p;
            uint64_t* kout = keyB.p;
            int32_t* vin = t.perm.p;
            int32_t* vout = permB.p;
            for (int pass = 0; pass < 8; ++pass) {
                int shift = pass * 8;
                k_radix_hist<<<nch, 256, 0, c.stream>>>(kin, d_rch.p, shift, hist.p);
                HMX_LAUNCHED();
                k_radix_offsets<<<(unsigned)rsegs.size(), 256, 0, c.stream>>>(hist.p, d_rseg.p, choff.p);
                HMX_LAUNCHED();
                k_radix_scatter<<<nch, 256, 0, c.stream>>>(kin, vin, d_rch.p, shift, choff.p, kout, vout);
                HMX_LAUNCHED();
                std::swap(kin, kout);
                std::swap(vin, vout);
            }
            // 8 passes: data is back in (keyA, t.perm)
            HMX_HIP(hipStreamSynchronize(c.stream));
        }
        if (!small.empty()) {
            DBuf<int32_t> d_small(small.size());
            HMX_HIP(hipMemcpyAsync(d_small.p, small.data(), small.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                                   c.stream));
            k_sort_small<<<(unsigned)small.size(), 256, 0, c.stream>>>(n, d_points, t.perm.p, d_small.p, d_off.p,
                                                                       splitDim.p);
            HMX_LAUNCHED();
            HMX_HIP(hipStreamSynchronize(c.stream));
        }
        if (p2l) {  // the next level's cuts come from the halved size
            HMX_HIP(hipStreamSynchronize(c.stream));
            continue;
        }
        // ---- split (fillOffsetsSort, src/kDTreeHelpers.cu:81-93)
        std::vector<int32_t> noff(2 * nseg + 1);
        for (uint32_t s = 0; s < nseg; ++s) {
            noff[2 * s] = off[s];
            noff[2 * s + 1] = (off[s + 1] - off[s] + 1) / 2 + off[s];
        }
        noff[2 * nseg] = (int32_t)n;
        off.swap(noff);
        HMX_HIP(hipStreamSynchronize(c.stream));  // per-level host vectors go out of scope
    }
    if (p2l) {
        off.assign(t.numSegments + 1, 0);
        for (uint32_t q = 0; q < t.numSegments; ++q) off[q] = (int32_t)(q * size);
        off[t.numSegments] = (int32_t)n;
    }
    HMX_HIP(hipMemcpyAsync(t.offsets.p, off.data(), off.size() * sizeof(int32_t), hipMemcpyHostToDevice, c.stream));
    t.h_offsets = off;
    t.h_bbmin.resize(nbox * d);
    t.h_bbmax.resize(nbox * d);
    HMX_HIP(hipMemcpyAsync(t.h_bbmin.data(), t.bbmin.p, nbox * d * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    HMX_HIP(hipMemcpyAsync(t.h_bbmax.data(), t.bbmax.p, nbox * d * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    HMX_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace hmx
