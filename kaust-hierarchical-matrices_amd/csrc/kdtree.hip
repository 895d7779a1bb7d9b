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
//
// Many clouds of the same dimension are built as one forest: their points
// side by side, every level's segments of every cloud sorted by the same
// launches, so a batch of clouds costs launches per level, not per cloud.
// Each segment is processed exactly as in a one-cloud build, so every tree
// is bitwise the tree its cloud gets alone.
#include <algorithm>
#include <cfloat>

#include "hmx_device.h"
#include "hmx_internal.h"

namespace hmx {

constexpr int kBoxChunk = 4096;
constexpr int kSmallSeg = 2048;
constexpr int kRadixChunk = 2048;

// partial min/max of one chunk of one segment, all dims (d <= 3)
__global__ void __launch_bounds__(256) k_bbox_partial(int64_t n, uint32_t d, const double* __restrict__ p,
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
                             const int64_t* __restrict__ node, double* __restrict__ bbmin, double* __restrict__ bbmax,
                             int32_t* __restrict__ splitDim) {
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
    const int64_t o = node[s] * d;  // the segment's box slot
    for (uint32_t j = 0; j < d; ++j) { bbmin[o + j] = mn[j]; bbmax[o + j] = mx[j]; }
    splitDim[s] = dim;
}

// ---- small segments: bitonic sort in LDS ---------------------------------
__global__ void __launch_bounds__(256) k_sort_small(int64_t n, const double* __restrict__ p,
                                                    int32_t* __restrict__ perm, const int4* __restrict__ segs,
                                                    const int32_t* __restrict__ splitDim) {
    __shared__ uint64_t key[kSmallSeg];
    __shared__ int32_t pos[kSmallSeg];
    __shared__ int32_t val[kSmallSeg];
    const int4 sg = segs[blockIdx.x];  // {start, end, segment, -}
    const int s = sg.z;
    int a = sg.x, e = sg.y;
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
__global__ void __launch_bounds__(256) k_make_keys(int64_t n, const double* __restrict__ p,
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

// the forest's points side by side: Pc[j * ntot + base_k + i] = p_k[j * n_k + i]
struct KdSrc {
    int64_t base;
    const double* p;
    uint32_t n;
};

__global__ void __launch_bounds__(256) k_forest_gather(int64_t ntot, uint32_t d, const KdSrc* __restrict__ src, int ns,
                                                       double* __restrict__ Pc) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= ntot) return;
    int lo = 0, hi = ns - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (src[mid].base <= e) lo = mid;
        else hi = mid - 1;
    }
    const KdSrc q = src[lo];
    const int64_t i = e - q.base;
    for (uint32_t j = 0; j < d; ++j) Pc[(size_t)j * ntot + e] = q.p[(size_t)j * q.n + i];
}

// identity permutation over the forest (initIndexMap, src/kDTreeHelpers.cu:12-26)
__global__ void __launch_bounds__(256) k_iota(int64_t n, int32_t* __restrict__ v) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e < n) v[e] = (int32_t)e;
}

// forest positions back to each cloud's own point numbers
__global__ void __launch_bounds__(256) k_forest_localize(int64_t ntot, const KdSrc* __restrict__ src, int ns,
                                                         int32_t* __restrict__ perm) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= ntot) return;
    int lo = 0, hi = ns - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (src[mid].base <= e) lo = mid;
        else hi = mid - 1;
    }
    perm[e] -= (int32_t)src[lo].base;
}

// One cloud of a forest: its cuts at the current level, in its own positions
struct KdJob {
    KDTree* t;
    const double* p;
    int64_t base = 0;      // first position in the forest
    int64_t nodeBase = 0;  // first box slot in the forest's box arrays
    uint32_t L = 0, size = 0;
    bool p2l = false;
    std::vector<int32_t> off;
};

// cuts of a cloud at a level: FULL_TREE / DIVIDE_IN_HALF split every segment
// at a + ceil((e-a)/2) (fillOffsetsSort, src/kDTreeHelpers.cu:81-93);
// POWER_OF_TWO_ON_LEFT cuts at multiples of a power-of-two size that starts
// at upperPowerOfTwo(n) and halves while it exceeds b (fillOffsets,
// src/kDTreeHelpers.cu:29-43)
static void p2l_cuts(uint32_t n, uint32_t lsize, std::vector<int32_t>& off) {
    const uint32_t ns = (n + lsize - 1) / lsize;
    off.assign(ns + 1, 0);
    for (uint32_t q = 0; q < ns; ++q) off[q] = (int32_t)(q * lsize);
    off[ns] = (int32_t)n;
}

static void kd_forest(Context& c, uint32_t d, std::vector<KdJob>& jobs) {
    const int nj = (int)jobs.size();
    int64_t ntot = 0, nodes = 0;
    uint32_t maxL = 0;
    for (KdJob& J : jobs) {
        KDTree& t = *J.t;
        HMX_CHECK(t.d == d, HMX_EINVAL, "a forest needs one dimension");
        HMX_CHECK(t.d >= 1 && t.d <= 3, HMX_EINVAL, "dimensionOfInputPoints must be 1, 2 or 3");
        HMX_CHECK(t.b >= 1, HMX_EINVAL, "bucketSize must be >= 1");
        HMX_CHECK(t.n >= 1, HMX_EINVAL, "numberOfInputPoints must be >= 1");
        HMX_CHECK(J.p != nullptr, HMX_EINVAL, "null points");
        J.p2l = t.divMethod == HMX_POWER_OF_TWO_ON_LEFT;
        if (J.p2l) {
            J.size = upper_pow2(t.n);
            J.L = 0;
            while (J.size > t.b) { J.size >>= 1; ++J.L; }
            t.numSegments = (t.n + J.size - 1) / J.size;
        } else {
            J.L = kd_num_levels(t.n, t.b);
            t.numSegments = 1u << J.L;
        }
        t.numLevels = J.L;
        J.base = ntot;
        J.nodeBase = nodes;
        J.off = {0, (int32_t)t.n};
        ntot += t.n;
        nodes += ((int64_t)2 << J.L) - 1;
        maxL = std::max(maxL, J.L);
    }
    HMX_CHECK(ntot < ((int64_t)1 << 31), HMX_EINVAL, "a forest holds fewer than 2^31 points");
    // the points: one cloud is read where it lies; several are gathered side by side
    std::vector<KdSrc> hsrc(nj);
    for (int k = 0; k < nj; ++k) hsrc[k] = KdSrc{jobs[k].base, jobs[k].p, jobs[k].t->n};
    DBuf<KdSrc> dsrc(nj);
    HMX_HIP(hipMemcpyAsync(dsrc.p, hsrc.data(), nj * sizeof(KdSrc), hipMemcpyHostToDevice, c.stream));
    DBuf<double> Pc;
    const double* P = jobs[0].p;
    if (nj > 1) {
        Pc.alloc((size_t)ntot * d);
        k_forest_gather<<<(unsigned)((ntot + 255) / 256), 256, 0, c.stream>>>(ntot, d, dsrc.p, nj, Pc.p);
        HMX_LAUNCHED();
        P = Pc.p;
    }
    auto permF = std::make_shared<DBuf<int32_t>>((size_t)ntot);
    auto bbminF = std::make_shared<DBuf<double>>((size_t)nodes * d);
    auto bbmaxF = std::make_shared<DBuf<double>>((size_t)nodes * d);
    k_iota<<<(unsigned)((ntot + 255) / 256), 256, 0, c.stream>>>(ntot, permF->p);
    HMX_LAUNCHED();
    int32_t* perm = permF->p;
    DBuf<int32_t> permB((size_t)ntot);
    DBuf<uint64_t> keyA((size_t)ntot), keyB((size_t)ntot);

    for (uint32_t level = 0; level < maxL; ++level) {
        // ---- this level's segments of every cloud still splitting
        std::vector<int2> segs;  // forest positions [a, e)
        std::vector<int64_t> node;
        for (KdJob& J : jobs) {
            if (level >= J.L) continue;
            if (J.p2l) p2l_cuts(J.t->n, J.size << (J.L - level), J.off);
            const int64_t nb = J.nodeBase + ((int64_t)1 << level) - 1;
            for (size_t s = 0; s + 1 < J.off.size(); ++s) {
                segs.push_back(make_int2((int)(J.base + J.off[s]), (int)(J.base + J.off[s + 1])));
                node.push_back(nb + (int64_t)s);
            }
        }
        const uint32_t nseg = (uint32_t)segs.size();
        // ---- bounding boxes
        std::vector<int2> bchunks;
        std::vector<int32_t> segChunk(nseg + 1);
        for (uint32_t s = 0; s < nseg; ++s) {
            segChunk[s] = (int32_t)bchunks.size();
            for (int32_t a = segs[s].x; a < segs[s].y; a += kBoxChunk)
                bchunks.push_back(make_int2(a, std::min(a + kBoxChunk, segs[s].y)));
        }
        segChunk[nseg] = (int32_t)bchunks.size();
        DBuf<int2> d_bch(std::max<size_t>(1, bchunks.size()));
        DBuf<int32_t> d_sch(nseg + 1), splitDim(std::max<uint32_t>(1, nseg));
        DBuf<int64_t> d_node(std::max<uint32_t>(1, nseg));
        DBuf<double> pmin(std::max<size_t>(1, bchunks.size() * d)), pmax(std::max<size_t>(1, bchunks.size() * d));
        HMX_HIP(hipMemcpyAsync(d_bch.p, bchunks.data(), bchunks.size() * sizeof(int2), hipMemcpyHostToDevice,
                               c.stream));
        HMX_HIP(hipMemcpyAsync(d_sch.p, segChunk.data(), (nseg + 1) * sizeof(int32_t), hipMemcpyHostToDevice,
                               c.stream));
        HMX_HIP(hipMemcpyAsync(d_node.p, node.data(), nseg * sizeof(int64_t), hipMemcpyHostToDevice, c.stream));
        if (!bchunks.empty()) {
            k_bbox_partial<<<(unsigned)bchunks.size(), 256, 0, c.stream>>>(ntot, d, P, perm, d_bch.p, pmin.p, pmax.p);
            HMX_LAUNCHED();
        }
        k_bbox_final<<<(nseg + 255) / 256, 256, 0, c.stream>>>(nseg, d, d_sch.p, pmin.p, pmax.p, d_node.p, bbminF->p,
                                                                bbmaxF->p, splitDim.p);
        HMX_LAUNCHED();

        // ---- sort each segment by its split coordinate
        std::vector<int4> small;
        std::vector<int4> rchunks;
        std::vector<int4> rsegs;
        for (uint32_t s = 0; s < nseg; ++s) {
            const int32_t a = segs[s].x, e = segs[s].y;
            if (e - a <= 1) continue;
            if (e - a <= kSmallSeg) {
                small.push_back(make_int4(a, e, (int)s, 0));
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
            k_make_keys<<<nch, 256, 0, c.stream>>>(ntot, P, perm, d_rch.p, splitDim.p, keyA.p);
            HMX_LAUNCHED();
            uint64_t* kin = keyA.p;
            uint64_t* kout = keyB.p;
            int32_t* vin = perm;
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
            // 8 passes: data is back in (keyA, perm)
            HMX_HIP(hipStreamSynchronize(c.stream));
        }
        if (!small.empty()) {
            DBuf<int4> d_small(small.size());
            HMX_HIP(hipMemcpyAsync(d_small.p, small.data(), small.size() * sizeof(int4), hipMemcpyHostToDevice,
                                   c.stream));
            k_sort_small<<<(unsigned)small.size(), 256, 0, c.stream>>>(ntot, P, perm, d_small.p, splitDim.p);
            HMX_LAUNCHED();
        }
        // ---- next level's cuts
        for (KdJob& J : jobs) {
            if (level >= J.L || J.p2l) continue;
            const size_t ns = J.off.size() - 1;
            std::vector<int32_t> noff(2 * ns + 1);
            for (size_t s = 0; s < ns; ++s) {
                noff[2 * s] = J.off[s];
                noff[2 * s + 1] = (J.off[s + 1] - J.off[s] + 1) / 2 + J.off[s];
            }
            noff[2 * ns] = (int32_t)J.t->n;
            J.off.swap(noff);
        }
        HMX_HIP(hipStreamSynchronize(c.stream));  // per-level host vectors go out of scope
    }
    if (nj > 1) {
        k_forest_localize<<<(unsigned)((ntot + 255) / 256), 256, 0, c.stream>>>(ntot, dsrc.p, nj, perm);
        HMX_LAUNCHED();
    }
    // leaf offsets of every tree in one block; boxes copied to the host once
    std::vector<int32_t> allOff;
    std::vector<int64_t> offBase(nj);
    for (int k = 0; k < nj; ++k) {
        KdJob& J = jobs[k];
        if (J.p2l) p2l_cuts(J.t->n, J.size, J.off);
        offBase[k] = (int64_t)allOff.size();
        allOff.insert(allOff.end(), J.off.begin(), J.off.end());
    }
    auto offF = std::make_shared<DBuf<int32_t>>(allOff.size());
    HMX_HIP(hipMemcpyAsync(offF->p, allOff.data(), allOff.size() * sizeof(int32_t), hipMemcpyHostToDevice, c.stream));
    std::vector<double> hmin((size_t)nodes * d), hmax((size_t)nodes * d);
    HMX_HIP(hipMemcpyAsync(hmin.data(), bbminF->p, hmin.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    HMX_HIP(hipMemcpyAsync(hmax.data(), bbmaxF->p, hmax.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    HMX_HIP(hipStreamSynchronize(c.stream));
    for (int k = 0; k < nj; ++k) {
        KdJob& J = jobs[k];
        KDTree& t = *J.t;
        const size_t nbox = ((size_t)2 << J.L) - 1;
        t.perm.view(permF, (size_t)J.base, t.n);
        t.offsets.view(offF, (size_t)offBase[k], J.off.size());
        t.bbmin.view(bbminF, (size_t)J.nodeBase * d, nbox * d);
        t.bbmax.view(bbmaxF, (size_t)J.nodeBase * d, nbox * d);
        t.h_offsets = J.off;
        t.h_bbmin.assign(hmin.begin() + J.nodeBase * d, hmin.begin() + (J.nodeBase + (int64_t)nbox) * d);
        t.h_bbmax.assign(hmax.begin() + J.nodeBase * d, hmax.begin() + (J.nodeBase + (int64_t)nbox) * d);
    }
}

void build_kdtree(Context& c, KDTree& t, const double* d_points) {
    HMX_CHECK(t.d >= 1 && t.d <= 3, HMX_EINVAL, "dimensionOfInputPoints must be 1, 2 or 3");
    std::vector<KdJob> jobs(1);
    jobs[0].t = &t;
    jobs[0].p = d_points;
    kd_forest(c, t.d, jobs);
}

// Trees of many clouds: one forest per point dimension
void build_kdtree_batch(Context& c, const std::vector<KDTree*>& ts, const std::vector<const double*>& pts) {
    HMX_CHECK(ts.size() == pts.size() && !ts.empty(), HMX_EINVAL, "batch arrays of different lengths");
    for (uint32_t d = 1; d <= 3; ++d) {
        std::vector<KdJob> jobs;
        for (size_t k = 0; k < ts.size(); ++k) {
            HMX_CHECK(ts[k]->d >= 1 && ts[k]->d <= 3, HMX_EINVAL, "dimensionOfInputPoints must be 1, 2 or 3");
            if (ts[k]->d != d) continue;
            jobs.emplace_back();
            jobs.back().t = ts[k];
            jobs.back().p = pts[k];
        }
        if (!jobs.empty()) kd_forest(c, d, jobs);
    }
}

}  // namespace hmx
