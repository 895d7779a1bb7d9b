// Hierarchical matrix assembly: allocateHMatrix (src/HMatrix.cu:57-103) and
// generateHMatrixFromStruct (src/constructHMatrixFromStruct.cu:13-103).
//
// The leaf level is gathered from the Morton-ordered TLR matrix.  Every upper
// level is recompressed by a batched adaptive randomized approximation whose
// operator is the implicit sum of the tile's bus x bus TLR sub-tiles
// (Y = sum_s U_s (V_s^T Omega), B = sum_s V_s (U_s^T Q)), i.e. the operator of
// lr_kblas_ara_batch.  The ARA is the same blocked algorithm as the TLR kernel
// (and the CPU oracle), split into one launch per step over all tiles of a
// level: sample, project (x2), column-by-column decisions, project the
// accepted columns, re-normalise, commit.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <unordered_map>

#include "hmx_device.h"
#include "hmx_internal.h"

namespace hmx {

double ara_threshold(double tol);

constexpr int kB16 = 16;

struct LrState {
    int k, small_run, converged, capped, na;
};

// Per-tile parameters of a batched ARA launch (tiles of one launch share m
// but may come from different levels or different H-matrices).
struct TileParam {
    double tau;       // acceptance threshold
    double nil;       // -1 / correlation length (direct mode)
    int64_t pbase;    // first padded position of the tile's H-matrix in P (direct mode)
    uint64_t opid;    // random stream id of Omega
    int32_t maxRank;
    int32_t code;     // Morton code of the tile within its level
};

struct LrLevel {
    int T, m, b, bus, r;
    uint64_t seed;
    const TileParam* tp;
    const int64_t* first;   // first Morton sub-tile of each tile
    const int32_t* ranksMO;
    const int64_t* offMO;
    const double* UMO;
    const double* VMO;
    LrState* st;
    double *Q, *Y, *Ya, *Om, *W;
    int capQ;
    // direct mode: the operator is the kernel itself, evaluated on the fly
    const double* P;        // padded KD-ordered points, (x, y, z, w) per position; w = 1 real, 0 padding
    const int64_t* boff;    // A^T Q of tile t goes to column boff[t] of the packed V factors
};

__device__ inline bool lr_active(const LrState& s, int maxRank) {
    return !(s.converged || s.capped || s.k >= maxRank);
}

__global__ void k_lr_init(int T, LrState* st) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    st[t].k = 0; st[t].small_run = 0; st[t].converged = 0; st[t].capped = 0; st[t].na = 0;
}

// Omega of the nb blocks it .. it+nb-1: Om[t][s*m + c] (transposed == 0,
// nb == 1) or Om[t][c*16nb + 16p + s] for sample s of block it+p
// (transposed == 1).  Entry (block, s, c) is the same number however many
// blocks one call generates.
__global__ void __launch_bounds__(256) k_lr_omega(LrLevel L, int it, int nb, int transposed) {
    int t = blockIdx.y;
    if (!lr_active(L.st[t], L.tp[t].maxRank)) return;
    const int wd = kB16 * nb;
    int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= L.m * wd) return;
    int j = transposed ? e % wd : e / L.m, c = transposed ? e / wd : e % L.m;
    L.Om[(size_t)t * L.m * wd + e] = gauss_at(L.seed, kTagOmega, L.tp[t].opid, (uint32_t)(it * kB16 + j), (uint32_t)c);
}

// Y rows of row-leaf i of tile t: sum over column leaves j of U_ij (V_ij^T Om_j)
template <int BMAX>
__global__ void __launch_bounds__(256) k_lr_sample(LrLevel L) {
    __shared__ double sV[BMAX * BMAX], sU[BMAX * BMAX], sO[BMAX * kB16], sT[BMAX * kB16];
    const int t = blockIdx.y, i = blockIdx.x, b = L.b, m = L.m;
    if (!lr_active(L.st[t], L.tp[t].maxRank)) return;
    const double* Om = L.Om + (size_t)t * m * kB16;
    double acc[(BMAX * kB16 + 255) / 256];
#pragma unroll
    for (int q = 0; q < (BMAX * kB16 + 255) / 256; ++q) acc[q] = 0.0;
    for (int j = 0; j < L.bus; ++j) {
        int64_t mo = L.first[t] + (int64_t)morton_xy((uint32_t)j, (uint32_t)i);
        int rk = L.ranksMO[mo];
        if (rk == 0) continue;
        const double* U = L.UMO + (size_t)L.offMO[mo] * b;
        const double* V = L.VMO + (size_t)L.offMO[mo] * b;
        for (int e = threadIdx.x; e < rk * b; e += 256) { sU[e] = U[e]; sV[e] = V[e]; }
        for (int e = threadIdx.x; e < b * kB16; e += 256) {
            int s = e / b, x = e % b;
            sO[e] = Om[(size_t)s * m + (size_t)j * b + x];
        }
        __syncthreads();
        for (int e = threadIdx.x; e < rk * kB16; e += 256) {  // sT[s*rk + a] = V_a . Om_s
            int s = e / rk, a = e % rk;
            double v = 0.0;
            for (int x = 0; x < b; ++x) v += sV[a * b + x] * sO[s * b + x];
            sT[e] = v;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < (BMAX * kB16 + 255) / 256; ++q) {
            int e = threadIdx.x + q * 256;
            if (e < b * kB16) {
                int s = e / b, x = e % b;
                double v = acc[q];
                for (int a = 0; a < rk; ++a) v += sU[a * b + x] * sT[s * rk + a];
                acc[q] = v;
            }
        }
        __syncthreads();
    }
    double* Y = L.Y + (size_t)t * m * kB16;
#pragma unroll
    for (int q = 0; q < (BMAX * kB16 + 255) / 256; ++q) {
        int e = threadIdx.x + q * 256;
        if (e < b * kB16) {
            int s = e / b, x = e % b;
            Y[(size_t)s * m + (size_t)i * b + x] = acc[q];
        }
    }
}

// ---- direct mode: samples from on-the-fly kernel evaluations ----------------
// Tile t of a level covers row cluster v and column cluster u (Morton code);
// its padded row q is KD position v*m + q of the padded point array.  Every
// wave owns 64 consecutive rows (sample) or columns (transposed sample) of
// one tile (m is a multiple of 64); the other side's points and the dense
// operand are staged through the wave's slice of LDS 64 at a time.

// padded points: P[(leaf*b + i)*4 + j]
struct PadCloud {
    int64_t leafBase;  // first leaf of the cloud in the padded leaf space
    const double* p;   // SoA coordinates p[j * n + i]
    const int32_t* perm;
    const int32_t* offsets;
    uint32_t n, d;
};

// The padded leaf space of one or more clouds: position leaf * b + i holds
// point i of the leaf as (x, y, z, 1), or zeros past the leaf's points; the
// clouds' leaves follow each other (cloud k from cl[k].leafBase on)
__global__ void __launch_bounds__(256) k_pad_points(int64_t total, int b, const PadCloud* __restrict__ cl, int ncl,
                                                    double* __restrict__ P) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= total) return;
    const int64_t gl = e / b;
    const int i = (int)(e % b);
    int lo = 0, hi = ncl - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cl[mid].leafBase <= gl) lo = mid;
        else hi = mid - 1;
    }
    const PadCloud q = cl[lo];
    const int leaf = (int)(gl - q.leafBase);
    const int a0 = q.offsets[leaf], cnt = q.offsets[leaf + 1] - a0;
    double x[4] = {0.0, 0.0, 0.0, 0.0};
    if (i < cnt) {
        for (uint32_t j = 0; j < q.d; ++j) x[j] = q.p[(size_t)j * q.n + q.perm[a0 + i]];
        x[3] = 1.0;
    }
    double4* o = reinterpret_cast<double4*>(P + e * 4);
    *o = make_double4(x[0], x[1], x[2], x[3]);
}

static void pad_points(Context& c, int b, const std::vector<PadCloud>& cl, int64_t leaves, double* P) {
    DBuf<PadCloud> d_cl(cl.size());
    HMX_HIP(hipMemcpyAsync(d_cl.p, cl.data(), cl.size() * sizeof(PadCloud), hipMemcpyHostToDevice, c.stream));
    const int64_t np = leaves * b;
    k_pad_points<<<(unsigned)((np + 255) / 256), 256, 0, c.stream>>>(np, b, d_cl.p, (int)cl.size(), P);
    HMX_LAUNCHED();
    HMX_HIP(hipStreamSynchronize(c.stream));  // the cloud table goes out of scope
}

// The samplers run on the FP64 matrix cores (v_mfma_f64_16x16x4f64).  A
// workgroup of nw = min(4, m/64) waves covers 64*nw consecutive rows (AΩ) or
// columns (AᵀQ) of one tile; its waves share the staged points of the other
// side and the staged rows of Ω or Q.  A wave owns 64 rows (columns) as four
// blocks of 16; per step of 4 along the inner index a lane evaluates one
// kernel entry per block -- the A operand, lane l holding A[l%16][l/16] -- and
// uses it for every 16-column block of the dense operand, read from LDS as
// the B operand (B[l/16][l%16]).  The 16x16 result of a block sits in
// register v of lane l at row 4v + l/16, column l%16.  Kernel evaluations
// (exp, sqrt on the FP64 vector units) dominate, so the more dense columns a
// sweep carries, the fewer evaluations a tile costs.
typedef double v4d_t __attribute__((ext_vector_type(4)));

// NB consecutive ARA blocks in one sweep over the tile's columns: block p of
// the pass is Y_p[t][s*m + q] = sum_c K(x_q, x_c) Om[t][c*16NB + 16p + s],
// written to Yout + p*bstride + slot*16m with slot = t (one split) or
// split*T + t (column splits, summed by k_dir_sum).  Per block the
// accumulation order over c is that of a one-block sweep, so every block's
// samples are the same bits whatever NB is.
template <int D, int NB>
__global__ void __launch_bounds__(256) k_dir_sample_mf(LrLevel L, int splits, double* __restrict__ Yout,
                                                       size_t bstride) {
    constexpr int NC = kB16 * NB;
    __shared__ double4 sP[64];
    __shared__ double sO[64 * NC];
    const int nw = blockDim.x >> 6;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, lg = lane >> 4, lc = lane & 15;
    const int m = L.m;
    const int wgPerTile = m / (64 * nw);
    const int t = blockIdx.x / wgPerTile;
    if (!lr_active(L.st[t], L.tp[t].maxRank)) return;  // uniform over the workgroup
    const int q0 = (blockIdx.x % wgPerTile) * 64 * nw + 64 * w;
    const TileParam tpt = L.tp[t];
    const uint32_t code = (uint32_t)tpt.code;
    const double nil = tpt.nil;
    const int64_t rowBase = tpt.pbase + (int64_t)compact16(code >> 1) * m;
    const int64_t colBase = tpt.pbase + (int64_t)compact16(code) * m;
    const double4* P4 = reinterpret_cast<const double4*>(L.P);
    double xr[4][3], wr[4];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
        const double4 p = P4[rowBase + q0 + 16 * rb + lc];
        xr[rb][0] = p.x; xr[rb][1] = p.y; xr[rb][2] = p.z; wr[rb] = p.w;
    }
    const int chunk = m / splits, c0 = blockIdx.y * chunk;
    const double* Om = L.Om + (size_t)t * m * NC;
    v4d_t acc[4][NB];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
#pragma unroll
        for (int p = 0; p < NB; ++p) acc[rb][p] = v4d_t{0.0, 0.0, 0.0, 0.0};
    for (int cb = c0; cb < c0 + chunk; cb += 64) {
        __syncthreads();
        if (threadIdx.x < 64) sP[threadIdx.x] = P4[colBase + cb + threadIdx.x];
        const double2* src = reinterpret_cast<const double2*>(Om + (size_t)cb * NC);
        for (int e = threadIdx.x; e < 32 * NC; e += blockDim.x) {
            const double2 v = src[e];
            sO[2 * e] = v.x;
            sO[2 * e + 1] = v.y;
        }
        __syncthreads();
        for (int c4 = 0; c4 < 64; c4 += 4) {
            const double4 xc = sP[c4 + lg];
            const double xca[3] = {xc.x, xc.y, xc.z};
            double om[NB];
#pragma unroll
            for (int p = 0; p < NB; ++p) om[p] = sO[(c4 + lg) * NC + 16 * p + lc];
#pragma unroll
            for (int rb = 0; rb < 4; ++rb) {
                const double a = kernel_eval_scaled<D>(xr[rb], xca, nil) * (xc.w * wr[rb]);
#pragma unroll
                for (int p = 0; p < NB; ++p) acc[rb][p] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, om[p], acc[rb][p], 0, 0, 0);
            }
        }
    }
    const size_t slot = splits > 1 ? (size_t)blockIdx.y * L.T + t : (size_t)t;
#pragma unroll
    for (int p = 0; p < NB; ++p) {
        double* Y = Yout + (size_t)p * bstride + slot * m * kB16;
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
#pragma unroll
            for (int v = 0; v < 4; ++v) Y[(size_t)lc * m + q0 + 16 * rb + 4 * v + lg] = acc[rb][p][v];
    }
}

// B[t][a*m + c] = sum_q K(x_q, x_c) Q[t][a*m + q] for the 16NB basis columns
// a0 = 16NB*blockIdx.y ... of every tile with k > a0; the blocks past the
// tile's k are skipped.  A wave covers 16CB tile columns: CB = 4 with up to
// 64 basis columns per pass, CB = 2 with 128 (the same accumulator registers;
// a tile of rank 65..128 is evaluated once instead of twice).  Q comes
// through LDS in blocks of QB rows.  Only tiles of rank kLo < k <= kHi are
// processed.  A column's sum runs over q in the same order for any NB, CB and
// QB, so B does not depend on them.
template <int D, int NB, int CB = 4>
__global__ void __launch_bounds__(256) k_dir_sampleT_mf(LrLevel L, double* __restrict__ Bout, int kLo, int kHi) {
    constexpr int NQ = kB16 * NB, QB = NB > 4 ? 32 : 64, WR = 16 * CB;
    __shared__ double4 sP[QB];
    __shared__ double sQ[QB][NQ + 1];
    const int nw = blockDim.x >> 6;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, lg = lane >> 4, lc = lane & 15;
    const int m = L.m;
    const int wgPerTile = m / (WR * nw);
    const int t = blockIdx.x / wgPerTile;
    const int k = L.st[t].k;
    const int a0 = blockIdx.y * NQ;
    if (a0 >= k || k <= kLo || k > kHi) return;  // uniform over the workgroup
    const int na = min(NQ, k - a0);
    const int cw = (blockIdx.x % wgPerTile) * WR * nw + WR * w;
    const TileParam tpt = L.tp[t];
    const uint32_t code = (uint32_t)tpt.code;
    const double nil = tpt.nil;
    const int64_t rowBase = tpt.pbase + (int64_t)compact16(code >> 1) * m;
    const int64_t colBase = tpt.pbase + (int64_t)compact16(code) * m;
    const double4* P4 = reinterpret_cast<const double4*>(L.P);
    double xc[CB][3], wc[CB];
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) {
        const double4 p = P4[colBase + cw + 16 * cb + lc];
        xc[cb][0] = p.x; xc[cb][1] = p.y; xc[cb][2] = p.z; wc[cb] = p.w;
    }
    const double* Q = L.Q + (size_t)t * m * L.capQ + (size_t)a0 * m;
    v4d_t acc[CB][NB];
#pragma unroll
    for (int cb = 0; cb < CB; ++cb)
#pragma unroll
        for (int p = 0; p < NB; ++p) acc[cb][p] = v4d_t{0.0, 0.0, 0.0, 0.0};
    for (int qb = 0; qb < m; qb += QB) {
        __syncthreads();
        if (threadIdx.x < QB) sP[threadIdx.x] = P4[rowBase + qb + threadIdx.x];
        for (int e = threadIdx.x; e < QB * NQ; e += blockDim.x) {
            const int a = e / QB, r = e % QB;
            sQ[r][a] = a < na ? Q[(size_t)a * m + qb + r] : 0.0;
        }
        __syncthreads();
        for (int q4 = 0; q4 < QB; q4 += 4) {
            const double4 xr = sP[q4 + lg];
            const double xra[3] = {xr.x, xr.y, xr.z};
            double bq[NB];
#pragma unroll
            for (int p = 0; p < NB; ++p) bq[p] = sQ[q4 + lg][16 * p + lc];
#pragma unroll
            for (int cb = 0; cb < CB; ++cb) {
                const double a = kernel_eval_scaled<D>(xra, xc[cb], nil) * (xr.w * wc[cb]);
#pragma unroll
                for (int p = 0; p < NB; ++p)
                    if (16 * p < na) acc[cb][p] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bq[p], acc[cb][p], 0, 0, 0);
            }
        }
    }
    double* B = Bout + (size_t)L.boff[t] * m;
#pragma unroll
    for (int cb = 0; cb < CB; ++cb)
#pragma unroll
        for (int p = 0; p < NB; ++p)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int a = 16 * p + lc;
                if (a < na) B[(size_t)(a0 + a) * m + cw + 16 * cb + 4 * v + lg] = acc[cb][p][v];
            }
}

// Y_p = sum over the splits, in split order, for the NB blocks of a pass
__global__ void __launch_bounds__(256) k_dir_sum(LrLevel L, int splits, const double* __restrict__ Ypart,
                                                 double* __restrict__ Yout, size_t bstride, size_t pstride) {
    const int t = blockIdx.y, p = blockIdx.z;
    if (!lr_active(L.st[t], L.tp[t].maxRank)) return;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t per = (int64_t)L.m * kB16;
    if (e >= per) return;
    const double* src = Ypart + (size_t)p * pstride;
    double v = 0.0;
    for (int sp = 0; sp < splits; ++sp) v += src[((size_t)sp * L.T + t) * per + e];
    Yout[(size_t)p * bstride + (size_t)t * per + e] = v;
}

// W[t][s*k + a] = sum_i Q_t[a*m+i] B_t[s*m+i], a < k_t, on the FP64 matrix
// cores: one workgroup per 16 basis columns of a tile, its four waves take
// a quarter of the rows each (lane l holds Q^T[a0 + l%16][i + l/16] and
// B[i + l/16][l%16] for the step at row i), and the four partial 16 x 16
// blocks are added in wave order.
// mode 0: B = Y (16 columns, tile still iterating); mode 2: B = accepted block (na columns)
__global__ void __launch_bounds__(256) k_lr_dot(LrLevel L, int mode) {
    __shared__ double red[3][64][4];
    const int t = blockIdx.y;
    const LrState& S = L.st[t];
    // before the decision step the tile must still be iterating; after it,
    // accepted columns are appended even when the tile converged in this block
    if (mode == 0 ? !lr_active(S, L.tp[t].maxRank) : S.na == 0) return;
    const int K = S.k;
    const int C = mode == 0 ? kB16 : S.na;
    const int a0 = blockIdx.x * 16;
    if (a0 >= K) return;
    const int m = L.m;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, lc = lane & 15, lg = lane >> 4;
    const bool av = a0 + lc < K, sv = lc < C;
    const double* qa = L.Q + (size_t)t * m * L.capQ + (size_t)(av ? a0 + lc : 0) * m;
    const double* bs = (mode == 0 ? L.Y : L.Ya) + (size_t)t * m * kB16 + (size_t)(sv ? lc : 0) * m;
    const int part = m / 4, i0 = w * part;  // m is a multiple of 64
    v4d_t acc = {0.0, 0.0, 0.0, 0.0};
    for (int i = i0; i < i0 + part; i += 4) {
        const double x = av ? qa[i + lg] : 0.0;
        const double y = sv ? bs[i + lg] : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, acc, 0, 0, 0);
    }
    if (w > 0)
#pragma unroll
        for (int v = 0; v < 4; ++v) red[w - 1][lane][v] = acc[v];
    __syncthreads();
    if (w > 0) return;
    double* out = L.W + (size_t)t * L.capQ * kB16;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const double r = ((acc[v] + red[0][lane][v]) + red[1][lane][v]) + red[2][lane][v];
        const int a = a0 + 4 * v + lg;  // output row 4v + l/16, column l%16
        if (a < K && lc < C) out[lc * K + a] = r;
    }
}

// Y[s*m+i] -= sum_a Q[a*m+i] W[s*k+a]  (mode 0: Y, 16 cols; mode 2: Ya, na cols)
// on the FP64 matrix cores: one wave per 16 rows (lane l holds
// Q[row0 + l%16][a + l/16] and -W[a + l/16][l%16] for the step at a).
__global__ void __launch_bounds__(256) k_lr_update(LrLevel L, int mode) {
    const int t = blockIdx.y;
    const LrState& S = L.st[t];
    if (mode == 0 ? !lr_active(S, L.tp[t].maxRank) : S.na == 0) return;
    const int k = S.k;
    const int C = mode == 0 ? kB16 : S.na;
    if (k == 0 || C == 0) return;
    const int m = L.m;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, lc = lane & 15, lg = lane >> 4;
    const int row0 = blockIdx.x * 64 + 16 * w;
    if (row0 >= m) return;
    double* Y = (mode == 0 ? L.Y : L.Ya) + (size_t)t * m * kB16;
    const double* Q = L.Q + (size_t)t * m * L.capQ;
    const double* W = L.W + (size_t)t * L.capQ * kB16;
    const bool sv = lc < C;
    v4d_t acc;
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[v] = sv ? Y[(size_t)lc * m + row0 + 4 * v + lg] : 0.0;
    for (int a = 0; a < k; a += 4) {
        const bool ok = a + lg < k;
        const double x = ok ? Q[(size_t)(a + lg) * m + row0 + lc] : 0.0;
        const double y = ok && sv ? -W[(size_t)lc * k + a + lg] : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, acc, 0, 0, 0);
    }
    if (sv)
#pragma unroll
        for (int v = 0; v < 4; ++v) Y[(size_t)lc * m + row0 + 4 * v + lg] = acc[v];
}

// sum over the block; every thread receives the identical value.  red holds
// two slots of 16: consecutive calls alternate (parity), so one barrier per
// call suffices -- a thread reaches the call after next only once every
// thread has passed the next call's barrier, i.e. has read this one's slot.
__device__ inline double block_sum(double v, double* red, int& parity) {
    v = wave_sum_dpp(v);
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    double* slot = red + 16 * parity;
    parity ^= 1;
    if ((threadIdx.x & 63) == 0) slot[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int q = 0; q < nw; ++q) s += slot[q];
    return s;
}

// One wave per tile for tiles of m <= 512 rows (R = m / 64 rows per lane):
// the block's accepted columns and the current column stay in registers and
// every reduction is a DPP wave sum -- no barriers, no re-reads.  The same
// decisions as k_lr_decide.
template <int R>
__global__ void __launch_bounds__(64) k_lr_decide_w(LrLevel L) {
    const int t = blockIdx.x, lane = threadIdx.x;
    LrState& S = L.st[t];
    if (!lr_active(S, L.tp[t].maxRank)) {
        if (lane == 0) S.na = 0;
        return;
    }
    const int m = L.m;
    const double* Y = L.Y + (size_t)t * m * kB16;
    double* Qb = L.Ya + (size_t)t * m * kB16;
    double qb[kB16][R];
    int na = 0, small = S.small_run, conv = 0, capped = 0;
    const int capleft = L.tp[t].maxRank - S.k;
    const double tau = L.tp[t].tau;
    for (int j = 0; j < kB16; ++j) {
        double y[R];
#pragma unroll
        for (int r = 0; r < R; ++r) y[r] = Y[(size_t)j * m + lane + 64 * r];
#pragma unroll
        for (int pass = 0; pass < 2; ++pass)
#pragma unroll
            for (int a = 0; a < kB16; ++a)
                if (a < na) {
                    double p = 0.0;
#pragma unroll
                    for (int r = 0; r < R; ++r) p += qb[a][r] * y[r];
                    const double d = wave_sum_dpp(p);
#pragma unroll
                    for (int r = 0; r < R; ++r) y[r] -= qb[a][r] * d;
                }
        double p = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r) p += y[r] * y[r];
        const double nu = sqrt(wave_sum_dpp(p));
        if (nu > tau) {
            if (na >= capleft) { capped = 1; break; }
#pragma unroll
            for (int a = 0; a < kB16; ++a)
                if (a == na)
#pragma unroll
                    for (int r = 0; r < R; ++r) qb[a][r] = y[r] / nu;
#pragma unroll
            for (int r = 0; r < R; ++r) Qb[(size_t)na * m + lane + 64 * r] = y[r] / nu;
            ++na;
            small = 0;
        } else if (++small >= L.r) {
            conv = 1;
            break;
        }
    }
    if (lane == 0) { S.na = na; S.small_run = small; S.converged = conv; S.capped = capped; }
}

// The same decisions for tiles of 1024 and 2048 rows: four waves, thread
// t owning rows t + 256 i (i < R = m / 256), the current column and the
// block's accepted columns in registers, one barrier per sum (block_sum)
template <int R>
__global__ void __launch_bounds__(256) k_lr_decide_b(LrLevel L) {
    __shared__ double red[32];
    int parity = 0;
    const int t = blockIdx.x, tid = threadIdx.x;
    LrState& S = L.st[t];
    if (!lr_active(S, L.tp[t].maxRank)) {
        if (tid == 0) S.na = 0;
        return;
    }
    const int m = L.m;
    const double* Y = L.Y + (size_t)t * m * kB16;
    double* Qb = L.Ya + (size_t)t * m * kB16;
    double qb[kB16][R];
    int na = 0, small = S.small_run, conv = 0, capped = 0;
    const int capleft = L.tp[t].maxRank - S.k;
    const double tau = L.tp[t].tau;
    for (int j = 0; j < kB16; ++j) {
        double y[R];
#pragma unroll
        for (int r = 0; r < R; ++r) y[r] = Y[(size_t)j * m + tid + 256 * r];
#pragma unroll
        for (int pass = 0; pass < 2; ++pass)
#pragma unroll
            for (int a = 0; a < kB16; ++a)
                if (a < na) {
                    double p = 0.0;
#pragma unroll
                    for (int r = 0; r < R; ++r) p += qb[a][r] * y[r];
                    const double d = block_sum(p, red, parity);
#pragma unroll
                    for (int r = 0; r < R; ++r) y[r] -= qb[a][r] * d;
                }
        double p = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r) p += y[r] * y[r];
        const double nu = sqrt(block_sum(p, red, parity));
        if (nu > tau) {
            if (na >= capleft) { capped = 1; break; }
#pragma unroll
            for (int a = 0; a < kB16; ++a)
                if (a == na)
#pragma unroll
                    for (int r = 0; r < R; ++r) qb[a][r] = y[r] / nu;
#pragma unroll
            for (int r = 0; r < R; ++r) Qb[(size_t)na * m + tid + 256 * r] = y[r] / nu;
            ++na;
            small = 0;
        } else if (++small >= L.r) {
            conv = 1;
            break;
        }
    }
    if (tid == 0) { S.na = na; S.small_run = small; S.converged = conv; S.capped = capped; }
}

// k_lr_renorm for tiles of 1024 and 2048 rows: four waves, registers
template <int R>
__global__ void __launch_bounds__(256) k_lr_renorm_b(LrLevel L) {
    __shared__ double red[32];
    int parity = 0;
    const int t = blockIdx.x, tid = threadIdx.x;
    const LrState& S = L.st[t];
    const int na = S.na;
    if (na == 0) return;
    const int m = L.m;
    double* Qb = L.Ya + (size_t)t * m * kB16;
    double* Q = L.Q + (size_t)t * m * L.capQ;
    double qb[kB16][R];
#pragma unroll
    for (int j = 0; j < kB16; ++j)
        if (j < na) {
            double q[R];
#pragma unroll
            for (int r = 0; r < R; ++r) q[r] = Qb[(size_t)j * m + tid + 256 * r];
#pragma unroll
            for (int a = 0; a < j; ++a) {
                double p = 0.0;
#pragma unroll
                for (int r = 0; r < R; ++r) p += qb[a][r] * q[r];
                const double d = block_sum(p, red, parity);
#pragma unroll
                for (int r = 0; r < R; ++r) q[r] -= qb[a][r] * d;
            }
            double p = 0.0;
#pragma unroll
            for (int r = 0; r < R; ++r) p += q[r] * q[r];
            const double nu = sqrt(block_sum(p, red, parity));
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const double v = q[r] / nu;
                qb[j][r] = v;
                Qb[(size_t)j * m + tid + 256 * r] = v;
                Q[(size_t)(S.k + j) * m + tid + 256 * r] = v;
            }
        }
}

// k_lr_renorm for tiles of m <= 512 rows, one wave per tile, registers only
template <int R>
__global__ void __launch_bounds__(64) k_lr_renorm_w(LrLevel L) {
    const int t = blockIdx.x, lane = threadIdx.x;
    const LrState& S = L.st[t];
    const int na = S.na;
    if (na == 0) return;
    const int m = L.m;
    double* Qb = L.Ya + (size_t)t * m * kB16;
    double* Q = L.Q + (size_t)t * m * L.capQ;
    double qb[kB16][R];
#pragma unroll
    for (int j = 0; j < kB16; ++j)
        if (j < na) {
            double q[R];
#pragma unroll
            for (int r = 0; r < R; ++r) q[r] = Qb[(size_t)j * m + lane + 64 * r];
#pragma unroll
            for (int a = 0; a < j; ++a) {
                double p = 0.0;
#pragma unroll
                for (int r = 0; r < R; ++r) p += qb[a][r] * q[r];
                const double d = wave_sum_dpp(p);
#pragma unroll
                for (int r = 0; r < R; ++r) q[r] -= qb[a][r] * d;
            }
            double p = 0.0;
#pragma unroll
            for (int r = 0; r < R; ++r) p += q[r] * q[r];
            const double nu = sqrt(wave_sum_dpp(p));
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const double v = q[r] / nu;
                qb[j][r] = v;
                Qb[(size_t)j * m + lane + 64 * r] = v;
                Q[(size_t)(S.k + j) * m + lane + 64 * r] = v;
            }
        }
}

// One workgroup per tile: column-by-column decisions on explicitly projected
// vectors (two modified Gram-Schmidt sweeps against the block's accepted
// columns, accept when the norm exceeds tau).  Threads own fixed rows, so
// only the reductions need the workgroup barrier.
__global__ void __launch_bounds__(1024) k_lr_decide(LrLevel L) {
    __shared__ double red[32];
    int parity = 0;
    const int t = blockIdx.x;
    LrState& S = L.st[t];
    if (!lr_active(S, L.tp[t].maxRank)) {
        if (threadIdx.x == 0) S.na = 0;
        return;
    }
    const int m = L.m, nthr = blockDim.x, tid = threadIdx.x;
    double* Y = L.Y + (size_t)t * m * kB16;
    double* Qb = L.Ya + (size_t)t * m * kB16;
    int na = 0, small = S.small_run, conv = 0, capped = 0;
    const int capleft = L.tp[t].maxRank - S.k;
    const double tau = L.tp[t].tau;
    for (int j = 0; j < kB16; ++j) {
        double* y = Y + (size_t)j * m;
        for (int pass = 0; pass < 2; ++pass)
            for (int a = 0; a < na; ++a) {
                const double* qa = Qb + (size_t)a * m;
                double p = 0.0;
                for (int i = tid; i < m; i += nthr) p += qa[i] * y[i];
                const double d = block_sum(p, red, parity);
                for (int i = tid; i < m; i += nthr) y[i] -= qa[i] * d;
            }
        double p = 0.0;
        for (int i = tid; i < m; i += nthr) p += y[i] * y[i];
        const double nu = sqrt(block_sum(p, red, parity));
        if (nu > tau) {
            if (na >= capleft) { capped = 1; break; }
            for (int i = tid; i < m; i += nthr) Qb[(size_t)na * m + i] = y[i] / nu;
            ++na;
            small = 0;
        } else if (++small >= L.r) {
            conv = 1;
            break;
        }
    }
    if (tid == 0) { S.na = na; S.small_run = small; S.converged = conv; S.capped = capped; }
}

// After the accepted columns were projected against Q: one more modified
// Gram-Schmidt sweep inside the block, normalise, append to Q.
__global__ void __launch_bounds__(1024) k_lr_renorm(LrLevel L) {
    __shared__ double red[32];
    int parity = 0;
    const int t = blockIdx.x;
    const LrState& S = L.st[t];
    const int na = S.na;
    if (na == 0) return;
    const int m = L.m, nthr = blockDim.x, tid = threadIdx.x;
    double* Qb = L.Ya + (size_t)t * m * kB16;
    double* Q = L.Q + (size_t)t * m * L.capQ;
    for (int j = 0; j < na; ++j) {
        double* q = Qb + (size_t)j * m;
        for (int a = 0; a < j; ++a) {
            const double* qa = Qb + (size_t)a * m;
            double p = 0.0;
            for (int i = tid; i < m; i += nthr) p += qa[i] * q[i];
            const double d = block_sum(p, red, parity);
            for (int i = tid; i < m; i += nthr) q[i] -= qa[i] * d;
        }
        double p = 0.0;
        for (int i = tid; i < m; i += nthr) p += q[i] * q[i];
        const double nu = sqrt(block_sum(p, red, parity));
        for (int i = tid; i < m; i += nthr) {
            const double v = q[i] / nu;
            q[i] = v;
            Q[(size_t)(S.k + j) * m + i] = v;
        }
    }
}

// k += na; report active count and max k
__global__ void k_lr_commit(LrLevel L, int* flags) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= L.T) return;
    LrState& S = L.st[t];
    S.k += S.na;
    S.na = 0;
    if (!S.converged && S.k >= L.tp[t].maxRank) S.capped = 1;
    if (!(S.converged || S.capped)) atomicAdd(&flags[0], 1);
    atomicMax(&flags[1], S.k);
}

// B[t][a*m + j*b + x] = sum_i V_ij (U_ij^T Q_i), 64 Q columns per block
template <int BMAX>
__global__ void __launch_bounds__(256) k_lr_sampleT(LrLevel L, double* Bout) {
    __shared__ double sV[BMAX * BMAX], sU[BMAX * BMAX], sQ[BMAX * 64], sT[BMAX * 64];
    const int t = blockIdx.z, j = blockIdx.y, c0 = blockIdx.x * 64, b = L.b, m = L.m;
    const int k = L.st[t].k;
    if (c0 >= k) return;
    const int nc = min(64, k - c0);
    const double* Q = L.Q + (size_t)t * m * L.capQ;
    constexpr int NACC = (BMAX * 64 + 255) / 256;
    double acc[NACC];
#pragma unroll
    for (int q = 0; q < NACC; ++q) acc[q] = 0.0;
    for (int i = 0; i < L.bus; ++i) {
        int64_t mo = L.first[t] + (int64_t)morton_xy((uint32_t)j, (uint32_t)i);
        int rk = L.ranksMO[mo];
        if (rk == 0) continue;
        const double* U = L.UMO + (size_t)L.offMO[mo] * b;
        const double* V = L.VMO + (size_t)L.offMO[mo] * b;
        for (int e = threadIdx.x; e < rk * b; e += 256) { sU[e] = U[e]; sV[e] = V[e]; }
        for (int e = threadIdx.x; e < b * nc; e += 256) {
            int cc = e / b, x = e % b;
            sQ[e] = Q[(size_t)(c0 + cc) * m + (size_t)i * b + x];
        }
        __syncthreads();
        for (int e = threadIdx.x; e < rk * nc; e += 256) {  // sT[cc*rk + a] = U_a . Q_cc
            int cc = e / rk, a = e % rk;
            double v = 0.0;
            for (int x = 0; x < b; ++x) v += sU[a * b + x] * sQ[cc * b + x];
            sT[e] = v;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NACC; ++q) {
            int e = threadIdx.x + q * 256;
            if (e < b * nc) {
                int cc = e / b, x = e % b;
                double v = acc[q];
                for (int a = 0; a < rk; ++a) v += sV[a * b + x] * sT[cc * rk + a];
                acc[q] = v;
            }
        }
        __syncthreads();
    }
    double* B = Bout + (size_t)L.boff[t] * m;
#pragma unroll
    for (int q = 0; q < NACC; ++q) {
        int e = threadIdx.x + q * 256;
        if (e < b * nc) {
            int cc = e / b, x = e % b;
            B[(size_t)(c0 + cc) * m + (size_t)j * b + x] = acc[q];
        }
    }
}

// grow Q capacity: copy each tile's first k columns
__global__ void __launch_bounds__(256) k_lr_regrow(int T, int m, int oldCap, int newCap, const LrState* st,
                                                   const double* __restrict__ Qold, double* __restrict__ Qnew) {
    const int t = blockIdx.y;
    const size_t cnt = (size_t)st[t].k * m;
    const double* s = Qold + (size_t)t * m * oldCap;
    double* d = Qnew + (size_t)t * m * newCap;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < cnt; e += (size_t)gridDim.x * 256) d[e] = s[e];
}

// pack tile factors: dst[off[t]*m ...] = src[t*m*cap ...] (k_t*m elements)
__global__ void __launch_bounds__(256) k_lr_pack(int T, int m, int cap, const LrState* st,
                                                 const int64_t* __restrict__ off, const double* __restrict__ src,
                                                 double* __restrict__ dst) {
    const int t = blockIdx.y;
    const size_t cnt = (size_t)st[t].k * m;
    const double* s = src + (size_t)t * m * cap;
    double* d = dst + (size_t)off[t] * m;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < cnt; e += (size_t)gridDim.x * 256) d[e] = s[e];
}

// mirrored tiles: tile j takes tile i's factors exchanged, U_j = V_i and
// V_j = U_i (k columns of m rows each); one grid row per mirrored tile
__global__ void __launch_bounds__(256) k_lr_mirror(int m, const int64_t* __restrict__ dstOff,
                                                   const int64_t* __restrict__ srcOff, const int32_t* __restrict__ k,
                                                   double* __restrict__ U, double* __restrict__ V) {
    const int j = blockIdx.y;
    const size_t cnt = (size_t)k[j] * m, d = (size_t)dstOff[j] * m, s = (size_t)srcOff[j] * m;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < cnt; e += (size_t)gridDim.x * 256) {
        U[d + e] = V[s + e];
        V[d + e] = U[s + e];
    }
}

// leaf level: gather TLR tiles by Morton tile index
__global__ void __launch_bounds__(256) k_leaf_gather(int T, int b, const int32_t* __restrict__ tiles,
                                                     const int32_t* __restrict__ ranksMO,
                                                     const int64_t* __restrict__ offMO, const int64_t* __restrict__ off,
                                                     const double* __restrict__ UMO, const double* __restrict__ VMO,
                                                     double* __restrict__ U, double* __restrict__ V) {
    int t = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= T) return;
    int lane = threadIdx.x & 63;
    int mo = tiles[t];
    size_t cnt = (size_t)ranksMO[mo] * b;
    const double* us = UMO + (size_t)offMO[mo] * b;
    const double* vs = VMO + (size_t)offMO[mo] * b;
    for (size_t e = lane; e < cnt; e += 64) {
        U[(size_t)off[t] * b + e] = us[e];
        V[(size_t)off[t] * b + e] = vs[e];
    }
}

__global__ void k_leaf_ranks(int T, const int32_t* tiles, const int32_t* ranksMO, int32_t* out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < T) out[t] = ranksMO[tiles[t]];
}

// Source of a level's operator: the Morton TLR sub-tiles, or (direct mode)
// the kernel evaluated on the fly from the padded points.
struct LevelSource {
    const TLR* mo = nullptr;
    const double* P = nullptr;
    double ell = 0.0;
    int d = 0;
};

template <int BMAX>
static void launch_sample_tlr(Context& c, const LrLevel& L, int bus, int T) {
    k_lr_sample<BMAX><<<dim3(bus, T), 256, 0, c.stream>>>(L);
    HMX_LAUNCHED();
}

template <int BMAX>
static void launch_sampleT_tlr(Context& c, const LrLevel& L, int kmax, int bus, int T, double* B) {
    dim3 g((unsigned)((kmax + 63) / 64), (unsigned)bus, (unsigned)T);
    k_lr_sampleT<BMAX><<<g, 256, 0, c.stream>>>(L, B);
    HMX_LAUNCHED();
}

// direct samplers: workgroups of nw waves, all in one tile (m is 64 * 2^j)
static int sampler_threads(int m) { return 64 * (m >= 256 ? 4 : m / 64); }

template <int D, int NB>
static void launch_sample_direct_nb(Context& c, const LrLevel& L, int splits, double* Ypre, double* Ypart) {
    const int thr = sampler_threads(L.m);
    const unsigned wgs = (unsigned)((int64_t)L.T * L.m / thr);
    const size_t per = (size_t)L.T * L.m * kB16;
    if (splits > 1) {
        k_dir_sample_mf<D, NB><<<dim3(wgs, (unsigned)splits), thr, 0, c.stream>>>(L, splits, Ypart, per * splits);
        HMX_LAUNCHED();
        k_dir_sum<<<dim3((unsigned)(((int64_t)L.m * kB16 + 255) / 256), (unsigned)L.T, (unsigned)NB), 256, 0,
                    c.stream>>>(L, splits, Ypart, Ypre, per, per * splits);
        HMX_LAUNCHED();
    } else {
        k_dir_sample_mf<D, NB><<<dim3(wgs, 1u), thr, 0, c.stream>>>(L, 1, Ypre, per);
        HMX_LAUNCHED();
    }
}

// nb consecutive ARA blocks from one sweep over every active tile
template <int D>
static void launch_sample_direct(Context& c, const LrLevel& L, int nb, int splits, double* Ypre, double* Ypart) {
    switch (nb) {
        case 1: launch_sample_direct_nb<D, 1>(c, L, splits, Ypre, Ypart); break;
        case 2: launch_sample_direct_nb<D, 2>(c, L, splits, Ypre, Ypart); break;
        case 3: launch_sample_direct_nb<D, 3>(c, L, splits, Ypre, Ypart); break;
        default: launch_sample_direct_nb<D, 4>(c, L, splits, Ypre, Ypart); break;
    }
}

template <int D, int NB, int CB = 4>
static void launch_sampleT_direct_nb(Context& c, const LrLevel& L, int kmax, double* B, int kLo, int kHi) {
    const int thr = sampler_threads(L.m);
    // a workgroup covers (thr/64) * 16CB tile columns
    const unsigned wgs = (unsigned)((int64_t)L.T * L.m / (thr / 64 * 16 * CB));
    k_dir_sampleT_mf<D, NB, CB>
        <<<dim3(wgs, (unsigned)((kmax + kB16 * NB - 1) / (kB16 * NB))), thr, 0, c.stream>>>(L, B, kLo, kHi);
    HMX_LAUNCHED();
}

// One pass carries up to 64 basis columns.  When some tile has more, the
// tiles of rank <= 64 take that pass and those of rank 65.. a second launch
// with 128 columns per pass (half the tile columns per wave): one evaluation
// of their entries instead of two, while the smaller tiles keep the shape
// that suits them.
template <int D>
static void launch_sampleT_direct(Context& c, const LrLevel& L, int kmax, double* B) {
    constexpr int kAll = 1 << 30;
    switch (std::min(5, (kmax + kB16 - 1) / kB16)) {
        case 1: launch_sampleT_direct_nb<D, 1>(c, L, kmax, B, 0, kAll); break;
        case 2: launch_sampleT_direct_nb<D, 2>(c, L, kmax, B, 0, kAll); break;
        case 3: launch_sampleT_direct_nb<D, 3>(c, L, kmax, B, 0, kAll); break;
        case 4: launch_sampleT_direct_nb<D, 4>(c, L, kmax, B, 0, kAll); break;
        default:
            launch_sampleT_direct_nb<D, 4>(c, L, 64, B, 0, 64);
            launch_sampleT_direct_nb<D, 8, 2>(c, L, kmax, B, 64, kAll);
            break;
    }
}

// progress lines on stderr when HMX_LOG is set (long builds)
static bool log_on() {
    static const bool on = std::getenv("HMX_LOG") != nullptr;
    return on;
}

// Scratch of the batched ARA.  One build keeps it across its levels and
// sizes it once for the largest of them: allocating tens of GB afresh for
// every level (right after freeing the previous level's) stalled the build
// for a second and more per level.
struct AraScratch {
    DBuf<double> Q, W, Ya, Om, Ypre, Ypart, Y;
    DBuf<LrState> st;
    DBuf<TileParam> tp;
    DBuf<int64_t> first, off;
    DBuf<int> flags;
    // direct-mode samplers (Y = A Omega, B = A^T Q from on-the-fly kernel
    // evaluations): algorithmic flops (2 m^2 per sample column and tile),
    // kernel evaluations, and their kernel time on the stream
    double samplerFlops = 0, kernelEvals = 0, samplerMs = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};
    ~AraScratch() {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    }
    void mark(Context& c, int i) {
        if (!ev[i]) HMX_HIP(hipEventCreate(&ev[i]));
        HMX_HIP(hipEventRecord(ev[i], c.stream));
    }
    void add_time() {  // after a synchronisation: ev[0] .. ev[1]
        float t = 0;
        HMX_HIP(hipEventElapsedTime(&t, ev[0], ev[1]));
        samplerMs += t;
    }
};

// Buffer sizes of one batched ARA launch sequence.  Direct mode: one
// sampler sweep yields up to nbMax consecutive blocks (each kernel entry
// evaluated once for all of them), kept in Ypre until the ARA reaches them;
// small launches split their columns over several workgroups.
struct AraShape {
    int capQ = 0, nbMax = 1, splits = 1;
    size_t blockElems = 0;
};

static AraShape ara_shape(int T, int m, int maxRank, bool direct) {
    AraShape s;
    s.capQ = std::max(std::min(maxRank, 128), std::min(maxRank, kB16));
    s.blockElems = (size_t)T * m * kB16;
    if (direct) {
        s.nbMax = (int)std::max<size_t>(1, std::min<size_t>(4, (size_t(16) << 30) / (2 * s.blockElems * 8)));
        const int64_t rows = (int64_t)T * m;
        while (s.splits * 2 <= m / 64 && rows * s.splits < ((int64_t)1 << 18)) s.splits *= 2;
    }
    return s;
}

struct AraSizes {
    size_t q = 0, w = 0, blk = 0, om = 0, ypre = 0, ypart = 0, y = 0, tiles = 0;
    void add(int T, int m, int maxRank, bool direct) {
        const AraShape s = ara_shape(T, m, maxRank, direct);
        q = std::max(q, (size_t)T * m * s.capQ);
        w = std::max(w, (size_t)T * s.capQ * kB16);
        blk = std::max(blk, s.blockElems);
        om = std::max(om, s.blockElems * s.nbMax);
        if (direct) ypre = std::max(ypre, s.blockElems * s.nbMax);
        else y = std::max(y, s.blockElems);
        if (s.splits > 1) ypart = std::max(ypart, (size_t)s.splits * s.blockElems * s.nbMax);
        tiles = std::max(tiles, (size_t)T);
    }
};

template <class X>
static X* need(DBuf<X>& b, size_t n) {
    if (b.n < n) b.alloc(n);
    return b.p;
}

static void ara_reserve(AraScratch& ws, const AraSizes& z) {
    need(ws.Q, z.q);
    need(ws.W, z.w);
    need(ws.Ya, z.blk);
    need(ws.Om, z.om);
    need(ws.Ypre, z.ypre);
    need(ws.Y, z.y);
    need(ws.Ypart, z.ypart);
    need(ws.st, z.tiles);
    need(ws.tp, z.tiles);
    need(ws.first, z.y ? z.tiles : 0);
    need(ws.off, z.tiles + 1);
    need(ws.flags, 2);
}

// Result of one batched ARA launch sequence: ranks, exclusive column
// offsets and the packed factors of its tiles, in tile order.
struct LrOut {
    std::vector<int32_t> ranks;
    std::vector<char> capped;
    std::vector<int64_t> offsets;
    DBuf<double> U, V;
    int64_t rankSum = 0;
    int caps = 0, kmax = 0, iters = 0, sweeps = 0;
};

// Batched ARA over T tiles of dimension m (bus = m/b leaves per side), each
// with its own threshold, rank cap, random stream and (direct mode) point
// base and correlation length.  Scratch comes from ws (grown if too small).
//
// outSrc (optional, one entry per output tile): output tile j holds
// compressed tile outSrc[j], or, for outSrc[j] = ~i < 0, tile i's factors
// with U and V exchanged (a mirrored tile).  Every compressed tile appears
// once unmirrored.  Without outSrc the outputs are the compressed tiles.
static void compress_batch(Context& c, int araR, uint64_t seed, const LevelSource& src, int b, int m,
                           const std::vector<TileParam>& tp, const std::vector<int64_t>& first, int firstBlocks,
                           AraScratch& ws, LrOut& out, const std::vector<int32_t>* outSrc = nullptr) {
    const bool direct = src.mo == nullptr;
    const int bus = m / b;
    const int T = (int)tp.size();
    const int Tout = outSrc ? (int)outSrc->size() : T;
    const auto tA = std::chrono::steady_clock::now();
    out.ranks.assign(Tout, 0);
    out.offsets.assign(Tout + 1, 0);
    if (T == 0) return;
    HMX_CHECK(!direct || m % 64 == 0, HMX_EUNSUPPORTED, "direct compression needs tile dimensions that are multiples of 64");
    int maxRank = 0;
    for (const TileParam& q : tp) maxRank = std::max(maxRank, (int)q.maxRank);
    const AraShape shape = ara_shape(T, m, maxRank, direct);
    {
        AraSizes z;
        z.add(T, m, maxRank, direct);
        ara_reserve(ws, z);
    }
    int capQ = shape.capQ;
    const size_t blockElems = shape.blockElems;
    const int nbMax = shape.nbMax, splits = shape.splits;
    HMX_HIP(hipMemcpyAsync(ws.tp.p, tp.data(), T * sizeof(TileParam), hipMemcpyHostToDevice, c.stream));
    if (!direct)
        HMX_HIP(hipMemcpyAsync(ws.first.p, first.data(), T * sizeof(int64_t), hipMemcpyHostToDevice, c.stream));
    int nbFirst = std::max(1, std::min(nbMax, firstBlocks)), nbLater = std::min(nbMax, 2);
    if (const char* e = std::getenv("HMX_SAMPLER_BLOCKS")) {  // tuning: "first,later"
        int f = 0, l = 0;
        if (std::sscanf(e, "%d,%d", &f, &l) == 2 && f >= 1 && l >= 1) {
            nbFirst = std::min(nbMax, f);
            nbLater = std::min(nbMax, l);
        }
    }
    LrLevel L;
    L.T = T; L.m = m; L.b = b; L.bus = bus; L.r = araR;
    L.seed = seed;
    L.tp = ws.tp.p;
    L.first = direct ? nullptr : ws.first.p;
    L.ranksMO = direct ? nullptr : src.mo->ranks.p;
    L.offMO = direct ? nullptr : src.mo->offsets.p;
    L.UMO = direct ? nullptr : src.mo->U.p;
    L.VMO = direct ? nullptr : src.mo->V.p;
    L.st = ws.st.p; L.Q = ws.Q.p; L.Y = direct ? ws.Ypre.p : ws.Y.p; L.Ya = ws.Ya.p; L.Om = ws.Om.p; L.W = ws.W.p;
    L.capQ = capQ;
    L.P = src.P;
    L.boff = nullptr;
    k_lr_init<<<(T + 255) / 256, 256, 0, c.stream>>>(T, ws.st.p);
    HMX_LAUNCHED();
    const auto tB = std::chrono::steady_clock::now();
    const unsigned rowBlocks = (unsigned)((m + 63) / 64);
    const unsigned wgThreads = m >= 4096 ? 1024u : 256u;
    int kmax = 0, preStart = 0, preN = 0, sweeps = 0;
    int active = T;  // tiles still iterating (the count k_lr_commit reports)
    const int maxIters = (2 * maxRank + araR) / kB16 + 2;
    for (int it = 0;; ++it) {
        bool timedSweep = false;
        HMX_CHECK(it <= maxIters, HMX_EDEVICE, "batched ARA did not converge (non-finite input?)");
        if (kmax + kB16 > capQ && capQ < maxRank) {  // grow the basis storage
            int nc = std::min(maxRank, std::max(2 * capQ, kmax + kB16));
            DBuf<double> Q2((size_t)T * m * nc), W2((size_t)T * nc * kB16);
            k_lr_regrow<<<dim3(std::min<unsigned>(1024, (unsigned)(((size_t)m * kmax + 255) / 256 + 1)), T), 256, 0,
                          c.stream>>>(T, m, capQ, nc, ws.st.p, ws.Q.p, Q2.p);
            HMX_LAUNCHED();
            HMX_HIP(hipStreamSynchronize(c.stream));
            ws.Q = std::move(Q2);
            ws.W = std::move(W2);
            capQ = nc;
            L.Q = ws.Q.p; L.W = ws.W.p; L.capQ = capQ;
        }
        if (direct) {
            if (it >= preStart + preN) {  // sweep for the next nb blocks
                const int nb = it == 0 ? nbFirst : nbLater;
                k_lr_omega<<<dim3((m * kB16 * nb + 255) / 256, T), 256, 0, c.stream>>>(L, it, nb, 1);
                HMX_LAUNCHED();
                ws.mark(c, 0);
                if (src.d == 1) launch_sample_direct<1>(c, L, nb, splits, ws.Ypre.p, ws.Ypart.p);
                else if (src.d == 2) launch_sample_direct<2>(c, L, nb, splits, ws.Ypre.p, ws.Ypart.p);
                else launch_sample_direct<3>(c, L, nb, splits, ws.Ypre.p, ws.Ypart.p);
                ws.mark(c, 1);
                timedSweep = true;
                const double mm = (double)m * m;
                ws.samplerFlops += (double)active * 2.0 * mm * kB16 * nb;
                ws.kernelEvals += (double)active * mm;
                preStart = it;
                preN = nb;
                ++sweeps;
            }
            L.Y = ws.Ypre.p + (size_t)(it - preStart) * blockElems;
        } else {
            k_lr_omega<<<dim3((m * kB16 + 255) / 256, T), 256, 0, c.stream>>>(L, it, 1, 0);
            HMX_LAUNCHED();
            if (b <= 32) launch_sample_tlr<32>(c, L, bus, T);
            else launch_sample_tlr<64>(c, L, bus, T);
        }
        const unsigned kb = (unsigned)std::max(1, (kmax + 15) / 16);
        for (int pass = 0; pass < 2 && kmax > 0; ++pass) {
            k_lr_dot<<<dim3(kb, T), 256, 0, c.stream>>>(L, 0);
            HMX_LAUNCHED();
            k_lr_update<<<dim3(rowBlocks, T), 256, 0, c.stream>>>(L, 0);
            HMX_LAUNCHED();
        }
        if (m == 64) k_lr_decide_w<1><<<T, 64, 0, c.stream>>>(L);
        else if (m == 128) k_lr_decide_w<2><<<T, 64, 0, c.stream>>>(L);
        else if (m == 256) k_lr_decide_w<4><<<T, 64, 0, c.stream>>>(L);
        else if (m == 512) k_lr_decide_w<8><<<T, 64, 0, c.stream>>>(L);
        else if (m == 1024) k_lr_decide_b<4><<<T, 256, 0, c.stream>>>(L);
        else if (m == 2048) k_lr_decide_b<8><<<T, 256, 0, c.stream>>>(L);
        else k_lr_decide<<<T, wgThreads, 0, c.stream>>>(L);
        HMX_LAUNCHED();
        if (kmax > 0) {
            k_lr_dot<<<dim3(kb, T), 256, 0, c.stream>>>(L, 2);
            HMX_LAUNCHED();
            k_lr_update<<<dim3(rowBlocks, T), 256, 0, c.stream>>>(L, 2);
            HMX_LAUNCHED();
        }
        if (m == 64) k_lr_renorm_w<1><<<T, 64, 0, c.stream>>>(L);
        else if (m == 128) k_lr_renorm_w<2><<<T, 64, 0, c.stream>>>(L);
        else if (m == 256) k_lr_renorm_w<4><<<T, 64, 0, c.stream>>>(L);
        else if (m == 512) k_lr_renorm_w<8><<<T, 64, 0, c.stream>>>(L);
        else if (m == 1024) k_lr_renorm_b<4><<<T, 256, 0, c.stream>>>(L);
        else if (m == 2048) k_lr_renorm_b<8><<<T, 256, 0, c.stream>>>(L);
        else k_lr_renorm<<<T, wgThreads, 0, c.stream>>>(L);
        HMX_LAUNCHED();
        HMX_HIP(hipMemsetAsync(ws.flags.p, 0, 2 * sizeof(int), c.stream));
        k_lr_commit<<<(T + 255) / 256, 256, 0, c.stream>>>(L, ws.flags.p);
        HMX_LAUNCHED();
        int hf[2];
        HMX_HIP(hipMemcpyAsync(hf, ws.flags.p, sizeof(hf), hipMemcpyDeviceToHost, c.stream));
        HMX_HIP(hipStreamSynchronize(c.stream));
        if (timedSweep) ws.add_time();
        kmax = hf[1];
        active = hf[0];
        out.iters = it + 1;
        if (hf[0] == 0) break;
    }
    const auto tC = std::chrono::steady_clock::now();
    out.sweeps = sweeps;
    // ranks and packed offsets, then B = A^T Q straight into the packed V
    std::vector<LrState> hst(T);
    HMX_HIP(hipMemcpyAsync(hst.data(), ws.st.p, T * sizeof(LrState), hipMemcpyDeviceToHost, c.stream));
    HMX_HIP(hipStreamSynchronize(c.stream));
    int64_t acc = 0;
    int caps = 0;
    out.capped.assign(Tout, 0);
    // offsets of the compressed tiles' own outputs (where A^T Q and the
    // packed Q go), and the mirrored outputs to fill from them
    std::vector<int64_t> primOff(T + 1, 0), mDst, mSrc;
    std::vector<int32_t> mK;
    for (int j = 0; j < Tout; ++j) {
        const int32_t q = outSrc ? (*outSrc)[j] : j;
        const int i = q >= 0 ? q : ~q;
        out.ranks[j] = hst[i].k;
        out.offsets[j] = acc;
        caps += hst[i].capped;
        out.capped[j] = (char)hst[i].capped;
        if (q >= 0) primOff[i] = acc;
        acc += hst[i].k;
    }
    for (int j = 0; j < Tout; ++j) {
        const int32_t q = outSrc ? (*outSrc)[j] : j;
        if (q < 0 && hst[~q].k > 0) {
            mDst.push_back(out.offsets[j]);
            mSrc.push_back(primOff[~q]);
            mK.push_back(hst[~q].k);
        }
    }
    primOff[T] = acc;
    out.offsets[Tout] = acc;
    out.rankSum = acc;
    out.caps = caps;
    out.kmax = kmax;
    out.U.alloc((size_t)acc * m);
    out.V.alloc((size_t)acc * m);
    const auto tAlloc = std::chrono::steady_clock::now();
    double sampleTms = 0;
    if (acc > 0) {
        HMX_HIP(hipMemcpyAsync(ws.off.p, primOff.data(), (T + 1) * sizeof(int64_t), hipMemcpyHostToDevice, c.stream));
        L.boff = ws.off.p;
        if (direct) {
            ws.mark(c, 0);
            if (src.d == 1) launch_sampleT_direct<1>(c, L, kmax, out.V.p);
            else if (src.d == 2) launch_sampleT_direct<2>(c, L, kmax, out.V.p);
            else launch_sampleT_direct<3>(c, L, kmax, out.V.p);
            ws.mark(c, 1);
            if (log_on()) {  // the AᵀQ sweep alone (the stream is synchronised below anyway)
                HMX_HIP(hipStreamSynchronize(c.stream));
                sampleTms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tAlloc).count();
            }
            const double mm = (double)m * m;
            for (int t = 0; t < T; ++t) {
                ws.samplerFlops += 2.0 * mm * hst[t].k;
                ws.kernelEvals += mm * (double)((hst[t].k + 63) / 64);
            }
        } else if (b <= 32) {
            launch_sampleT_tlr<32>(c, L, kmax, bus, T, out.V.p);
        } else {
            launch_sampleT_tlr<64>(c, L, kmax, bus, T, out.V.p);
        }
        dim3 g(std::min<unsigned>(1024, (unsigned)(((size_t)m * kmax + 255) / 256 + 1)), T);
        k_lr_pack<<<g, 256, 0, c.stream>>>(T, m, capQ, ws.st.p, ws.off.p, ws.Q.p, out.U.p);
        HMX_LAUNCHED();
        if (!mDst.empty()) {
            const int nm = (int)mDst.size();
            DBuf<int64_t> dd(nm), ds(nm);
            DBuf<int32_t> dk(nm);
            HMX_HIP(hipMemcpyAsync(dd.p, mDst.data(), nm * sizeof(int64_t), hipMemcpyHostToDevice, c.stream));
            HMX_HIP(hipMemcpyAsync(ds.p, mSrc.data(), nm * sizeof(int64_t), hipMemcpyHostToDevice, c.stream));
            HMX_HIP(hipMemcpyAsync(dk.p, mK.data(), nm * sizeof(int32_t), hipMemcpyHostToDevice, c.stream));
            const int kmaxM = *std::max_element(mK.begin(), mK.end());
            dim3 gm(std::min<unsigned>(1024, (unsigned)(((size_t)m * kmaxM + 255) / 256)), (unsigned)nm);
            k_lr_mirror<<<gm, 256, 0, c.stream>>>(m, dd.p, ds.p, dk.p, out.U.p, out.V.p);
            HMX_LAUNCHED();
            HMX_HIP(hipStreamSynchronize(c.stream));  // before dd, ds, dk go
        }
        HMX_HIP(hipStreamSynchronize(c.stream));
        if (direct) ws.add_time();
    }
    if (log_on()) {
        auto ms = [](std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
            return std::chrono::duration<double, std::milli>(b - a).count();
        };
        const auto tD = std::chrono::steady_clock::now();
        std::fprintf(stderr,
                     "[hmx]   T=%d m=%d: setup %.1f ms, ARA loop %.1f ms, A^T Q + pack %.1f ms "
                     "(ranks and factor allocation %.1f ms, A^T Q %.1f ms)\n",
                     T, m, ms(tA, tB), ms(tB, tC), ms(tC, tD), ms(tC, tAlloc), sampleTms);
    }
}

static TileParam tile_param(double tol, double ell, int64_t pbase, int level, int t, int maxRank, int32_t code) {
    TileParam q;
    q.tau = ara_threshold(tol);
    q.nil = ell > 0 ? -1.0 / ell : 0.0;
    q.pbase = pbase;
    q.opid = ((uint64_t)(level + 1) << 40) | (uint64_t)t;
    q.maxRank = maxRank;
    q.code = code;
    return q;
}

// Blocks a tile of the next level up is likely to need in its first sweep:
// the ARA stops r samples after its last accepted column, so a tile of rank
// k takes ceil((k + r)/16) blocks; the median rank of the level below stands
// in for k (ranks grow slowly towards the root).  A block costs one more
// product per kernel entry, a second sweep another evaluation of every
// entry, so covering the median tile in one sweep pays: at N = 2^20 3D the
// samplers take 8.9 s against 9.6 s from the 10th-percentile rank rounded
// down (14% fewer kernel evaluations; the ranks are the same).
static int first_sweep_blocks(std::vector<int32_t> ranks, int araR) {
    if (ranks.empty()) return 2;
    const size_t q = ranks.size() / 2;
    std::nth_element(ranks.begin(), ranks.begin() + q, ranks.end());
    return std::max(1, std::min(4, (ranks[q] + araR + kB16 - 1) / kB16));
}

// Symmetric kernels and admissibility: tile (u, v) of a level (column
// cluster u, row cluster v) is the transpose of tile (v, u).  The tiles to
// compress are those with u < v, those with u == v and any whose partner is
// missing (prim: their indices in the level); outSrc maps every tile of the
// level to its compressed tile (compress_batch), the partners mirrored.
void mirror_plan(const std::vector<int32_t>& tiles, std::vector<int>& prim, std::vector<int32_t>& outSrc) {
    const int T = (int)tiles.size();
    std::unordered_map<int32_t, int> pos;
    pos.reserve((size_t)T * 2);
    for (int t = 0; t < T; ++t) pos[tiles[t]] = t;
    auto partner = [&](int t) {
        const uint32_t code = (uint32_t)tiles[t];
        const auto it = pos.find((int32_t)morton_xy(compact16(code >> 1), compact16(code)));
        return it == pos.end() ? -1 : it->second;
    };
    std::vector<int> idx(T, -1);
    prim.clear();
    for (int t = 0; t < T; ++t) {
        const uint32_t code = (uint32_t)tiles[t];
        if (compact16(code) > compact16(code >> 1) && partner(t) >= 0) continue;  // mirrored
        idx[t] = (int)prim.size();
        prim.push_back(t);
    }
    outSrc.assign(T, 0);
    for (int t = 0; t < T; ++t) outSrc[t] = idx[t] >= 0 ? idx[t] : ~idx[partner(t)];
}

static void compress_level(Context& c, const hmx_config& cfg, const LevelSource& src, int b,
                           const std::vector<int32_t>& tiles, int level, int numLevels, float tol, int maxRank,
                           int firstBlocks, AraScratch& ws, HLevel& out) {
    const bool direct = src.mo == nullptr;
    const auto tstart = std::chrono::steady_clock::now();
    const int bus = 1 << (numLevels - (level + 1));
    const int m = bus * b;
    const int T = (int)tiles.size();
    out.level = level;
    out.numTiles = T;
    out.m = m;
    out.maxRank = maxRank;
    out.tol = tol;
    out.h_ranks.assign(T, 0);
    out.h_offsets.assign(T + 1, 0);
    if (T == 0) return;
    std::vector<TileParam> tp;
    std::vector<int64_t> first;
    std::vector<int> prim;
    std::vector<int32_t> outSrc;
    const bool mirror = direct && cfg.symmetricTiles;
    if (mirror) {
        mirror_plan(tiles, prim, outSrc);
    } else {
        prim.resize(T);
        std::iota(prim.begin(), prim.end(), 0);
    }
    for (int t : prim) {  // a compressed tile keeps its own random stream (its index t in the level)
        tp.push_back(tile_param((double)tol, src.ell, 0, level, t, maxRank, tiles[t]));
        first.push_back((int64_t)tiles[t] * bus * bus);
    }
    LrOut r;
    compress_batch(c, cfg.araR, cfg.seed, src, b, m, tp, first, firstBlocks, ws, r, mirror ? &outSrc : nullptr);
    out.h_ranks = std::move(r.ranks);
    out.h_offsets = std::move(r.offsets);
    out.rankSum = r.rankSum;
    out.capHits = r.caps;
    out.U = std::move(r.U);
    out.V = std::move(r.V);
    if (log_on()) {
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tstart).count();
        std::fprintf(stderr, "[hmx] level %d (%s): %d tiles m=%d rank sum %lld max %d, %d ARA blocks in %d sweeps, %.1f ms\n",
                     level, direct ? "direct" : "TLR sum", T, m, (long long)out.rankSum, r.kmax, r.iters, r.sweeps, ms);
        std::fflush(stderr);
    }
}

// upper levels (generateHMatrixFromStruct): tolerance doubles per non-empty
// level (src/constructHMatrixFromStruct.cu:54), rank cap from
// generateMaxRanks (src/helperFunctions.cu:67-74)
static void compress_upper_levels(Context& c, const hmx_config& cfg, const Structure& s, const LevelSource& src, int b,
                                  HMatrix& h) {
    const int L = s.numLevels;
    std::vector<uint32_t> maxRanks(L > 2 ? L - 2 : 1, 0);
    hmx_generateMaxRanks((uint32_t)L, (uint32_t)b, maxRanks.data());
    auto level_cap = [&](int level) {
        const int m = (1 << (L - (level + 1))) * b;
        const int maxRank = cfg.maxRankPolicy == HMX_MAXRANK_REF ? (int)maxRanks[L - level - 2] : m;
        return std::min(maxRank, m);
    };
    AraScratch ws;
    {
        AraSizes z;
        for (int level = L - 2; level > 0; --level)
            if (s.numTiles[level - 1] > 0)
                z.add(s.numTiles[level - 1], (1 << (L - (level + 1))) * b, level_cap(level), src.mo == nullptr);
        ara_reserve(ws, z);
    }
    float tol = cfg.lowestLevelTolerance;
    const std::vector<int32_t>* below = &h.levels[L - 2].h_ranks;
    for (int level = L - 2; level > 0; --level) {
        HLevel& lv = h.levels[level - 1];
        const int T = s.numTiles[level - 1];
        const int m = (1 << (L - (level + 1))) * b;
        if (T == 0) {
            lv.level = level; lv.numTiles = 0; lv.m = m;
            lv.h_ranks.clear(); lv.h_offsets.assign(1, 0);
            continue;
        }
        tol *= 2;
        compress_level(c, cfg, src, b, s.tileIndices[level - 1], level, L, tol, level_cap(level),
                       first_sweep_blocks(*below, cfg.araR), ws, lv);
        below = &lv.h_ranks;
    }
    h.samplerFlops = ws.samplerFlops;
    h.kernelEvals = ws.kernelEvals;
    h.samplerMs = ws.samplerMs;
}

static void finish_accounting(HMatrix& h) {
    h.lowRankBytes = 0;
    h.factorEntries = 0;
    for (auto& lv : h.levels) {
        h.lowRankBytes += 2ull * (uint64_t)lv.rankSum * lv.m * sizeof(double);
        h.factorEntries += 2ull * (uint64_t)lv.rankSum * lv.m;
    }
    h.denseBytes = (uint64_t)h.S * h.b * h.b * sizeof(double);
    h.factorEntries += (uint64_t)h.S * h.b * h.b;
}

void build_hmatrix(Context& c, const hmx_config& cfg, const KDTree& t, const Structure& s, const TLR& mo,
                   HMatrix& h) {
    HMX_CHECK(mo.ordering == HMX_MORTON, HMX_ESTATE, "generateHMatrixFromStruct: TLR matrix must be Morton-ordered");
    HMX_CHECK(mo.S == t.numSegments && mo.b == t.b, HMX_EINVAL, "TLR matrix and KD-tree do not match");
    HMX_CHECK(s.numLevels >= 2 && (uint32_t)(1u << (s.numLevels - 1)) == t.numSegments, HMX_EINVAL,
              "structure and KD-tree do not match");
    HMX_CHECK(cfg.araBlockSize == kB16, HMX_EUNSUPPORTED, "ARA block size must be 16");
    HMX_CHECK(t.b <= 64, HMX_EUNSUPPORTED, "hierarchical recompression: bucketSize must be <= 64");
    const int L = s.numLevels, b = (int)t.b;
    h.n = t.n; h.d = t.d; h.b = t.b; h.S = t.numSegments; h.numLevels = L;
    h.leafOffsets = t.h_offsets;
    h.structureCopy = s;
    h.structure = &h.structureCopy;
    h.levels.clear();
    h.levels.resize(L - 1);
    h.diag.alloc(mo.diagonal.n);
    HMX_HIP(hipMemcpyAsync(h.diag.p, mo.diagonal.p, mo.diagonal.bytes(), hipMemcpyDeviceToDevice, c.stream));
    // leaf level (allocateHMatrix, src/HMatrix.cu:57-103)
    {
        HLevel& lv = h.levels[L - 2];
        const std::vector<int32_t>& tiles = s.tileIndices[L - 2];
        const int T = (int)tiles.size();
        lv.level = L - 1; lv.numTiles = T; lv.m = b;
        lv.maxRank = cfg.maxRankPolicy == HMX_MAXRANK_REF ? b / 2 : b;
        lv.tol = cfg.lowestLevelTolerance;
        lv.h_ranks.assign(T, 0);
        lv.h_offsets.assign(T + 1, 0);
        if (T > 0) {
            DBuf<int32_t> d_tiles(T), d_r(T);
            DBuf<int64_t> d_off(T + 1);
            HMX_HIP(hipMemcpyAsync(d_tiles.p, tiles.data(), T * sizeof(int32_t), hipMemcpyHostToDevice, c.stream));
            k_leaf_ranks<<<(T + 255) / 256, 256, 0, c.stream>>>(T, d_tiles.p, mo.ranks.p, d_r.p);
            HMX_LAUNCHED();
            HMX_HIP(hipMemcpyAsync(lv.h_ranks.data(), d_r.p, T * sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
            HMX_HIP(hipStreamSynchronize(c.stream));
            int64_t acc = 0;
            for (int q = 0; q < T; ++q) { lv.h_offsets[q] = acc; acc += lv.h_ranks[q]; }
            lv.h_offsets[T] = acc;
            lv.rankSum = acc;
            lv.U.alloc((size_t)acc * b);
            lv.V.alloc((size_t)acc * b);
            HMX_HIP(hipMemcpyAsync(d_off.p, lv.h_offsets.data(), (T + 1) * sizeof(int64_t), hipMemcpyHostToDevice,
                                   c.stream));
            if (acc > 0) {
                k_leaf_gather<<<(T + 3) / 4, 256, 0, c.stream>>>(T, b, d_tiles.p, mo.ranks.p, mo.offsets.p, d_off.p,
                                                                 mo.U.p, mo.V.p, lv.U.p, lv.V.p);
                HMX_LAUNCHED();
            }
            HMX_HIP(hipStreamSynchronize(c.stream));
        }
    }
    LevelSource src;
    src.mo = &mo;
    compress_upper_levels(c, cfg, s, src, b, h);
    finish_accounting(h);
    // the matvec plan is built on the first HMatrixVecMult call
}

// Direct assembly for sizes where the TLR matrix of all b x b tiles does not
// fit in memory (S^2 tiles): the dense diagonal blocks and the leaf level are
// produced exactly as the TLR path produces them (same kernel, same random
// streams), every upper level is compressed by the same batched ARA and
// tolerance/rank schedule, sampling the kernel on the fly instead of the
// TLR sub-tiles.  Replaces createColumnMajorLRMatrix + convertColumnMajorToMorton
// + generateHMatrixFromStruct (src/constructTLRMatrix.cu:13-151,
// src/helperFunctions.cu:8-57, src/constructHMatrixFromStruct.cu:13-103).
void build_hmatrix_direct(Context& c, const hmx_config& cfg, const KDTree& t, const Structure& s,
                          const double* d_points, HMatrix& h) {
    HMX_CHECK(s.numLevels >= 2 && (uint32_t)(1u << (s.numLevels - 1)) == t.numSegments, HMX_EINVAL,
              "structure and KD-tree do not match");
    HMX_CHECK(cfg.araBlockSize == kB16, HMX_EUNSUPPORTED, "ARA block size must be 16");
    HMX_CHECK(t.b == 32 || t.b == 64, HMX_EUNSUPPORTED, "direct assembly: bucketSize must be 32 or 64");
    HMX_CHECK(t.d >= 1 && t.d <= 3, HMX_EUNSUPPORTED, "direct assembly: dimension must be 1, 2 or 3");
    HMX_CHECK(d_points != nullptr, HMX_EINVAL, "null points");
    const int L = s.numLevels, b = (int)t.b;
    h.n = t.n; h.d = t.d; h.b = t.b; h.S = t.numSegments; h.numLevels = L;
    h.leafOffsets = t.h_offsets;
    h.structureCopy = s;
    h.structure = &h.structureCopy;
    h.levels.clear();
    h.levels.resize(L - 1);
    DBuf<double> P((size_t)t.numSegments * b * 4);
    pad_points(c, b, {PadCloud{0, d_points, t.perm.p, t.offsets.p, t.n, t.d}}, t.numSegments, P.p);
    h.diag.alloc((size_t)t.numSegments * b * b);
    HLevel& leafLevel = h.levels[L - 2];
    leafLevel.level = L - 1;
    const std::vector<LeafJob> job = {LeafJob{&cfg, 0, t.numSegments, &s.tileIndices[L - 2], &leafLevel}};
    build_dense_diagonal_padded(c, t.b, t.d, P.p, job, h.diag.p);
    {
        const auto t0 = std::chrono::steady_clock::now();
        compress_leaf_tiles_padded(c, t.b, t.d, P.p, job);
        if (log_on()) {
            const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            std::fprintf(stderr, "[hmx] level %d (leaf tiles): %d tiles rank sum %lld, %.1f ms\n", L - 1,
                         leafLevel.numTiles, (long long)leafLevel.rankSum, ms);
            std::fflush(stderr);
        }
    }
    LevelSource src;
    src.P = P.p;
    src.ell = cfg.correlationLength > 0 ? cfg.correlationLength : default_ell(t.d);
    src.d = (int)t.d;
    compress_upper_levels(c, cfg, s, src, b, h);
    HMX_HIP(hipStreamSynchronize(c.stream));
    h.pts = std::move(P);
    h.ell = src.ell;
    finish_accounting(h);
}

// Many H-matrices assembled directly at once: per instance the diagonal
// blocks, the leaf level and the padded points; then the upper levels of all
// instances together, one batched ARA per (tile dimension, point dimension)
// group, each tile with its own instance's threshold schedule, rank cap,
// random stream and correlation length.  Per tile the same algorithm as
// build_hmatrix_direct; only the column split of the sampler (and so the
// summation order of the samples) can differ.
void build_hmatrix_direct_batch(Context& c, const std::vector<const hmx_config*>& cfgs,
                                const std::vector<const KDTree*>& trees, const std::vector<const Structure*>& ss,
                                const std::vector<const double*>& pts, const std::vector<HMatrix*>& hs) {
    const size_t N = hs.size();
    HMX_CHECK(N >= 1 && cfgs.size() == N && trees.size() == N && ss.size() == N && pts.size() == N, HMX_EINVAL,
              "batch arrays of different lengths");
    const int b = (int)trees[0]->b;
    std::vector<int64_t> pbase(N);
    int64_t ptot = 0;
    for (size_t i = 0; i < N; ++i) {
        const KDTree& t = *trees[i];
        const Structure& s = *ss[i];
        const hmx_config& cfg = *cfgs[i];
        HMX_CHECK((int)t.b == b, HMX_EINVAL, "all instances of a batch need the same bucketSize");
        HMX_CHECK(cfg.araR == cfgs[0]->araR && cfg.seed == cfgs[0]->seed && cfg.araBlockSize == kB16, HMX_EINVAL,
                  "all instances of a batch need the same araR, seed and ARA block size 16");
        HMX_CHECK(s.numLevels >= 2 && (uint32_t)(1u << (s.numLevels - 1)) == t.numSegments, HMX_EINVAL,
                  "structure and KD-tree do not match");
        HMX_CHECK(b == 32 || b == 64, HMX_EUNSUPPORTED, "direct assembly: bucketSize must be 32 or 64");
        HMX_CHECK(t.d >= 1 && t.d <= 3, HMX_EUNSUPPORTED, "direct assembly: dimension must be 1, 2 or 3");
        HMX_CHECK(pts[i] != nullptr, HMX_EINVAL, "null points");
        pbase[i] = ptot;
        ptot += (int64_t)t.numSegments * b;
    }
    // the padded points, the diagonal blocks and the leaf level of every
    // instance: one launch each (the leaf level one launch sequence per point
    // dimension); every instance gets its range of the shared blocks as a view
    const int64_t leaves = ptot / b;
    auto P = std::make_shared<DBuf<double>>((size_t)ptot * 4);
    auto Dg = std::make_shared<DBuf<double>>((size_t)leaves * b * b);
    {
        std::vector<PadCloud> pc(N);
        for (size_t i = 0; i < N; ++i) {
            const KDTree& t = *trees[i];
            pc[i] = PadCloud{pbase[i] / b, pts[i], t.perm.p, t.offsets.p, t.n, t.d};
        }
        pad_points(c, b, pc, leaves, P->p);
    }
    std::vector<LeafJob> jobsByD[4];
    for (size_t i = 0; i < N; ++i) {
        const KDTree& t = *trees[i];
        const Structure& s = *ss[i];
        HMatrix& h = *hs[i];
        const int L = s.numLevels;
        h.n = t.n; h.d = t.d; h.b = t.b; h.S = t.numSegments; h.numLevels = L;
        h.leafOffsets = t.h_offsets;
        h.structureCopy = s;
        h.structure = &h.structureCopy;
        h.levels.clear();
        h.levels.resize(L - 1);
        h.levels[L - 2].level = L - 1;
        h.diag.view(Dg, (size_t)pbase[i] * b, (size_t)t.numSegments * b * b);
        h.pts.view(P, (size_t)pbase[i] * 4, (size_t)t.numSegments * b * 4);
        h.ell = cfgs[i]->correlationLength > 0 ? cfgs[i]->correlationLength : default_ell(t.d);
        jobsByD[t.d].push_back(
            LeafJob{cfgs[i], pbase[i] / b, t.numSegments, &hs[i]->structure->tileIndices[L - 2], &h.levels[L - 2]});
    }
    for (uint32_t d = 1; d <= 3; ++d) {
        if (jobsByD[d].empty()) continue;
        build_dense_diagonal_padded(c, (uint32_t)b, d, P->p, jobsByD[d], Dg->p);
        compress_leaf_tiles_padded(c, (uint32_t)b, d, P->p, jobsByD[d]);
    }
    // upper levels: the reference's schedule per instance, grouped across instances
    std::vector<int32_t> leafRanks;
    for (size_t i = 0; i < N; ++i) {
        const auto& r = hs[i]->levels[ss[i]->numLevels - 2].h_ranks;
        leafRanks.insert(leafRanks.end(), r.begin(), r.end());
    }
    struct Job {
        int inst, level;
    };
    std::vector<std::pair<std::pair<int, int>, std::vector<Job>>> groups;  // ((m, d), jobs)
    for (size_t i = 0; i < N; ++i) {
        const Structure& s = *ss[i];
        const hmx_config& cfg = *cfgs[i];
        HMatrix& h = *hs[i];
        const int L = s.numLevels;
        std::vector<uint32_t> maxRanks(L > 2 ? L - 2 : 1, 0);
        hmx_generateMaxRanks((uint32_t)L, (uint32_t)b, maxRanks.data());
        float tol = cfg.lowestLevelTolerance;
        for (int level = L - 2; level > 0; --level) {
            HLevel& lv = h.levels[level - 1];
            const int T = s.numTiles[level - 1];
            const int m = (1 << (L - (level + 1))) * b;
            lv.level = level;
            lv.numTiles = T;
            lv.m = m;
            lv.h_ranks.assign(T, 0);
            lv.h_offsets.assign(T + 1, 0);
            if (T == 0) continue;
            tol *= 2;
            int maxRank = cfg.maxRankPolicy == HMX_MAXRANK_REF ? (int)maxRanks[L - level - 2] : m;
            lv.maxRank = std::min(maxRank, m);
            lv.tol = tol;
            const std::pair<int, int> key(m, (int)h.d);
            size_t g = 0;
            while (g < groups.size() && groups[g].first != key) ++g;
            if (g == groups.size()) groups.push_back({key, {}});
            groups[g].second.push_back(Job{(int)i, level});
        }
    }
    AraScratch ws;
    {
        AraSizes z;
        for (auto& grp : groups) {
            int T = 0, cap = 0;
            for (const Job& j : grp.second) {
                const HLevel& lv = hs[j.inst]->levels[j.level - 1];
                T += lv.numTiles;
                cap = std::max(cap, lv.maxRank);
            }
            z.add(T, grp.first.first, cap, true);
        }
        ara_reserve(ws, z);
    }
    for (auto& grp : groups) {
        const int m = grp.first.first, d = grp.first.second;
        const auto tstart = std::chrono::steady_clock::now();
        std::vector<TileParam> tp;
        std::vector<int32_t> outSrc;  // the level tiles of all jobs, in job order
        bool anyMirror = false;
        for (const Job& j : grp.second) {
            const hmx_config& cfg = *cfgs[j.inst];
            const HLevel& lv = hs[j.inst]->levels[j.level - 1];
            const double ell = cfg.correlationLength > 0 ? cfg.correlationLength : default_ell(trees[j.inst]->d);
            const auto& tiles = ss[j.inst]->tileIndices[j.level - 1];
            std::vector<int> prim;
            std::vector<int32_t> src1;
            if (cfg.symmetricTiles) {
                mirror_plan(tiles, prim, src1);
                anyMirror = true;
            } else {
                prim.resize(tiles.size());
                std::iota(prim.begin(), prim.end(), 0);
                src1.resize(tiles.size());
                std::iota(src1.begin(), src1.end(), 0);
            }
            const int32_t cb = (int32_t)tp.size();
            for (int32_t q : src1) outSrc.push_back(q >= 0 ? q + cb : ~(~q + cb));
            for (int t : prim)
                tp.push_back(tile_param((double)lv.tol, ell, pbase[j.inst], j.level, t, lv.maxRank, tiles[t]));
        }
        LevelSource src;
        src.P = P->p;
        src.d = d;
        LrOut r;
        compress_batch(c, cfgs[0]->araR, cfgs[0]->seed, src, b, m, tp, {}, first_sweep_blocks(leafRanks, cfgs[0]->araR),
                       ws, r, anyMirror ? &outSrc : nullptr);
        auto gU = std::make_shared<DBuf<double>>(std::move(r.U));
        auto gV = std::make_shared<DBuf<double>>(std::move(r.V));
        size_t t0 = 0;
        for (const Job& j : grp.second) {
            HLevel& lv = hs[j.inst]->levels[j.level - 1];
            const int T = lv.numTiles;
            int64_t acc = 0;
            int caps = 0;
            for (int t = 0; t < T; ++t) {
                lv.h_ranks[t] = r.ranks[t0 + t];
                lv.h_offsets[t] = acc;
                acc += r.ranks[t0 + t];
                caps += r.capped[t0 + t];
            }
            lv.h_offsets[T] = acc;
            lv.rankSum = acc;
            lv.capHits = caps;
            lv.U.view(gU, (size_t)r.offsets[t0] * m, (size_t)acc * m);
            lv.V.view(gV, (size_t)r.offsets[t0] * m, (size_t)acc * m);
            t0 += T;
        }
        HMX_HIP(hipStreamSynchronize(c.stream));
        if (log_on()) {
            const double ms =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tstart).count();
            std::fprintf(stderr, "[hmx] batch group m=%d d=%d: %zu tiles from %zu levels, %d ARA blocks in %d sweeps, %.1f ms\n",
                         m, d, tp.size(), grp.second.size(), r.iters, r.sweeps, ms);
            std::fflush(stderr);
        }
    }
    for (HMatrix* h : hs) finish_accounting(*h);
}

}  // namespace hmx
