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
    double* Ypart;          // column-split partial samples [split][T][16][m]
};

__device__ inline bool lr_active(const LrState& s, int maxRank) {
    return !(s.converged || s.capped || s.k >= maxRank);
}

__global__ void k_lr_init(int T, LrState* st) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    st[t].k = 0; st[t].small_run = 0; st[t].converged = 0; st[t].capped = 0; st[t].na = 0;
}

// Om[t][s*m + c] (transposed == 0) or Om[t][c*16 + s] (transposed == 1)
__global__ void __launch_bounds__(256) k_lr_omega(LrLevel L, int it, int transposed) {
    int t = blockIdx.y;
    if (!lr_active(L.st[t], L.tp[t].maxRank)) return;
    int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= L.m * kB16) return;
    int s = transposed ? e % kB16 : e / L.m, c = transposed ? e / kB16 : e % L.m;
    L.Om[(size_t)t * L.m * kB16 + e] = gauss_at(L.seed, kTagOmega, L.tp[t].opid, (uint32_t)(it * kB16 + s), (uint32_t)c);
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
__global__ void __launch_bounds__(256) k_pad_points(uint32_t n, uint32_t d, int b, uint32_t S,
                                                    const double* __restrict__ p, const int32_t* __restrict__ perm,
                                                    const int32_t* __restrict__ offsets, double* __restrict__ P) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)S * b) return;
    const int leaf = (int)(e / b), i = (int)(e % b);
    const int a0 = offsets[leaf], cnt = offsets[leaf + 1] - a0;
    double x[4] = {0.0, 0.0, 0.0, 0.0};
    if (i < cnt) {
        for (uint32_t j = 0; j < d; ++j) x[j] = p[(size_t)j * n + perm[a0 + i]];
        x[3] = 1.0;
    }
    double4* o = reinterpret_cast<double4*>(P + e * 4);
    *o = make_double4(x[0], x[1], x[2], x[3]);
}

// Y[t][s*m + q] (or the split's partial) = sum_c K(x_q, x_c) Om[t][c*16 + s]
template <int D>
__global__ void __launch_bounds__(256) k_dir_sample(LrLevel L, int splits) {
    __shared__ double4 sP[4][64];
    __shared__ double sO[4][64][kB16];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int m = L.m;
    const int64_t gr = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool inb = gr < (int64_t)L.T * m;  // uniform per wave (T*m is a multiple of 64)
    const int t = inb ? (int)(gr / m) : 0, q = (int)(gr % m);
    const bool act = inb && lr_active(L.st[t], L.tp[t].maxRank);
    const TileParam tpt = L.tp[t];
    const uint32_t code = (uint32_t)tpt.code;
    const double nil = tpt.nil;
    const int64_t rowBase = tpt.pbase + (int64_t)compact16(code >> 1) * m;
    const int64_t colBase = tpt.pbase + (int64_t)compact16(code) * m;
    const double4 xr = act ? reinterpret_cast<const double4*>(L.P)[rowBase + q] : make_double4(0, 0, 0, 0);
    const double xra[3] = {xr.x, xr.y, xr.z};
    const int chunk = m / splits, c0 = blockIdx.y * chunk;
    const double* Om = L.Om + (size_t)t * m * kB16;
    double acc[kB16];
#pragma unroll
    for (int s = 0; s < kB16; ++s) acc[s] = 0.0;
    for (int cb = c0; cb < c0 + chunk; cb += 64) {
        __syncthreads();
        if (act) {
            sP[w][lane] = reinterpret_cast<const double4*>(L.P)[colBase + cb + lane];
            const double2* src = reinterpret_cast<const double2*>(Om + (size_t)(cb + lane) * kB16);
#pragma unroll
            for (int s2 = 0; s2 < kB16 / 2; ++s2) {
                const double2 v = src[s2];
                sO[w][lane][2 * s2] = v.x;
                sO[w][lane][2 * s2 + 1] = v.y;
            }
        }
        __syncthreads();
        if (act) {
            for (int cc = 0; cc < 64; ++cc) {
                const double4 xc = sP[w][cc];
                const double xca[3] = {xc.x, xc.y, xc.z};
                const double kv = kernel_eval_scaled<D>(xra, xca, nil) * xc.w;
#pragma unroll
                for (int s = 0; s < kB16; ++s) acc[s] = fma(kv, sO[w][cc][s], acc[s]);
            }
        }
    }
    if (!act) return;
    double* Y = splits > 1 ? L.Ypart + ((size_t)blockIdx.y * L.T + t) * m * kB16 : L.Y + (size_t)t * m * kB16;
#pragma unroll
    for (int s = 0; s < kB16; ++s) Y[(size_t)s * m + q] = acc[s] * xr.w;
}

// The same two products on the FP64 matrix cores (v_mfma_f64_16x16x4f64).
// A wave owns 64 rows (AΩ) or 64 columns (AᵀQ) of one tile as four blocks of
// 16; per step of 4 along the inner index a lane evaluates one kernel entry
// per block -- the A operand, lane l holding A[l%16][l/16] -- and reads one
// entry of Ω or Q from LDS as the B operand (B[l/16][l%16]).  The 16x16
// result of a block sits in register v of lane l at row 4v + l/16, column l%16.
typedef double v4d_t __attribute__((ext_vector_type(4)));

template <int D>
__global__ void __launch_bounds__(256) k_dir_sample_mf(LrLevel L, int splits) {
    __shared__ double4 sP[4][64];
    __shared__ double sO[4][64][kB16];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, lg = lane >> 4, lc = lane & 15;
    const int m = L.m;
    const int64_t r0g = ((int64_t)blockIdx.x * 4 + w) * 64;  // first row of this wave in the level's row space
    const bool inb = r0g < (int64_t)L.T * m;                 // m is a multiple of 64
    const int t = inb ? (int)(r0g / m) : 0, q0 = (int)(r0g % m);
    const bool act = inb && lr_active(L.st[t], L.tp[t].maxRank);
    const TileParam tpt = L.tp[t];
    const uint32_t code = (uint32_t)tpt.code;
    const double nil = tpt.nil;
    const int64_t rowBase = tpt.pbase + (int64_t)compact16(code >> 1) * m;
    const int64_t colBase = tpt.pbase + (int64_t)compact16(code) * m;
    double xr[4][3], wr[4];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
        const double4 p = act ? reinterpret_cast<const double4*>(L.P)[rowBase + q0 + 16 * rb + lc]
                              : make_double4(0, 0, 0, 0);
        xr[rb][0] = p.x; xr[rb][1] = p.y; xr[rb][2] = p.z; wr[rb] = p.w;
    }
    const int chunk = m / splits, c0 = blockIdx.y * chunk;
    const double* Om = L.Om + (size_t)t * m * kB16;
    v4d_t acc[4];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) acc[rb] = v4d_t{0.0, 0.0, 0.0, 0.0};
    for (int cb = c0; cb < c0 + chunk; cb += 64) {
        __syncthreads();
        if (act) {
            sP[w][lane] = reinterpret_cast<const double4*>(L.P)[colBase + cb + lane];
            const double2* src = reinterpret_cast<const double2*>(Om + (size_t)(cb + lane) * kB16);
#pragma unroll
            for (int s2 = 0; s2 < kB16 / 2; ++s2) {
                const double2 v = src[s2];
                sO[w][lane][2 * s2] = v.x;
                sO[w][lane][2 * s2 + 1] = v.y;
            }
        }
        __syncthreads();
        if (act) {
            for (int c4 = 0; c4 < 64; c4 += 4) {
                const double4 xc = sP[w][c4 + lg];
                const double xca[3] = {xc.x, xc.y, xc.z};
                const double om = sO[w][c4 + lg][lc];
#pragma unroll
                for (int rb = 0; rb < 4; ++rb) {
                    const double a = kernel_eval_scaled<D>(xr[rb], xca, nil) * (xc.w * wr[rb]);
                    acc[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, om, acc[rb], 0, 0, 0);
                }
            }
        }
    }
    if (!act) return;
    double* Y = splits > 1 ? L.Ypart + ((size_t)blockIdx.y * L.T + t) * m * kB16 : L.Y + (size_t)t * m * kB16;
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
#pragma unroll
        for (int v = 0; v < 4; ++v) Y[(size_t)lc * m + q0 + 16 * rb + 4 * v + lg] = acc[rb][v];
}

// B[t][a*m + c] = sum_q K(x_q, x_c) Q[t][a*m + q] for 32 basis columns a0.. per launch row
template <int D>
__global__ void __launch_bounds__(256) k_dir_sampleT_mf(LrLevel L, double* __restrict__ Bout) {
    constexpr int NQ = 32;
    __shared__ double4 sP[4][64];
    __shared__ double sQ[4][64][NQ];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, lg = lane >> 4, lc = lane & 15;
    const int m = L.m;
    const int64_t c0g = ((int64_t)blockIdx.x * 4 + w) * 64;
    const bool inb = c0g < (int64_t)L.T * m;
    const int t = inb ? (int)(c0g / m) : 0, cw = (int)(c0g % m);
    const int k = inb ? L.st[t].k : 0;
    const int a0 = blockIdx.y * NQ;
    const bool act = inb && a0 < k;
    const int na = act ? min(NQ, k - a0) : 0;
    const TileParam tpt = L.tp[t];
    const uint32_t code = (uint32_t)tpt.code;
    const double nil = tpt.nil;
    const int64_t rowBase = tpt.pbase + (int64_t)compact16(code >> 1) * m;
    const int64_t colBase = tpt.pbase + (int64_t)compact16(code) * m;
    double xc[4][3], wc[4];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
        const double4 p = act ? reinterpret_cast<const double4*>(L.P)[colBase + cw + 16 * cb + lc]
                              : make_double4(0, 0, 0, 0);
        xc[cb][0] = p.x; xc[cb][1] = p.y; xc[cb][2] = p.z; wc[cb] = p.w;
    }
    const double* Q = L.Q + (size_t)t * m * L.capQ;
    v4d_t acc[4][2];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) { acc[cb][0] = v4d_t{0.0, 0.0, 0.0, 0.0}; acc[cb][1] = acc[cb][0]; }
    for (int qb = 0; qb < m; qb += 64) {
        __syncthreads();
        if (act) {
            sP[w][lane] = reinterpret_cast<const double4*>(L.P)[rowBase + qb + lane];
#pragma unroll
            for (int a = 0; a < NQ; ++a) sQ[w][lane][a] = a < na ? Q[(size_t)(a0 + a) * m + qb + lane] : 0.0;
        }
        __syncthreads();
        if (act) {
            for (int q4 = 0; q4 < 64; q4 += 4) {
                const double4 xr = sP[w][q4 + lg];
                const double xra[3] = {xr.x, xr.y, xr.z};
                const double b0 = sQ[w][q4 + lg][lc], b1 = sQ[w][q4 + lg][16 + lc];
#pragma unroll
                for (int cb = 0; cb < 4; ++cb) {
                    const double a = kernel_eval_scaled<D>(xra, xc[cb], nil) * (xr.w * wc[cb]);
                    acc[cb][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b0, acc[cb][0], 0, 0, 0);
                    acc[cb][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b1, acc[cb][1], 0, 0, 0);
                }
            }
        }
    }
    if (!act) return;
    double* B = Bout + (size_t)t * m * L.capQ;
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
        for (int ab = 0; ab < 2; ++ab)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int a = 16 * ab + lc;
                if (a < na) B[(size_t)(a0 + a) * m + cw + 16 * cb + 4 * v + lg] = acc[cb][ab][v];
            }
}

// Y = sum over the splits, in split order
__global__ void __launch_bounds__(256) k_dir_sum(LrLevel L, int splits) {
    const int t = blockIdx.y;
    if (!lr_active(L.st[t], L.tp[t].maxRank)) return;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t per = (int64_t)L.m * kB16;
    if (e >= per) return;
    double v = 0.0;
    for (int sp = 0; sp < splits; ++sp) v += L.Ypart[((size_t)sp * L.T + t) * per + e];
    L.Y[(size_t)t * per + e] = v;
}

// B[t][a*m + c] = sum_q K(x_q, x_c) Q[t][a*m + q], NQ basis columns per launch row
template <int D, int NQ>
__global__ void __launch_bounds__(256) k_dir_sampleT(LrLevel L, double* __restrict__ Bout) {
    __shared__ double4 sP[4][64];
    __shared__ double sQ[4][64][NQ];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int m = L.m;
    const int64_t gc = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool inb = gc < (int64_t)L.T * m;
    const int t = inb ? (int)(gc / m) : 0, c = (int)(gc % m);
    const int k = inb ? L.st[t].k : 0;
    const int a0 = blockIdx.y * NQ;
    const bool act = inb && a0 < k;
    const int na = act ? min(NQ, k - a0) : 0;
    const TileParam tpt = L.tp[t];
    const uint32_t code = (uint32_t)tpt.code;
    const double nil = tpt.nil;
    const int64_t rowBase = tpt.pbase + (int64_t)compact16(code >> 1) * m;
    const int64_t colBase = tpt.pbase + (int64_t)compact16(code) * m;
    const double4 xc = act ? reinterpret_cast<const double4*>(L.P)[colBase + c] : make_double4(0, 0, 0, 0);
    const double xca[3] = {xc.x, xc.y, xc.z};
    const double* Q = L.Q + (size_t)t * m * L.capQ;
    double acc[NQ];
#pragma unroll
    for (int a = 0; a < NQ; ++a) acc[a] = 0.0;
    for (int qb = 0; qb < m; qb += 64) {
        __syncthreads();
        if (act) {
            sP[w][lane] = reinterpret_cast<const double4*>(L.P)[rowBase + qb + lane];
#pragma unroll
            for (int a = 0; a < NQ; ++a) sQ[w][lane][a] = a < na ? Q[(size_t)(a0 + a) * m + qb + lane] : 0.0;
        }
        __syncthreads();
        if (act) {
            for (int qq = 0; qq < 64; ++qq) {
                const double4 xr = sP[w][qq];
                const double xra[3] = {xr.x, xr.y, xr.z};
                const double kv = kernel_eval_scaled<D>(xra, xca, nil) * xr.w;
#pragma unroll
                for (int a = 0; a < NQ; ++a) acc[a] = fma(kv, sQ[w][qq][a], acc[a]);
            }
        }
    }
    if (!act) return;
    double* B = Bout + (size_t)t * m * L.capQ;
#pragma unroll
    for (int a = 0; a < NQ; ++a)
        if (a < na) B[(size_t)(a0 + a) * m + c] = acc[a] * xc.w;
}

// W[t][s*k + a] = sum_i Q_t[a*m+i] B_t[s*m+i], a < k_t:
// mode 0: B = Y (16 columns, tile still iterating); mode 2: B = accepted block (na columns)
__global__ void __launch_bounds__(256) k_lr_dot(LrLevel L, int mode) {
    __shared__ double sA[16][65], sB[16][65];
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
    const double* A = L.Q + (size_t)t * m * L.capQ;
    const double* Bm = (mode == 0 ? L.Y : L.Ya) + (size_t)t * m * kB16;
    double* out = L.W + (size_t)t * L.capQ * kB16;
    const int ta = threadIdx.x & 15, ts = threadIdx.x >> 4;
    double acc = 0.0;
    for (int i0 = 0; i0 < m; i0 += 64) {
        for (int e = threadIdx.x; e < 16 * 64; e += 256) {
            int c = e >> 6, x = e & 63;
            int i = i0 + x;
            sA[c][x] = (a0 + c < K && i < m) ? A[(size_t)(a0 + c) * m + i] : 0.0;
            sB[c][x] = (c < C && i < m) ? Bm[(size_t)c * m + i] : 0.0;
        }
        __syncthreads();
        int lim = min(64, m - i0);
        for (int x = 0; x < lim; ++x) acc += sA[ta][x] * sB[ts][x];
        __syncthreads();
    }
    if (a0 + ta < K && ts < C) out[ts * K + a0 + ta] = acc;
}

// Y[s*m+i] -= sum_a Q[a*m+i] W[s*k+a]  (mode 0: Y, 16 cols; mode 2: Ya, na cols)
__global__ void __launch_bounds__(256) k_lr_update(LrLevel L, int mode) {
    __shared__ double sW[64 * kB16];
    const int t = blockIdx.y;
    const LrState& S = L.st[t];
    if (mode == 0 ? !lr_active(S, L.tp[t].maxRank) : S.na == 0) return;
    const int k = S.k;
    const int C = mode == 0 ? kB16 : S.na;
    if (k == 0 || C == 0) return;
    const int m = L.m;
    const int i = blockIdx.x * 256 + threadIdx.x;
    double* Y = (mode == 0 ? L.Y : L.Ya) + (size_t)t * m * kB16;
    const double* Q = L.Q + (size_t)t * m * L.capQ;
    const double* W = L.W + (size_t)t * L.capQ * kB16;
    double acc[kB16];
#pragma unroll
    for (int s = 0; s < kB16; ++s) acc[s] = (s < C && i < m) ? Y[(size_t)s * m + i] : 0.0;
    for (int a0 = 0; a0 < k; a0 += 64) {
        int na = min(64, k - a0);
        __syncthreads();
        for (int e = threadIdx.x; e < na * C; e += 256) {
            int s = e / na, a = e % na;
            sW[a * kB16 + s] = W[s * k + a0 + a];
        }
        __syncthreads();
        if (i < m) {
            for (int a = 0; a < na; ++a) {
                double q = Q[(size_t)(a0 + a) * m + i];
#pragma unroll
                for (int s = 0; s < kB16; ++s)
                    if (s < C) acc[s] -= q * sW[a * kB16 + s];
            }
        }
    }
    if (i < m)
#pragma unroll
        for (int s = 0; s < kB16; ++s)
            if (s < C) Y[(size_t)s * m + i] = acc[s];
}

__device__ inline double wave_sum64(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// sum over the block; every thread receives the identical value
__device__ inline double block_sum(double v, double* red) {
    v = wave_sum64(v);
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int q = 0; q < nw; ++q) s += red[q];
    return s;
}

// One workgroup per tile: column-by-column decisions on explicitly projected
// vectors (two modified Gram-Schmidt sweeps against the block's accepted
// columns, accept when the norm exceeds tau).  Threads own fixed rows, so
// only the reductions need the workgroup barrier.
__global__ void __launch_bounds__(1024) k_lr_decide(LrLevel L) {
    __shared__ double red[16];
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
                const double d = block_sum(p, red);
                for (int i = tid; i < m; i += nthr) y[i] -= qa[i] * d;
            }
        double p = 0.0;
        for (int i = tid; i < m; i += nthr) p += y[i] * y[i];
        const double nu = sqrt(block_sum(p, red));
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
    __shared__ double red[16];
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
            const double d = block_sum(p, red);
            for (int i = tid; i < m; i += nthr) q[i] -= qa[i] * d;
        }
        double p = 0.0;
        for (int i = tid; i < m; i += nthr) p += q[i] * q[i];
        const double nu = sqrt(block_sum(p, red));
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
    double* B = Bout + (size_t)t * m * L.capQ;
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

template <int D>
static void launch_sample_direct(Context& c, const LrLevel& L, int splits) {
    const int64_t rows = (int64_t)L.T * L.m;
    k_dir_sample_mf<D><<<dim3((unsigned)((rows + 255) / 256), (unsigned)splits), 256, 0, c.stream>>>(L, splits);
    HMX_LAUNCHED();
    if (splits > 1) {
        k_dir_sum<<<dim3((unsigned)(((int64_t)L.m * kB16 + 255) / 256), (unsigned)L.T), 256, 0, c.stream>>>(L, splits);
        HMX_LAUNCHED();
    }
}

template <int D>
static void launch_sampleT_direct(Context& c, const LrLevel& L, int kmax, double* B) {
    const int64_t cols = (int64_t)L.T * L.m;
    k_dir_sampleT_mf<D><<<dim3((unsigned)((cols + 255) / 256), (unsigned)((kmax + 31) / 32)), 256, 0, c.stream>>>(L, B);
    HMX_LAUNCHED();
}

// progress lines on stderr when HMX_LOG is set (long builds)
static bool log_on() {
    static const bool on = std::getenv("HMX_LOG") != nullptr;
    return on;
}

// Result of one batched ARA launch sequence: ranks, exclusive column
// offsets and the packed factors of its tiles, in tile order.
struct LrOut {
    std::vector<int32_t> ranks;
    std::vector<char> capped;
    std::vector<int64_t> offsets;
    DBuf<double> U, V;
    int64_t rankSum = 0;
    int caps = 0, kmax = 0, iters = 0;
};

// Batched ARA over T tiles of dimension m (bus = m/b leaves per side), each
// with its own threshold, rank cap, random stream and (direct mode) point
// base and correlation length.
static void compress_batch(Context& c, int araR, uint64_t seed, const LevelSource& src, int b, int m,
                           const std::vector<TileParam>& tp, const std::vector<int64_t>& first, LrOut& out) {
    const bool direct = src.mo == nullptr;
    const int bus = m / b;
    const int T = (int)tp.size();
    out.ranks.assign(T, 0);
    out.offsets.assign(T + 1, 0);
    if (T == 0) return;
    HMX_CHECK(!direct || m % 64 == 0, HMX_EUNSUPPORTED, "direct compression needs tile dimensions that are multiples of 64");
    int maxRank = 0;
    for (const TileParam& q : tp) maxRank = std::max(maxRank, (int)q.maxRank);
    DBuf<TileParam> d_tp(T);
    HMX_HIP(hipMemcpyAsync(d_tp.p, tp.data(), T * sizeof(TileParam), hipMemcpyHostToDevice, c.stream));
    DBuf<int64_t> d_first(direct ? 0 : T);
    if (!direct)
        HMX_HIP(hipMemcpyAsync(d_first.p, first.data(), T * sizeof(int64_t), hipMemcpyHostToDevice, c.stream));
    DBuf<LrState> st(T);
    int capQ = std::min(maxRank, 128);
    capQ = std::max(capQ, std::min(maxRank, kB16));
    DBuf<double> Q((size_t)T * m * capQ), W((size_t)T * capQ * kB16);
    DBuf<double> Y((size_t)T * m * kB16), Ya((size_t)T * m * kB16), Om((size_t)T * m * kB16);
    // direct mode: split the columns of small launches over several workgroups
    int splits = 1;
    if (direct) {
        const int64_t rows = (int64_t)T * m;
        while (splits * 2 <= m / 64 && rows * splits < ((int64_t)1 << 18)) splits *= 2;
    }
    DBuf<double> Ypart(splits > 1 ? (size_t)splits * T * m * kB16 : 0);
    DBuf<int> flags(2);
    LrLevel L;
    L.T = T; L.m = m; L.b = b; L.bus = bus; L.r = araR;
    L.seed = seed;
    L.tp = d_tp.p;
    L.first = d_first.p;
    L.ranksMO = direct ? nullptr : src.mo->ranks.p;
    L.offMO = direct ? nullptr : src.mo->offsets.p;
    L.UMO = direct ? nullptr : src.mo->U.p;
    L.VMO = direct ? nullptr : src.mo->V.p;
    L.st = st.p; L.Q = Q.p; L.Y = Y.p; L.Ya = Ya.p; L.Om = Om.p; L.W = W.p; L.capQ = capQ;
    L.P = src.P; L.Ypart = Ypart.p;
    k_lr_init<<<(T + 255) / 256, 256, 0, c.stream>>>(T, st.p);
    HMX_LAUNCHED();
    const unsigned rowBlocks = (unsigned)((m + 255) / 256);
    const unsigned wgThreads = m >= 4096 ? 1024u : 256u;
    int kmax = 0;
    const int maxIters = (2 * maxRank + araR) / kB16 + 2;
    for (int it = 0;; ++it) {
        HMX_CHECK(it <= maxIters, HMX_EDEVICE, "batched ARA did not converge (non-finite input?)");
        if (kmax + kB16 > capQ && capQ < maxRank) {  // grow the basis storage
            int nc = std::min(maxRank, std::max(2 * capQ, kmax + kB16));
            DBuf<double> Q2((size_t)T * m * nc), W2((size_t)T * nc * kB16);
            k_lr_regrow<<<dim3(std::min<unsigned>(1024, (unsigned)(((size_t)m * kmax + 255) / 256 + 1)), T), 256, 0,
                          c.stream>>>(T, m, capQ, nc, st.p, Q.p, Q2.p);
            HMX_LAUNCHED();
            HMX_HIP(hipStreamSynchronize(c.stream));
            Q = std::move(Q2);
            W = std::move(W2);
            capQ = nc;
            L.Q = Q.p; L.W = W.p; L.capQ = capQ;
        }
        k_lr_omega<<<dim3((m * kB16 + 255) / 256, T), 256, 0, c.stream>>>(L, it, direct ? 1 : 0);
        HMX_LAUNCHED();
        if (direct) {
            if (src.d == 1) launch_sample_direct<1>(c, L, splits);
            else if (src.d == 2) launch_sample_direct<2>(c, L, splits);
            else launch_sample_direct<3>(c, L, splits);
        } else if (b <= 32) {
            launch_sample_tlr<32>(c, L, bus, T);
        } else {
            launch_sample_tlr<64>(c, L, bus, T);
        }
        const unsigned kb = (unsigned)std::max(1, (kmax + 15) / 16);
        for (int pass = 0; pass < 2 && kmax > 0; ++pass) {
            k_lr_dot<<<dim3(kb, T), 256, 0, c.stream>>>(L, 0);
            HMX_LAUNCHED();
            k_lr_update<<<dim3(rowBlocks, T), 256, 0, c.stream>>>(L, 0);
            HMX_LAUNCHED();
        }
        k_lr_decide<<<T, wgThreads, 0, c.stream>>>(L);
        HMX_LAUNCHED();
        if (kmax > 0) {
            k_lr_dot<<<dim3(kb, T), 256, 0, c.stream>>>(L, 2);
            HMX_LAUNCHED();
            k_lr_update<<<dim3(rowBlocks, T), 256, 0, c.stream>>>(L, 2);
            HMX_LAUNCHED();
        }
        k_lr_renorm<<<T, wgThreads, 0, c.stream>>>(L);
        HMX_LAUNCHED();
        HMX_HIP(hipMemsetAsync(flags.p, 0, 2 * sizeof(int), c.stream));
        k_lr_commit<<<(T + 255) / 256, 256, 0, c.stream>>>(L, flags.p);
        HMX_LAUNCHED();
        int hf[2];
        HMX_HIP(hipMemcpyAsync(hf, flags.p, sizeof(hf), hipMemcpyDeviceToHost, c.stream));
        HMX_HIP(hipStreamSynchronize(c.stream));
        kmax = hf[1];
        out.iters = it + 1;
        if (hf[0] == 0) break;
    }
    Ypart.release();
    Y.release();
    Ya.release();
    Om.release();
    // B = A^T Q
    DBuf<double> Bf((size_t)T * m * capQ);
    if (kmax > 0) {
        if (direct) {
            if (src.d == 1) launch_sampleT_direct<1>(c, L, kmax, Bf.p);
            else if (src.d == 2) launch_sampleT_direct<2>(c, L, kmax, Bf.p);
            else launch_sampleT_direct<3>(c, L, kmax, Bf.p);
        } else if (b <= 32) {
            launch_sampleT_tlr<32>(c, L, kmax, bus, T, Bf.p);
        } else {
            launch_sampleT_tlr<64>(c, L, kmax, bus, T, Bf.p);
        }
    }
    std::vector<LrState> hst(T);
    HMX_HIP(hipMemcpyAsync(hst.data(), st.p, T * sizeof(LrState), hipMemcpyDeviceToHost, c.stream));
    HMX_HIP(hipStreamSynchronize(c.stream));
    int64_t acc = 0;
    int caps = 0;
    out.capped.assign(T, 0);
    for (int t = 0; t < T; ++t) {
        out.ranks[t] = hst[t].k;
        out.offsets[t] = acc;
        acc += hst[t].k;
        caps += hst[t].capped;
        out.capped[t] = (char)hst[t].capped;
    }
    out.offsets[T] = acc;
    out.rankSum = acc;
    out.caps = caps;
    out.kmax = kmax;
    out.U.alloc((size_t)acc * m);
    out.V.alloc((size_t)acc * m);
    if (acc > 0) {
        DBuf<int64_t> d_off(T + 1);
        HMX_HIP(hipMemcpyAsync(d_off.p, out.offsets.data(), (T + 1) * sizeof(int64_t), hipMemcpyHostToDevice,
                               c.stream));
        dim3 g(std::min<unsigned>(1024, (unsigned)(((size_t)m * kmax + 255) / 256 + 1)), T);
        k_lr_pack<<<g, 256, 0, c.stream>>>(T, m, capQ, st.p, d_off.p, Q.p, out.U.p);
        HMX_LAUNCHED();
        k_lr_pack<<<g, 256, 0, c.stream>>>(T, m, capQ, st.p, d_off.p, Bf.p, out.V.p);
        HMX_LAUNCHED();
        HMX_HIP(hipStreamSynchronize(c.stream));
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

static void compress_level(Context& c, const hmx_config& cfg, const LevelSource& src, int b,
                           const std::vector<int32_t>& tiles, int level, int numLevels, float tol, int maxRank,
                           HLevel& out) {
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
    std::vector<TileParam> tp(T);
    std::vector<int64_t> first(T);
    for (int t = 0; t < T; ++t) {
        tp[t] = tile_param((double)tol, src.ell, 0, level, t, maxRank, tiles[t]);
        first[t] = (int64_t)tiles[t] * bus * bus;
    }
    LrOut r;
    compress_batch(c, cfg.araR, cfg.seed, src, b, m, tp, first, r);
    out.h_ranks = std::move(r.ranks);
    out.h_offsets = std::move(r.offsets);
    out.rankSum = r.rankSum;
    out.capHits = r.caps;
    out.U = std::move(r.U);
    out.V = std::move(r.V);
    if (log_on()) {
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tstart).count();
        std::fprintf(stderr, "[hmx] level %d (%s): %d tiles m=%d rank sum %lld max %d, %d ARA blocks, %.1f ms\n", level,
                     direct ? "direct" : "TLR sum", T, m, (long long)out.rankSum, r.kmax, r.iters, ms);
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
    float tol = cfg.lowestLevelTolerance;
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
        int maxRank = cfg.maxRankPolicy == HMX_MAXRANK_REF ? (int)maxRanks[L - level - 2] : m;
        maxRank = std::min(maxRank, m);
        compress_level(c, cfg, src, b, s.tileIndices[level - 1], level, L, tol, maxRank, lv);
    }
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
    h.diag.alloc((size_t)t.numSegments * b * b);
    build_dense_diagonal(c, cfg, t, d_points, h.diag.p);
    {
        HLevel& lv = h.levels[L - 2];
        lv.level = L - 1;
        lv.maxRank = cfg.maxRankPolicy == HMX_MAXRANK_REF ? b / 2 : b;
        lv.tol = cfg.lowestLevelTolerance;
        const auto t0 = std::chrono::steady_clock::now();
        compress_leaf_tiles(c, cfg, t, d_points, s.tileIndices[L - 2], lv);
        if (log_on()) {
            const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            std::fprintf(stderr, "[hmx] level %d (leaf tiles): %d tiles rank sum %lld, %.1f ms\n", L - 1, lv.numTiles,
                         (long long)lv.rankSum, ms);
            std::fflush(stderr);
        }
    }
    DBuf<double> P((size_t)t.numSegments * b * 4);
    const int64_t np = (int64_t)t.numSegments * b;
    k_pad_points<<<(unsigned)((np + 255) / 256), 256, 0, c.stream>>>(t.n, t.d, b, t.numSegments, d_points, t.perm.p,
                                                                      t.offsets.p, P.p);
    HMX_LAUNCHED();
    LevelSource src;
    src.P = P.p;
    src.ell = cfg.correlationLength > 0 ? cfg.correlationLength : default_ell(t.d);
    src.d = (int)t.d;
    compress_upper_levels(c, cfg, s, src, b, h);
    HMX_HIP(hipStreamSynchronize(c.stream));
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
    DBuf<double> P((size_t)ptot * 4);
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
        h.diag.alloc((size_t)t.numSegments * b * b);
        build_dense_diagonal(c, *cfgs[i], t, pts[i], h.diag.p);
        HLevel& lv = h.levels[L - 2];
        lv.level = L - 1;
        lv.maxRank = cfgs[i]->maxRankPolicy == HMX_MAXRANK_REF ? b / 2 : b;
        lv.tol = cfgs[i]->lowestLevelTolerance;
        compress_leaf_tiles(c, *cfgs[i], t, pts[i], s.tileIndices[L - 2], lv);
        const int64_t np = (int64_t)t.numSegments * b;
        k_pad_points<<<(unsigned)((np + 255) / 256), 256, 0, c.stream>>>(t.n, t.d, b, t.numSegments, pts[i],
                                                                          t.perm.p, t.offsets.p, P.p + pbase[i] * 4);
        HMX_LAUNCHED();
    }
    // upper levels: the reference's schedule per instance, grouped across instances
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
    for (auto& grp : groups) {
        const int m = grp.first.first, d = grp.first.second;
        const auto tstart = std::chrono::steady_clock::now();
        std::vector<TileParam> tp;
        for (const Job& j : grp.second) {
            const hmx_config& cfg = *cfgs[j.inst];
            const HLevel& lv = hs[j.inst]->levels[j.level - 1];
            const double ell = cfg.correlationLength > 0 ? cfg.correlationLength : default_ell(trees[j.inst]->d);
            const auto& tiles = ss[j.inst]->tileIndices[j.level - 1];
            for (int t = 0; t < (int)tiles.size(); ++t)
                tp.push_back(tile_param((double)lv.tol, ell, pbase[j.inst], j.level, t, lv.maxRank, tiles[t]));
        }
        LevelSource src;
        src.P = P.p;
        src.d = d;
        LrOut r;
        compress_batch(c, cfgs[0]->araR, cfgs[0]->seed, src, b, m, tp, {}, r);
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
            lv.U.alloc((size_t)acc * m);
            lv.V.alloc((size_t)acc * m);
            if (acc > 0) {
                const size_t o = (size_t)r.offsets[t0] * m;
                HMX_HIP(hipMemcpyAsync(lv.U.p, r.U.p + o, (size_t)acc * m * sizeof(double), hipMemcpyDeviceToDevice,
                                       c.stream));
                HMX_HIP(hipMemcpyAsync(lv.V.p, r.V.p + o, (size_t)acc * m * sizeof(double), hipMemcpyDeviceToDevice,
                                       c.stream));
            }
            t0 += T;
        }
        HMX_HIP(hipStreamSynchronize(c.stream));
        if (log_on()) {
            const double ms =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tstart).count();
            std::fprintf(stderr, "[hmx] batch group m=%d d=%d: %zu tiles from %zu levels, %d ARA blocks, %.1f ms\n", m,
                         d, tp.size(), grp.second.size(), r.iters, ms);
            std::fflush(stderr);
        }
    }
    for (HMatrix* h : hs) finish_accounting(*h);
}

}  // namespace hmx
