// Gradients of a loss L(Y), Y = H X, with respect to the low-rank factors
// (no reference counterpart; the backward of HMatrixVecMult):
//   dU_t = dY[rows_t] (V_t^T X[cols_t])^T      (m x r)
//   dV_t = X[cols_t] (U_t^T dY[rows_t])^T      (n x r)
// Z_t = V_t^T X and G_t = U_t^T dY come from phase 1 of the forward and of the
// transposed matvec plans (same tile/column order), then one kernel forms the
// per-tile outer products, one thread per factor row.
#include <algorithm>

#include "hmx_device.h"
#include "hmx_internal.h"

namespace hmx {

struct GradTile {
    double* dU;        // gradient pool position of the tile (m x r column-major)
    double* dV;
    int32_t m, r, rowStart, colStart;
    int32_t zbase;     // first global column in the plans' z
    int32_t pad;
};

// blockIdx.y = tile, threads over rows; K right-hand sides in this pass
__global__ void __launch_bounds__(256) k_grad_outer(const GradTile* __restrict__ tiles, int K, int ldx,
                                                    const double* __restrict__ X, const double* __restrict__ dY,
                                                    const double* __restrict__ Z, const double* __restrict__ G,
                                                    int accumulate) {
    const GradTile t = tiles[blockIdx.y];
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= t.m) return;
    for (int a = 0; a < t.r; ++a) {
        const double* z = Z + (size_t)(t.zbase + a) * K;
        const double* g = G + (size_t)(t.zbase + a) * K;
        double su = 0.0, sv = 0.0;
        for (int s = 0; s < K; ++s) {
            su += dY[(size_t)s * ldx + t.rowStart + i] * z[s];
            sv += X[(size_t)s * ldx + t.colStart + i] * g[s];
        }
        const size_t o = (size_t)a * t.m + i;
        if (accumulate) { t.dU[o] += su; t.dV[o] += sv; }
        else { t.dU[o] = su; t.dV[o] = sv; }
    }
}

void hmatvec_grad_factors(Context& c, HMatrix& h, uint32_t k, const double* X, const double* dY,
                          double* const* dU, double* const* dV) {
    HMX_CHECK(k >= 1, HMX_EINVAL, "vectorWidth must be >= 1");
    HMX_CHECK(X && dY && dU && dV, HMX_EINVAL, "null pointer");
    std::vector<GradTile> tiles;
    int64_t g = 0;
    for (int l = 1; l < h.numLevels; ++l) {
        const HLevel& lv = h.levels[l - 1];
        const auto& ti = h.structure->tileIndices[l - 1];
        for (int t = 0; t < lv.numTiles; ++t) {
            const int r = lv.h_ranks[t];
            if (!r) continue;
            HMX_CHECK(dU[l - 1] && dV[l - 1], HMX_EINVAL, "null gradient buffer for a non-empty level");
            const uint32_t u = compact16((uint32_t)ti[t]), v = compact16((uint32_t)ti[t] >> 1);
            tiles.push_back(GradTile{dU[l - 1] + (size_t)lv.h_offsets[t] * lv.m, dV[l - 1] + (size_t)lv.h_offsets[t] * lv.m,
                                     lv.m, r, (int32_t)(v * lv.m), (int32_t)(u * lv.m), (int32_t)g, 0});
            g += r;
        }
    }
    if (tiles.empty()) return;
    DBuf<GradTile> d_tiles(tiles.size());
    HMX_HIP(hipMemcpyAsync(d_tiles.p, tiles.data(), tiles.size() * sizeof(GradTile), hipMemcpyHostToDevice, c.stream));
    int maxm = 0;
    for (auto& t : tiles) maxm = std::max(maxm, t.m);
    // Z of X (forward plan) and G of dY (transposed plan), in passes of 16, 8,
    // 4 or 1 right-hand sides.  Tile rows are positions of the padded leaf
    // space: with partial leaves the outer products read the gathered copies
    // the two matvecs leave in the plans (zero on padding rows, so the
    // padding rows of dU and dV come out zero as the factors' own are).
    // the gradient plans run every tile through z, in the tile order above
    if (!h.gplan.built) build_matvec_plan(c, h, false, h.gplan, false);
    if (!h.gplanT.built) build_matvec_plan(c, h, true, h.gplanT, false);
    DBuf<double> scratch((size_t)h.n * pass_width(k));
    for (uint32_t s0 = 0; s0 < k;) {
        const uint32_t K = pass_width(k - s0);
        const double* Xs = X + (size_t)s0 * h.n;
        const double* Ws = dY + (size_t)s0 * h.n;
        run_matvec_plan(c, h.gplan, (int)h.b, K, Xs, (int64_t)h.n, scratch.p, nullptr);   // Z in gplan.z
        run_matvec_plan(c, h.gplanT, (int)h.b, K, Ws, (int64_t)h.n, scratch.p, nullptr);  // G in gplanT.z
        const bool pad = h.gplan.gather;
        const double* Xg = pad ? h.gplan.Xp.p : Xs;
        const double* Wg = pad ? h.gplanT.Xp.p : Ws;
        const int ld = pad ? (int)h.gplan.npad : (int)h.n;
        dim3 grid((unsigned)((maxm + 255) / 256), (unsigned)tiles.size());
        k_grad_outer<<<grid, 256, 0, c.stream>>>(d_tiles.p, (int)K, ld, Xg, Wg, h.gplan.z.p, h.gplanT.z.p,
                                                 s0 > 0 ? 1 : 0);
        HMX_LAUNCHED();
        s0 += K;
    }
    HMX_HIP(hipStreamSynchronize(c.stream));
}

// overwrite one level's factors (device buffers in the level's layout)
void hmatrix_set_level(Context& c, HMatrix& h, int level, const double* U, const double* V) {
    HMX_CHECK(level >= 1 && level < h.numLevels, HMX_EINVAL, "level out of range");
    HLevel& lv = h.levels[level - 1];
    if (lv.U.n) {
        HMX_CHECK(U && V, HMX_EINVAL, "null factor buffer");
        HMX_HIP(hipMemcpyAsync(lv.U.p, U, lv.U.bytes(), hipMemcpyDeviceToDevice, c.stream));
        HMX_HIP(hipMemcpyAsync(lv.V.p, V, lv.V.bytes(), hipMemcpyDeviceToDevice, c.stream));
    }
    HMX_HIP(hipStreamSynchronize(c.stream));
    // converted copies used by the matrix-core paths, and the streaming
    // plans' dense near-field blocks, are stale now
    ++h.factorGen;
    h.mf64 = MfPlan();
    h.mfMixed = MfPlan();
    h.plan = MatvecPlan();
    h.planT = MatvecPlan();
}

}  // namespace hmx
