// The H-matrix structure pass on the GPU (src/HMatrixStructure.cu:9-90),
// level-synchronous instead of recursive.  The recursion visits the four
// children of a node pair in a fixed order and appends a pair to its level's
// tile list when it reaches it; so a level's list is the level's pairs in the
// order of their parents in the level above, four children per parent in
// that fixed order.  One pass per level classifies the children of every
// active pair (none / tile / active again), and two stable compactions (an
// exclusive scan each) give the tile list and the next level's active pairs
// in exactly that order.  The admissibility takes the host pass's
// operations in the same order with nothing contracted, on the KD-tree's
// device boxes, so the lists are the host pass's bit for bit.
#include "hmx_device.h"
#include "hmx_internal.h"

namespace hmx {

// admissible(level, u, v) of the host pass: BBoxCenterDistance /
// BBoxDiameter (src/admissibilityCondition.cu:7-52)
__device__ bool struct_admissible(int adm, float eta, uint32_t d, const double* __restrict__ bbmin,
                                  const double* __restrict__ bbmax, int level, uint32_t u, uint32_t v) {
#pragma clang fp contract(off)
    if (u == v) return false;
    if (adm == HMX_WEAK_ADMISSIBILITY) return true;
    const size_t base = ((size_t)1 << level) - 1;
    const double* mxu = bbmax + (base + u) * d;
    const double* mnu = bbmin + (base + u) * d;
    const double* mxv = bbmax + (base + v) * d;
    const double* mnv = bbmin + (base + v) * d;
    double distance = 0, du = 0, dv = 0;
    for (uint32_t i = 0; i < d; ++i) {
        const double c = 0.5 * ((mxu[i] + mnu[i]) - (mxv[i] + mnv[i]));
        distance += c * c;
    }
    distance = sqrt(distance);
    for (uint32_t i = 0; i < d; ++i) {
        const double t = mxu[i] - mnu[i];
        du += t * t;
    }
    du = sqrt(du);
    for (uint32_t i = 0; i < d; ++i) {
        const double t = mxv[i] - mnv[i];
        dv += t * t;
    }
    dv = sqrt(dv);
    if (du == 0 || dv == 0) return false;
    if (adm == HMX_BOX_CENTER_ADMISSIBILITY) return 0.5 * (du + dv) <= (double)eta * distance;
    return (du > dv ? du : dv) <= (double)eta * distance;
}

// child c of active pair p (c = 0..3 in the recursion's order: (2u, 2v),
// (2u+1, 2v), (2u, 2v+1), (2u+1, 2v+1)) at level `level`: a tile (leaf level
// off the diagonal, or admissible), active (to be split further) or nothing
// (a diagonal pair at the leaf level)
__global__ void __launch_bounds__(256) k_struct_children(int64_t np, const uint32_t* __restrict__ pu,
                                                         const uint32_t* __restrict__ pv, int level, int leaf,
                                                         int adm, float eta, uint32_t d,
                                                         const double* __restrict__ bbmin,
                                                         const double* __restrict__ bbmax,
                                                         int32_t* __restrict__ isTile, int32_t* __restrict__ isActive) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= 4 * np) return;
    const int64_t p = e >> 2;
    const int c = (int)(e & 3);
    const uint32_t u = 2 * pu[p] + (c & 1), v = 2 * pv[p] + (c >> 1);
    int t = 0, a = 0;
    if (!(u == v && leaf)) {
        if (leaf || struct_admissible(adm, eta, d, bbmin, bbmax, level, u, v)) t = 1;
        else a = 1;
    }
    isTile[e] = t;
    isActive[e] = a;
}

__global__ void __launch_bounds__(256) k_struct_scatter(int64_t np, const uint32_t* __restrict__ pu,
                                                        const uint32_t* __restrict__ pv,
                                                        const int64_t* __restrict__ tpos,
                                                        const int64_t* __restrict__ apos, int32_t* __restrict__ tiles,
                                                        uint32_t* __restrict__ nu, uint32_t* __restrict__ nv) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= 4 * np) return;
    const int64_t p = e >> 2;
    const int c = (int)(e & 3);
    const uint32_t u = 2 * pu[p] + (c & 1), v = 2 * pv[p] + (c >> 1);
    if (tpos[e + 1] > tpos[e]) tiles[tpos[e]] = (int32_t)morton_xy(u, v);  // CMIndextoMOIndex
    if (apos[e + 1] > apos[e]) {
        nu[apos[e]] = u;
        nv[apos[e]] = v;
    }
}

void build_structure_device(Context& c, Structure& s, uint32_t n, uint32_t d, uint32_t b, int adm, float eta,
                            const double* d_bbmin, const double* d_bbmax) {
    structure_shape(s, n, b, adm);
    const int maxDepth = s.numLevels - 1;
    if (maxDepth >= 1) {
        DBuf<uint32_t> pu(1), pv(1);
        const uint32_t zero = 0;
        HMX_HIP(hipMemcpyAsync(pu.p, &zero, 4, hipMemcpyHostToDevice, c.stream));
        HMX_HIP(hipMemcpyAsync(pv.p, &zero, 4, hipMemcpyHostToDevice, c.stream));
        int64_t np = 1;  // the root pair (0, 0): never a tile, always split
        for (int level = 1; level <= maxDepth && np > 0; ++level) {
            const int64_t nc = 4 * np;
            DBuf<int32_t> isTile(nc), isActive(nc);
            DBuf<int64_t> tpos(nc + 1), apos(nc + 1);
            const unsigned g = (unsigned)((nc + 255) / 256);
            k_struct_children<<<g, 256, 0, c.stream>>>(np, pu.p, pv.p, level, level == maxDepth ? 1 : 0, adm, eta, d,
                                                        d_bbmin, d_bbmax, isTile.p, isActive.p);
            HMX_LAUNCHED();
            exclusive_scan_i32_i64(c, isTile.p, tpos.p, (size_t)nc);
            exclusive_scan_i32_i64(c, isActive.p, apos.p, (size_t)nc);
            int64_t nt = 0, na = 0;
            HMX_HIP(hipMemcpyAsync(&nt, tpos.p + nc, 8, hipMemcpyDeviceToHost, c.stream));
            HMX_HIP(hipMemcpyAsync(&na, apos.p + nc, 8, hipMemcpyDeviceToHost, c.stream));
            HMX_HIP(hipStreamSynchronize(c.stream));
            HMX_CHECK(nt < ((int64_t)1 << 31), HMX_EUNSUPPORTED, "structure: more than 2^31 tiles in one level");
            DBuf<int32_t> tiles(std::max<int64_t>(1, nt));
            DBuf<uint32_t> nu(std::max<int64_t>(1, na)), nv(std::max<int64_t>(1, na));
            k_struct_scatter<<<g, 256, 0, c.stream>>>(np, pu.p, pv.p, tpos.p, apos.p, tiles.p, nu.p, nv.p);
            HMX_LAUNCHED();
            std::vector<int32_t>& out = s.tileIndices[level - 1];
            out.resize((size_t)nt);
            if (nt) HMX_HIP(hipMemcpyAsync(out.data(), tiles.p, (size_t)nt * 4, hipMemcpyDeviceToHost, c.stream));
            HMX_HIP(hipStreamSynchronize(c.stream));
            pu = std::move(nu);
            pv = std::move(nv);
            np = na;
        }
    }
    structure_finish(s);
}

}  // namespace hmx
