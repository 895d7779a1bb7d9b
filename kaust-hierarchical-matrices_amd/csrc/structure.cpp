// H-matrix structure (src/HMatrixStructure.cu:9-160), admissibility
// conditions (src/admissibilityCondition.cu:7-68) and per-level JDS layout
// (src/JDS.cuh:7-15).  Host code: the pass visits O(#tiles) node pairs and
// reads the KD-tree boxes that the build copies back once.
#include <cmath>
#include <cstring>

#include "hmx_internal.h"

namespace hmx {

namespace {

inline uint32_t spread(uint32_t x) {
    x = (x | (x << 8)) & 0x00FF00FFu;
    x = (x | (x << 4)) & 0x0F0F0F0Fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}
inline uint32_t compact(uint32_t x) {
    x &= 0x55555555u;
    x = (x | (x >> 1)) & 0x33333333u;
    x = (x | (x >> 2)) & 0x0F0F0F0Fu;
    x = (x | (x >> 4)) & 0x00FF00FFu;
    x = (x | (x >> 8)) & 0x0000FFFFu;
    return x;
}

struct Walker {
    uint32_t d;
    int maxDepth;
    int adm;
    float eta;
    const double *bbmin, *bbmax;
    Structure* s;

    // BBoxCenterDistance / BBoxDiameter: the reference's operation order
    double center_distance(const double* mxu, const double* mnu, const double* mxv, const double* mnv) const {
        double distance = 0;
        for (uint32_t i = 0; i < d; ++i) {
            double c = 0.5 * ((mxu[i] + mnu[i]) - (mxv[i] + mnv[i]));
            distance += c * c;
        }
        return std::sqrt(distance);
    }
    double diameter(const double* mx, const double* mn) const {
        double diam = 0;
        for (uint32_t i = 0; i < d; ++i) {
            double t = mx[i] - mn[i];
            diam += t * t;
        }
        return std::sqrt(diam);
    }
    bool admissible(int level, uint32_t u, uint32_t v) const {
        if (u == v) return false;
        if (adm == HMX_WEAK_ADMISSIBILITY) return true;
        size_t base = ((size_t)1 << level) - 1;
        const double* mxu = bbmax + (base + u) * d;
        const double* mnu = bbmin + (base + u) * d;
        const double* mxv = bbmax + (base + v) * d;
        const double* mnv = bbmin + (base + v) * d;
        double distance = center_distance(mxu, mnu, mxv, mnv);
        double du = diameter(mxu, mnu), dv = diameter(mxv, mnv);
        if (du == 0 || dv == 0) return false;
        if (adm == HMX_BOX_CENTER_ADMISSIBILITY) return 0.5 * (du + dv) <= eta * distance;
        return (du > dv ? du : dv) <= eta * distance;
    }
    void rec(int level, uint32_t u, uint32_t v) {
        bool diag = (u == v), leaf = (level == maxDepth);
        if (diag && leaf) return;
        if (leaf || admissible(level, u, v)) {
            uint32_t numRows = 1u << level;
            uint32_t cm = u * numRows + v;  // CMIndextoMOIndex(numRows, u*numRows+v)
            uint32_t mo = spread(cm / numRows) | (spread(cm % numRows) << 1);
            s->tileIndices[level - 1].push_back((int32_t)mo);
            return;
        }
        rec(level + 1, 2 * u, 2 * v);
        rec(level + 1, 2 * u + 1, 2 * v);
        rec(level + 1, 2 * u, 2 * v + 1);
        rec(level + 1, 2 * u + 1, 2 * v + 1);
    }
};

}  // namespace

void build_jds(Structure::JDS& j, int level, const std::vector<int32_t>& tiles) {
    const uint32_t numRows = 1u << level;
    std::vector<int32_t> cnt(numRows, 0), start(numRows + 1, 0), fill(numRows, 0), rowTiles(tiles.size());
    for (int32_t t : tiles) cnt[compact((uint32_t)t >> 1)]++;
    for (uint32_t r = 0; r < numRows; ++r) start[r + 1] = start[r] + cnt[r];
    for (size_t t = 0; t < tiles.size(); ++t) {
        uint32_t r = compact((uint32_t)tiles[t] >> 1);
        rowTiles[start[r] + fill[r]++] = (int32_t)t;
    }
    int32_t maxc = 0;
    for (uint32_t r = 0; r < numRows; ++r) maxc = cnt[r] > maxc ? cnt[r] : maxc;
    j.rowIndices.clear();
    j.rowIndices.reserve(numRows);
    for (int32_t c = maxc; c >= 0; --c)
        for (uint32_t r = 0; r < numRows; ++r)
            if (cnt[r] == c) j.rowIndices.push_back((int32_t)r);
    j.columnIndices.assign(tiles.size(), 0);
    j.tilePos.assign(tiles.size(), 0);
    j.iterPtr.assign(maxc + 1, 0);
    int32_t pos = 0;
    for (int32_t dg = 0; dg < maxc; ++dg) {
        j.iterPtr[dg] = pos;
        for (uint32_t i = 0; i < numRows; ++i) {
            int32_t r = j.rowIndices[i];
            if (cnt[r] <= dg) break;
            int32_t t = rowTiles[start[r] + dg];
            j.columnIndices[pos] = (int32_t)compact((uint32_t)tiles[t]);
            j.tilePos[pos] = t;
            ++pos;
        }
    }
    j.iterPtr[maxc] = pos;
}

namespace {

}  // namespace

// argument checks, depth and empty tile lists (constructHMatrixStructure,
// src/HMatrixStructure.cu:92-105); shared by the host and the device pass
void structure_shape(Structure& s, uint32_t n, uint32_t b, int adm) {
    HMX_CHECK(b >= 1 && n >= 1, HMX_EINVAL, "structure: n and bucketSize must be positive");
    HMX_CHECK(adm >= 0 && adm <= 2, HMX_EINVAL, "structure: unknown admissibility condition");
    uint32_t n2 = upper_pow2(n);
    HMX_CHECK(n2 / b >= 1 && ((n2 / b) & (n2 / b - 1)) == 0, HMX_EINVAL,
              "structure: upperPowerOfTwo(n)/bucketSize must be a power of two");
    s.n = n;
    s.b = b;
    s.numLevels = ctz32(n2 / b) + 1;
    s.numTiles.assign(s.numLevels, 0);
    s.tileIndices.assign(s.numLevels, {});
}

// tile counts and the per-level JDS once the tile lists are filled
void structure_finish(Structure& s) {
    for (int l = 0; l < s.numLevels; ++l) s.numTiles[l] = (int32_t)s.tileIndices[l].size();
    s.jds.assign(s.numLevels > 1 ? s.numLevels - 1 : 0, {});
    for (int l = 1; l < s.numLevels; ++l) build_jds(s.jds[l - 1], l, s.tileIndices[l - 1]);
}

void build_structure(Structure& s, uint32_t n, uint32_t d, uint32_t b, int adm, float eta, const double* bbmin,
                     const double* bbmax) {
    structure_shape(s, n, b, adm);
    if (s.numLevels > 1) {
        Walker w{d, s.numLevels - 1, adm, eta, bbmin, bbmax, &s};
        w.rec(0, 0, 0);
    }
    structure_finish(s);
}

}  // namespace hmx
