// Binary H-matrix files: an assembled H-matrix written once and loaded back
// for matvecs without the O(N^2) build (the reference can only print text
// dumps of its intermediates, src/helperFunctions.cu:76-169).
//
// Layout (little-endian): "HMXH", version 1, n, d, b, S (uint32),
// numLevels (int32), leaf offsets int32[S+1]; per level l = 1..numLevels-1 the
// structure's tile list (count, int32 Morton codes); dense diagonal blocks
// double[S*b*b]; per level: level, numTiles, m, maxRank, capHits (int32),
// tolerance (float), rank sum (int64), ranks int32[numTiles], U and V
// double[rankSum*m] in the library's layout.
#include <cstdio>
#include <cstring>
#include <memory>

#include "hmx_internal.h"

namespace hmx {

void build_jds(Structure::JDS& j, int level, const std::vector<int32_t>& tiles);

namespace {

constexpr uint32_t kMagic = 0x48584D48u;  // "HMXH"
constexpr uint32_t kVersion = 1;

struct File {
    FILE* f = nullptr;
    ~File() {
        if (f) std::fclose(f);
    }
};

void wr(File& F, const void* p, size_t bytes) {
    if (bytes && std::fwrite(p, 1, bytes, F.f) != bytes) throw Error(HMX_EDEVICE, "write failed");
}
void rd(File& F, void* p, size_t bytes) {
    if (bytes && std::fread(p, 1, bytes, F.f) != bytes) throw Error(HMX_EINVAL, "truncated H-matrix file");
}
template <class T>
void wv(File& F, const T& v) { wr(F, &v, sizeof(T)); }
template <class T>
T rv(File& F) {
    T v;
    rd(F, &v, sizeof(T));
    return v;
}

// device buffer <-> file through a bounded host staging buffer
void dev_write(Context& c, File& F, const double* d, size_t n) {
    const size_t chunk = (size_t)1 << 24;
    std::vector<double> h(std::min(n, chunk));
    for (size_t o = 0; o < n; o += chunk) {
        const size_t m = std::min(chunk, n - o);
        HMX_HIP(hipMemcpyAsync(h.data(), d + o, m * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        HMX_HIP(hipStreamSynchronize(c.stream));
        wr(F, h.data(), m * sizeof(double));
    }
}
void dev_read(Context& c, File& F, double* d, size_t n) {
    const size_t chunk = (size_t)1 << 24;
    std::vector<double> h(std::min(n, chunk));
    for (size_t o = 0; o < n; o += chunk) {
        const size_t m = std::min(chunk, n - o);
        rd(F, h.data(), m * sizeof(double));
        HMX_HIP(hipMemcpyAsync(d + o, h.data(), m * sizeof(double), hipMemcpyHostToDevice, c.stream));
        HMX_HIP(hipStreamSynchronize(c.stream));
    }
}

}  // namespace

void hmatrix_save(Context& c, const HMatrix& h, const char* path) {
    HMX_CHECK(path != nullptr, HMX_EINVAL, "null path");
    File F;
    F.f = std::fopen(path, "wb");
    HMX_CHECK(F.f != nullptr, HMX_EINVAL, std::string("cannot open for writing: ") + path);
    wv(F, kMagic);
    wv(F, kVersion);
    wv(F, h.n); wv(F, h.d); wv(F, h.b); wv(F, h.S);
    wv(F, h.numLevels);
    HMX_CHECK(h.leafOffsets.size() == (size_t)h.S + 1, HMX_ESTATE, "H-matrix without leaf offsets");
    wr(F, h.leafOffsets.data(), h.leafOffsets.size() * sizeof(int32_t));
    for (int l = 1; l < h.numLevels; ++l) {
        const auto& t = h.structure->tileIndices[l - 1];
        wv(F, (int32_t)t.size());
        wr(F, t.data(), t.size() * sizeof(int32_t));
    }
    dev_write(c, F, h.diag.p, h.diag.n);
    for (int l = 1; l < h.numLevels; ++l) {
        const HLevel& lv = h.levels[l - 1];
        wv(F, (int32_t)l); wv(F, lv.numTiles); wv(F, lv.m); wv(F, lv.maxRank); wv(F, lv.capHits);
        wv(F, lv.tol);
        wv(F, lv.rankSum);
        wr(F, lv.h_ranks.data(), lv.h_ranks.size() * sizeof(int32_t));
        dev_write(c, F, lv.U.p, (size_t)lv.rankSum * lv.m);
        dev_write(c, F, lv.V.p, (size_t)lv.rankSum * lv.m);
    }
    HMX_CHECK(std::fflush(F.f) == 0, HMX_EDEVICE, "write failed");
}

void hmatrix_load(Context& c, const char* path, HMatrix& h) {
    HMX_CHECK(path != nullptr, HMX_EINVAL, "null path");
    File F;
    F.f = std::fopen(path, "rb");
    HMX_CHECK(F.f != nullptr, HMX_EINVAL, std::string("cannot open: ") + path);
    HMX_CHECK(rv<uint32_t>(F) == kMagic, HMX_EINVAL, "not an H-matrix file");
    HMX_CHECK(rv<uint32_t>(F) == kVersion, HMX_EUNSUPPORTED, "unsupported H-matrix file version");
    h.n = rv<uint32_t>(F); h.d = rv<uint32_t>(F); h.b = rv<uint32_t>(F); h.S = rv<uint32_t>(F);
    h.numLevels = rv<int32_t>(F);
    HMX_CHECK(h.b >= 1 && h.b <= 64 && h.S >= 1 && (h.S & (h.S - 1)) == 0 && h.numLevels >= 2 &&
                  (uint32_t)(1u << (h.numLevels - 1)) == h.S && h.n <= h.S * h.b,
              HMX_EINVAL, "inconsistent H-matrix file header");
    h.leafOffsets.resize(h.S + 1);
    rd(F, h.leafOffsets.data(), h.leafOffsets.size() * sizeof(int32_t));
    HMX_CHECK(h.leafOffsets[0] == 0 && h.leafOffsets[h.S] == (int32_t)h.n, HMX_EINVAL, "bad leaf offsets");
    for (uint32_t q = 0; q < h.S; ++q)
        HMX_CHECK(h.leafOffsets[q + 1] >= h.leafOffsets[q] && h.leafOffsets[q + 1] - h.leafOffsets[q] <= (int32_t)h.b,
                  HMX_EINVAL, "bad leaf offsets");
    Structure& s = h.structureCopy;
    s.numLevels = h.numLevels;
    s.n = h.n;
    s.b = h.b;
    s.numTiles.assign(h.numLevels, 0);
    s.tileIndices.assign(h.numLevels, {});
    s.jds.assign(h.numLevels - 1, {});
    for (int l = 1; l < h.numLevels; ++l) {
        const int32_t T = rv<int32_t>(F);
        HMX_CHECK(T >= 0 && (int64_t)T <= ((int64_t)1 << (2 * l)), HMX_EINVAL, "bad tile count");
        s.numTiles[l - 1] = T;
        s.tileIndices[l - 1].resize(T);
        rd(F, s.tileIndices[l - 1].data(), (size_t)T * sizeof(int32_t));
        for (int32_t code : s.tileIndices[l - 1])
            HMX_CHECK(code >= 0 && (int64_t)code < ((int64_t)1 << (2 * l)), HMX_EINVAL, "bad tile index");
        build_jds(s.jds[l - 1], l, s.tileIndices[l - 1]);
    }
    h.structure = &h.structureCopy;
    h.diag.alloc((size_t)h.S * h.b * h.b);
    dev_read(c, F, h.diag.p, h.diag.n);
    h.levels.clear();
    h.levels.resize(h.numLevels - 1);
    for (int l = 1; l < h.numLevels; ++l) {
        HLevel& lv = h.levels[l - 1];
        HMX_CHECK(rv<int32_t>(F) == l, HMX_EINVAL, "level records out of order");
        lv.level = l;
        lv.numTiles = rv<int32_t>(F); lv.m = rv<int32_t>(F); lv.maxRank = rv<int32_t>(F); lv.capHits = rv<int32_t>(F);
        lv.tol = rv<float>(F);
        lv.rankSum = rv<int64_t>(F);
        HMX_CHECK(lv.numTiles == s.numTiles[l - 1] && lv.m == (int32_t)((1u << (h.numLevels - 1 - l)) * h.b) &&
                      lv.rankSum >= 0,
                  HMX_EINVAL, "level record does not match the structure");
        lv.h_ranks.resize(lv.numTiles);
        rd(F, lv.h_ranks.data(), (size_t)lv.numTiles * sizeof(int32_t));
        lv.h_offsets.assign(lv.numTiles + 1, 0);
        int64_t acc = 0;
        for (int t = 0; t < lv.numTiles; ++t) {
            HMX_CHECK(lv.h_ranks[t] >= 0 && lv.h_ranks[t] <= lv.m, HMX_EINVAL, "bad tile rank");
            lv.h_offsets[t] = acc;
            acc += lv.h_ranks[t];
        }
        lv.h_offsets[lv.numTiles] = acc;
        HMX_CHECK(acc == lv.rankSum, HMX_EINVAL, "rank sum does not match the ranks");
        lv.U.alloc((size_t)acc * lv.m);
        lv.V.alloc((size_t)acc * lv.m);
        dev_read(c, F, lv.U.p, (size_t)acc * lv.m);
        dev_read(c, F, lv.V.p, (size_t)acc * lv.m);
    }
    h.lowRankBytes = 0;
    h.factorEntries = 0;
    for (auto& lv : h.levels) {
        h.lowRankBytes += 2ull * (uint64_t)lv.rankSum * lv.m * sizeof(double);
        h.factorEntries += 2ull * (uint64_t)lv.rankSum * lv.m;
    }
    h.denseBytes = (uint64_t)h.S * h.b * h.b * sizeof(double);
    h.factorEntries += (uint64_t)h.S * h.b * h.b;
}

}  // namespace hmx
