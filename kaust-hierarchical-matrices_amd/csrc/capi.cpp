// C ABI (include/hmx.h): argument checks, exception -> status translation,
// and the reference-named entry points.
#include <cstring>
#include <memory>
#include <new>

#include "hmx_internal.h"

using namespace hmx;

namespace hmx {
std::atomic<int64_t> g_allocCount{0}, g_allocFailAt{-1}, g_liveBytes{0};
}

namespace {

thread_local std::string g_noctx_error;

// A failed HIP call leaves its error as the thread's last HIP error; it is
// taken off here, so the launch checks of later calls (hipGetLastError) do
// not report it again.
template <class F>
hmx_status guard(hmx_context* ctx, F&& f) {
    try {
        f();
        if (ctx) ctx->last_error.clear();
        return HMX_OK;
    } catch (const Error& e) {
        (void)hipGetLastError();
        if (ctx) ctx->last_error = e.what();
        else g_noctx_error = e.what();
        return e.status;
    } catch (const std::bad_alloc&) {
        (void)hipGetLastError();
        if (ctx) ctx->last_error = "host allocation failed";
        return HMX_ENOMEM;
    } catch (const std::exception& e) {
        (void)hipGetLastError();
        if (ctx) ctx->last_error = e.what();
        return HMX_EDEVICE;
    }
}

#define NEED(p)                                                                  \
    do {                                                                         \
        if (!(p)) throw Error(HMX_EINVAL, std::string("null argument: ") + #p); \
    } while (0)

void check_config(const hmx_config* cfg) {
    NEED(cfg);
    HMX_CHECK(cfg->divMethod == HMX_FULL_TREE || cfg->divMethod == HMX_DIVIDE_IN_HALF ||
                  cfg->divMethod == HMX_POWER_OF_TWO_ON_LEFT,
              HMX_EINVAL, "divMethod must be POWER_OF_TWO_ON_LEFT, DIVIDE_IN_HALF or FULL_TREE");
    HMX_CHECK(cfg->lowestLevelTolerance > 0.f, HMX_EINVAL, "lowestLevelTolerance must be > 0");
    HMX_CHECK(cfg->maxRankPolicy == HMX_MAXRANK_REF || cfg->maxRankPolicy == HMX_MAXRANK_FULL, HMX_EINVAL,
              "maxRankPolicy must be 0 (reference) or 1 (full)");
    HMX_CHECK(cfg->araR >= 1, HMX_EINVAL, "araR must be >= 1");
    HMX_CHECK(cfg->strictRankCap == 0 || cfg->strictRankCap == 1, HMX_EINVAL, "strictRankCap must be 0 or 1");
    HMX_CHECK(cfg->symmetricTiles == 0 || cfg->symmetricTiles == 1, HMX_EINVAL, "symmetricTiles must be 0 or 1");
}

// strictRankCap: a build whose ARA stopped at a rank cap below the tile
// dimension fails.  Uncapped ranks (HMX_MAXRANK_FULL) stop at the dimension
// at most, where the factors are exact, so only the reference schedule can
// truncate.
void check_cap_hits(const hmx_config& cfg, uint64_t hits, const char* what) {
    HMX_CHECK(!cfg.strictRankCap || cfg.maxRankPolicy != HMX_MAXRANK_REF || hits == 0, HMX_ERANKCAP,
              std::to_string(hits) + " " + what + " reached the rank cap before the ARA converged (strictRankCap)");
}

uint64_t level_cap_hits(const HMatrix& h) {
    uint64_t s = 0;
    for (const HLevel& lv : h.levels) s += (uint64_t)lv.capHits;
    return s;
}

}  // namespace

extern "C" {

hmx_status hmx_create(hmx_context** ctx, void* stream) {
    if (!ctx) return HMX_EINVAL;
    *ctx = nullptr;
    hmx_context* c = new (std::nothrow) hmx_context();
    if (!c) return HMX_ENOMEM;
    c->stream = (hipStream_t)stream;
    hmx_status st = guard(c, [&] {
        HMX_HIP(hipGetDevice(&c->device));
        noalloc_init(c->device);  // this device's copy of the no-allocation matvec's scratch array
    });
    if (st != HMX_OK) {
        g_noctx_error = c->last_error;
        delete c;
        return st;
    }
    *ctx = c;
    return HMX_OK;
}

hmx_status hmx_destroy(hmx_context* ctx) {
    delete ctx;
    return HMX_OK;
}

const char* hmx_last_error(const hmx_context* ctx) { return ctx ? ctx->last_error.c_str() : g_noctx_error.c_str(); }

void hmx_config_default(hmx_config* cfg) {
    if (!cfg) return;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->divMethod = HMX_FULL_TREE;
    cfg->admissibilityCondition = HMX_WEAK_ADMISSIBILITY;
    cfg->numberOfInputPoints = 1024;
    cfg->dimensionOfInputPoints = 2;
    cfg->bucketSize = 32;
    cfg->vectorWidth = 16;
    cfg->lowestLevelTolerance = 1e-5f;
    cfg->eta = 1.0f;
    cfg->correlationLength = 0.0;
    cfg->araBlockSize = 16;
    cfg->araR = 10;
    cfg->maxRankPolicy = HMX_MAXRANK_REF;
    cfg->seed = 1234;
}

const char* hmx_version(void) { return "hmx 0.1 gfx950"; }

void hmx_debug_set_alloc_fault(int64_t failAt) {
    g_allocFailAt.store(failAt < 0 ? -1 : failAt);
    g_allocCount.store(0);
}

int64_t hmx_debug_alloc_count(void) { return g_allocCount.load(); }

int64_t hmx_debug_live_bytes(void) { return g_liveBytes.load(); }

hmx_status hmx_generateDataset(hmx_context* ctx, uint32_t n, uint32_t d, uint64_t seed, double* d_points) {
    if (!ctx) return HMX_EINVAL;
    return guard(ctx, [&] {
        NEED(d_points);
        HMX_CHECK(d >= 1 && d <= 3, HMX_EINVAL, "dimensionOfInputPoints must be 1..3");
        generate_dataset(*ctx, n, d, seed, d_points);
    });
}

hmx_status hmx_generateDatasetBatch(hmx_context* ctx, int32_t count, const uint32_t* n, const uint32_t* d,
                                    const uint64_t* seed, double* const* d_points) {
    if (!ctx) return HMX_EINVAL;
    return guard(ctx, [&] {
        NEED(n);
        NEED(d);
        NEED(seed);
        NEED(d_points);
        HMX_CHECK(count >= 1, HMX_EINVAL, "batch needs at least one cloud");
        std::vector<uint64_t> counts(count), seeds(count);
        std::vector<double*> outs(count);
        for (int32_t k = 0; k < count; ++k) {
            HMX_CHECK(d[k] >= 1 && d[k] <= 3, HMX_EINVAL, "dimensionOfInputPoints must be 1..3");
            HMX_CHECK(d_points[k] != nullptr || (uint64_t)n[k] * d[k] == 0, HMX_EINVAL, "null points");
            counts[k] = (uint64_t)n[k] * d[k];
            seeds[k] = seed[k];
            outs[k] = d_points[k];
        }
        generate_dataset_batch(*ctx, counts, seeds, outs);
    });
}

hmx_status hmx_generateRandomVector(hmx_context* ctx, uint32_t vectorWidth, uint32_t n, uint64_t seed,
                                    double* d_out) {
    if (!ctx) return HMX_EINVAL;
    return guard(ctx, [&] {
        NEED(d_out);
        generate_random_vector(*ctx, vectorWidth, n, seed, d_out);
    });
}

hmx_status hmx_reorderVectors(hmx_context* ctx, const hmx_kdtree* tree, uint32_t vectorWidth, int32_t toOriginal,
                              const double* d_in, double* d_out) {
    if (!ctx) return HMX_EINVAL;
    return guard(ctx, [&] {
        NEED(tree);
        NEED(d_in);
        NEED(d_out);
        HMX_CHECK(d_in != d_out, HMX_EINVAL, "reorderVectors works out of place");
        reorder_vectors(*ctx, tree->n, vectorWidth, toOriginal ? 1 : 0, tree->perm.p, d_in, d_out);
    });
}

hmx_status hmx_constructKDTree(hmx_context* ctx, uint32_t n, uint32_t d, uint32_t bucketSize, int32_t divMethod,
                               const double* d_points, hmx_kdtree** tree) {
    if (!ctx) return HMX_EINVAL;
    return guard(ctx, [&] {
        NEED(tree);
        NEED(d_points);
        *tree = nullptr;
        HMX_CHECK(divMethod == HMX_FULL_TREE || divMethod == HMX_DIVIDE_IN_HALF || divMethod == HMX_POWER_OF_TWO_ON_LEFT,
                  HMX_EINVAL, "divMethod must be POWER_OF_TWO_ON_LEFT, DIVIDE_IN_HALF or FULL_TREE");
        hmx_kdtree* t = new hmx_kdtree();
        t->n = n;
        t->d = d;
        t->b = bucketSize;
        t->divMethod = divMethod;
        try {
            build_kdtree(*ctx, *t, d_points);
        } catch (...) {
            delete t;
            throw;
        }
        *tree = t;
    });
}

hmx_status hmx_constructKDTreeBatch(hmx_context* ctx, int32_t count, const uint32_t* n, const uint32_t* d,
                                    uint32_t bucketSize, int32_t divMethod, const double* const* d_points,
                                    hmx_kdtree** trees) {
    if (!ctx) return HMX_EINVAL;
    return guard(ctx, [&] {
        NEED(n);
        NEED(d);
        NEED(d_points);
        NEED(trees);
        HMX_CHECK(count >= 1, HMX_EINVAL, "batch needs at least one cloud");
        HMX_CHECK(divMethod == HMX_FULL_TREE || divMethod == HMX_DIVIDE_IN_HALF || divMethod == HMX_POWER_OF_TWO_ON_LEFT,
                  HMX_EINVAL, "divMethod must be POWER_OF_TWO_ON_LEFT, DIVIDE_IN_HALF or FULL_TREE");
        for (int32_t k = 0; k < count; ++k) trees[k] = nullptr;
        std::vector<std::unique_ptr<hmx_kdtree>> own;
        std::vector<KDTree*> ts;
        std::vector<const double*> ps;
        for (int32_t k = 0; k < count; ++k) {
            HMX_CHECK(d_points[k] != nullptr, HMX_EINVAL, "null points");
            own.emplace_back(new hmx_kdtree());
            hmx_kdtree& t = *own.back();
            t.n = n[k];
            t.d = d[k];
            t.b = bucketSize;
            t.divMethod = divMethod;
            ts.push_back(&t);
            ps.push_back(d_points[k]);
        }
        build_kdtree_batch(*ctx, ts, ps);
        for (int32_t k = 0; k < count; ++k) trees[k] = own[k].release();
    });
}

hmx_status hmx_kdtree_info(const hmx_kdtree* tree, uint32_t* numSegments, uint32_t* segmentSize,
                           uint32_t* numLevels) {
    if (!tree) return HMX_EINVAL;
    if (numSegments) *numSegments = tree->numSegments;
    if (segmentSize) *segmentSize = tree->b;
    if (numLevels) *numLevels = tree->numLevels;
    return HMX_OK;
}

hmx_status hmx_kdtree_copy(hmx_context* ctx, const hmx_kdtree* tree, int32_t* perm, int32_t* offsets, double* bbmin,
                           double* bbmax) {
    if (!ctx || !tree) return HMX_EINVAL;
    return guard(ctx, [&] {
        if (perm)
            HMX_HIP(hipMemcpyAsync(perm, tree->perm.p, tree->perm.bytes(), hipMemcpyDeviceToHost, ctx->stream));
        if (offsets) std::memcpy(offsets, tree->h_offsets.data(), tree->h_offsets.size() * sizeof(int32_t));
        if (bbmin) std::memcpy(bbmin, tree->h_bbmin.data(), tree->h_bbmin.size() * sizeof(double));
        if (bbmax) std::memcpy(bbmax, tree->h_bbmax.data(), tree->h_bbmax.size() * sizeof(double));
        HMX_HIP(hipStreamSynchronize(ctx->stream));
    });
}

const int32_t* hmx_kdtree_device_perm(const hmx_kdtree* tree) { return tree ? tree->perm.p : nullptr; }

hmx_status hmx_freeKDTree(hmx_kdtree* tree) {
    delete tree;
    return HMX_OK;
}

hmx_status hmx_constructHMatrixStructure(hmx_context* ctx, uint32_t n, uint32_t d, uint32_t bucketSize,
                                         int32_t admissibilityCondition, float eta, const hmx_kdtree* tree,
                                         hmx_structure** structure) {
    if (!ctx) return HMX_EINVAL;
    return guard(ctx, [&] {
        NEED(tree);
        NEED(structure);
        *structure = nullptr;
        HMX_CHECK(tree->d == d && tree->b == bucketSize, HMX_EINVAL, "structure arguments do not match the tree");
        HMX_CHECK(upper_pow2(n) == upper_pow2(tree->n), HMX_EINVAL, "structure n does not match the tree");
        HMX_CHECK(tree->numSegments == (1u << tree->numLevels), HMX_EUNSUPPORTED,
                  "the H-matrix structure needs 2^levels leaves (a POWER_OF_TWO_ON_LEFT tree whose last "
                  "power-of-two block is partly empty has fewer)");
        std::unique_ptr<hmx_structure> s(new hmx_structure());
        build_structure(*s, n, d, bucketSize, admissibilityCondition, eta, tree->h_bbmin.data(),
                        tree->h_bbmax.data());
        HMX_CHECK((uint32_t)(1u << (s->numLevels - 1)) == tree->numSegments, HMX_EINVAL,
                  "structure depth does not match the tree");
        *structure = s.release();
    });
}

hmx_status hmx_constructHMatrixStructureDevice(hmx_context* ctx, uint32_t n, uint32_t d, uint32_t bucketSize,
                                               int32_t admissibilityCondition, float eta, const hmx_kdtree* tree,
                                               hmx_structure** structure) {
    if (!ctx) return HMX_EINVAL;
    return guard(ctx, [&] {
        NEED(tree);
        NEED(structure);
        *structure = nullptr;
        HMX_CHECK(tree->d == d && tree->b == bucketSize, HMX_EINVAL, "structure arguments do not match the tree");
        HMX_CHECK(upper_pow2(n) == upper_pow2(tree->n), HMX_EINVAL, "structure n does not match the tree");
        HMX_CHECK(tree->numSegments == (1u << tree->numLevels), HMX_EUNSUPPORTED,
                  "the H-matrix structure needs 2^levels leaves (a POWER_OF_TWO_ON_LEFT tree whose last "
                  "power-of-two block is partly empty has fewer)");
        std::unique_ptr<hmx_structure> s(new hmx_structure());
        build_structure_device(*ctx, *s, n, d, bucketSize, admissibilityCondition, eta, tree->bbmin.p,
                               tree->bbmax.p);
        HMX_CHECK((uint32_t)(1u << (s->numLevels - 1)) == tree->numSegments, HMX_EINVAL,
                  "structure depth does not match the tree");
        *structure = s.release();
    });
}

hmx_status hmx_structure_from_boxes(uint32_t n, uint32_t d, uint32_t bucketSize, int32_t admissibilityCondition,
                                    float eta, const double* bbmin, const double* bbmax, hmx_structure** structure) {
    return guard(nullptr, [&] {
        NEED(structure);
        NEED(bbmin);
        NEED(bbmax);
        *structure = nullptr;
        hmx_structure* s = new hmx_structure();
        try {
            build_structure(*s, n, d, bucketSize, admissibilityCondition, eta, bbmin, bbmax);
        } catch (...) {
            delete s;
            throw;
        }
        *structure = s;
    });
}

hmx_status hmx_structure_info(const hmx_structure* s, int32_t* numLevels, int32_t* numTiles) {
    if (!s) return HMX_EINVAL;
    if (numLevels) *numLevels = s->numLevels;
    if (numTiles) std::memcpy(numTiles, s->numTiles.data(), s->numTiles.size() * sizeof(int32_t));
    return HMX_OK;
}

hmx_status hmx_structure_tiles(const hmx_structure* s, int32_t level, int32_t* tileIndices) {
    if (!s || !tileIndices || level < 1 || level >= s->numLevels) return HMX_EINVAL;
    const auto& v = s->tileIndices[level - 1];
    std::memcpy(tileIndices, v.data(), v.size() * sizeof(int32_t));
    return HMX_OK;
}

hmx_status hmx_structure_jds(const hmx_structure* s, int32_t level, int32_t* rowIndices, int32_t* columnIndices,
                             int32_t* tilePos, int32_t* iterPtr, int32_t* numDiagonals) {
    if (!s || level < 1 || level >= s->numLevels) return HMX_EINVAL;
    const auto& j = s->jds[level - 1];
    if (rowIndices) std::memcpy(rowIndices, j.rowIndices.data(), j.rowIndices.size() * sizeof(int32_t));
    if (columnIndices) std::memcpy(columnIndices, j.columnIndices.data(), j.columnIndices.size() * sizeof(int32_t));
    if (tilePos) std::memcpy(tilePos, j.tilePos.data(), j.tilePos.size() * sizeof(int32_t));
    if (iterPtr) std::memcpy(iterPtr, j.iterPtr.data(), j.iterPtr.size() * sizeof(int32_t));
    if (numDiagonals) *numDiagonals = (int32_t)j.iterPtr.size() - 1;
    return HMX_OK;
}

hmx_status hmx_freeHMatrixStructure(hmx_structure* s) {
    delete s;
    return HMX_OK;
}

hmx_status hmx_createColumnMajorLRMatrix(hmx_context* ctx, const hmx_config* cfg, const hmx_kdtree* tree,
                                         const double* d_points, hmx_tlr** matrix, uint64_t* rankSum) {
    if (!ctx) return HMX_EINVAL;
    return guard(ctx, [&] {
        check_config(cfg);
        NEED(tree);
        NEED(d_points);
        NEED(matrix);
        *matrix = nullptr;
        hmx_tlr* m = new hmx_tlr();
        try {
            build_tlr_cm(*ctx, *cfg, *tree, d_points, *m);
            check_cap_hits(*cfg, m->capHits, "TLR tiles");
        } catch (...) {
            delete m;
            throw;
        }
        *matrix = m;
        if (rankSum) *rankSum = m->rankSum;
    });
}

hmx_status hmx_createColumnMajorLRMatrixPiece(hmx_context* ctx, const hmx_config* cfg, const hmx_kdtree* tree,
                                              const double* d_points, uint32_t pieces, uint32_t rowPiece,
                                              uint32_t colPiece, hmx_tlr** matrix, uint64_t* rankSum) {
    if (!ctx) return HMX_EINVAL;
    return guard(ctx, [&] {
        check_config(cfg);
        NEED(tree);
        NEED(d_points);
        NEED(matrix);
        *matrix = nullptr;
        hmx_tlr* m = new hmx_tlr();
        try {
            build_tlr_piece(*ctx, *cfg, *tree, d_points, pieces, rowPiece, colPiece, *m);
            check_cap_hits(*cfg, m->capHits, "TLR tiles");
        } catch (...) {
            delete m;
            throw;
        }
        *matrix = m;
        if (rankSum) *rankSum = m->rankSum;
    });
}

hmx_status hmx_convertColumnMajorToMorton(hmx_context* ctx, const hmx_tlr* matrix, hmx_tlr** mortonMatrix) {
    if (!ctx) return HMX_EINVAL;
    return guard(ctx, [&] {
        NEED(matrix);
        NEED(mortonMatrix);
        *mortonMatrix = nullptr;
        HMX_CHECK(matrix->pieces == 1, HMX_EINVAL, "a TLR piece has no Morton form (convert the whole matrix)");
        hmx_tlr* m = new hmx_tlr();
        try {
            convert_to_morton(*ctx, *matrix, *m);
        } catch (...) {
            delete m;
            throw;
        }
        *mortonMatrix = m;
    });
}

hmx_status hmx_tlr_info(const hmx_tlr* m, uint32_t* numSegments, uint32_t* segmentSize, int32_t* ordering,
                        uint64_t* rankSum, uint64_t* capHits) {
    if (!m) return HMX_EINVAL;
    if (numSegments) *numSegments = m->S;
    if (segmentSize) *segmentSize = m->b;
    if (ordering) *ordering = m->ordering;
    if (rankSum) *rankSum = m->rankSum;
    if (capHits) *capHits = m->capHits;
    return HMX_OK;
}

hmx_status hmx_tlr_copy(hmx_context* ctx, const hmx_tlr* m, int32_t* blockRanks, int64_t* blockOffsets, double* U,
                        double* V, double* diagonal) {
    if (!ctx || !m) return HMX_EINVAL;
    return guard(ctx, [&] {
        auto cp = [&](void* dst, const void* src, size_t bytes) {
            if (dst && bytes) HMX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        };
        cp(blockRanks, m->ranks.p, m->ranks.bytes());
        cp(blockOffsets, m->offsets.p, m->offsets.bytes());
        cp(U, m->U.p, m->U.bytes());
        cp(V, m->V.p, m->V.bytes());
        cp(diagonal, m->diagonal.p, m->diagonal.bytes());
        HMX_HIP(hipStreamSynchronize(ctx->stream));
    });
}

hmx_status hmx_freeMatrix(hmx_tlr* m) {
    delete m;
    return HMX_OK;
}

void hmx_generateMaxRanks(uint32_t numLevels, uint32_t bucketSize, uint32_t* maxRanks) {
    if (!maxRanks || numLevels < 3) return;
    for (uint32_t i = 0; i < numLevels - 2; ++i) {
        maxRanks[i] = bucketSize * (1u << i);
        if (i > 5) maxRanks[i] /= 4;
    }
}

hmx_status hmx_generateHMatrixFromStruct(hmx_context* ctx, const hmx_config* cfg, const hmx_kdtree* tree,
                                         const hmx_structure* structure, const hmx_tlr* mortonMatrix,
                                         hmx_hmatrix** hmatrix) {
    if (!ctx) return HMX_EINVAL;
    return guard(ctx, [&] {
        check_config(cfg);
        NEED(tree);
        NEED(structure);
        NEED(mortonMatrix);
        NEED(hmatrix);
        *hmatrix = nullptr;
        HMX_CHECK(mortonMatrix->pieces == 1, HMX_EINVAL, "an H-matrix needs the whole TLR matrix, not a piece");
        hmx_hmatrix* h = new hmx_hmatrix();
        try {
            build_hmatrix(*ctx, *cfg, *tree, *structure, *mortonMatrix, *h);
            check_cap_hits(*cfg, level_cap_hits(*h), "H-matrix tiles");
        } catch (...) {
            delete h;
            throw;
        }
        *hmatrix = h;
    });
}

hmx_status hmx_generateHMatrixDirect(hmx_context* ctx, const hmx_config* cfg, const hmx_kdtree* tree,
                                     const hmx_structure* structure, const double* d_points, hmx_hmatrix** hmatrix) {
    if (!ctx) return HMX_EINVAL;
    return guard(ctx, [&] {
        check_config(cfg);
        NEED(tree);
        NEED(structure);
        NEED(d_points);
        NEED(hmatrix);
        *hmatrix = nullptr;
        HMX_CHECK(tree->n == cfg->numberOfInputPoints && tree->d == cfg->dimensionOfInputPoints &&
                      tree->b == cfg->bucketSize,
                  HMX_EINVAL, "configuration and KD-tree do not match");
        hmx_hmatrix* h = new hmx_hmatrix();
        try {
            build_hmatrix_direct(*ctx, *cfg, *tree, *structure, d_points, *h);
            check_cap_hits(*cfg, level_cap_hits(*h), "H-matrix tiles");
        } catch (...) {
            delete h;
            throw;
        }
        *hmatrix = h;
    });
}

hmx_status hmx_hmatrix_build_info(const hmx_hmatrix* h, double* samplerFlops, double* kernelEvals, double* samplerMs) {
    if (!h) return HMX_EINVAL;
    if (samplerFlops) *samplerFlops = h->samplerFlops;
    if (kernelEvals) *kernelEvals = h->kernelEvals;
    if (samplerMs) *samplerMs = h->samplerMs;
    return HMX_OK;
}

hmx_status hmx_hmatrix_level_info(const hmx_hmatrix* h, int32_t level, int32_t* numTiles, int64_t* rankSum,
                                  int32_t* maxRank, float* tolerance, int32_t* capHits) {
    if (!h || level < 1 || level >= h->numLevels) return HMX_EINVAL;
    const HLevel& lv = h->levels[level - 1];
    if (numTiles) *numTiles = lv.numTiles;
    if (rankSum) *rankSum = lv.rankSum;
    if (maxRank) *maxRank = lv.maxRank;
    if (tolerance) *tolerance = lv.tol;
    if (capHits) *capHits = lv.capHits;
    return HMX_OK;
}

hmx_status hmx_hmatrix_level_copy(hmx_context* ctx, const hmx_hmatrix* h, int32_t level, int32_t* ranks, double* U,
                                  double* V) {
    if (!ctx || !h || level < 1 || level >= h->numLevels) return HMX_EINVAL;
    return guard(ctx, [&] {
        const HLevel& lv = h->levels[level - 1];
        if (ranks) std::memcpy(ranks, lv.h_ranks.data(), lv.h_ranks.size() * sizeof(int32_t));
        if (U && lv.U.n) HMX_HIP(hipMemcpyAsync(U, lv.U.p, lv.U.bytes(), hipMemcpyDeviceToHost, ctx->stream));
        if (V && lv.V.n) HMX_HIP(hipMemcpyAsync(V, lv.V.p, lv.V.bytes(), hipMemcpyDeviceToHost, ctx->stream));
        HMX_HIP(hipStreamSynchronize(ctx->stream));
    });
}

hmx_status hmx_hmatrix_diag_copy(hmx_context* ctx, const hmx_hmatrix* h, double* diag) {
    if (!ctx || !h || !diag) return HMX_EINVAL;
    return guard(ctx, [&] {
        HMX_HIP(hipMemcpyAsync(diag, h->diag.p, h->diag.bytes(), hipMemcpyDeviceToHost, ctx->stream));
        HMX_HIP(hipStreamSynchronize(ctx->stream));
    });
}

hmx_status hmx_hmatrix_bytes(const hmx_hmatrix* h, uint64_t* lowRankBytes, uint64_t* denseBytes,
                             uint64_t* factorEntries) {
    if (!h) return HMX_EINVAL;
    if (lowRankBytes) *lowRankBytes = h->lowRankBytes;
    if (denseBytes) *denseBytes = h->denseBytes;
    if (factorEntries) *factorEntries = h->factorEntries;
    return HMX_OK;
}

hmx_status hmx_freeHMatrix(hmx_hmatrix* h) {
    delete h;
    return HMX_OK;
}

hmx_status hmx_HMatrixVecMult(hmx_context* ctx, const hmx_hmatrix* h, uint32_t vectorWidth, const double* d_in,
                              double* d_out) {
    if (!ctx) return HMX_EINVAL;
    return guard(ctx, [&] {
        NEED(h);
        // two or more right-hand sides: the FP64 matrix cores (passes of 16
        // or 64, partial leaves through the padded space); one: the
        // streaming kernels
        if (vectorWidth >= 2)
            hmatvec_multi(*ctx, const_cast<hmx_hmatrix&>(*h), vectorWidth, d_in, d_out, nullptr, 0);
        else
            hmatvec(*ctx, const_cast<hmx_hmatrix&>(*h), vectorWidth, d_in, d_out, false, nullptr, 0);
    });
}

hmx_status hmx_HMatrixVecMultNoAlloc(hmx_context* ctx, const hmx_hmatrix* h, uint32_t vectorWidth,
                                     const double* d_in, double* d_out) {
    if (!ctx) return HMX_EINVAL;
    return guard(ctx, [&] {
        NEED(h);
        hmatvec_noalloc(*ctx, *h, vectorWidth, d_in, d_out);
    });
}

hmx_status hmx_matvec_profile_multi(hmx_context* ctx, const hmx_hmatrix* h, uint32_t vectorWidth, const void* d_in,
                                    void* d_out, int32_t iters, int32_t mixed, float* ms) {
    if (!ctx) return HMX_EINVAL;
    return guard(ctx, [&] {
        NEED(h);
        NEED(ms);
        HMX_CHECK(vectorWidth == 16 || vectorWidth == 64, HMX_EINVAL, "profiling needs vectorWidth 16 or 64");
        auto& H = const_cast<hmx_hmatrix&>(*h);
        if (mixed) hmatvec_mixed(*ctx, H, vectorWidth, (const float*)d_in, (float*)d_out, ms, iters);
        else hmatvec_multi(*ctx, H, vectorWidth, (const double*)d_in, (double*)d_out, ms, iters);
    });
}

hmx_status hmx_hmatrix_make_mixed(hmx_context* ctx, hmx_hmatrix* h, double targetRelErr) {
    if (!ctx) return HMX_EINVAL;
    return guard(ctx, [&] {
        NEED(h);
        hmatrix_make_mixed(*ctx, *h, targetRelErr);
    });
}

hmx_status hmx_hmatrix_mixed_info(const hmx_hmatrix* h, int64_t* hiCols, int64_t* loCols, double* frobH2,
                                  double* loEnergy, uint64_t* bytes) {
    if (!h || !h->mfMixed.built) return HMX_ESTATE;
    if (hiCols) *hiCols = h->mfMixed.hiCols;
    if (loCols) *loCols = h->mfMixed.loCols;
    if (frobH2) *frobH2 = h->mfMixed.frobH2;
    if (loEnergy) *loEnergy = h->mfMixed.loEnergy;
    if (bytes) *bytes = h->mfMixed.bytes;
    return HMX_OK;
}

hmx_status hmx_HMatrixVecMultMixed(hmx_context* ctx, const hmx_hmatrix* h, uint32_t vectorWidth, const float* d_in,
                                   float* d_out) {
    if (!ctx) return HMX_EINVAL;
    return guard(ctx, [&] {
        NEED(h);
        hmatvec_mixed(*ctx, const_cast<hmx_hmatrix&>(*h), vectorWidth, d_in, d_out, nullptr, 0);
    });
}

hmx_status hmx_hmatvec_grad_factors(hmx_context* ctx, const hmx_hmatrix* h, uint32_t vectorWidth, const double* d_X,
                                    const double* d_dY, double* const* d_dU, double* const* d_dV) {
    if (!ctx) return HMX_EINVAL;
    return guard(ctx, [&] {
        NEED(h);
        hmatvec_grad_factors(*ctx, const_cast<hmx_hmatrix&>(*h), vectorWidth, d_X, d_dY, d_dU, d_dV);
    });
}

hmx_status hmx_hmatrix_level_get(hmx_context* ctx, const hmx_hmatrix* h, int32_t level, double* d_U, double* d_V) {
    if (!ctx || !h || level < 1 || level >= h->numLevels) return HMX_EINVAL;
    return guard(ctx, [&] {
        const HLevel& lv = h->levels[level - 1];
        if (lv.U.n) {
            NEED(d_U);
            NEED(d_V);
            HMX_HIP(hipMemcpyAsync(d_U, lv.U.p, lv.U.bytes(), hipMemcpyDeviceToDevice, ctx->stream));
            HMX_HIP(hipMemcpyAsync(d_V, lv.V.p, lv.V.bytes(), hipMemcpyDeviceToDevice, ctx->stream));
        }
        HMX_HIP(hipStreamSynchronize(ctx->stream));
    });
}

hmx_status hmx_hmatrix_set_level(hmx_context* ctx, hmx_hmatrix* h, int32_t level, const double* d_U,
                                 const double* d_V) {
    if (!ctx) return HMX_EINVAL;
    return guard(ctx, [&] {
        NEED(h);
        hmatrix_set_level(*ctx, *h, level, d_U, d_V);
    });
}

hmx_status hmx_stream_read_probe(hmx_context* ctx, const void* d_buf, uint64_t bytes, int32_t iters, float* ms) {
    if (!ctx) return HMX_EINVAL;
    return guard(ctx, [&] {
        NEED(d_buf);
        NEED(ms);
        stream_read(*ctx, d_buf, bytes, iters, ms);
    });
}

hmx_status hmx_fp64_peak_probe(hmx_context* ctx, int32_t unit, int32_t iters, float* tflops) {
    if (!ctx) return HMX_EINVAL;
    return guard(ctx, [&] {
        NEED(tflops);
        fp64_peak(*ctx, unit, iters, tflops);
    });
}

hmx_status hmx_selftest_mfma(hmx_context* ctx, int32_t type, const double* d_A, const double* d_B, double* d_D) {
    if (!ctx) return HMX_EINVAL;
    return guard(ctx, [&] {
        NEED(d_A);
        NEED(d_B);
        NEED(d_D);
        mfma_selftest(*ctx, type, d_A, d_B, d_D);
    });
}

hmx_status hmx_HMatrixVecMultTranspose(hmx_context* ctx, const hmx_hmatrix* h, uint32_t vectorWidth,
                                       const double* d_in, double* d_out) {
    if (!ctx) return HMX_EINVAL;
    return guard(ctx, [&] {
        NEED(h);
        hmatvec(*ctx, const_cast<hmx_hmatrix&>(*h), vectorWidth, d_in, d_out, true, nullptr, 0);
    });
}

hmx_status hmx_matvec_profile(hmx_context* ctx, const hmx_hmatrix* h, uint32_t vectorWidth, const double* d_in,
                              double* d_out, int32_t iters, float* ms) {
    if (!ctx) return HMX_EINVAL;
    return guard(ctx, [&] {
        NEED(h);
        NEED(ms);
        HMX_CHECK(vectorWidth == 1 || vectorWidth == 4 || vectorWidth == 8 || vectorWidth == 16, HMX_EINVAL,
                  "profiling needs vectorWidth in {1,4,8,16}");
        hmatvec(*ctx, const_cast<hmx_hmatrix&>(*h), vectorWidth, d_in, d_out, false, ms, iters);
    });
}

hmx_status hmx_denseRowsXVector(hmx_context* ctx, uint32_t n, uint32_t d, const double* d_points,
                                const int32_t* d_perm, double correlationLength, uint32_t nrows,
                                const int32_t* d_rows, uint32_t k, const double* d_X, double* d_Y) {
    if (!ctx) return HMX_EINVAL;
    return guard(ctx, [&] {
        NEED(d_points);
        NEED(d_perm);
        NEED(d_rows);
        NEED(d_X);
        NEED(d_Y);
        dense_rows(*ctx, n, d, d_points, d_perm, correlationLength > 0 ? correlationLength : default_ell(d), nrows,
                   d_rows, k, d_X, d_Y);
    });
}

hmx_status hmx_hmatrix_save(hmx_context* ctx, const hmx_hmatrix* h, const char* path) {
    if (!ctx) return HMX_EINVAL;
    return guard(ctx, [&] {
        NEED(h);
        hmatrix_save(*ctx, *h, path);
    });
}

hmx_status hmx_hmatrix_load(hmx_context* ctx, const char* path, hmx_hmatrix** hmatrix) {
    if (!ctx) return HMX_EINVAL;
    return guard(ctx, [&] {
        NEED(hmatrix);
        *hmatrix = nullptr;
        hmx_hmatrix* h = new hmx_hmatrix();
        try {
            hmatrix_load(*ctx, path, *h);
        } catch (...) {
            delete h;
            throw;
        }
        *hmatrix = h;
    });
}

hmx_status hmx_hmatrix_set_plan_options(hmx_hmatrix* h, int32_t denseNear, int32_t directZ, int32_t locality) {
    if (!h) return HMX_EINVAL;
    if ((denseNear & ~1) || (directZ & ~1) || (locality & ~1)) return HMX_EINVAL;
    h->planDenseNear = denseNear;
    h->planDirectZ = directZ;
    h->planLocality = locality;
    h->mf64 = MfPlan();
    h->mfMixed = MfPlan();
    h->plan = MatvecPlan();
    h->planT = MatvecPlan();
    return HMX_OK;
}

hmx_status hmx_hmatrix_set_fp64_layout(hmx_hmatrix* h, int32_t layout) {
    if (!h || layout < 0 || layout > 2) return HMX_EINVAL;
    h->planFp64Layout = layout;
    h->mf64 = MfPlan();
    return HMX_OK;
}

hmx_status hmx_hmatrix_plan_info(const hmx_hmatrix* h, int32_t plan, uint64_t* bytes, int64_t* denseBlocks,
                                 int64_t* lowRankColumns) {
    if (!h || plan < 0 || plan > 2) return HMX_EINVAL;
    if (plan == 2) {  // the streaming (one right-hand side) plan
        if (!h->plan.built) return HMX_ESTATE;
        if (bytes) *bytes = h->plan.bytes;
        if (denseBlocks) *denseBlocks = h->plan.numNear;
        if (lowRankColumns) *lowRankColumns = h->plan.numCols;
        return HMX_OK;
    }
    const MfPlan& P = plan == 0 ? h->mf64 : h->mfMixed;
    if (!P.built) return HMX_ESTATE;
    if (bytes) *bytes = P.bytes;
    if (denseBlocks) *denseBlocks = P.numNear;
    if (lowRankColumns) *lowRankColumns = plan == 0 ? P.g[0].numColsPad : P.hiCols + P.loCols;
    return HMX_OK;
}

hmx_status hmx_hmatrix_plan_phase_bytes(const hmx_hmatrix* h, int32_t plan, uint64_t* phase1Bytes,
                                        uint64_t* phase2Bytes) {
    if (!h || plan < 0 || plan > 2) return HMX_EINVAL;
    uint64_t tot = 0, v = 0;
    if (plan == 2) {
        if (!h->plan.built) return HMX_ESTATE;
        tot = h->plan.bytes;
        v = h->plan.lowRankBytes / 2;
    } else {
        const MfPlan& P = plan == 0 ? h->mf64 : h->mfMixed;
        if (!P.built) return HMX_ESTATE;
        tot = P.bytes;
        v = (P.g[0].factorBytes + (plan == 1 ? P.g[1].factorBytes : 0)) / 2;
    }
    if (phase1Bytes) *phase1Bytes = v;
    if (phase2Bytes) *phase2Bytes = tot - v;
    return HMX_OK;
}

hmx_status hmx_generateHMatrixDirectBatch(hmx_context* ctx, int32_t count, const hmx_config* cfgs,
                                          hmx_kdtree* const* trees, hmx_structure* const* structures,
                                          const double* const* d_points, hmx_hmatrix** hmatrices) {
    if (!ctx) return HMX_EINVAL;
    return guard(ctx, [&] {
        NEED(cfgs);
        NEED(trees);
        NEED(structures);
        NEED(d_points);
        NEED(hmatrices);
        HMX_CHECK(count >= 1, HMX_EINVAL, "batch needs at least one instance");
        for (int32_t i = 0; i < count; ++i) hmatrices[i] = nullptr;
        std::vector<const hmx_config*> cf;
        std::vector<const KDTree*> tr;
        std::vector<const Structure*> st;
        std::vector<const double*> pt;
        for (int32_t i = 0; i < count; ++i) {
            check_config(&cfgs[i]);
            HMX_CHECK(trees[i] && structures[i] && d_points[i], HMX_EINVAL, "null instance argument");
            HMX_CHECK(trees[i]->n == cfgs[i].numberOfInputPoints && trees[i]->d == cfgs[i].dimensionOfInputPoints &&
                          trees[i]->b == cfgs[i].bucketSize,
                      HMX_EINVAL, "configuration and KD-tree do not match");
            cf.push_back(&cfgs[i]);
            tr.push_back(trees[i]);
            st.push_back(structures[i]);
            pt.push_back(d_points[i]);
        }
        std::vector<std::unique_ptr<hmx_hmatrix>> own;
        std::vector<HMatrix*> hs;
        for (int32_t i = 0; i < count; ++i) {
            own.emplace_back(new hmx_hmatrix());
            hs.push_back(own.back().get());
        }
        build_hmatrix_direct_batch(*ctx, cf, tr, st, pt, hs);
        for (int32_t i = 0; i < count; ++i) check_cap_hits(cfgs[i], level_cap_hits(*hs[i]), "H-matrix tiles");
        for (int32_t i = 0; i < count; ++i) hmatrices[i] = own[i].release();
    });
}

hmx_status hmx_batch_create(hmx_context* ctx, int32_t count, hmx_hmatrix* const* hmatrices, hmx_batch** batch) {
    if (!ctx) return HMX_EINVAL;
    return guard(ctx, [&] {
        NEED(hmatrices);
        NEED(batch);
        HMX_CHECK(count >= 1, HMX_EINVAL, "batch needs at least one H-matrix");
        *batch = nullptr;
        std::vector<HMatrix*> hs;
        for (int32_t i = 0; i < count; ++i) {
            HMX_CHECK(hmatrices[i] != nullptr, HMX_EINVAL, "null H-matrix in batch");
            hs.push_back(hmatrices[i]);
        }
        hmx_batch* b = new hmx_batch();
        try {
            batch_create(*ctx, hs, *b);
        } catch (...) {
            delete b;
            throw;
        }
        *batch = b;
    });
}

hmx_status hmx_batch_info(const hmx_batch* batch, int64_t* totalPoints, int64_t* totalLeaves, int32_t* chunk) {
    if (!batch) return HMX_EINVAL;
    if (totalPoints) *totalPoints = batch->P.n;
    if (totalLeaves) *totalLeaves = batch->P.numLeaves;
    if (chunk) *chunk = batch->P.chunk;
    return HMX_OK;
}

hmx_status hmx_batch_matvec(hmx_context* ctx, hmx_batch* batch, uint32_t vectorWidth, const double* d_in,
                            double* d_out, int32_t useGraph) {
    if (!ctx) return HMX_EINVAL;
    return guard(ctx, [&] {
        NEED(batch);
        batch_matvec(*ctx, *batch, vectorWidth, d_in, d_out, useGraph != 0);
    });
}

hmx_status hmx_batch_tune(hmx_context* ctx, hmx_batch* batch, uint32_t vectorWidth, int32_t* bestChunk,
                          float* bestMs) {
    if (!ctx) return HMX_EINVAL;
    return guard(ctx, [&] {
        NEED(batch);
        HMX_CHECK(vectorWidth >= 1, HMX_EINVAL, "vectorWidth must be >= 1");
        int bc = 0;
        batch_tune(*ctx, *batch, vectorWidth, &bc, bestMs);
        if (bestChunk) *bestChunk = bc;
    });
}

hmx_status hmx_batch_free(hmx_batch* batch) {
    delete batch;
    return HMX_OK;
}

}  // extern "C"
