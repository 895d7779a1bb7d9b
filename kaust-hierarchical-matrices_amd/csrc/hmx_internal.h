// Host-side internals of the H-matrix library: contexts, device buffers and
// the objects behind the opaque C-ABI handles.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/hmx.h"

namespace hmx {

// Error carrying an hmx_status; thrown by host code, caught at the C-ABI edge.
struct Error : std::runtime_error {
    hmx_status status;
    Error(hmx_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

#define HMX_HIP(call)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            throw ::hmx::Error(e_ == hipErrorOutOfMemory ? HMX_ENOMEM : HMX_EDEVICE,                   \
                               std::string(#call) + ": " + hipGetErrorString(e_) + " at " + __FILE__ + \
                                   ":" + std::to_string(__LINE__));                                    \
    } while (0)

#define HMX_CHECK(cond, status, msg)                              \
    do {                                                          \
        if (!(cond)) throw ::hmx::Error((status), std::string(msg)); \
    } while (0)

// Launch check: catches launch-configuration errors right after a launch.
#define HMX_LAUNCHED() HMX_HIP(hipGetLastError())

// Forced allocation failure for the error-path tests (hmx_debug_set_alloc_fault):
// allocation number g_allocFailAt (counted by g_allocCount) fails.
extern std::atomic<int64_t> g_allocCount, g_allocFailAt, g_liveBytes;
inline void alloc_fault_point() {
    const int64_t k = g_allocCount.fetch_add(1);
    if (k == g_allocFailAt.load())
        throw Error(HMX_ENOMEM, "forced device allocation failure (allocation " + std::to_string(k) + ")");
}

// Owning device buffer (hipMalloc / hipFree), movable, not copyable.
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    std::shared_ptr<void> keep;  // set for a view: the shared block this range lies in
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), n(o.n), keep(std::move(o.keep)) { o.p = nullptr; o.n = 0; }
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p; n = o.n; keep = std::move(o.keep);
            o.p = nullptr; o.n = 0;
        }
        return *this;
    }
    ~DBuf() { release(); }
    // A view of count elements at offset off of a shared block: no copy; the
    // block is freed with its last view (batched builds hand every instance
    // its range of one allocation this way)
    void view(const std::shared_ptr<DBuf<T>>& block, size_t off, size_t count) {
        release();
        if (count) {
            keep = block;
            p = block->p + off;
        }
        n = count;
    }
    void alloc(size_t count) {
        release();
        if (count) {
            alloc_fault_point();
            HMX_HIP(hipMalloc((void**)&p, count * sizeof(T)));
            g_liveBytes.fetch_add((int64_t)(count * sizeof(T)));
        }
        n = count;
    }
    void release() {
        if (keep) {
            keep.reset();
        } else if (p) {
            (void)hipFree(p);
            g_liveBytes.fetch_sub((int64_t)(n * sizeof(T)));
        }
        p = nullptr;
        n = 0;
    }
    size_t bytes() const { return n * sizeof(T); }
};

struct Context {
    hipStream_t stream = nullptr;
    int device = 0;
    std::string last_error;
};

// KD-tree (src/kDTree.cuh:8-14) with device permutation/offsets and
// per-level bounding boxes (src/boundingBoxes.h:8-23).
struct KDTree {
    uint32_t n = 0, d = 0, b = 0;
    int32_t divMethod = HMX_FULL_TREE;
    uint32_t numLevels = 0;    // number of split levels (getMaxSegmentSize().second)
    uint32_t numSegments = 0;  // 2^numLevels leaves
    DBuf<int32_t> perm;        // segmentIndices
    DBuf<int32_t> offsets;     // segmentOffsets, numSegments+1
    DBuf<double> bbmin, bbmax; // [(2^(numLevels+1)-1)][d]
    std::vector<int32_t> h_offsets;
    std::vector<double> h_bbmin, h_bbmax;  // host mirror for the structure pass
};

// HMatrixStructure (src/HMatrixStructure.cuh:8-13) plus the JDS of each level.
struct Structure {
    int32_t numLevels = 0;
    uint32_t n = 0, b = 0;
    std::vector<int32_t> numTiles;                 // [numLevels]
    std::vector<std::vector<int32_t>> tileIndices; // [numLevels]; level l at l-1
    struct JDS {
        std::vector<int32_t> rowIndices, columnIndices, tilePos, iterPtr;
    };
    std::vector<JDS> jds;  // [numLevels-1]; level l at l-1
};

// TLR_Matrix (src/TLRMatrix.cuh:10-21) with 64-bit offsets.
struct TLR {
    int32_t ordering = HMX_COLUMN_MAJOR;
    uint32_t S = 0, b = 0, n = 0;
    uint32_t pieces = 1, rowPiece = 0, colPiece = 0;  // a piece: S = leaves per piece side
    uint64_t rankSum = 0, capHits = 0;
    DBuf<int32_t> ranks;     // blockRanks [S*S]
    DBuf<int64_t> offsets;   // blockOffsets [S*S+1] (exclusive scan, in columns)
    DBuf<double> U, V;       // [rankSum*b]
    DBuf<double> diagonal;   // [S*b*b]
};

// One level of the hierarchical matrix (src/HMatrix.cuh:9-18).
struct HLevel {
    int32_t level = 0, numTiles = 0, m = 0, maxRank = 0, capHits = 0;
    float tol = 0.f;
    int64_t rankSum = 0;
    std::vector<int32_t> h_ranks;
    std::vector<int64_t> h_offsets;  // exclusive scan over tiles (columns)
    DBuf<double> U, V;               // [rankSum*m]
};

// Phase-1 work item: a contiguous run of one tile's input-side factor
// (V for H*x, U for H^T*x).  colLen == 0: the run is one chunk of a single
// column (len elements, one partial result); colLen > 0: the run holds
// len/colLen whole columns of length colLen, one result per column.
struct MvItem {
    const double* v;
    int32_t xoff, len, colLen, out;
};
// Phase-2 entry: the leaf's rows of one tile's output-side factor.
struct MvEntry {
    const double* u;  // first row of this leaf inside the tile's factor
    int32_t m;        // column stride (tile dimension)
    int32_t rank;
    int32_t zbase;    // first global column (index into z)
    int32_t pad;
};

// Leaf-level tile applied as a dense block by the streaming plan: the block
// is U V^T (V U^T for the transposed plan), outLeaf += block * x[inLeaf].
struct NearRef {
    const double* U;
    const double* V;
    int32_t r, outLeaf, inLeaf, pad;
};

// Host form of a matvec plan over the padded vector space (position
// leaf*b + i for point i of a leaf; padded positions hold zeros).
struct PlanHost {
    int b = 0;
    std::vector<NearRef> nears;     // leaf-level tiles applied as dense blocks
    uint64_t lowRankBytes = 0;      // U and V bytes of the tiles applied as factors
    std::vector<MvItem> items;
    std::vector<int64_t> colPart;
    std::vector<int32_t> colNpart;
    std::vector<int32_t> leafPtr;   // [S+1]
    std::vector<MvEntry> entries;
    // tiles of dimension >= rowBlock (0: none), per rowBlock-row block of the
    // padded space instead of per leaf
    int rowBlock = 0;
    std::vector<int32_t> blkPtr;    // [S*b/rowBlock + 1]
    std::vector<MvEntry> blkEntries;
    int64_t numCols = 0, numPart = 0;
};

// Per-call workspaces of the streaming matvec kernels: chunk partial sums,
// z, and (partial leaves) x and y in the padded space.
struct MvWork {
    double *zpart, *z, *Xp, *Yp;
};

// Device-side tables used by the matvec kernels (built once per H-matrix).
struct MatvecPlan {
    DBuf<MvItem> items;
    int32_t numItems = 0;
    DBuf<int64_t> colPart;   // per global column: first partial slot
    DBuf<int32_t> colNpart;  // per global column: number of partial slots
    int64_t numCols = 0, numPart = 0;
    int32_t numLeaves = 0;
    DBuf<int32_t> leafPtr;   // [S+1] CSR over leaves
    DBuf<MvEntry> entries;
    int rowBlock = 0;        // tiles of dimension >= rowBlock in a second phase-2 pass (0: none)
    int32_t numBlocks = 0;
    // leaf-level tiles of rank above b/2 as dense column-major b x b blocks
    int64_t numNear = 0;
    DBuf<int32_t> nearPtr, nearCol;  // CSR over output leaves; the block's input leaf
    DBuf<double> nearD;
    uint64_t bytes = 0;              // factor and dense-block bytes one pass reads
    uint64_t lowRankBytes = 0;       // of which U and V of the tiles applied as factors
    DBuf<int32_t> blkPtr;    // CSR over rowBlock-row blocks
    DBuf<MvEntry> blkEntries;
    const double* diag = nullptr;  // D (or D^T for the transpose plan)
    DBuf<double> diagT;            // owned transposed copy
    int chunk = 2048;              // elements of one single-column phase-1 item
    bool built = false;
    // partial leaves: vectors are gathered into / scattered from the padded space
    int64_t n = 0, npad = 0;
    bool gather = false;
    DBuf<int32_t> gmap;            // padded position -> vector index, -1 for padding
    // per-call workspaces (grown on demand)
    DBuf<double> zpart, z, Xp, Yp;
    uint32_t wsK = 0;
};

// ---- multi right-hand-side (MFMA) matvec -----------------------------------
// Phase-1 item: up to 64 columns x a chunk of rows of one tile's V
// (column-major, column stride ld), all right-hand sides at once.
struct MfItem {
    const void* v;
    int32_t xoff, rows, ld, ncols;
    int64_t out;   // partial slot of the item's first column (tiles split into several row chunks)
    int64_t zidx;  // the item's first column in the group's column order (index into zmap)
    int32_t nch;   // row chunks per column of this tile
    int32_t direct;  // 1: whole columns, the item stores z itself; 0: partial sums for the reduce
};
// Phase-2 entry: the leaf's rows of one tile's U, either a row-major copy with
// the tile's padded rank as row stride, or (FP64) the level's column-major U
// in place with column stride ld.
struct MfEntry {
    const void* u;
    int32_t rpad;
    int32_t zcol;  // first (padded) column of the tile in the group's z
    int32_t rank;  // columns present in U (column-major form)
    int32_t ld;    // column stride of the column-major form
};

// One precision group (f64, f32 or f16) of low-rank factors.
struct MfGroup {
    int type = 0;                  // 0 f64, 1 f32, 2 f16
    int64_t numColsPad = 0, numPart = 0;
    DBuf<unsigned char> V, Urm;    // owned converted storage (V empty when it points into the FP64 levels)
    DBuf<float> colScale;          // per padded column (f16 group: su*sv), else empty
    DBuf<float> colScaleV;         // the same by V column (phase 1 reads it beside zmap)
    DBuf<MfItem> items;
    int32_t numItems = 0;
    DBuf<int32_t> zmap;            // group column (V order) -> padded z column
    int64_t numRed = 0;            // columns summed from row-chunk partials
    DBuf<int64_t> redCol;          // their padded column
    DBuf<int64_t> colPart;         // their first partial slot
    DBuf<int32_t> colNpart;        // their partial count
    DBuf<int32_t> leafPtr;         // [S+1]
    DBuf<MfEntry> entries;
    bool grouped = false;          // 64-row leaves in groups of 4: entries hold only tiles of at most two
                                   // leaves, tiles spanning a group are in grpPtr/grpEntries
    DBuf<int32_t> grpPtr;          // [S/4+1]
    DBuf<MfEntry> grpEntries;      // U at the group's first row
    uint64_t factorBytes = 0;      // V + U bytes of this group (unpadded entries)
    bool vfrag = false;            // V copied in fragment order (else FP64 V in place, column-major)
    int split = 0;                 // FP32 group as FP16 hi + lo: 1 V (phase 1 on the FP16 matrix cores),
                                   // 2 V, U and z (both phases)
    int64_t vLoOff = 0, uLoOff = 0; // split: the lo copies' offsets from the hi copies (FP16 elements)
    DBuf<unsigned char> zpart, zt; // workspaces
    size_t wsNR = 0;               // pass width the workspaces are sized for
    int ztNR = 0;                  // pass width whose z padding columns were last cleared
};

// Leaf-level tile stored dense in a matrix-core plan (rank above b/2: the
// b x b block is smaller than its two factors and needs no z).
struct MfNear {
    const void* d;     // row-major b x b block (plan's dense type)
    int32_t colLeaf;   // leaf of the block's columns
    int32_t pad;
};

struct MfPlan {
    bool built = false;
    int dtype = 0;                 // 0: FP64 plan; 1: mixed FP32/FP16 plan
    bool ucm = false;              // FP64 plan reads the levels' column-major U and V in place
    MfGroup g[2];                  // FP64: g[0]; mixed: g[0] = f32 hi, g[1] = f16 lo
    DBuf<unsigned char> Drm;       // diagonal blocks, row-major (f64 or f32)
    DBuf<unsigned char> nearStore; // dense leaf-level blocks, row-major (f64 or f32)
    DBuf<int32_t> nearPtr;         // [S+1] CSR over row leaves
    DBuf<MfNear> near;
    DBuf<int32_t> gnearPtr;        // nearOtf: [S/4+1] CSR over groups of 4 leaves
    DBuf<MfNear> gnear;            // the group's column leaves (diagonals included), pad = mask of the
                                   // group's leaves that have the block
    int64_t numNear = 0;
    // leaf space of the plan (one H-matrix, or a batch's block-diagonal
    // operator): S leaves of b rows, npad = S*b positions for n vector
    // entries; gmap maps a position to its entry (-1 on padding) when gather
    int S = 0, b = 0;
    int64_t n = 0, npad = 0;
    bool gather = false;
    DBuf<int32_t> gmap;
    DBuf<unsigned char> Yp;        // product in the padded space before the scatter
    int xcdChunk = -1;             // phase-1 workgroups per XCD chunk (-1: the library default)
    DBuf<unsigned char> Xb;        // right-hand sides in the blocked layout (plan's dense type)
    DBuf<_Float16> X16;            // the same in FP16, scaled per right-hand side (mixed)
    DBuf<_Float16> X16lo;          // X/scale - X16 in FP16 (mixed with near-field blocks on the fly)
    bool nearOtf = false;          // mixed: diagonal and near-field blocks evaluated from the points
    DBuf<float> ptsL;              // nearOtf: points relative to their leaf's first point, FP32 (x, y, z, 0)
    DBuf<double> origin;           // nearOtf: each leaf's first point, FP64 (x, y, z, w)
    float nlog2e = 0.f;            // -log2(e) / correlation length (nearOtf)
    DBuf<float> xscale, xpart;     // per right-hand-side scale and its per-slice partial maxima
    size_t x16n = 0;
    float loMul = 1.f;             // 2^-E: FP16 z holds V^T x scaled by 2^E (mixed)
    float hiMul = 1.f;             // 2^(E - Ea): a split FP32 group's z holds V^T x scaled by 2^Ea
    // selection statistics (mixed)
    double frobH2 = 0, loEnergy = 0, target = 0;
    int64_t hiCols = 0, loCols = 0;
    uint64_t bytes = 0;            // factor + dense bytes read by one call
};

struct HMatrix {
    uint32_t n = 0, d = 0, b = 0, S = 0;
    int32_t numLevels = 0;
    uint64_t factorGen = 0;            // bumped whenever factors are overwritten (plans holding copies rebuild)
    std::vector<int32_t> leafOffsets;  // KD-tree segment offsets [S+1]
    DBuf<double> diag;              // [S*b*b]
    std::vector<HLevel> levels;     // index l-1
    const Structure* structure = nullptr;
    Structure structureCopy;        // owned copy (the caller may free theirs)
    MatvecPlan plan, planT;
    MatvecPlan gplan, gplanT;        // factor gradients: every tile through z (no dense near-field blocks)
    MfPlan mf64, mfMixed;
    // matrix-core plan choices (hmx_hmatrix_set_plan_options)
    int planDenseNear = 1;   // leaf-level tiles of rank > b/2 as dense blocks
    int planDirectZ = 1;     // phase 1 stores whole-column z itself (else partials + reduce)
    int planLocality = 1;    // phase-1 items ordered by right-hand-side block
    int planFp64Layout = 0;  // FP64 plan factors: 0 copy when it fits in half the free memory, 1 in place, 2 copy
    uint64_t lowRankBytes = 0, denseBytes = 0, factorEntries = 0;
    // direct assembly: the upper-level samplers' algorithmic flops, kernel
    // evaluations and kernel time (hmx_hmatrix_build_info)
    double samplerFlops = 0, kernelEvals = 0, samplerMs = 0;
    // direct assembly: the points in the padded leaf space, (x, y, z, w) per
    // position (w = 1 for a point, 0 for padding), and the kernel's
    // correlation length -- what the mixed product evaluates its diagonal and
    // near-field blocks from
    DBuf<double> pts;
    double ell = 0;
};

// Many H-matrices applied as one block-diagonal operator: one concatenated
// plan, its own stream for hipGraph capture and replay.
struct Batch {
    std::vector<HMatrix*> hs;
    MatvecPlan P;                // streaming kernels (one right-hand side)
    std::vector<uint64_t> pGen;  // instances' factorGen when P was built (its dense blocks are copies)
    MfPlan mf;                   // FP64 matrix cores (two or more), built on first use
    std::vector<uint64_t> mfGen; // instances' factorGen when mf was built
    DBuf<double> diag;
    int b = 0;
    hipStream_t stream = nullptr;
    hipEvent_t evIn = nullptr, evOut = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const double* gX = nullptr;  // arguments the graph was captured with
    double* gY = nullptr;
    uint32_t gk = 0;
    uint64_t epoch = 0;          // bumped whenever a plan or workspace the launches use changes
    uint64_t gEpoch = 0;         // epoch the graph was captured in
    void drop_graph();
    ~Batch();
};

// ---- internal entry points (implemented in the .hip/.cpp files) ----------
void generate_dataset(Context& c, uint32_t n, uint32_t d, uint64_t seed, double* d_points);
void generate_dataset_batch(Context& c, const std::vector<uint64_t>& counts, const std::vector<uint64_t>& seeds,
                            const std::vector<double*>& outs);
void generate_random_vector(Context& c, uint32_t k, uint32_t n, uint64_t seed, double* d_out);
void stream_read(Context& c, const void* buf, uint64_t bytes, int iters, float* ms);
void fp64_peak(Context& c, int unit, int iters, float* tflops);
void reorder_vectors(Context& c, uint32_t n, uint32_t k, int toOriginal, const int32_t* d_perm, const double* d_in,
                     double* d_out);
void build_kdtree(Context& c, KDTree& t, const double* d_points);
void build_kdtree_batch(Context& c, const std::vector<KDTree*>& ts, const std::vector<const double*>& pts);
void build_structure(Structure& s, uint32_t n, uint32_t d, uint32_t b, int adm, float eta, const double* bbmin,
                     const double* bbmax);
void build_structure_device(Context& c, Structure& s, uint32_t n, uint32_t d, uint32_t b, int adm, float eta,
                            const double* d_bbmin, const double* d_bbmax);
void structure_shape(Structure& s, uint32_t n, uint32_t b, int adm);
void structure_finish(Structure& s);
void build_jds(Structure::JDS& j, int level, const std::vector<int32_t>& tiles);
void build_tlr_cm(Context& c, const hmx_config& cfg, const KDTree& t, const double* d_points, TLR& out);
void build_tlr_piece(Context& c, const hmx_config& cfg, const KDTree& t, const double* d_points, uint32_t pieces,
                     uint32_t rowPiece, uint32_t colPiece, TLR& out);
void convert_to_morton(Context& c, const TLR& cm, TLR& mo);
void build_hmatrix(Context& c, const hmx_config& cfg, const KDTree& t, const Structure& s, const TLR& mo, HMatrix& h);
void build_hmatrix_direct(Context& c, const hmx_config& cfg, const KDTree& t, const Structure& s,
                          const double* d_points, HMatrix& h);
void build_hmatrix_direct_batch(Context& c, const std::vector<const hmx_config*>& cfgs,
                                const std::vector<const KDTree*>& trees, const std::vector<const Structure*>& ss,
                                const std::vector<const double*>& pts, const std::vector<HMatrix*>& hs);
// One cloud's share of a padded-leaf-space build: its leaves start at
// leafBase; tiles is its leaf level's Morton tile list; out receives the level
struct LeafJob {
    const hmx_config* cfg;
    int64_t leafBase;
    uint32_t S;
    const std::vector<int32_t>* tiles;
    HLevel* out;
};
void compress_leaf_tiles_padded(Context& c, uint32_t b, uint32_t d, const double* P, const std::vector<LeafJob>& jobs);
// symmetric tiles: the tiles of a level to compress (prim) and, per tile, its
// compressed tile (outSrc[t] = index into prim, or ~index for the transpose)
void mirror_plan(const std::vector<int32_t>& tiles, std::vector<int>& prim, std::vector<int32_t>& outSrc);
void build_dense_diagonal_padded(Context& c, uint32_t b, uint32_t d, const double* P, const std::vector<LeafJob>& jobs,
                                 double* diag);
void hmatrix_save(Context& c, const HMatrix& h, const char* path);
void hmatrix_load(Context& c, const char* path, HMatrix& h);
void build_matvec_plan(Context& c, HMatrix& h, bool transpose, MatvecPlan& P, bool denseNear);
PlanHost plan_host(const HMatrix& h, bool transpose, int chunk, int rowBlock, bool denseNear);
void upload_plan(Context& c, MatvecPlan& P, const PlanHost& H);
bool ensure_matvec_ws(MatvecPlan& P, uint32_t K);
// widest streaming-kernel pass (16, 8, 4 or 1 right-hand sides) for rem columns
uint32_t pass_width(uint32_t rem);
void run_matvec_plan(Context& c, MatvecPlan& P, int b, uint32_t k, const double* X, int64_t ldx, double* Y,
                     hipEvent_t* ev);
size_t matvec_ws_elems(const MatvecPlan& P, uint32_t K);
void run_matvec_plan_ws(Context& c, const MatvecPlan& P, const MvWork& W, int b, uint32_t k, const double* X, int64_t ldx,
                        double* Y, hipEvent_t* ev);
void build_batch_plan(Context& c, const std::vector<HMatrix*>& hs, int chunk, MatvecPlan& P, DBuf<double>& diag);
void batch_create(Context& c, const std::vector<HMatrix*>& hs, Batch& B);
void batch_matvec(Context& c, Batch& B, uint32_t k, const double* X, double* Y, bool useGraph);
void batch_tune(Context& c, Batch& B, uint32_t k, int* bestChunk, float* bestMs);
void hmatvec(Context& c, HMatrix& h, uint32_t k, const double* d_in, double* d_out, bool transpose, float* ms,
             int iters);
void dense_rows(Context& c, uint32_t n, uint32_t d, const double* p, const int32_t* perm, double ell, uint32_t nr,
                const int32_t* rows, uint32_t k, const double* X, double* Y);
void hmatvec_multi(Context& c, HMatrix& h, uint32_t k, const double* d_in, double* d_out, float* ms, int iters);
std::vector<int32_t> padding_map(const HMatrix& h, int64_t base);
bool batch_mf_stale(const Batch& B);
bool batch_mf_prepare(Context& c, Batch& B, uint32_t k);
void batch_mf_run(Context& c, Batch& B, uint32_t k, const double* d_in, double* d_out);
int default_xcd_chunk();
void hmatrix_make_mixed(Context& c, HMatrix& h, double targetRelErr);
void hmatvec_mixed(Context& c, HMatrix& h, uint32_t k, const float* d_in, float* d_out, float* ms, int iters);
void mfma_selftest(Context& c, int type, const double* A, const double* B, double* D);
void hmatvec_grad_factors(Context& c, HMatrix& h, uint32_t k, const double* X, const double* dY, double* const* dU,
                          double* const* dV);
void hmatrix_set_level(Context& c, HMatrix& h, int level, const double* U, const double* V);
// the no-allocation H-matvec: per-device scratch set-up (hmx_create) and the product
void noalloc_init(int device);
void hmatvec_noalloc(Context& c, const HMatrix& h, uint32_t k, const double* X, double* Y);
// device exclusive scan int32 -> int64 (out has count+1 entries)
void exclusive_scan_i32_i64(Context& c, const int32_t* in, int64_t* out, size_t count);

inline double default_ell(uint32_t d) { return d == 3 ? 0.2 : 0.1; }
inline int ctz32(uint32_t v) { return __builtin_ctz(v); }
inline uint32_t upper_pow2(uint32_t v) {
    v--; v |= v >> 1; v |= v >> 2; v |= v >> 4; v |= v >> 8; v |= v >> 16; v++;
    return v;
}
inline uint32_t kd_num_levels(uint32_t n, uint32_t b) {
    uint32_t it = 0;
    while (n > b) { n = (n + 1) / 2; ++it; }
    return it;
}

}  // namespace hmx

// opaque C handles are the internal objects
struct hmx_context : hmx::Context {};
struct hmx_kdtree : hmx::KDTree {};
struct hmx_structure : hmx::Structure {};
struct hmx_tlr : hmx::TLR {};
struct hmx_hmatrix : hmx::HMatrix {};
struct hmx_batch : hmx::Batch {};
