/*
 * hmx.h -- C ABI of the MI355X-native H-matrix builder and H-matvec.
 *
 * Every entry point below replaces one free function of the reference
 * (AdnanJaljuli/KAUST-Hierarchical-Matrices); the reference interface each
 * one stands in for is cited as file:line.  Plain pointers and sizes only:
 * device buffers passed in (points, vectors) are caller-owned device memory;
 * objects returned through an out-handle are owned by the library and
 * released with the matching hmx_free* call.
 *
 * Threading: one hmx_context per stream; calls on one context are not
 * re-entrant, distinct contexts may be used from distinct threads.
 * Errors: every call returns hmx_status; hmx_last_error() describes the most
 * recent failure on that context (the reference asserts/exits instead:
 * src/constructTLRMatrix.cu:75, src/helperFunctions.cuh:9-17).
 */
#ifndef HMX_H
#define HMX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum hmx_status {
    HMX_OK = 0,
    HMX_EINVAL = 1,       /* bad argument (sizes, null pointers, unsupported value) */
    HMX_ENOMEM = 2,       /* device or host allocation failed */
    HMX_EDEVICE = 3,      /* HIP runtime error */
    HMX_ERANKCAP = 4,     /* strictRankCap: a tile's ARA stopped at its rank cap before converging */
    HMX_EUNSUPPORTED = 5, /* valid in the reference but not provided (see hmx_last_error) */
    HMX_ESTATE = 6        /* object in the wrong state (e.g. TLR not in Morton order) */
} hmx_status;

/* DIVISION_METHOD (src/config.h:48-52).  FULL_TREE (the default) and
   DIVIDE_IN_HALF run the same construction, as in the reference;
   POWER_OF_TWO_ON_LEFT cuts at multiples of halving power-of-two sizes.  The
   H-matrix structure needs 2^levels leaves. */
enum { HMX_POWER_OF_TWO_ON_LEFT = 0, HMX_DIVIDE_IN_HALF = 1, HMX_FULL_TREE = 2 };
/* ADMISSIBILITY_CONDITION (src/config.h:88-91) plus the max-diameter rule */
enum { HMX_WEAK_ADMISSIBILITY = 0, HMX_BOX_CENTER_ADMISSIBILITY = 1, HMX_MAX_DIAM_ADMISSIBILITY = 2 };
/* rank caps: 0 = reference schedule (b/2 for TLR tiles, generateMaxRanks for
   upper levels); 1 = uncapped (tile dimension) */
enum { HMX_MAXRANK_REF = 0, HMX_MAXRANK_FULL = 1 };
/* ordering of a TLR matrix (src/TLRMatrix.cuh:8) */
enum { HMX_COLUMN_MAJOR = 0, HMX_MORTON = 1 };

/* Config (src/config.h:109-117) with the reference defaults (src/config.h:119-127)
   plus the parameters the reference hard-codes (src/main.cu:107,
   src/constructTLRMatrix.cu:73, src/HMatrixStructure.cu:7, src/TLRMatrixHelpers.cu:40-42). */
typedef struct hmx_config {
    int32_t divMethod;               /* HMX_FULL_TREE (default), HMX_DIVIDE_IN_HALF, HMX_POWER_OF_TWO_ON_LEFT */
    int32_t admissibilityCondition;  /* HMX_WEAK_ADMISSIBILITY */
    uint32_t numberOfInputPoints;    /* 1024 */
    uint32_t dimensionOfInputPoints; /* 2 */
    uint32_t bucketSize;             /* 32 */
    uint32_t vectorWidth;            /* 16 */
    float lowestLevelTolerance;      /* 1e-5 (float, as in the reference) */
    float eta;                       /* 1.0 (DEFAULT_ETA) */
    double correlationLength;        /* <= 0: 0.2 if d == 3 else 0.1 */
    int32_t araBlockSize;            /* 16 samples per ARA block */
    int32_t araR;                    /* 10 (ARA_R) */
    int32_t maxRankPolicy;           /* HMX_MAXRANK_REF */
    uint64_t seed;                   /* 1234 */
    /* 0 (default): a tile whose ARA reaches its rank cap before converging
       keeps the capped factors and is counted (capHits of hmx_tlr_info and
       hmx_hmatrix_level_info).  1: with the reference's rank caps
       (HMX_MAXRANK_REF) such a build fails with HMX_ERANKCAP and returns no
       object -- the reference asserts on a failed ARA instead
       (src/constructTLRMatrix.cu:75, src/constructHMatrixFromStruct.cu:80);
       uncapped ranks end at the tile dimension, where the factors are exact. */
    int32_t strictRankCap;
    /* 0 (default): every admissible tile is compressed on its own, as in the
       reference.  1: direct assembly (hmx_generateHMatrixDirect and its batch
       form) uses the symmetry of the kernel and of the admissibility -- tile
       (v, u) is the transpose of tile (u, v) -- and compresses only the tiles
       with u < v (column cluster before row cluster); tile (v, u) takes the
       same rank with U and V exchanged (its V then has the orthonormal
       columns).  Half the sampling work; the same error bound per tile.  No
       counterpart in the reference. */
    int32_t symmetricTiles;
} hmx_config;

typedef struct hmx_context hmx_context;
typedef struct hmx_kdtree hmx_kdtree;
typedef struct hmx_structure hmx_structure;
typedef struct hmx_tlr hmx_tlr;
typedef struct hmx_batch hmx_batch;
typedef struct hmx_hmatrix hmx_hmatrix;

/* ---- lifecycle -------------------------------------------------------- */
/* stream: a hipStream_t (NULL = default stream).  Replaces the implicit
   device-0/default-stream state of src/config.h:167-168 and magma_init
   (src/main.cu:62). */
hmx_status hmx_create(hmx_context** ctx, void* stream);
hmx_status hmx_destroy(hmx_context* ctx);
const char* hmx_last_error(const hmx_context* ctx);
/* parseArgs defaults, src/config.h:119-127 */
void hmx_config_default(hmx_config* cfg);
/* library version string */
const char* hmx_version(void);

/* ---- dataset ---------------------------------------------------------- */
/* generateDataset, src/generateDataset.cu:15-19.  d_points: n*d doubles, SoA
   (p[j*n+i]); uniform [0,1) from counter-based Philox4x32-10 keyed by seed. */
hmx_status hmx_generateDataset(hmx_context* ctx, uint32_t n, uint32_t d, uint64_t seed, double* d_points);
/* Many clouds at once: cloud k gets n[k]*d[k] values from seed[k], exactly
   what hmx_generateDataset gives it alone, all in one launch.  No reference
   counterpart: the reference generates one cloud per run
   (src/main.cu:51-56). */
hmx_status hmx_generateDatasetBatch(hmx_context* ctx, int32_t count, const uint32_t* n, const uint32_t* d,
                                    const uint64_t* seed, double* const* d_points);
/* generateRandomVector, src/helperFunctions.cu:59-65: n x vectorWidth
   column-major N(0,1) doubles (Philox + Box-Muller) */
hmx_status hmx_generateRandomVector(hmx_context* ctx, uint32_t vectorWidth, uint32_t n, uint64_t seed,
                                    double* d_out);

/* ---- KD-tree ---------------------------------------------------------- */
/* allocateKDTree + constructKDTree, src/kDTree.cu:6-31 and
   src/kDTreeConstruction.cu:14-231 */
hmx_status hmx_constructKDTree(hmx_context* ctx, uint32_t n, uint32_t d, uint32_t bucketSize, int32_t divMethod,
                               const double* d_points, hmx_kdtree** tree);
/* KD-trees of many clouds (one bucketSize and divMethod): the clouds of each
   point dimension are built as one forest, every level of every tree by the
   same launches; trees[k] equals hmx_constructKDTree on cloud k bit for bit
   and is freed by hmx_freeKDTree (its device arrays share blocks with the
   other trees of the call; a block is released with its last tree).  No
   reference counterpart: constructKDTree (src/kDTreeConstruction.cu:14-231)
   builds one tree per run. */
hmx_status hmx_constructKDTreeBatch(hmx_context* ctx, int32_t count, const uint32_t* n, const uint32_t* d,
                                    uint32_t bucketSize, int32_t divMethod, const double* const* d_points,
                                    hmx_kdtree** trees);
/* numSegments / segmentSize (src/kDTree.cuh:8-14) and the number of split levels */
hmx_status hmx_kdtree_info(const hmx_kdtree* tree, uint32_t* numSegments, uint32_t* segmentSize,
                           uint32_t* numLevels);
/* host copies: perm[n] (segmentIndices), offsets[numSegments+1] (segmentOffsets),
   bbmin/bbmax[(2^(numLevels+1)-1)*d] laid out [level][node][dim] (level l at node 2^l-1);
   any pointer may be NULL */
hmx_status hmx_kdtree_copy(hmx_context* ctx, const hmx_kdtree* tree, int32_t* perm, int32_t* offsets,
                           double* bbmin, double* bbmax);
/* device pointer to segmentIndices (valid while the tree lives) */
const int32_t* hmx_kdtree_device_perm(const hmx_kdtree* tree);
/* vectorWidth vectors (n x vectorWidth, column-major, ld n) between the
   caller's original point order and the KD-tree order every other entry
   point uses (position i of the KD order holds original point
   segmentIndices[i]): toOriginal = 0 gathers (out[i] = in[perm[i]]),
   toOriginal = 1 scatters (out[perm[i]] = in[i]).  Out of place.  The
   reference takes vectors in KD order only (src/main.cu:158-173); this is the
   original-order option of the drop-in contract. */
hmx_status hmx_reorderVectors(hmx_context* ctx, const hmx_kdtree* tree, uint32_t vectorWidth, int32_t toOriginal,
                              const double* d_in, double* d_out);
/* freeKDTree, src/kDTree.cu:33-36 */
hmx_status hmx_freeKDTree(hmx_kdtree* tree);

/* ---- H-matrix structure ----------------------------------------------- */
/* constructHMatrixStructure, src/HMatrixStructure.cu:133-160, called with
   upperPowerOfTwo(n) as in src/main.cu:89-96 */
hmx_status hmx_constructHMatrixStructure(hmx_context* ctx, uint32_t n, uint32_t d, uint32_t bucketSize,
                                         int32_t admissibilityCondition, float eta, const hmx_kdtree* tree,
                                         hmx_structure** structure);
/* The same structure from a pass on the GPU (the reference's recursion,
   src/HMatrixStructure.cu:9-90, run level by level over the KD-tree's device
   boxes): the same tile lists in the same order and the same JDS, bit for
   bit. */
hmx_status hmx_constructHMatrixStructureDevice(hmx_context* ctx, uint32_t n, uint32_t d, uint32_t bucketSize,
                                               int32_t admissibilityCondition, float eta, const hmx_kdtree* tree,
                                               hmx_structure** structure);
/* Same, from host bounding boxes (host only, no device needed) */
hmx_status hmx_structure_from_boxes(uint32_t n, uint32_t d, uint32_t bucketSize, int32_t admissibilityCondition,
                                    float eta, const double* bbmin, const double* bbmax,
                                    hmx_structure** structure);
/* numLevels and numTiles[numLevels] (src/HMatrixStructure.cuh:8-13); numTiles[l-1] = tiles at level l */
hmx_status hmx_structure_info(const hmx_structure* s, int32_t* numLevels, int32_t* numTiles);
/* tileIndices[level-1] (Morton codes, column in even bits) */
hmx_status hmx_structure_tiles(const hmx_structure* s, int32_t level, int32_t* tileIndices);
/* JDS of one level (src/JDS.cuh:7-15): rowIndices[2^level], columnIndices[numTiles],
   tilePos[numTiles], iterPtr[numDiagonals+1] (capacity 2^level+1) */
hmx_status hmx_structure_jds(const hmx_structure* s, int32_t level, int32_t* rowIndices, int32_t* columnIndices,
                             int32_t* tilePos, int32_t* iterPtr, int32_t* numDiagonals);
/* freeHMatrixStructure, src/HMatrixStructure.cu:162-168 */
hmx_status hmx_freeHMatrixStructure(hmx_structure* s);

/* ---- TLR matrix ------------------------------------------------------- */
/* createColumnMajorLRMatrix, src/constructTLRMatrix.cu:13-151: batched ARA of
   all off-diagonal bucketSize x bucketSize tiles, kernel tiles evaluated on
   the fly; dense diagonal blocks.  rankSum may be NULL. */
hmx_status hmx_createColumnMajorLRMatrix(hmx_context* ctx, const hmx_config* cfg, const hmx_kdtree* tree,
                                         const double* d_points, hmx_tlr** matrix, uint64_t* rankSum);
/* One piece of a pieces x pieces partition of the TLR matrix: the tiles of
   row leaves [rowPiece*S/pieces, (rowPiece+1)*S/pieces) and column leaves
   [colPiece*S/pieces, ...), column-major within the piece, each equal bit for
   bit to that tile of hmx_createColumnMajorLRMatrix (same kernel and random
   stream); a diagonal piece (rowPiece == colPiece) carries its S/pieces
   diagonal blocks.  numSegments of the piece is S/pieces.  For matrices whose
   S x S TLR does not fit at once: the reference's large-matrices branch idea
   (README.md:4; the branch is not in the snapshot).  A piece has no Morton
   form and builds no H-matrix (HMX_EINVAL there). */
hmx_status hmx_createColumnMajorLRMatrixPiece(hmx_context* ctx, const hmx_config* cfg, const hmx_kdtree* tree,
                                              const double* d_points, uint32_t pieces, uint32_t rowPiece,
                                              uint32_t colPiece, hmx_tlr** matrix, uint64_t* rankSum);
/* convertColumnMajorToMorton, src/helperFunctions.cu:8-49 */
hmx_status hmx_convertColumnMajorToMorton(hmx_context* ctx, const hmx_tlr* matrix, hmx_tlr** mortonMatrix);
/* numSegments, segmentSize, ordering, rankSum, number of rank-cap hits */
hmx_status hmx_tlr_info(const hmx_tlr* m, uint32_t* numSegments, uint32_t* segmentSize, int32_t* ordering,
                        uint64_t* rankSum, uint64_t* capHits);
/* host copies of blockRanks[S*S], blockOffsets[S*S+1] (int64), U/V [rankSum*b],
   diagonal[S*b*b]; any pointer may be NULL */
hmx_status hmx_tlr_copy(hmx_context* ctx, const hmx_tlr* m, int32_t* blockRanks, int64_t* blockOffsets, double* U,
                        double* V, double* diagonal);
/* freeMatrix, src/TLRMatrix.cu:5-11 */
hmx_status hmx_freeMatrix(hmx_tlr* m);

/* ---- hierarchical matrix ---------------------------------------------- */
/* generateMaxRanks, src/helperFunctions.cu:67-74 (maxRanks has numLevels-2 entries) */
void hmx_generateMaxRanks(uint32_t numLevels, uint32_t bucketSize, uint32_t* maxRanks);
/* allocateHMatrix + generateHMatrixFromStruct, src/HMatrix.cu:57-103 and
   src/constructHMatrixFromStruct.cu:13-103: leaf level copied from the Morton
   TLR, upper levels recompressed by batched low-rank-sum ARA. */
hmx_status hmx_generateHMatrixFromStruct(hmx_context* ctx, const hmx_config* cfg, const hmx_kdtree* tree,
                                         const hmx_structure* structure, const hmx_tlr* mortonMatrix,
                                         hmx_hmatrix** hmatrix);
/* createColumnMajorLRMatrix + convertColumnMajorToMorton + generateHMatrixFromStruct
   (src/constructTLRMatrix.cu:13-151, src/helperFunctions.cu:8-57,
   src/constructHMatrixFromStruct.cu:13-103) without the S x S TLR matrix, for
   sizes where it does not fit: diagonal blocks and leaf-level tiles are the
   TLR path's (same kernel, same random streams, bit for bit), upper levels are
   compressed by the same ARA and tolerance/rank schedule from on-the-fly
   kernel evaluations.  bucketSize 32 or 64.  d_points as for the KD-tree. */
hmx_status hmx_generateHMatrixDirect(hmx_context* ctx, const hmx_config* cfg, const hmx_kdtree* tree,
                                     const hmx_structure* structure, const double* d_points, hmx_hmatrix** hmatrix);
/* Binary H-matrix file (no counterpart in the reference, which only prints
   text dumps, src/helperFunctions.cu:76-169): save an assembled H-matrix,
   load it back on this context's device for matvecs without rebuilding.
   The loaded handle owns its copy of the structure. */
hmx_status hmx_hmatrix_save(hmx_context* ctx, const hmx_hmatrix* h, const char* path);
hmx_status hmx_hmatrix_load(hmx_context* ctx, const char* path, hmx_hmatrix** hmatrix);
/* per level l (1..numLevels-1): tile count, rank sum, max rank, tolerance used, cap hits */
hmx_status hmx_hmatrix_level_info(const hmx_hmatrix* h, int32_t level, int32_t* numTiles, int64_t* rankSum,
                                  int32_t* maxRank, float* tolerance, int32_t* capHits);
/* host copies of one level: ranks[numTiles], U/V[rankSum*m] (m = bus*b rows per column) */
hmx_status hmx_hmatrix_level_copy(hmx_context* ctx, const hmx_hmatrix* h, int32_t level, int32_t* ranks, double* U,
                                  double* V);
/* diagonal blocks [S*b*b] (HMatrix.diagonalBlocks, src/HMatrix.cuh:30-34) */
hmx_status hmx_hmatrix_diag_copy(hmx_context* ctx, const hmx_hmatrix* h, double* diag);
/* bytes of factor storage read by one matvec (low-rank U+V and dense blocks) */
hmx_status hmx_hmatrix_bytes(const hmx_hmatrix* h, uint64_t* lowRankBytes, uint64_t* denseBytes,
                             uint64_t* factorEntries);
/* Build statistics of hmx_generateHMatrixDirect (no reference counterpart):
   algorithmic flops of the upper-level samplers (2 m^2 per sample or basis
   column of an m x m tile: Y = A Omega and B = A^T Q on the FP64 matrix
   cores), kernel evaluations they made, and their kernel time in ms (hipEvents
   on the context stream).  Zeros for the other builds. */
hmx_status hmx_hmatrix_build_info(const hmx_hmatrix* h, double* samplerFlops, double* kernelEvals, double* samplerMs);
/* freeHMatrix, src/HMatrix.cu:105-110 */
hmx_status hmx_freeHMatrix(hmx_hmatrix* h);

/* ---- H-matvec ---------------------------------------------------------- */
/* HMatrixVecMult, src/HMatrixVectorMultiplication.cu:10-26: out = H * in,
   in/out are n x vectorWidth column-major (ld n) FP64 device buffers in
   KD-permuted order; out is overwritten.  Any n: with partial leaves (the
   reference needs n = S*b, src/cutlassDiagonalXVector.cu:59-65) the vectors
   are gathered into the padded leaf space and the product scattered back.
   One right-hand side runs the FP64 streaming kernels, two or more the FP64
   matrix cores (passes of 16, or 64 when vectorWidth is a multiple of 64). */
hmx_status hmx_HMatrixVecMult(hmx_context* ctx, const hmx_hmatrix* h, uint32_t vectorWidth, const double* d_in,
                              double* d_out);
/* No counterpart in the reference: for callers that cannot allocate device
   memory (code under HIP graph capture, tight serving loops).  out = H * in
   as hmx_HMatrixVecMult computes it with one right-hand side (the FP64
   streaming kernels; vectorWidth >= 2 runs them in passes of 16, 8, 4 or 1,
   not the matrix cores), calling no hipMalloc, hipFree or other allocating
   routine.  Its workspaces live in a fixed 64 MiB device array held per GPU
   until the process ends; an H-matrix whose passes need more is rejected
   with HMX_EINVAL (hmx_last_error names the size).  The H-matrix's streaming
   plan must exist: the first hmx_HMatrixVecMult call with vectorWidth 1
   builds it, else HMX_ESTATE.  Calls on one GPU share the array: calls that
   overlap (other streams or threads) race. */
hmx_status hmx_HMatrixVecMultNoAlloc(hmx_context* ctx, const hmx_hmatrix* h, uint32_t vectorWidth,
                                     const double* d_in, double* d_out);
/* out = H^T * in (no reference counterpart; the backward of HMatrixVecMult);
   any vectorWidth >= 1 and any n, on the streaming kernels in passes of 16,
   8, 4 or 1 right-hand sides */
hmx_status hmx_HMatrixVecMultTranspose(hmx_context* ctx, const hmx_hmatrix* h, uint32_t vectorWidth,
                                       const double* d_in, double* d_out);
/* Times the matvec kernels with hipEvents on the context stream over iters
   calls: ms[0] = phase 1 (V^T x), ms[1] = partial reduction, ms[2] = phase 2
   (row owner U z + D x), ms[3] = whole call.  Averages per call. */
hmx_status hmx_matvec_profile(hmx_context* ctx, const hmx_hmatrix* h, uint32_t vectorWidth, const double* d_in,
                              double* d_out, int32_t iters, float* ms);

/* ---- many right-hand sides on the matrix cores ------------------------- */
/* hmx_HMatrixVecMult with vectorWidth >= 2 runs on the FP64 matrix cores
   (passes of 64 right-hand sides when vectorWidth is a multiple of 64, else
   of 16, the last pass padded with zero columns).  Times the multi-RHS
   kernels like hmx_matvec_profile (ms[0] phase 1, ms[1] reduce, ms[2] phase
   2, ms[3] call) for vectorWidth 16 or 64; mixed != 0 times
   hmx_HMatrixVecMultMixed (in/out then point to float). */
hmx_status hmx_matvec_profile_multi(hmx_context* ctx, const hmx_hmatrix* h, uint32_t vectorWidth, const void* d_in,
                                    void* d_out, int32_t iters, int32_t mixed, float* ms);
/* Mixed precision (no reference counterpart; the precision extension of
   HMatrixVecMult): stores every low-rank column either in FP32 or in FP16
   with per-column FP32 scales, choosing FP16 for the columns of least energy
   such that the storage error stays below targetRelErr relative to ||H||_F;
   dense diagonal blocks in FP32. */
hmx_status hmx_hmatrix_make_mixed(hmx_context* ctx, hmx_hmatrix* h, double targetRelErr);
/* columns kept in FP32 / stored in FP16, squared Frobenius norm of H, energy
   of the FP16 columns, bytes read per call */
hmx_status hmx_hmatrix_mixed_info(const hmx_hmatrix* h, int64_t* hiCols, int64_t* loCols, double* frobH2,
                                  double* loEnergy, uint64_t* bytes);
/* Matvec plan choices (no counterpart in the reference); each 0 or 1, all 1
   by default (the fastest combination measured at N = 2^20, 3D).  denseNear:
   leaf-level tiles of rank above bucketSize/2 applied as dense blocks (the
   streaming and the matrix-core plans).  directZ: matrix-core phase 1 stores
   whole-column z itself instead of partial sums for the reduce.  locality:
   matrix-core phase-1 work ordered by right-hand-side block.  Drops the built
   plans. */
hmx_status hmx_hmatrix_set_plan_options(hmx_hmatrix* h, int32_t denseNear, int32_t directZ, int32_t locality);
/* Storage of the FP64 matrix-core plan's factors (no counterpart in the
   reference): 0 (default) copies them into the plan's run layout when the copy
   fits in half the free device memory, else reads the levels in place; 1 in
   place; 2 copy.  The copy costs the low-rank factor bytes again and makes
   phase 1 and phase 2 work on longer runs of columns.  A batch takes the first
   instance's choice.  Drops the built FP64 plan. */
hmx_status hmx_hmatrix_set_fp64_layout(hmx_hmatrix* h, int32_t layout);
/* Matvec plan summary (plan 0: FP64 matrix cores, built on the first call
   with two or more right-hand sides; plan 1: mixed, built by
   hmx_hmatrix_make_mixed; plan 2: the FP64 streaming kernels, built on the
   first call with one right-hand side): bytes read per call or pass (factors
   + dense blocks), leaf-level tiles applied as dense blocks (rank above
   bucketSize/2), low-rank columns.  HMX_ESTATE if not built. */
hmx_status hmx_hmatrix_plan_info(const hmx_hmatrix* h, int32_t plan, uint64_t* bytes, int64_t* denseBlocks,
                                 int64_t* lowRankColumns);
/* The same plan's bytes split by phase: phase1Bytes = the V factors it
   reads (Z = V^T X), phase2Bytes = the U factors, the dense diagonal and
   near-field blocks (or, evaluated on the fly, the points) and the column
   scales (Y = D X + near + U Z); right-hand sides and z excluded.  No
   reference counterpart (the reference runs the two phases as CUTLASS
   grouped GEMMs, src/cutlassHMatrixXVector.cu:152-328). */
hmx_status hmx_hmatrix_plan_phase_bytes(const hmx_hmatrix* h, int32_t plan, uint64_t* phase1Bytes,
                                        uint64_t* phase2Bytes);
/* Y = H X with X, Y n x vectorWidth column-major FP32, any vectorWidth >= 1
   (passes of 64 or 16 right-hand sides, the last padded) and any n (partial
   leaves through the padded leaf space); FP16/FP32 matrix cores with FP32
   accumulation.  Each right-hand side is converted to FP16 divided by its
   max-abs, and that scale is applied again in FP32, so the relative error
   does not depend on the magnitude of the input. */
hmx_status hmx_HMatrixVecMultMixed(hmx_context* ctx, const hmx_hmatrix* h, uint32_t vectorWidth, const float* d_in,
                                   float* d_out);
/* One MFMA of each element type with the lane layout the kernels assume:
   D(16x16) = A(16xK) B(Kx16), K = 4 (type 0 f64, 1 f32), 16 (type 2 f16) or
   32 (type 3 f16);
   A, B, D row-major FP64 device buffers.  A self-check of the matrix-core
   operand layout. */
hmx_status hmx_selftest_mfma(hmx_context* ctx, int32_t type, const double* d_A, const double* d_B, double* d_D);
/* Diagnostic, no reference counterpart: average ms of one pass that reads
   the caller's device buffer (bytes >= 1 MiB, a multiple of 16) once with
   16-byte non-temporal loads -- the measured streaming-read ceiling the
   H-matvec bandwidth is compared against beside the spec peak. */
hmx_status hmx_stream_read_probe(hmx_context* ctx, const void* d_buf, uint64_t bytes, int32_t iters, float* ms);
/* Diagnostic, no reference counterpart: measured FP64 arithmetic rate of
   this GPU in TFLOP/s, unit 0 = the matrix cores (v_mfma_f64_16x16x4f64),
   1 = vector FMA, from `iters` rounds of independent register-only chains in
   8 waves per SIMD -- the ceiling the FP64 matrix-core products and the
   samplers are compared against beside the spec peak. */
hmx_status hmx_fp64_peak_probe(hmx_context* ctx, int32_t unit, int32_t iters, float* tflops);

/* ---- differentiation (no reference counterpart: the backward of
   HMatrixVecMult, for autograd callers) ------------------------------------ */
/* Gradients of a loss with output gradient dY (Y = H X) with respect to every
   tile's factors, written into per-level buffers laid out like the level's
   U/V (dU[l-1], dV[l-1] for level l; hmx_hmatrix_level_info gives sizes).
   X, dY: n x vectorWidth column-major FP64 device buffers, any vectorWidth
   >= 1 and any n (rows of the padding positions of partial leaves get zero
   gradients).  The input gradient is hmx_HMatrixVecMultTranspose(dY). */
hmx_status hmx_hmatvec_grad_factors(hmx_context* ctx, const hmx_hmatrix* h, uint32_t vectorWidth, const double* d_X,
                                    const double* d_dY, double* const* d_dU, double* const* d_dV);
/* Copy one level's factors into device buffers (same layout) */
hmx_status hmx_hmatrix_level_get(hmx_context* ctx, const hmx_hmatrix* h, int32_t level, double* d_U, double* d_V);
/* Overwrite one level's factors from device buffers in the level's layout
   (e.g. after an optimizer step on them). */
hmx_status hmx_hmatrix_set_level(hmx_context* ctx, hmx_hmatrix* h, int32_t level, const double* d_U,
                                 const double* d_V);

/* ---- dense check (EXPAND_MATRIX oracle on sampled rows) ----------------- */
/* checkErrorInHmatrixVecMult / ReferenceGemm_kernel, src/expandMatrix.cu:174-234,
   restricted to nrows sampled rows: Y[r,:] = sum_j K(x_perm[rows[r]], x_perm[j]) X[j,:]
   (Y nrows x k, X n x k, column-major, device buffers). */
hmx_status hmx_denseRowsXVector(hmx_context* ctx, uint32_t n, uint32_t d, const double* d_points,
                                const int32_t* d_perm, double correlationLength, uint32_t nrows,
                                const int32_t* d_rows, uint32_t k, const double* d_X, double* d_Y);

/* ---- many point clouds (no counterpart in the reference, whose driver
   handles one point cloud per run, src/main.cu:46-173) ------------------- */
/* Assemble `count` H-matrices directly (as hmx_generateHMatrixDirect) in one
   go: the padded points and the diagonal blocks of all instances in one
   launch each, the leaf levels in one launch sequence per point dimension,
   then the upper levels of all instances together -- one batched ARA per
   (tile dimension, point dimension) group, each tile with its own instance's
   tolerance schedule, rank cap, random stream and correlation length.  The
   instances' factors share device blocks (a block is released with its last
   instance).  cfgs, trees, structures, d_points and hmatrices are arrays of
   `count`; araR and seed must agree across the batch. */
hmx_status hmx_generateHMatrixDirectBatch(hmx_context* ctx, int32_t count, const hmx_config* cfgs,
                                          hmx_kdtree* const* trees, hmx_structure* const* structures,
                                          const double* const* d_points, hmx_hmatrix** hmatrices);
/* Concatenate the matvec plans of `count` H-matrices (same bucketSize, any n):
   the batch applies Y = blockdiag(H_0, H_1, ...) X with one launch per phase.
   Instance i's entries sit at positions [sum_{j<i} n_j, sum_{j<=i} n_j) of every
   right-hand side (length sum n_j, KD order of each instance).  The H-matrices
   must outlive the batch. */
hmx_status hmx_batch_create(hmx_context* ctx, int32_t count, hmx_hmatrix* const* hmatrices, hmx_batch** batch);
hmx_status hmx_batch_info(const hmx_batch* batch, int64_t* totalPoints, int64_t* totalLeaves, int32_t* chunk);
/* FP64, any vectorWidth >= 1 (two or more run on the matrix cores, in passes
   of 16 or 64 columns).  useGraph != 0: the launch
   sequence is captured in a hipGraph on first use (again when d_in, d_out or
   vectorWidth change) and replayed, ordered after prior work on the context
   stream. */
hmx_status hmx_batch_matvec(hmx_context* ctx, hmx_batch* batch, uint32_t vectorWidth, const double* d_in,
                            double* d_out, int32_t useGraph);
/* Autotuner for the path vectorWidth takes; keeps the fastest candidate and
   returns it in bestChunk.  vectorWidth 1: the streaming phase-1 chunk length
   (512, 1024, 2048 or 4096; also reported by hmx_batch_info).  vectorWidth >= 2:
   the matrix-core phase-1 workgroups per XCD chunk (0 = plain order, 8, 16, 32). */
hmx_status hmx_batch_tune(hmx_context* ctx, hmx_batch* batch, uint32_t vectorWidth, int32_t* bestChunk,
                          float* bestMs);
hmx_status hmx_batch_free(hmx_batch* batch);

/* ---- test hooks (no reference counterpart) ----------------------------- */
/* Forced device-allocation failure for the error-path tests: after this call
   the failAt-th device allocation the library makes (0-based, counted from
   this call, all contexts and threads) fails as an out-of-memory would, with
   HMX_ENOMEM from the entry point that made it; failAt < 0 fails none and
   only counts.  hmx_debug_alloc_count returns the allocations made since the
   last hmx_debug_set_alloc_fault call, hmx_debug_live_bytes the device bytes
   the library holds right now (all objects, all contexts). */
void hmx_debug_set_alloc_fault(int64_t failAt);
int64_t hmx_debug_alloc_count(void);
int64_t hmx_debug_live_bytes(void);

#ifdef __cplusplus
}
#endif

#endif /* HMX_H */
