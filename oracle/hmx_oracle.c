/*
 * hmx_oracle.c -- CPU restatement of the reference H-matrix pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it may be loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * by nothing else.  The product path (libhmx.so) never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference checkout).  Pinning of each function (what independent check
 * fixes it) is listed in the header of oracle/oracle.py.
 *
 * Layout conventions (shared with the HIP library):
 *   points   : SoA, p[j*n + i]                       (src/generateDataset.cu:5-13)
 *   perm     : position -> original point index      (src/kDTree.cuh:11)
 *   offsets  : leaf segment offsets, S+1 entries     (src/kDTree.cuh:12)
 *   bb       : [level][node][dim], level l starts at node (2^l - 1)
 *   TLR tile : b x b, column-major, rows padded to b (src/TLRMatrixHelpers.cu:60-101)
 *   H level  : tiles in structure order, U/V columns of length bus*b
 *              (src/HMatrix.cu:12-48)
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------- */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as    */
/* easy as 1, 2, 3", SC'11).  Replaces cuRAND (src/generateDataset.cu:9-11).  */
/* ------------------------------------------------------------------------- */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

#define TAG_POINTS 0x504F0000u
#define TAG_OMEGA 0x4F4D0000u
#define TAG_RHS 0x52480000u

OR_API void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += PHILOX_W0; k1 += PHILOX_W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline double u53(uint32_t a, uint32_t b) {
    return (double)(((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6)) * 0x1.0p-53;
}

/* standard normal from one Philox block: Box-Muller, cosine branch */
static inline double gauss_at(uint64_t seed, uint32_t tag, uint64_t a, uint32_t b, uint32_t c) {
    uint32_t ctr[4] = {c, b, (uint32_t)a, tag ^ (uint32_t)(a >> 32)};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    or_philox4x32_10(ctr, key, o);
    double u1 = u53(o[0], o[1]) + 0x1.0p-53; /* (0, 1] */
    double u2 = u53(o[2], o[3]);             /* [0, 1) */
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

OR_API double or_gauss(uint64_t seed, uint32_t tag, uint64_t a, uint32_t b, uint32_t c) {
    return gauss_at(seed, tag, a, b, c);
}

/* src/generateDataset.cu:5-19 -- i.i.d. uniform coordinates, SoA p[j*n+i]. */
OR_API void or_generate_points(int n, int d, uint64_t seed, double* p) {
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (uint64_t e = 0; e < (uint64_t)n * d; ++e) {
        uint32_t ctr[4] = {(uint32_t)e, (uint32_t)(e >> 32), 0u, TAG_POINTS};
        uint32_t o[4];
        or_philox4x32_10(ctr, key, o);
        p[e] = u53(o[0], o[1]);
    }
}

/* N x k column-major N(0,1) right-hand sides (replaces src/helperFunctions.cu:59-65). */
OR_API void or_generate_rhs(int n, int k, uint64_t seed, double* x) {
    for (int s = 0; s < k; ++s)
        for (int i = 0; i < n; ++i) x[(size_t)s * n + i] = gauss_at(seed, TAG_RHS, 0, (uint32_t)s, (uint32_t)i);
}

/* ------------------------------------------------------------------------- */
/* Small helpers (src/kDTreeHelpers.cu:107-128)                              */
/* ------------------------------------------------------------------------- */
OR_API int or_upper_pow2(int v) {
    v--; v |= v >> 1; v |= v >> 2; v |= v >> 4; v |= v >> 8; v |= v >> 16; v++;
    return v;
}

/* getMaxSegmentSize(n, b).second (src/kDTreeHelpers.cu:118-128) */
OR_API int or_kd_num_levels(int n, int b) {
    int it = 0;
    while (n > b) { n = (n + 1) / 2; ++it; }
    return it;
}

/* ------------------------------------------------------------------------- */
/* Morton helpers (src/helperKernels.cuh:87-125)                             */
/* ------------------------------------------------------------------------- */
static uint32_t spread16(uint32_t x) {
    x = (x | (x << 8)) & 0x00FF00FFu;
    x = (x | (x << 4)) & 0x0F0F0F0Fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}
static uint32_t compact16(uint32_t x) {
    x &= 0x55555555u;
    x = (x | (x >> 1)) & 0x33333333u;
    x = (x | (x >> 2)) & 0x0F0F0F0Fu;
    x = (x | (x >> 4)) & 0x00FF00FFu;
    x = (x | (x >> 8)) & 0x0000FFFFu;
    return x;
}
/* getMOfromXY: x (column) in even bits, y (row) in odd bits */
OR_API int or_morton_from_xy(uint32_t x, uint32_t y) { return (int)(spread16(x) | (spread16(y) << 1)); }
/* CMIndextoMOIndex: n -> (row=n%S, col=n/S) */
OR_API int or_cm_to_mo(int S, int n) { return or_morton_from_xy((uint32_t)(n / S), (uint32_t)(n % S)); }
/* mortonToCM */
OR_API void or_mo_to_xy(uint32_t d, uint32_t* x, uint32_t* y) { *x = compact16(d); *y = compact16(d >> 1); }

/* ------------------------------------------------------------------------- */
/* KD-tree, FULL_TREE division (src/kDTreeConstruction.cu:14-231)            */
/* ------------------------------------------------------------------------- */
/* CUB radix-sort key twiddle for doubles: orders like IEEE total order       */
static inline uint64_t dkey(double v) {
    uint64_t u;
    memcpy(&u, &v, 8);
    return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}

typedef struct { uint64_t key; int pos; int val; } kd_item;
static int kd_cmp(const void* a, const void* b) {
    const kd_item* x = (const kd_item*)a;
    const kd_item* y = (const kd_item*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->pos - y->pos; /* stability */
}

/*
 * perm[n], offsets[S+1] (S = 2^levels), bbmin/bbmax [(2^(levels+1)-1) * d]
 * (BB levels 0..levels-1 are filled, the leaf level is left untouched,
 *  exactly as src/boundingBoxes.cu:23-36 + src/kDTreeConstruction.cu:99-204).
 * split_dims[2^levels - 1] receives the split dimension of each inner node.
 * Returns the number of levels.
 */
OR_API int or_kdtree(int n, int d, int b, const double* p, int* perm, int* offsets, double* bbmin, double* bbmax,
                     int* split_dims) {
    int levels = or_kd_num_levels(n, b);
    int S = 1 << levels;
    int* off = (int*)malloc(sizeof(int) * (S + 1));
    int* noff = (int*)malloc(sizeof(int) * (S + 1));
    kd_item* items = (kd_item*)malloc(sizeof(kd_item) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; ++i) perm[i] = i; /* initIndexMap, src/kDTreeHelpers.cu:12-26 */
    off[0] = 0; off[1] = n;
    int nseg = 1;
    for (int level = 0; level < levels; ++level) {
        int base = (1 << level) - 1;
        for (int s = 0; s < nseg; ++s) {
            int a = off[s], e = off[s + 1];
            /* segmented Min/Max (src/kDTreeConstruction.cu:113-127); empty -> identity */
            double* mn = bbmin + (size_t)(base + s) * d;
            double* mx = bbmax + (size_t)(base + s) * d;
            for (int j = 0; j < d; ++j) { mn[j] = DBL_MAX; mx[j] = -DBL_MAX; }
            for (int i = a; i < e; ++i)
                for (int j = 0; j < d; ++j) {
                    double v = p[(size_t)j * n + perm[i]];
                    if (v < mn[j]) mn[j] = v;
                    if (v > mx[j]) mx[j] = v;
                }
            /* findSpan + ArgMax, first maximum wins (src/kDTreeHelpers.cu:52-64) */
            int dim = 0;
            double best = mx[0] - mn[0];
            for (int j = 1; j < d; ++j) {
                double sp = mx[j] - mn[j];
                if (sp > best) { best = sp; dim = j; }
            }
            if (split_dims) split_dims[base + s] = dim;
            /* stable ascending segmented sort (src/kDTreeConstruction.cu:165-174) */
            for (int i = a; i < e; ++i) {
                items[i - a].key = dkey(p[(size_t)dim * n + perm[i]]);
                items[i - a].pos = i - a;
                items[i - a].val = perm[i];
            }
            qsort(items, (size_t)(e - a), sizeof(kd_item), kd_cmp);
            for (int i = a; i < e; ++i) perm[i] = items[i - a].val;
            /* fillOffsetsSort: left child takes the ceil half (src/kDTreeHelpers.cu:81-93) */
            noff[2 * s] = a;
            noff[2 * s + 1] = (e - a + 1) / 2 + a;
        }
        noff[2 * nseg] = n;
        nseg *= 2;
        memcpy(off, noff, sizeof(int) * (nseg + 1));
    }
    memcpy(offsets, off, sizeof(int) * (nseg + 1));
    free(off); free(noff); free(items);
    return levels;
}

/*
 * POWER_OF_TWO_ON_LEFT division (src/kDTreeConstruction.cu:77-82, 90-95,
 * 101-105, 154-156, 180-183, 206-208; fillOffsets src/kDTreeHelpers.cu:29-43):
 * the segments of a level are [s*size, min((s+1)*size, n)) with size starting
 * at upperPowerOfTwo(n) and halving while size > b; each segment is sorted by
 * its widest coordinate as in FULL_TREE.  Returns the number of levels; the
 * leaf count ceil(n/size) goes to *numSegments.
 */
OR_API int or_kdtree_pow2left(int n, int d, int b, const double* p, int* perm, int* offsets, double* bbmin,
                              double* bbmax, int* split_dims, int* numSegments) {
    kd_item* items = (kd_item*)malloc(sizeof(kd_item) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; ++i) perm[i] = i;
    int size = or_upper_pow2(n);
    int level = 0;
    while (size > b) {
        int nseg = (n + size - 1) / size;
        int base = (1 << level) - 1;
        for (int s = 0; s < nseg; ++s) {
            int a = s * size, e = (s + 1) * size < n ? (s + 1) * size : n;
            double* mn = bbmin + (size_t)(base + s) * d;
            double* mx = bbmax + (size_t)(base + s) * d;
            for (int j = 0; j < d; ++j) { mn[j] = DBL_MAX; mx[j] = -DBL_MAX; }
            for (int i = a; i < e; ++i)
                for (int j = 0; j < d; ++j) {
                    double v = p[(size_t)j * n + perm[i]];
                    if (v < mn[j]) mn[j] = v;
                    if (v > mx[j]) mx[j] = v;
                }
            int dim = 0;
            double best = mx[0] - mn[0];
            for (int j = 1; j < d; ++j) {
                double sp = mx[j] - mn[j];
                if (sp > best) { best = sp; dim = j; }
            }
            if (split_dims) split_dims[base + s] = dim;
            for (int i = a; i < e; ++i) {
                items[i - a].key = dkey(p[(size_t)dim * n + perm[i]]);
                items[i - a].pos = i - a;
                items[i - a].val = perm[i];
            }
            qsort(items, (size_t)(e - a), sizeof(kd_item), kd_cmp);
            for (int i = a; i < e; ++i) perm[i] = items[i - a].val;
        }
        size >>= 1;
        ++level;
    }
    int nseg = (n + size - 1) / size;
    for (int s = 0; s < nseg; ++s) offsets[s] = s * size;
    offsets[nseg] = n;
    *numSegments = nseg;
    free(items);
    return level;
}

/* ------------------------------------------------------------------------- */
/* Admissibility (src/admissibilityCondition.cu:7-68)                        */
/* ------------------------------------------------------------------------- */
enum { ADM_WEAK = 0, ADM_BOX_CENTER = 1, ADM_MAX_DIAM = 2 };

static double bb_center_distance(const double* mxu, const double* mnu, const double* mxv, const double* mnv, int d) {
    double distance = 0;
    for (int i = 0; i < d; ++i) {
        double c = 0.5 * ((mxu[i] + mnu[i]) - (mxv[i] + mnv[i]));
        distance += c * c;
    }
    return sqrt(distance);
}
static double bb_diameter(const double* mx, const double* mn, int d) {
    double diameter = 0;
    for (int i = 0; i < d; ++i) {
        double t = mx[i] - mn[i];
        diameter += t * t;
    }
    return sqrt(diameter);
}

typedef struct {
    int d, maxDepth, adm;
    float eta;
    const double *bbmin, *bbmax;
    int* numTiles;
    int** tiles;
} st_ctx;

static int admissible(const st_ctx* c, int level, int u, int v) {
    if (u == v) return 0;
    if (c->adm == ADM_WEAK) return 1;
    int base = (1 << level) - 1;
    const double *mxu = c->bbmax + (size_t)(base + u) * c->d, *mnu = c->bbmin + (size_t)(base + u) * c->d;
    const double *mxv = c->bbmax + (size_t)(base + v) * c->d, *mnv = c->bbmin + (size_t)(base + v) * c->d;
    double distance = bb_center_distance(mxu, mnu, mxv, mnv, c->d);
    double du = bb_diameter(mxu, mnu, c->d);
    double dv = bb_diameter(mxv, mnv, c->d);
    if (du == 0 || dv == 0) return 0;
    if (c->adm == ADM_BOX_CENTER) return 0.5 * (du + dv) <= c->eta * distance;
    return (du > dv ? du : dv) <= c->eta * distance; /* maxDiam extension */
}

/* src/HMatrixStructure.cu:9-90 */
static void st_rec(st_ctx* c, int level, int u, int v) {
    int diag = (u == v);
    int leaf = (level == c->maxDepth);
    if (diag && leaf) return;
    if (leaf || admissible(c, level, u, v)) {
        int numRows = 1 << level;
        int ti = or_cm_to_mo(numRows, u * numRows + v);
        c->tiles[level - 1][c->numTiles[level - 1]++] = ti;
        return;
    }
    st_rec(c, level + 1, 2 * u, 2 * v);
    st_rec(c, level + 1, 2 * u + 1, 2 * v);
    st_rec(c, level + 1, 2 * u, 2 * v + 1);
    st_rec(c, level + 1, 2 * u + 1, 2 * v + 1);
}

/*
 * constructHMatrixStructure (src/HMatrixStructure.cu:133-160) called with
 * upperPowerOfTwo(N) (src/main.cu:89-96).  numTiles[numLevels-1]; tiles is a
 * caller buffer with per-level capacity 2^(l+1)*(2^(l+1)-1) laid out
 * consecutively (level l-1 at tile_base[l-1]).  Returns numLevels.
 */
OR_API int or_structure(int n, int d, int b, int adm, float eta, const double* bbmin, const double* bbmax,
                        int* numTiles, int* tiles, long long* tile_base) {
    int n2 = or_upper_pow2(n);
    int numLevels = __builtin_ctz((unsigned)(n2 / b)) + 1;
    int** lv = (int**)malloc(sizeof(int*) * numLevels);
    long long acc = 0;
    for (int l = 0; l < numLevels; ++l) {
        numTiles[l] = 0;
        long long cap = (long long)(1 << (l + 1)) * ((1 << (l + 1)) - 1);
        tile_base[l] = acc;
        lv[l] = tiles + acc;
        acc += cap;
    }
    st_ctx c = {d, numLevels - 1, adm, eta, bbmin, bbmax, numTiles, lv};
    if (numLevels > 1) st_rec(&c, 0, 0, 0);
    free(lv);
    return numLevels;
}

/*
 * JDS per level (src/JDS.cuh:7-15 declares rowIndices/columnIndices/iterPtr
 * but src/JDS.cu:5-16 only fills rowIndices).  Definition used here: block
 * rows sorted by descending tile count (stable); entry i of jagged diagonal j
 * is the j-th tile (ascending column) of row rowIndices[i]; iterPtr[j] is the
 * start of diagonal j; tilePos gives the tile's position in the level list.
 * Buffers sized numRows (rowIndices), ntiles (columnIndices, tilePos),
 * numRows+1 (iterPtr).  Returns the number of jagged diagonals.
 */
OR_API int or_jds(int level, int ntiles, const int* tiles, int* rowIndices, int* columnIndices, int* tilePos,
                  int* iterPtr) {
    int numRows = 1 << level;
    int* cnt = (int*)calloc(numRows, sizeof(int));
    int* start = (int*)calloc(numRows + 1, sizeof(int));
    int* rowTiles = (int*)malloc(sizeof(int) * (ntiles > 0 ? ntiles : 1));
    for (int t = 0; t < ntiles; ++t) {
        uint32_t x, y;
        or_mo_to_xy((uint32_t)tiles[t], &x, &y);
        cnt[y]++;
    }
    for (int r = 0; r < numRows; ++r) start[r + 1] = start[r] + cnt[r];
    int* fill = (int*)calloc(numRows, sizeof(int));
    for (int t = 0; t < ntiles; ++t) { /* structure order is ascending Morton = ascending column per row */
        uint32_t x, y;
        or_mo_to_xy((uint32_t)tiles[t], &x, &y);
        rowTiles[start[y] + fill[y]++] = t;
    }
    /* stable sort rows by descending count (counting sort, counts <= numRows) */
    int maxc = 0;
    for (int r = 0; r < numRows; ++r) if (cnt[r] > maxc) maxc = cnt[r];
    int k = 0;
    for (int c = maxc; c >= 0; --c)
        for (int r = 0; r < numRows; ++r) if (cnt[r] == c) rowIndices[k++] = r;
    int pos = 0;
    for (int j = 0; j < maxc; ++j) {
        iterPtr[j] = pos;
        for (int i = 0; i < numRows; ++i) {
            int r = rowIndices[i];
            if (cnt[r] <= j) break;
            int t = rowTiles[start[r] + j];
            uint32_t x, y;
            or_mo_to_xy((uint32_t)tiles[t], &x, &y);
            columnIndices[pos] = (int)x;
            tilePos[pos] = t;
            ++pos;
        }
    }
    iterPtr[maxc] = pos;
    free(cnt); free(start); free(rowTiles); free(fill);
    return maxc;
}

/* ------------------------------------------------------------------------- */
/* Kernel function (src/TLRMatrixHelpers.cu:40-58); its                     */
/* diff += (x - y)*(x - y) is one fused multiply-add under nvcc's default    */
/* contraction (--fmad=true), so it is written as fma here                   */
/* ------------------------------------------------------------------------- */
static inline double kfun(const double* p, int n, int d, int a, int b2, double ell) {
    double diff = 0;
    for (int j = 0; j < d; ++j) {
        double x = p[(size_t)j * n + a], y = p[(size_t)j * n + b2];
        diff = fma(x - y, x - y, diff);
    }
    return exp(-sqrt(diff) / ell);
}

OR_API double or_default_ell(int d) { return d == 3 ? 0.2 : 0.1; }

/* ------------------------------------------------------------------------- */
/* Adaptive randomized approximation (replaces kblas_ara_batch,              */
/* src/constructTLRMatrix.cu:70-75, and lr_kblas_ara_batch,                  */
/* src/constructHMatrixFromStruct.cu:76-80).  Restated from Halko, Martinsson */
/* & Tropp, SIAM Rev. 2011, Alg. 4.2, blocked with bs samples per iteration   */
/* and Cholesky-QR orthogonalisation (Boukaram, Turkiyyah & Keyes, SISC 2019)*/
/* ------------------------------------------------------------------------- */
typedef struct {
    int m, n;
    void (*sample)(void* ctx, const double* Om, int bs, double* Y);  /* Y(m x bs) = A Om(n x bs) */
    void (*sampleT)(void* ctx, const double* Q, int k, double* B);   /* B(n x k)  = A^T Q(m x k) */
    void* ctx;
} ara_op;

OR_API double or_ara_threshold(double tol) { return tol / (10.0 * sqrt(2.0 / 3.141592653589793)); }

/* Q(m x k) ; W(k x c) = Q^T Y(m x c) ; Y -= Q W */
static void project_out(int m, int k, const double* Q, double* Y, int c, double* W) {
    if (k == 0) return;
    for (int s = 0; s < c; ++s)
        for (int a = 0; a < k; ++a) {
            double t = 0;
            for (int i = 0; i < m; ++i) t += Q[(size_t)a * m + i] * Y[(size_t)s * m + i];
            W[s * k + a] = t;
        }
    for (int s = 0; s < c; ++s)
        for (int a = 0; a < k; ++a) {
            double w = W[s * k + a];
            for (int i = 0; i < m; ++i) Y[(size_t)s * m + i] -= Q[(size_t)a * m + i] * w;
        }
}

/* y -= b_a (b_a^T y) for a = 0..nb-1 in turn (modified Gram-Schmidt step) */
static void mgs_step(int m, const double* B, int nb, double* y) {
    for (int a = 0; a < nb; ++a) {
        double t = 0;
        for (int i = 0; i < m; ++i) t += B[(size_t)a * m + i] * y[i];
        for (int i = 0; i < m; ++i) y[i] -= B[(size_t)a * m + i] * t;
    }
}

static double norm2(int m, const double* y) {
    double t = 0;
    for (int i = 0; i < m; ++i) t += y[i] * y[i];
    return sqrt(t);
}

/*
 * generic ARA; U (m x maxRank), V (n x maxRank) column-major.  Returns rank.
 * Per block of bs samples: Y = A Omega; Y -= Q(Q^T Y) twice; then column by
 * column, project out the columns already accepted from this block (two
 * modified Gram-Schmidt sweeps)
 * and accept the column when its residual norm exceeds tau (HMT Alg. 4.2
 * test), the run of consecutive small columns carrying across blocks and
 * ending the ARA at r; the accepted columns are projected against Q once
 * more and re-normalised before being appended.  Norms are taken of the
 * explicitly projected vectors (no Gram matrix), so a decision is only as
 * uncertain as eps*|y| rather than sqrt(eps)*|y|.
 */
static int ara_run(const ara_op* op, uint64_t seed, uint64_t opid, double tau, int bs, int r, int maxRank, double* U,
                   double* V, int* capped_out) {
    int m = op->m, n = op->n;
    double* Om = (double*)malloc(sizeof(double) * (size_t)n * bs);
    double* Y = (double*)malloc(sizeof(double) * (size_t)m * bs);
    double* Qb = (double*)malloc(sizeof(double) * (size_t)m * bs);
    double* W = (double*)malloc(sizeof(double) * (size_t)(maxRank > 0 ? maxRank : 1) * bs);
    int k = 0, small_run = 0, converged = 0, capped = 0, it = 0;
    const int maxIters = (2 * maxRank + r) / bs + 2;
    while (!converged && !capped && k < maxRank && it < maxIters) {
        for (int s = 0; s < bs; ++s)
            for (int c = 0; c < n; ++c)
                Om[(size_t)s * n + c] = gauss_at(seed, TAG_OMEGA, opid, (uint32_t)(it * bs + s), (uint32_t)c);
        op->sample(op->ctx, Om, bs, Y);
        project_out(m, k, U, Y, bs, W);
        project_out(m, k, U, Y, bs, W);
        int na = 0, capleft = maxRank - k;
        for (int j = 0; j < bs; ++j) {
            double* y = Y + (size_t)j * m;
            mgs_step(m, Qb, na, y);
            mgs_step(m, Qb, na, y);
            double nu = norm2(m, y);
            if (nu > tau) {
                if (na >= capleft) { capped = 1; break; }
                for (int i = 0; i < m; ++i) Qb[(size_t)na * m + i] = y[i] / nu;
                ++na;
                small_run = 0;
            } else if (++small_run >= r) {
                converged = 1;
                break;
            }
        }
        if (na > 0) {
            project_out(m, k, U, Qb, na, W);
            for (int j = 0; j < na; ++j) {
                double* q = Qb + (size_t)j * m;
                mgs_step(m, Qb, j, q);
                double nu = norm2(m, q);
                for (int i = 0; i < m; ++i) {
                    q[i] /= nu;
                    U[(size_t)(k + j) * m + i] = q[i];
                }
            }
            k += na;
        }
        ++it;
    }
    if (!converged && k >= maxRank) capped = 1;
    if (k > 0) op->sampleT(op->ctx, U, k, V);
    if (capped_out) *capped_out = capped;
    free(Om); free(Y); free(Qb); free(W);
    return k;
}

/* dense explicit matrix sampler (used for the fixture check and tiles) */
typedef struct { int m, n; const double* A; } dense_ctx; /* A column-major m x n */
static void dense_sample(void* c, const double* Om, int bs, double* Y) {
    dense_ctx* x = (dense_ctx*)c;
    for (int s = 0; s < bs; ++s)
        for (int i = 0; i < x->m; ++i) {
            double t = 0;
            for (int j = 0; j < x->n; ++j) t += x->A[(size_t)j * x->m + i] * Om[(size_t)s * x->n + j];
            Y[(size_t)s * x->m + i] = t;
        }
}
static void dense_sampleT(void* c, const double* Q, int k, double* B) {
    dense_ctx* x = (dense_ctx*)c;
    for (int s = 0; s < k; ++s)
        for (int j = 0; j < x->n; ++j) {
            double t = 0;
            for (int i = 0; i < x->m; ++i) t += x->A[(size_t)j * x->m + i] * Q[(size_t)s * x->m + i];
            B[(size_t)s * x->n + j] = t;
        }
}

/* ARA of an explicit column-major m x n matrix.  tol is the user tolerance. */
OR_API int or_ara_dense(int m, int n, const double* A, double tol, int bs, int r, int maxRank, uint64_t seed,
                        uint64_t opid, double* U, double* V, int* capped) {
    dense_ctx c = {m, n, A};
    ara_op op = {m, n, dense_sample, dense_sampleT, &c};
    return ara_run(&op, seed, opid, or_ara_threshold(tol), bs, r, maxRank, U, V, capped);
}

/* ------------------------------------------------------------------------- */
/* The pipeline handle                                                       */
/* ------------------------------------------------------------------------- */
typedef struct {
    int n, d, b, S, kdLevels, numLevels, adm, bs, r, maxRankPolicy;
    float eta, tol;
    double ell;
    uint64_t seed;
    double* p;
    int *perm, *offsets, *splitDims;
    double *bbmin, *bbmax;
    int* numTiles;
    int* tiles;
    long long* tileBase;
    /* TLR (column-major ordering) */
    int* tlrRanksCM;
    long long* tlrOffCM;
    double *tlrUCM, *tlrVCM, *diag;
    int tlrCapHits;
    /* TLR (Morton ordering) */
    int* tlrRanksMO;
    long long* tlrOffMO;
    double *tlrUMO, *tlrVMO;
    /* H levels, index l-1 for l = 1..numLevels-1 */
    int** lvRanks;
    long long** lvOff; /* exclusive offsets in columns */
    double **lvU, **lvV;
    int* lvMaxRank;
    float* lvTol;
    int* lvCapHits;
} or_hm;

OR_API int or_hm_sizeof(void) { return (int)sizeof(or_hm); }

static void* xcalloc(size_t n, size_t s) { return calloc(n ? n : 1, s ? s : 1); }

/* generateMaxRanks (src/helperFunctions.cu:67-74) */
OR_API void or_generate_max_ranks(int numLevels, int b, unsigned* maxRanks) {
    for (int i = 0; i < numLevels - 2; ++i) {
        maxRanks[i] = (unsigned)b * (1u << i);
        if (i > 5) maxRanks[i] /= 4;
    }
}

OR_API or_hm* or_hm_create(int n, int d, int b, int adm, float eta, float tol, double ell, int bs, int r,
                           int maxRankPolicy, uint64_t seed, const double* points) {
    or_hm* h = (or_hm*)xcalloc(1, sizeof(or_hm));
    h->n = n; h->d = d; h->b = b; h->adm = adm; h->eta = eta; h->tol = tol; h->ell = ell; h->bs = bs; h->r = r;
    h->maxRankPolicy = maxRankPolicy; h->seed = seed;
    h->p = (double*)malloc(sizeof(double) * (size_t)n * d);
    memcpy(h->p, points, sizeof(double) * (size_t)n * d);
    return h;
}

OR_API void or_hm_destroy(or_hm* h) {
    if (!h) return;
    free(h->p); free(h->perm); free(h->offsets); free(h->splitDims); free(h->bbmin); free(h->bbmax);
    free(h->numTiles); free(h->tiles); free(h->tileBase);
    free(h->tlrRanksCM); free(h->tlrOffCM); free(h->tlrUCM); free(h->tlrVCM); free(h->diag);
    free(h->tlrRanksMO); free(h->tlrOffMO); free(h->tlrUMO); free(h->tlrVMO);
    if (h->lvRanks)
        for (int l = 0; l < h->numLevels - 1; ++l) {
            free(h->lvRanks[l]); free(h->lvOff[l]); free(h->lvU[l]); free(h->lvV[l]);
        }
    free(h->lvRanks); free(h->lvOff); free(h->lvU); free(h->lvV); free(h->lvMaxRank); free(h->lvTol);
    free(h->lvCapHits);
    free(h);
}

/* allocateKDTree + constructKDTree (src/kDTree.cu:6-31, src/kDTreeConstruction.cu:14-231) */
OR_API int or_hm_kdtree(or_hm* h) {
    int levels = or_kd_num_levels(h->n, h->b);
    h->kdLevels = levels;
    h->S = 1 << levels;
    int nb = (1 << (levels + 1)) - 1;
    h->perm = (int*)xcalloc(h->n, sizeof(int));
    h->offsets = (int*)xcalloc(h->S + 1, sizeof(int));
    h->splitDims = (int*)xcalloc(nb, sizeof(int));
    h->bbmin = (double*)xcalloc((size_t)nb * h->d, sizeof(double));
    h->bbmax = (double*)xcalloc((size_t)nb * h->d, sizeof(double));
    or_kdtree(h->n, h->d, h->b, h->p, h->perm, h->offsets, h->bbmin, h->bbmax, h->splitDims);
    return levels;
}

OR_API int or_hm_structure(or_hm* h) {
    int n2 = or_upper_pow2(h->n);
    int numLevels = __builtin_ctz((unsigned)(n2 / h->b)) + 1;
    h->numLevels = numLevels;
    long long cap = 0;
    for (int l = 0; l < numLevels; ++l) cap += (long long)(1 << (l + 1)) * ((1 << (l + 1)) - 1);
    h->numTiles = (int*)xcalloc(numLevels, sizeof(int));
    h->tileBase = (long long*)xcalloc(numLevels, sizeof(long long));
    h->tiles = (int*)xcalloc((size_t)cap, sizeof(int));
    or_structure(h->n, h->d, h->b, h->adm, h->eta, h->bbmin, h->bbmax, h->numTiles, h->tiles, h->tileBase);
    return numLevels;
}

/* point index of padded position q inside leaf 'leaf', or -1 for padding */
static inline int leaf_point(const or_hm* h, int leaf, int q) {
    int a = h->offsets[leaf], e = h->offsets[leaf + 1];
    return (q < e - a) ? h->perm[a + q] : -1;
}

/* TLR tile sampler: tile (row leaf v, col leaf u), entries evaluated on demand
   (generateDenseBlockColumn, src/TLRMatrixHelpers.cu:60-101) */
typedef struct { const or_hm* h; int u, v; double* A; } tile_ctx;

static void tile_fill(const or_hm* h, int u, int v, double* A) {
    int b = h->b;
    for (int c = 0; c < b; ++c)
        for (int q = 0; q < b; ++q) {
            int pc = leaf_point(h, u, c), pr = leaf_point(h, v, q);
            double val;
            if (pc < 0 || pr < 0) val = (u == v && c == q) ? 1.0 : 0.0;
            else val = kfun(h->p, h->n, h->d, pc, pr, h->ell);
            A[(size_t)c * b + q] = val;
        }
}

/* createColumnMajorLRMatrix (src/constructTLRMatrix.cu:13-151) */
OR_API long long or_hm_tlr(or_hm* h) {
    int S = h->S, b = h->b;
    int maxRank = h->maxRankPolicy == 0 ? b / 2 : b;
    size_t nt = (size_t)S * S;
    h->tlrRanksCM = (int*)xcalloc(nt, sizeof(int));
    h->tlrOffCM = (long long*)xcalloc(nt + 1, sizeof(long long));
    h->diag = (double*)xcalloc((size_t)S * b * b, sizeof(double));
    double* Us = (double*)xcalloc(nt * (size_t)b * maxRank, sizeof(double));
    double* Vs = (double*)xcalloc(nt * (size_t)b * maxRank, sizeof(double));
    double tau = or_ara_threshold((double)h->tol);
    int caps = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : caps)
    for (long long t = 0; t < (long long)nt; ++t) {
        int u = (int)(t / S), v = (int)(t % S);
        double* A = (double*)malloc(sizeof(double) * b * b);
        if (u == v) {
            tile_fill(h, u, v, h->diag + (size_t)u * b * b);
        } else {
            tile_fill(h, u, v, A);
            dense_ctx dc = {b, b, A};
            ara_op op = {b, b, dense_sample, dense_sampleT, &dc};
            int capped = 0;
            h->tlrRanksCM[t] = ara_run(&op, h->seed, (uint64_t)t, tau, h->bs, h->r, maxRank, Us + t * b * maxRank,
                                       Vs + t * b * maxRank, &capped);
            caps += capped;
        }
        free(A);
    }
    h->tlrCapHits = caps;
    long long acc = 0;
    for (size_t t = 0; t < nt; ++t) { h->tlrOffCM[t] = acc; acc += h->tlrRanksCM[t]; }
    h->tlrOffCM[nt] = acc;
    h->tlrUCM = (double*)xcalloc((size_t)acc * b, sizeof(double));
    h->tlrVCM = (double*)xcalloc((size_t)acc * b, sizeof(double));
    for (size_t t = 0; t < nt; ++t) {
        size_t rk = (size_t)h->tlrRanksCM[t];
        memcpy(h->tlrUCM + (size_t)h->tlrOffCM[t] * b, Us + t * b * maxRank, sizeof(double) * rk * b);
        memcpy(h->tlrVCM + (size_t)h->tlrOffCM[t] * b, Vs + t * b * maxRank, sizeof(double) * rk * b);
    }
    free(Us); free(Vs);
    return acc;
}

/* convertColumnMajorToMorton (src/helperFunctions.cu:8-57) */
OR_API void or_hm_to_morton(or_hm* h) {
    int S = h->S, b = h->b;
    size_t nt = (size_t)S * S;
    h->tlrRanksMO = (int*)xcalloc(nt, sizeof(int));
    h->tlrOffMO = (long long*)xcalloc(nt + 1, sizeof(long long));
    for (size_t i = 0; i < nt; ++i) h->tlrRanksMO[or_cm_to_mo(S, (int)i)] = h->tlrRanksCM[i];
    long long acc = 0;
    for (size_t t = 0; t < nt; ++t) { h->tlrOffMO[t] = acc; acc += h->tlrRanksMO[t]; }
    h->tlrOffMO[nt] = acc;
    h->tlrUMO = (double*)xcalloc((size_t)acc * b, sizeof(double));
    h->tlrVMO = (double*)xcalloc((size_t)acc * b, sizeof(double));
    for (size_t i = 0; i < nt; ++i) {
        int mo = or_cm_to_mo(S, (int)i);
        size_t rk = (size_t)h->tlrRanksCM[i];
        memcpy(h->tlrUMO + (size_t)h->tlrOffMO[mo] * b, h->tlrUCM + (size_t)h->tlrOffCM[i] * b, sizeof(double) * rk * b);
        memcpy(h->tlrVMO + (size_t)h->tlrOffMO[mo] * b, h->tlrVCM + (size_t)h->tlrOffCM[i] * b, sizeof(double) * rk * b);
    }
}

/* low-rank-sum sampler over bus x bus Morton sub-tiles (lr_kblas_ara_batch
   operator; sub-tile addressing src/HMatrixHelpers.cu:6-17) */
typedef struct { const or_hm* h; long long first; int bus; } lr_ctx;

static void lr_sample(void* cv, const double* Om, int bs, double* Y) {
    lr_ctx* c = (lr_ctx*)cv;
    const or_hm* h = c->h;
    int b = h->b, m = c->bus * b, n = m;
    double* t = (double*)malloc(sizeof(double) * (h->maxRankPolicy == 0 ? b : b) * bs + 1);
    memset(Y, 0, sizeof(double) * (size_t)m * bs);
    long long nsub = (long long)c->bus * c->bus;
    for (long long s = 0; s < nsub; ++s) {
        uint32_t j, i;
        or_mo_to_xy((uint32_t)s, &j, &i); /* j: column sub-block, i: row sub-block */
        long long mo = c->first + s;
        int rk = h->tlrRanksMO[mo];
        if (!rk) continue;
        const double* U = h->tlrUMO + (size_t)h->tlrOffMO[mo] * b;
        const double* V = h->tlrVMO + (size_t)h->tlrOffMO[mo] * b;
        for (int q = 0; q < bs; ++q)
            for (int a = 0; a < rk; ++a) {
                double acc = 0;
                for (int x = 0; x < b; ++x) acc += V[(size_t)a * b + x] * Om[(size_t)q * n + (size_t)j * b + x];
                t[q * rk + a] = acc;
            }
        for (int q = 0; q < bs; ++q)
            for (int a = 0; a < rk; ++a) {
                double w = t[q * rk + a];
                for (int x = 0; x < b; ++x) Y[(size_t)q * m + (size_t)i * b + x] += U[(size_t)a * b + x] * w;
            }
    }
    free(t);
}

static void lr_sampleT(void* cv, const double* Q, int k, double* B) {
    lr_ctx* c = (lr_ctx*)cv;
    const or_hm* h = c->h;
    int b = h->b, m = c->bus * b, n = m;
    double* t = (double*)malloc(sizeof(double) * b * (size_t)k + 1);
    memset(B, 0, sizeof(double) * (size_t)n * k);
    long long nsub = (long long)c->bus * c->bus;
    for (long long s = 0; s < nsub; ++s) {
        uint32_t j, i;
        or_mo_to_xy((uint32_t)s, &j, &i);
        long long mo = c->first + s;
        int rk = h->tlrRanksMO[mo];
        if (!rk) continue;
        const double* U = h->tlrUMO + (size_t)h->tlrOffMO[mo] * b;
        const double* V = h->tlrVMO + (size_t)h->tlrOffMO[mo] * b;
        for (int q = 0; q < k; ++q)
            for (int a = 0; a < rk; ++a) {
                double acc = 0;
                for (int x = 0; x < b; ++x) acc += U[(size_t)a * b + x] * Q[(size_t)q * m + (size_t)i * b + x];
                t[q * rk + a] = acc;
            }
        for (int q = 0; q < k; ++q)
            for (int a = 0; a < rk; ++a) {
                double w = t[q * rk + a];
                for (int x = 0; x < b; ++x) B[(size_t)q * n + (size_t)j * b + x] += V[(size_t)a * b + x] * w;
            }
    }
    free(t);
}

/* op id of H level tile: distinct from TLR tile ids (< 2^32) */
static inline uint64_t lv_opid(int level, int t) { return ((uint64_t)(level + 1) << 40) | (uint64_t)t; }

/* allocateHMatrix (src/HMatrix.cu:57-103) + generateHMatrixFromStruct
   (src/constructHMatrixFromStruct.cu:13-103) */
OR_API void or_hm_hmatrix(or_hm* h) {
    int L = h->numLevels, b = h->b;
    int nl = L - 1;
    h->lvRanks = (int**)xcalloc(nl, sizeof(int*));
    h->lvOff = (long long**)xcalloc(nl, sizeof(long long*));
    h->lvU = (double**)xcalloc(nl, sizeof(double*));
    h->lvV = (double**)xcalloc(nl, sizeof(double*));
    h->lvMaxRank = (int*)xcalloc(nl, sizeof(int));
    h->lvTol = (float*)xcalloc(nl, sizeof(float));
    h->lvCapHits = (int*)xcalloc(nl, sizeof(int));
    /* leaf level copied from the Morton TLR */
    {
        int l = L - 1;
        int nt = h->numTiles[l - 1];
        const int* tiles = h->tiles + h->tileBase[l - 1];
        h->lvRanks[l - 1] = (int*)xcalloc(nt, sizeof(int));
        h->lvOff[l - 1] = (long long*)xcalloc(nt + 1, sizeof(long long));
        long long acc = 0;
        for (int t = 0; t < nt; ++t) {
            h->lvRanks[l - 1][t] = h->tlrRanksMO[tiles[t]];
            h->lvOff[l - 1][t] = acc;
            acc += h->lvRanks[l - 1][t];
        }
        h->lvOff[l - 1][nt] = acc;
        h->lvU[l - 1] = (double*)xcalloc((size_t)acc * b, sizeof(double));
        h->lvV[l - 1] = (double*)xcalloc((size_t)acc * b, sizeof(double));
        for (int t = 0; t < nt; ++t) {
            size_t rk = (size_t)h->lvRanks[l - 1][t];
            memcpy(h->lvU[l - 1] + (size_t)h->lvOff[l - 1][t] * b, h->tlrUMO + (size_t)h->tlrOffMO[tiles[t]] * b,
                   sizeof(double) * rk * b);
            memcpy(h->lvV[l - 1] + (size_t)h->lvOff[l - 1][t] * b, h->tlrVMO + (size_t)h->tlrOffMO[tiles[t]] * b,
                   sizeof(double) * rk * b);
        }
        h->lvMaxRank[l - 1] = h->maxRankPolicy == 0 ? b / 2 : b;
        h->lvTol[l - 1] = h->tol;
    }
    unsigned* maxRanks = (unsigned*)xcalloc(L > 2 ? L - 2 : 1, sizeof(unsigned));
    or_generate_max_ranks(L, b, maxRanks);
    float tol = h->tol;
    for (int level = L - 2; level > 0; --level) {
        int nt = h->numTiles[level - 1];
        h->lvRanks[level - 1] = (int*)xcalloc(nt, sizeof(int));
        h->lvOff[level - 1] = (long long*)xcalloc(nt + 1, sizeof(long long));
        if (nt == 0) continue;
        int bus = 1 << (L - (level + 1));
        tol *= 2; /* src/constructHMatrixFromStruct.cu:54 */
        int m = bus * b;
        int maxRank = h->maxRankPolicy == 0 ? (int)maxRanks[L - level - 2] : m;
        if (maxRank > m) maxRank = m;
        h->lvMaxRank[level - 1] = maxRank;
        h->lvTol[level - 1] = tol;
        const int* tiles = h->tiles + h->tileBase[level - 1];
        double* Us = (double*)xcalloc((size_t)nt * m * maxRank, sizeof(double));
        double* Vs = (double*)xcalloc((size_t)nt * m * maxRank, sizeof(double));
        double tau = or_ara_threshold((double)tol);
        int caps = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : caps)
        for (int t = 0; t < nt; ++t) {
            lr_ctx c = {h, (long long)tiles[t] * bus * bus, bus};
            ara_op op = {m, m, lr_sample, lr_sampleT, &c};
            int capped = 0;
            h->lvRanks[level - 1][t] = ara_run(&op, h->seed, lv_opid(level, t), tau, h->bs, h->r, maxRank,
                                               Us + (size_t)t * m * maxRank, Vs + (size_t)t * m * maxRank, &capped);
            caps += capped;
        }
        h->lvCapHits[level - 1] = caps;
        long long acc = 0;
        for (int t = 0; t < nt; ++t) { h->lvOff[level - 1][t] = acc; acc += h->lvRanks[level - 1][t]; }
        h->lvOff[level - 1][nt] = acc;
        h->lvU[level - 1] = (double*)xcalloc((size_t)acc * m, sizeof(double));
        h->lvV[level - 1] = (double*)xcalloc((size_t)acc * m, sizeof(double));
        for (int t = 0; t < nt; ++t) {
            size_t rk = (size_t)h->lvRanks[level - 1][t];
            memcpy(h->lvU[level - 1] + (size_t)h->lvOff[level - 1][t] * m, Us + (size_t)t * m * maxRank,
                   sizeof(double) * rk * m);
            memcpy(h->lvV[level - 1] + (size_t)h->lvOff[level - 1][t] * m, Vs + (size_t)t * m * maxRank,
                   sizeof(double) * rk * m);
        }
        free(Us); free(Vs);
    }
    free(maxRanks);
}

/* ---- accessors for tests ---- */
OR_API int or_hm_get_int(const or_hm* h, int what) {
    switch (what) {
        case 0: return h->S;
        case 1: return h->kdLevels;
        case 2: return h->numLevels;
        case 3: return h->tlrCapHits;
        default: return -1;
    }
}
OR_API const int* or_hm_perm(const or_hm* h) { return h->perm; }
OR_API const int* or_hm_offsets(const or_hm* h) { return h->offsets; }
OR_API const double* or_hm_bbmin(const or_hm* h) { return h->bbmin; }
OR_API const double* or_hm_bbmax(const or_hm* h) { return h->bbmax; }
OR_API const int* or_hm_split_dims(const or_hm* h) { return h->splitDims; }
OR_API const int* or_hm_num_tiles(const or_hm* h) { return h->numTiles; }
OR_API const int* or_hm_tiles(const or_hm* h, int level) { return h->tiles + h->tileBase[level - 1]; }
OR_API const int* or_hm_tlr_ranks(const or_hm* h, int morton) { return morton ? h->tlrRanksMO : h->tlrRanksCM; }
OR_API const long long* or_hm_tlr_offsets(const or_hm* h, int morton) { return morton ? h->tlrOffMO : h->tlrOffCM; }
OR_API const double* or_hm_tlr_U(const or_hm* h, int morton) { return morton ? h->tlrUMO : h->tlrUCM; }
OR_API const double* or_hm_tlr_V(const or_hm* h, int morton) { return morton ? h->tlrVMO : h->tlrVCM; }
OR_API const double* or_hm_diag(const or_hm* h) { return h->diag; }
OR_API const int* or_hm_level_ranks(const or_hm* h, int level) { return h->lvRanks[level - 1]; }
OR_API const long long* or_hm_level_offsets(const or_hm* h, int level) { return h->lvOff[level - 1]; }
OR_API const double* or_hm_level_U(const or_hm* h, int level) { return h->lvU[level - 1]; }
OR_API const double* or_hm_level_V(const or_hm* h, int level) { return h->lvV[level - 1]; }
OR_API int or_hm_level_max_rank(const or_hm* h, int level) { return h->lvMaxRank[level - 1]; }
OR_API float or_hm_level_tol(const or_hm* h, int level) { return h->lvTol[level - 1]; }
OR_API int or_hm_level_cap_hits(const or_hm* h, int level) { return h->lvCapHits[level - 1]; }

/* ------------------------------------------------------------------------- */
/* H-matvec (HMatrixVecMult, src/HMatrixVectorMultiplication.cu:10-26 with   */
/* cutlassDiagonalXVec and cutlassHierarchicalXVec semantics): Y = H X,       */
/* X/Y are n x k column-major (ld n) in KD-permuted order.  Generalised to    */
/* any structure and any leaf occupancy.  transpose != 0 computes H^T X.      */
/* ------------------------------------------------------------------------- */
OR_API void or_hm_matvec(const or_hm* h, int k, const double* X, double* Y, int transpose) {
    int n = h->n, b = h->b, L = h->numLevels, S = h->S;
    memset(Y, 0, sizeof(double) * (size_t)n * k);
    /* diagonal blocks: y_leaf = D_leaf x_leaf (D stored column-major, ld b) */
    for (int leaf = 0; leaf < S; ++leaf) {
        int a = h->offsets[leaf], sz = h->offsets[leaf + 1] - a;
        const double* D = h->diag + (size_t)leaf * b * b;
        for (int s = 0; s < k; ++s)
            for (int q = 0; q < sz; ++q) {
                double t = 0;
                for (int c = 0; c < sz; ++c)
                    t += (transpose ? D[(size_t)q * b + c] : D[(size_t)c * b + q]) * X[(size_t)s * n + a + c];
                Y[(size_t)s * n + a + q] += t;
            }
    }
    for (int level = 1; level < L; ++level) {
        int nt = h->numTiles[level - 1];
        if (!nt) continue;
        int bus = 1 << (L - 1 - level);
        int m = bus * b;
        const int* tiles = h->tiles + h->tileBase[level - 1];
        double* z = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1) * k);
        for (int t = 0; t < nt; ++t) {
            uint32_t u, v;
            or_mo_to_xy((uint32_t)tiles[t], &u, &v); /* u: column cluster, v: row cluster */
            int rk = h->lvRanks[level - 1][t];
            if (!rk) continue;
            const double* U = h->lvU[level - 1] + (size_t)h->lvOff[level - 1][t] * m;
            const double* V = h->lvV[level - 1] + (size_t)h->lvOff[level - 1][t] * m;
            const double* Pin = transpose ? U : V;   /* factor applied to the input */
            const double* Pout = transpose ? V : U;  /* factor producing the output */
            int inCl = transpose ? (int)v : (int)u, outCl = transpose ? (int)u : (int)v;
            for (int s = 0; s < k; ++s)
                for (int a = 0; a < rk; ++a) {
                    double acc = 0;
                    for (int x = 0; x < m; ++x) {
                        int pt = (inCl * bus + x / b);
                        int q = x % b;
                        int off = h->offsets[pt], sz = h->offsets[pt + 1] - off;
                        if (q < sz) acc += Pin[(size_t)a * m + x] * X[(size_t)s * n + off + q];
                    }
                    z[(size_t)s * rk + a] = acc;
                }
            for (int s = 0; s < k; ++s)
                for (int x = 0; x < m; ++x) {
                    int pt = (outCl * bus + x / b), q = x % b;
                    int off = h->offsets[pt], sz = h->offsets[pt + 1] - off;
                    if (q >= sz) continue;
                    double acc = 0;
                    for (int a = 0; a < rk; ++a) acc += Pout[(size_t)a * m + x] * z[(size_t)s * rk + a];
                    Y[(size_t)s * n + off + q] += acc;
                }
        }
        free(z);
    }
}

/* ------------------------------------------------------------------------- */
/* Dense oracle (src/expandMatrix.cu:142-234) restricted to sampled rows:    */
/* Y[r, :] = sum_j K(x_perm[rows[r]], x_perm[j]) X[j, :] ; Y is nr x k.       */
/* ------------------------------------------------------------------------- */
OR_API void or_sampled_rows(int n, int d, const double* p, const int* perm, double ell, int nr, const int* rows, int k,
                            const double* X, double* Y) {
#pragma omp parallel for schedule(static)
    for (int r = 0; r < nr; ++r) {
        int pr = perm[rows[r]];
        for (int s = 0; s < k; ++s) Y[(size_t)s * nr + r] = 0;
        for (int j = 0; j < n; ++j) {
            double kv = kfun(p, n, d, pr, perm[j], ell);
            for (int s = 0; s < k; ++s) Y[(size_t)s * nr + r] += kv * X[(size_t)s * n + j];
        }
    }
}

/* dense kernel matrix in KD order (n x n column-major), small n only */
OR_API void or_dense_matrix(int n, int d, const double* p, const int* perm, double ell, double* K) {
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) K[(size_t)j * n + i] = kfun(p, n, d, perm[i], perm[j], ell);
}

OR_API int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------- */
/* CPU H-matvec on externally supplied factors (the GPU-built H-matrix copied */
/* to the host); same semantics as or_hm_matvec, row-owner order, OpenMP.     */
/* Full leaves (n = S*b).  levels: numLevels-1 entries, level l at index l-1. */
/* ------------------------------------------------------------------------- */
OR_API void or_matvec_factors(int n, int b, int numLevels, const int* numTiles, const int* const* tiles,
                              const int* const* ranks, const long long* const* offs, const double* const* U,
                              const double* const* V, const double* diag, int k, const double* X, double* Y) {
    int S = n / b, L = numLevels;
    /* phase 1: z = V^T x for every tile */
    long long* zoff = (long long*)calloc((size_t)L, sizeof(long long));
    long long tot = 0;
    for (int l = 1; l < L; ++l) { zoff[l - 1] = tot; tot += offs[l - 1][numTiles[l - 1]]; }
    double* z = (double*)calloc((size_t)(tot > 0 ? tot : 1) * k, sizeof(double));
    for (int l = 1; l < L; ++l) {
        int m = (1 << (L - 1 - l)) * b;
#pragma omp parallel for schedule(dynamic, 1)
        for (int t = 0; t < numTiles[l - 1]; ++t) {
            uint32_t u, v;
            or_mo_to_xy((uint32_t)tiles[l - 1][t], &u, &v);
            int rk = ranks[l - 1][t];
            const double* Vt = V[l - 1] + (size_t)offs[l - 1][t] * m;
            for (int a = 0; a < rk; ++a)
                for (int s = 0; s < k; ++s) {
                    const double* xs = X + (size_t)s * n + (size_t)u * m;
                    const double* va = Vt + (size_t)a * m;
                    double acc = 0;
                    for (int x = 0; x < m; ++x) acc += va[x] * xs[x];
                    z[(size_t)(zoff[l - 1] + offs[l - 1][t] + a) * k + s] = acc;
                }
        }
    }
    /* phase 2: each leaf row block sums D x and its rows of every U_t z_t */
    int** rowTiles = (int**)calloc((size_t)L, sizeof(int*));
    int** rowPtr = (int**)calloc((size_t)L, sizeof(int*));
    for (int l = 1; l < L; ++l) {
        int nr = 1 << l;
        rowPtr[l - 1] = (int*)calloc((size_t)nr + 1, sizeof(int));
        rowTiles[l - 1] = (int*)calloc((size_t)(numTiles[l - 1] > 0 ? numTiles[l - 1] : 1), sizeof(int));
        for (int t = 0; t < numTiles[l - 1]; ++t) { uint32_t u, v; or_mo_to_xy((uint32_t)tiles[l - 1][t], &u, &v); rowPtr[l - 1][v + 1]++; }
        for (int r = 0; r < nr; ++r) rowPtr[l - 1][r + 1] += rowPtr[l - 1][r];
        int* fill = (int*)calloc((size_t)nr, sizeof(int));
        for (int t = 0; t < numTiles[l - 1]; ++t) {
            uint32_t u, v;
            or_mo_to_xy((uint32_t)tiles[l - 1][t], &u, &v);
            rowTiles[l - 1][rowPtr[l - 1][v] + fill[v]++] = t;
        }
        free(fill);
    }
#pragma omp parallel for schedule(static)
    for (int leaf = 0; leaf < S; ++leaf) {
        double acc[64 * 16];
        for (int s = 0; s < k; ++s) {
            const double* D = diag + (size_t)leaf * b * b;
            const double* xs = X + (size_t)s * n + (size_t)leaf * b;
            for (int q = 0; q < b; ++q) acc[q] = 0;
            for (int c = 0; c < b; ++c)
                for (int q = 0; q < b; ++q) acc[q] += D[(size_t)c * b + q] * xs[c];
            for (int l = 1; l < L; ++l) {
                int bus = 1 << (L - 1 - l), m = bus * b;
                int row = leaf / bus, q0 = (leaf % bus) * b;
                for (int i = rowPtr[l - 1][row]; i < rowPtr[l - 1][row + 1]; ++i) {
                    int t = rowTiles[l - 1][i];
                    int rk = ranks[l - 1][t];
                    const double* Ut = U[l - 1] + (size_t)offs[l - 1][t] * m + q0;
                    const double* zt = z + (size_t)(zoff[l - 1] + offs[l - 1][t]) * k + s;
                    for (int a = 0; a < rk; ++a) {
                        double w = zt[(size_t)a * k];
                        const double* ua = Ut + (size_t)a * m;
                        for (int q = 0; q < b; ++q) acc[q] += ua[q] * w;
                    }
                }
            }
            for (int q = 0; q < b; ++q) Y[(size_t)s * n + (size_t)leaf * b + q] = acc[q];
        }
    }
    for (int l = 1; l < L; ++l) { free(rowTiles[l - 1]); free(rowPtr[l - 1]); }
    free(rowTiles); free(rowPtr); free(z); free(zoff);
}
