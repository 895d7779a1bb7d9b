/*
 * cpu_blas.c -- CPU baseline of the H-matvec for the benchmark's cpu_baseline
 * leg: the same two phases as the GPU product (z_t = V_t^T X for every tile,
 * then every leaf's rows Y = D X + sum_t U_t z_t), with every block product
 * handed to OpenBLAS dgemm (numpy's bundled ILP64 OpenBLAS 0.3.x, loaded at
 * run time; OpenBLAS single-threaded per call) and OpenMP across tiles and
 * leaves.  Test infrastructure, like the rest of oracle/: never linked into
 * or called by the library.
 *
 * Factor layout (per level l, tiles in Morton order, column-major U/V with
 * columns of m = 2^(L-1-l) * b rows) and the KD-ordered vectors follow the
 * reference's HMatrix (src/HMatrix.cuh:9-34) and HMatrixVecMult
 * (src/HMatrixVectorMultiplication.cu:10-26) as or_matvec_factors does.
 */
#include <dlfcn.h>
#include <stdint.h>
#include <stdlib.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_API __attribute__((visibility("default")))

/* CBLAS enumerators (column-major, no transpose, transpose) */
enum { kColMajor = 102, kNoTrans = 111, kTrans = 112 };

typedef void (*dgemm_fn)(int order, int ta, int tb, int64_t m, int64_t n, int64_t k, double alpha, const double* A,
                         int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc);
typedef void (*setthreads_fn)(int n);

static dgemm_fn g_dgemm = 0;

/* Loads the OpenBLAS library at path; 0 on success */
OR_API int cb_open(const char* path) {
    if (g_dgemm) return 0;
    void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return 1;
    g_dgemm = (dgemm_fn)dlsym(h, "scipy_cblas_dgemm64_");
    setthreads_fn st = (setthreads_fn)dlsym(h, "scipy_openblas_set_num_threads64_");
    if (!g_dgemm || !st) { g_dgemm = 0; return 2; }
    st(1); /* parallelism comes from the OpenMP loops over tiles and leaves */
    return 0;
}

static void mo_to_xy(uint32_t d, uint32_t* x, uint32_t* y) {
    uint32_t a = d & 0x55555555u, c = (d >> 1) & 0x55555555u;
    a = (a | (a >> 1)) & 0x33333333u; a = (a | (a >> 2)) & 0x0F0F0F0Fu;
    a = (a | (a >> 4)) & 0x00FF00FFu; a = (a | (a >> 8)) & 0x0000FFFFu;
    c = (c | (c >> 1)) & 0x33333333u; c = (c | (c >> 2)) & 0x0F0F0F0Fu;
    c = (c | (c >> 4)) & 0x00FF00FFu; c = (c | (c >> 8)) & 0x0000FFFFu;
    *x = a; /* column cluster (even bits) */
    *y = c; /* row cluster (odd bits) */
}

/* Y (n x k, ld n) = H X with the factors of numLevels-1 levels and S = n/b
   dense diagonal blocks (column-major b x b).  Returns 0, or 1 when OpenBLAS
   is not loaded. */
OR_API int cb_matvec_factors(int n, int b, int numLevels, const int* numTiles, const int* const* tiles,
                             const int* const* ranks, const long long* const* offs, const double* const* U,
                             const double* const* V, const double* diag, int k, const double* X, double* Y) {
    if (!g_dgemm) return 1;
    const int S = n / b, L = numLevels;
    long long* zoff = (long long*)calloc((size_t)L, sizeof(long long));
    long long tot = 0;
    for (int l = 1; l < L; ++l) { zoff[l - 1] = tot; tot += offs[l - 1][numTiles[l - 1]]; }
    double* z = (double*)malloc((size_t)(tot > 0 ? tot : 1) * k * sizeof(double));
    /* phase 1: z_t (rank x k, column-major) = V_t^T X[column cluster] */
    for (int l = 1; l < L; ++l) {
        const int m = (1 << (L - 1 - l)) * b;
#pragma omp parallel for schedule(dynamic, 4)
        for (int t = 0; t < numTiles[l - 1]; ++t) {
            uint32_t u, v;
            mo_to_xy((uint32_t)tiles[l - 1][t], &u, &v);
            const int rk = ranks[l - 1][t];
            if (!rk) continue;
            g_dgemm(kColMajor, kTrans, kNoTrans, rk, k, m, 1.0, V[l - 1] + (size_t)offs[l - 1][t] * m, m,
                    X + (size_t)u * m, n, 0.0, z + (size_t)(zoff[l - 1] + offs[l - 1][t]) * k, rk);
        }
    }
    /* tiles by row cluster, per level */
    int** rowTiles = (int**)calloc((size_t)L, sizeof(int*));
    int** rowPtr = (int**)calloc((size_t)L, sizeof(int*));
    for (int l = 1; l < L; ++l) {
        const int nr = 1 << l;
        rowPtr[l - 1] = (int*)calloc((size_t)nr + 1, sizeof(int));
        rowTiles[l - 1] = (int*)calloc((size_t)(numTiles[l - 1] > 0 ? numTiles[l - 1] : 1), sizeof(int));
        for (int t = 0; t < numTiles[l - 1]; ++t) {
            uint32_t u, v;
            mo_to_xy((uint32_t)tiles[l - 1][t], &u, &v);
            rowPtr[l - 1][v + 1]++;
        }
        for (int r = 0; r < nr; ++r) rowPtr[l - 1][r + 1] += rowPtr[l - 1][r];
        int* fill = (int*)calloc((size_t)nr, sizeof(int));
        for (int t = 0; t < numTiles[l - 1]; ++t) {
            uint32_t u, v;
            mo_to_xy((uint32_t)tiles[l - 1][t], &u, &v);
            rowTiles[l - 1][rowPtr[l - 1][v] + fill[v]++] = t;
        }
        free(fill);
    }
    /* phase 2: Y rows of each leaf = D X_leaf + sum of U_t[leaf rows] z_t */
#pragma omp parallel for schedule(static)
    for (int leaf = 0; leaf < S; ++leaf) {
        double* yl = Y + (size_t)leaf * b;
        g_dgemm(kColMajor, kNoTrans, kNoTrans, b, k, b, 1.0, diag + (size_t)leaf * b * b, b, X + (size_t)leaf * b, n,
                0.0, yl, n);
        for (int l = 1; l < L; ++l) {
            const int bus = 1 << (L - 1 - l), m = bus * b;
            const int row = leaf / bus, q0 = (leaf % bus) * b;
            for (int i = rowPtr[l - 1][row]; i < rowPtr[l - 1][row + 1]; ++i) {
                const int t = rowTiles[l - 1][i];
                const int rk = ranks[l - 1][t];
                if (!rk) continue;
                g_dgemm(kColMajor, kNoTrans, kNoTrans, b, k, rk, 1.0, U[l - 1] + (size_t)offs[l - 1][t] * m + q0, m,
                        z + (size_t)(zoff[l - 1] + offs[l - 1][t]) * k, rk, 1.0, yl, n);
            }
        }
    }
    for (int l = 1; l < L; ++l) { free(rowPtr[l - 1]); free(rowTiles[l - 1]); }
    free(rowPtr);
    free(rowTiles);
    free(z);
    free(zoff);
    return 0;
}
