// A C++ caller of libhmx: the reference's driver sequence (src/main.cu:46-186)
// through the C ABI -- dataset, KD-tree, structure, TLR matrix (column-major,
// then Morton), H-matrix from the structure, H-matvec -- followed by the
// reference's matvec error check (checkErrorInHmatrixVecMult,
// src/expandMatrix.cu:202-234) on sampled rows of the dense kernel.
//
//   hmx_main [N] [d] [bucketSize] [tolerance] [direct 0|1] [vectorWidth]
//
// Prints one line of key=value pairs and exits 0 when every call succeeded
// and the relative error is below 10 * tolerance, 1 otherwise.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "hmx.h"

#define CALL(expr)                                                                            \
    do {                                                                                      \
        hmx_status s_ = (expr);                                                               \
        if (s_ != HMX_OK) {                                                                   \
            std::fprintf(stderr, "%s failed: status %d: %s\n", #expr, (int)s_, hmx_last_error(ctx)); \
            return 1;                                                                         \
        }                                                                                     \
    } while (0)

#define HIPCALL(expr)                                                                  \
    do {                                                                               \
        if ((expr) != hipSuccess) {                                                    \
            std::fprintf(stderr, "%s failed\n", #expr);                                \
            return 1;                                                                  \
        }                                                                              \
    } while (0)

int main(int argc, char** argv) {
    hmx_config cfg;
    hmx_config_default(&cfg);
    cfg.numberOfInputPoints = argc > 1 ? (uint32_t)std::atoi(argv[1]) : 16384;
    cfg.dimensionOfInputPoints = argc > 2 ? (uint32_t)std::atoi(argv[2]) : 2;
    cfg.bucketSize = argc > 3 ? (uint32_t)std::atoi(argv[3]) : 32;
    cfg.lowestLevelTolerance = argc > 4 ? (float)std::atof(argv[4]) : 1e-6f;
    const int direct = argc > 5 ? std::atoi(argv[5]) : 0;
    const uint32_t vw = argc > 6 ? (uint32_t)std::atoi(argv[6]) : 1;
    cfg.maxRankPolicy = HMX_MAXRANK_FULL;
    const uint32_t N = cfg.numberOfInputPoints, d = cfg.dimensionOfInputPoints, b = cfg.bucketSize;

    hmx_context* ctx = nullptr;
    if (hmx_create(&ctx, nullptr) != HMX_OK) {
        std::fprintf(stderr, "hmx_create failed: %s\n", hmx_last_error(nullptr));
        return 1;
    }
    double *p = nullptr, *x = nullptr, *y = nullptr, *yd = nullptr;
    int32_t* rows = nullptr;
    HIPCALL(hipMalloc(&p, sizeof(double) * N * d));
    HIPCALL(hipMalloc(&x, sizeof(double) * N * vw));
    HIPCALL(hipMalloc(&y, sizeof(double) * N * vw));
    CALL(hmx_generateDataset(ctx, N, d, cfg.seed, p));

    hmx_kdtree* tree = nullptr;
    CALL(hmx_constructKDTree(ctx, N, d, b, cfg.divMethod, p, &tree));
    hmx_structure* st = nullptr;
    CALL(hmx_constructHMatrixStructure(ctx, N, d, b, cfg.admissibilityCondition, cfg.eta, tree, &st));
    hmx_hmatrix* H = nullptr;
    uint64_t rankSum = 0;
    if (direct) {
        CALL(hmx_generateHMatrixDirect(ctx, &cfg, tree, st, p, &H));
    } else {
        hmx_tlr *cm = nullptr, *mo = nullptr;
        CALL(hmx_createColumnMajorLRMatrix(ctx, &cfg, tree, p, &cm, &rankSum));
        CALL(hmx_convertColumnMajorToMorton(ctx, cm, &mo));
        CALL(hmx_freeMatrix(cm));
        CALL(hmx_generateHMatrixFromStruct(ctx, &cfg, tree, st, mo, &H));
        CALL(hmx_freeMatrix(mo));
    }
    CALL(hmx_generateRandomVector(ctx, vw, N, 1234, x));
    CALL(hmx_HMatrixVecMult(ctx, H, vw, x, y));

    // error on sampled rows of the dense kernel (rows i * N / nr)
    const uint32_t nr = N < 512 ? N : 512;
    std::vector<int32_t> hrows(nr);
    for (uint32_t i = 0; i < nr; ++i) hrows[i] = (int32_t)((uint64_t)i * N / nr);
    HIPCALL(hipMalloc(&rows, sizeof(int32_t) * nr));
    HIPCALL(hipMalloc(&yd, sizeof(double) * nr * vw));
    HIPCALL(hipMemcpy(rows, hrows.data(), sizeof(int32_t) * nr, hipMemcpyHostToDevice));
    CALL(hmx_denseRowsXVector(ctx, N, d, p, hmx_kdtree_device_perm(tree), cfg.correlationLength, nr, rows, vw, x, yd));
    std::vector<double> hy((size_t)N * vw), hyd((size_t)nr * vw);
    HIPCALL(hipMemcpy(hy.data(), y, sizeof(double) * hy.size(), hipMemcpyDeviceToHost));
    HIPCALL(hipMemcpy(hyd.data(), yd, sizeof(double) * hyd.size(), hipMemcpyDeviceToHost));
    double num = 0, den = 0;
    for (uint32_t s = 0; s < vw; ++s)
        for (uint32_t i = 0; i < nr; ++i) {
            const double a = hy[(size_t)s * N + hrows[i]], r = hyd[(size_t)s * nr + i];
            num += (a - r) * (a - r);
            den += r * r;
        }
    const double err = std::sqrt(num / den);
    uint64_t lowRank = 0, dense = 0, entries = 0;
    CALL(hmx_hmatrix_bytes(H, &lowRank, &dense, &entries));
    std::printf("hmx_main N=%u d=%u b=%u tol=%g direct=%d vectorWidth=%u rankSum=%llu factorBytes=%llu "
                "relErr=%.3e\n",
                N, d, b, (double)cfg.lowestLevelTolerance, direct, vw, (unsigned long long)rankSum,
                (unsigned long long)(lowRank + dense), err);

    CALL(hmx_freeHMatrix(H));
    CALL(hmx_freeHMatrixStructure(st));
    CALL(hmx_freeKDTree(tree));
    (void)hipFree(p);
    (void)hipFree(x);
    (void)hipFree(y);
    (void)hipFree(yd);
    (void)hipFree(rows);
    CALL(hmx_destroy(ctx));
    return err < 10.0 * cfg.lowestLevelTolerance ? 0 : 1;
}
