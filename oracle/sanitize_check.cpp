// Host-side sanitizer run (AddressSanitizer + UndefinedBehaviorSanitizer),
// test infrastructure: the CPU oracle's whole pipeline (KD-tree, structure,
// JDS, TLR ARA, Morton order, H-matrix from the structure, H-matvec) and the
// library's host C++ structure pass (kaust-hierarchical-matrices_amd/csrc/
// structure.cpp) on small inputs.  The structure and JDS the library builds
// from the oracle's boxes must equal the oracle's.  Exits 0 when they do and
// no sanitizer report stopped the run.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../kaust-hierarchical-matrices_amd/csrc/hmx_internal.h"

extern "C" {
void or_generate_points(int n, int d, uint64_t seed, double* p);
void or_generate_rhs(int n, int k, uint64_t seed, double* x);
int or_kd_num_levels(int n, int b);
int or_upper_pow2(int v);
int or_kdtree(int n, int d, int b, const double* p, int* perm, int* offsets, double* bbmin, double* bbmax,
              int* splitDims);
int or_structure(int n, int d, int b, int adm, float eta, const double* bbmin, const double* bbmax, int* numTiles,
                 int* tiles, long long* tile_base);
int or_jds(int level, int ntiles, const int* tiles, int* rowIndices, int* columnIndices, int* tilePos, int* iterPtr);
struct or_hm;
or_hm* or_hm_create(int n, int d, int b, int adm, float eta, float tol, double ell, int bs, int r, int policy,
                    uint64_t seed, const double* p);
void or_hm_destroy(or_hm* h);
int or_hm_kdtree(or_hm* h);
int or_hm_structure(or_hm* h);
long long or_hm_tlr(or_hm* h);
void or_hm_to_morton(or_hm* h);
void or_hm_hmatrix(or_hm* h);
void or_hm_matvec(const or_hm* h, int k, const double* X, double* Y, int transpose);
void or_sampled_rows(int n, int d, const double* p, const int* perm, double ell, int nr, const int* rows, int k,
                     const double* X, double* Y);
const int* or_hm_perm(const or_hm* h);
}

static int structure_case(int n, int d, int b, int adm) {
    std::vector<double> p((size_t)n * d);
    or_generate_points(n, d, 1234 + n, p.data());
    const int lv = or_kd_num_levels(n, b), nb = (1 << (lv + 1)) - 1;
    std::vector<int> perm(n), off((1 << lv) + 1), sd(nb);
    std::vector<double> bmin((size_t)nb * d), bmax((size_t)nb * d);
    or_kdtree(n, d, b, p.data(), perm.data(), off.data(), bmin.data(), bmax.data(), sd.data());
    const int n2 = or_upper_pow2(n);
    int L = 0;
    while ((1 << L) < n2 / b) ++L;
    L += 1;
    std::vector<int> nt(L);
    std::vector<long long> base(L);
    long long cap = 0;
    for (int l = 0; l < L; ++l) cap += (long long)(1 << (l + 1)) * ((1 << (l + 1)) - 1);
    std::vector<int> tiles(cap > 0 ? cap : 1);
    or_structure(n, d, b, adm, 1.0f, bmin.data(), bmax.data(), nt.data(), tiles.data(), base.data());
    hmx::Structure s;
    hmx::build_structure(s, (uint32_t)n, (uint32_t)d, (uint32_t)b, adm, 1.0f, bmin.data(), bmax.data());
    if (s.numLevels != L) return 1;
    for (int l = 1; l < L; ++l) {
        const std::vector<int32_t>& t = s.tileIndices[l - 1];
        if ((int)t.size() != nt[l - 1]) return 2;
        for (int i = 0; i < nt[l - 1]; ++i)
            if (t[i] != tiles[base[l - 1] + i]) return 3;
        const int rows = 1 << l;
        std::vector<int> ri(rows), ci(t.size() + 1), tp(t.size() + 1), ip(rows + 1);
        const int ndg = or_jds(l, (int)t.size(), t.data(), ri.data(), ci.data(), tp.data(), ip.data());
        const hmx::Structure::JDS& j = s.jds[l - 1];
        for (int i = 0; i < rows; ++i)
            if (j.rowIndices[i] != ri[i]) return 4;
        for (size_t i = 0; i < t.size(); ++i)
            if (j.columnIndices[i] != ci[i] || j.tilePos[i] != tp[i]) return 5;
        for (int i = 0; i <= ndg; ++i)
            if (j.iterPtr[i] != ip[i]) return 6;
    }
    return 0;
}

static int pipeline_case(int n, int d, int b, int adm, int policy) {
    std::vector<double> p((size_t)n * d);
    or_generate_points(n, d, 99, p.data());
    const double ell = d == 3 ? 0.2 : 0.1;
    or_hm* h = or_hm_create(n, d, b, adm, 1.0f, 1e-6f, ell, 16, 10, policy, 1234, p.data());
    or_hm_kdtree(h);
    or_hm_structure(h);
    or_hm_tlr(h);
    or_hm_to_morton(h);
    or_hm_hmatrix(h);
    std::vector<double> x((size_t)n * 2), y((size_t)n * 2), yt((size_t)n * 2);
    or_generate_rhs(n, 2, 5, x.data());
    or_hm_matvec(h, 2, x.data(), y.data(), 0);
    or_hm_matvec(h, 2, x.data(), yt.data(), 1);
    // the H-matvec against sampled dense rows
    const int nr = 64;
    std::vector<int> rows(nr);
    for (int i = 0; i < nr; ++i) rows[i] = (int)((long long)i * n / nr);
    std::vector<double> yd((size_t)nr * 2);
    or_sampled_rows(n, d, p.data(), or_hm_perm(h), ell, nr, rows.data(), 2, x.data(), yd.data());
    double num = 0, den = 0;
    for (int s = 0; s < 2; ++s)
        for (int i = 0; i < nr; ++i) {
            const double a = y[(size_t)s * n + rows[i]], r = yd[(size_t)s * nr + i];
            num += (a - r) * (a - r);
            den += r * r;
        }
    or_hm_destroy(h);
    return std::sqrt(num / den) < 1e-4 ? 0 : 10;
}

int main() {
    const int sc[][4] = {{1000, 2, 16, 0}, {4096, 2, 32, 1}, {3000, 3, 32, 2}, {5000, 3, 64, 1}, {777, 2, 8, 2}};
    for (const auto& c : sc) {
        const int r = structure_case(c[0], c[1], c[2], c[3]);
        if (r) {
            std::printf("structure case n=%d d=%d b=%d adm=%d: mismatch %d\n", c[0], c[1], c[2], c[3], r);
            return 1;
        }
    }
    const int pc[][5] = {{1024, 2, 32, 0, 1}, {1500, 3, 32, 1, 1}, {2048, 2, 64, 2, 1}, {2048, 2, 32, 0, 0}};
    for (const auto& c : pc) {
        const int r = pipeline_case(c[0], c[1], c[2], c[3], c[4]);
        if (r) {
            std::printf("pipeline case n=%d d=%d b=%d: error %d\n", c[0], c[1], c[2], r);
            return 1;
        }
    }
    std::printf("sanitize check ok\n");
    return 0;
}
