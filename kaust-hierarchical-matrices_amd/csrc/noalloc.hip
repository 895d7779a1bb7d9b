// H-matvec that allocates nothing: the streaming kernels of hmx_HMatrixVecMult
// with their per-call workspaces (phase-1 partial sums, z, padded copies of x
// and y) in one fixed file-scope device array instead of buffers the library
// allocates.  For callers that cannot call hipMalloc (code under HIP graph
// capture, tight serving loops).  No counterpart in the reference.
//
// The array exists once per GPU the library runs on and is held until the
// process ends.  Every call on a GPU shares that GPU's copy and leaves its
// contents behind; the kernels overwrite every workspace element they read,
// so nothing needs clearing between calls.  Calls that overlap on one GPU
// (two streams, two threads) race on the array.
#include <mutex>

#include "hmx_internal.h"

namespace hmx {

constexpr size_t kScratchBytes = 64ull << 20;
constexpr int kMaxDevices = 64;
constexpr size_t kAlignElems = 32;  // sub-buffers start at 256-byte boundaries

static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

namespace {
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};  // per device: the address of that device's copy
}  // namespace

// Resolves this device's copy of the array (hipGetSymbolAddress with the
// device current); called by hmx_create for the context's device.  The
// array's name is a host-side stand-in, not a device address: only the
// resolved pointer goes to kernels.
void noalloc_init(int device) {
    HMX_CHECK(device >= 0 && device < kMaxDevices, HMX_EUNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[device]) return;
    int cur = -1;
    HMX_HIP(hipGetDevice(&cur));
    HMX_CHECK(cur == device, HMX_ESTATE, "scratch set-up needs the context's device current");
    void* p = nullptr;
    HMX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    HMX_CHECK(p != nullptr, HMX_EDEVICE, "no device address for the matvec scratch array");
    g_scratch[device] = p;
}

namespace {
size_t round_up(size_t v) { return (v + kAlignElems - 1) / kAlignElems * kAlignElems; }
}  // namespace

void hmatvec_noalloc(Context& c, const HMatrix& h, uint32_t k, const double* X, double* Y) {
    HMX_CHECK(k >= 1, HMX_EINVAL, "vectorWidth must be >= 1");
    HMX_CHECK(X && Y, HMX_EINVAL, "null vector pointer");
    HMX_CHECK(X != Y, HMX_EINVAL, "in and out must not alias");
    const MatvecPlan& P = h.plan;
    HMX_CHECK(P.built, HMX_ESTATE,
              "the H-matrix has no streaming matvec plan yet (one hmx_HMatrixVecMult call with vectorWidth 1 "
              "builds it)");
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (c.device >= 0 && c.device < kMaxDevices) base = g_scratch[c.device];
    }
    HMX_CHECK(base != nullptr, HMX_ESTATE, "matvec scratch not set up on this context's device");
    // the layout run_matvec_plan_ws reads: passes of at most K right-hand sides
    const uint32_t K = pass_width(k);
    const size_t nPart = round_up((size_t)std::max<int64_t>(1, P.numPart) * K);
    const size_t nZ = round_up((size_t)std::max<int64_t>(1, P.numCols) * K);
    const size_t nPad = P.gather ? round_up((size_t)P.npad * K) : 0;
    const size_t need = (nPart + nZ + 2 * nPad) * sizeof(double);
    HMX_CHECK(need <= kScratchBytes, HMX_EINVAL,
              "the no-allocation H-matvec needs " + std::to_string(need >> 20) + " MiB of scratch for passes of " +
                  std::to_string(K) + " right-hand sides of this H-matrix; its fixed array holds " +
                  std::to_string(kScratchBytes >> 20) + " MiB");
    double* s = static_cast<double*>(base);
    MvWork W{s, s + nPart, P.gather ? s + nPart + nZ : nullptr, P.gather ? s + nPart + nZ + nPad : nullptr};
    run_matvec_plan_ws(c, P, W, (int)h.b, k, X, (int64_t)h.n, Y, nullptr);
}

}  // namespace hmx
