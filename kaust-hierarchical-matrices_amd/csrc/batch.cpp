// Batched H-matvec over many independent H-matrices (point clouds of uneven
// sizes): one concatenated plan applied by one launch per phase, optionally
// captured once in a hipGraph and replayed, with the phase-1 chunk length
// picked by timing.  No counterpart in the reference, whose driver handles one
// point cloud per run (src/main.cu:46-173).
#include <algorithm>

#include "hmx_internal.h"

namespace hmx {

void Batch::drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    gX = nullptr;
    gY = nullptr;
    gk = 0;
}

Batch::~Batch() {
    drop_graph();
    if (evIn) (void)hipEventDestroy(evIn);
    if (evOut) (void)hipEventDestroy(evOut);
    if (stream) (void)hipStreamDestroy(stream);
}

static void note_gens(Batch& B) {
    B.pGen.clear();
    for (const HMatrix* h : B.hs) B.pGen.push_back(h->factorGen);
}

static bool p_stale(const Batch& B) {
    for (size_t i = 0; i < B.hs.size(); ++i)
        if (B.hs[i]->factorGen != B.pGen[i]) return true;
    return false;
}

void batch_create(Context& c, const std::vector<HMatrix*>& hs, Batch& B) {
    for (HMatrix* h : hs) HMX_CHECK(h != nullptr, HMX_EINVAL, "null H-matrix in batch");
    B.hs = hs;
    B.b = (int)hs.at(0)->b;
    build_batch_plan(c, hs, 2048, B.P, B.diag);
    note_gens(B);
    HMX_HIP(hipStreamCreate(&B.stream));  // default priority
    HMX_HIP(hipEventCreateWithFlags(&B.evIn, hipEventDisableTiming));
    HMX_HIP(hipEventCreateWithFlags(&B.evOut, hipEventDisableTiming));
}

// One right-hand side: the streaming kernels.  Two or more: the FP64 matrix
// cores over the batch's block-diagonal operator (passes of 16 or 64, the
// last padded with zero columns).  Plans and workspaces are made ready here,
// never inside a capture; when any of them was rebuilt, reallocated or
// cleared the batch's epoch moves on, and a graph captured in an earlier
// epoch is captured again before its next replay.
static void prepare(Context& c, Batch& B, uint32_t k) {
    bool changed = false;
    if (k >= 2) {
        changed = batch_mf_prepare(c, B, k);
    } else {
        if (p_stale(B)) {  // an instance's factors were overwritten: its dense blocks are stale
            build_batch_plan(c, B.hs, B.P.chunk, B.P, B.diag);
            note_gens(B);
            changed = true;
        }
        changed = ensure_matvec_ws(B.P, 1) || changed;
    }
    if (changed) ++B.epoch;
}

static void run_batch(Context& c, Batch& B, uint32_t k, const double* X, double* Y) {
    if (k >= 2) batch_mf_run(c, B, k, X, Y);
    else run_matvec_plan(c, B.P, B.b, k, X, B.P.n, Y, nullptr);
}

void batch_matvec(Context& c, Batch& B, uint32_t k, const double* X, double* Y, bool useGraph) {
    HMX_CHECK(k >= 1, HMX_EINVAL, "vectorWidth must be >= 1");
    HMX_CHECK(X && Y && X != Y, HMX_EINVAL, "bad vector pointers");
    prepare(c, B, k);
    if (!useGraph) {
        run_batch(c, B, k, X, Y);
        return;
    }
    if (!B.exec || B.gX != X || B.gY != Y || B.gk != k || B.gEpoch != B.epoch) {
        B.drop_graph();
        // the captured sequence holds kernel launches only (gather or
        // blocked copy, phase 1, reduce, phase 2, scatter per pass)
        HMX_HIP(hipStreamSynchronize(c.stream));
        Context cap;
        cap.stream = B.stream;
        cap.device = c.device;
        HMX_HIP(hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal));
        try {
            run_batch(cap, B, k, X, Y);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(B.stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        HMX_HIP(hipStreamEndCapture(B.stream, &B.graph));
        HMX_HIP(hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0));
        B.gX = X;
        B.gY = Y;
        B.gk = k;
        B.gEpoch = B.epoch;
    }
    // order the replay after the caller's prior work and the caller's later work after it
    HMX_HIP(hipEventRecord(B.evIn, c.stream));
    HMX_HIP(hipStreamWaitEvent(B.stream, B.evIn, 0));
    HMX_HIP(hipGraphLaunch(B.exec, B.stream));
    HMX_HIP(hipEventRecord(B.evOut, B.stream));
    HMX_HIP(hipStreamWaitEvent(c.stream, B.evOut, 0));
}

// Times the candidates of the path k right-hand sides take and keeps the
// fastest: one right-hand side, the streaming phase-1 chunk length (512, 1024,
// 2048, 4096); two or more, the matrix-core phase-1 workgroups per XCD chunk
// (0 = plain order, 8, 16, 32).
void batch_tune(Context& c, Batch& B, uint32_t k, int* bestChunk, float* bestMs) {
    DBuf<double> X((size_t)B.P.n * k), Y((size_t)B.P.n * k);
    HMX_HIP(hipMemsetAsync(X.p, 0, X.bytes(), c.stream));
    hipEvent_t e0, e1;
    HMX_HIP(hipEventCreate(&e0));
    HMX_HIP(hipEventCreate(&e1));
    auto time_it = [&](auto&& run) {
        for (int w = 0; w < 2; ++w) run();
        HMX_HIP(hipEventRecord(e0, c.stream));
        const int reps = 5;
        for (int r = 0; r < reps; ++r) run();
        HMX_HIP(hipEventRecord(e1, c.stream));
        HMX_HIP(hipEventSynchronize(e1));
        float t = 0;
        HMX_HIP(hipEventElapsedTime(&t, e0, e1));
        return t / reps;
    };
    int best = 0;
    float bms = 1e30f;
    try {
        B.drop_graph();
        if (k == 1) {
            for (int ch : {512, 1024, 2048, 4096}) {
                build_batch_plan(c, B.hs, ch, B.P, B.diag);
                ensure_matvec_ws(B.P, 1);
                float t = time_it([&] { run_matvec_plan(c, B.P, B.b, 1, X.p, B.P.n, Y.p, nullptr); });
                if (t < bms) { bms = t; best = ch; }
            }
            build_batch_plan(c, B.hs, best, B.P, B.diag);
            note_gens(B);
        } else {
            batch_mf_prepare(c, B, k);
            for (int q : {0, 8, 16, 32}) {
                B.mf.xcdChunk = q;
                float t = time_it([&] { batch_mf_run(c, B, k, X.p, Y.p); });
                if (t < bms) { bms = t; best = q; }
            }
            B.mf.xcdChunk = best;
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ++B.epoch;  // plans rebuilt, tuned launch shapes
    if (bestChunk) *bestChunk = best;
    if (bestMs) *bestMs = bms;
}

}  // namespace hmx
