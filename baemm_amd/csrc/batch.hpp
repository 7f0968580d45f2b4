// batch.hpp -- batched boundary-operator applies over several handles in one
// persistent launch per arithmetic mode (batch.hip).
#pragma once

#include <vector>

#include "internal.hpp"

namespace baemm {

struct BatchEntry {
    Handle* h;
    ApplyArgs a;
};

// Tile shapes of the persistent launches (set by the autotuner or the caller).
struct BatchTuning {
    int mixed_steps = 256;      // 8-source steps per item of the MFMA path (2048 sources)
    int plain_blocks = 16;      // 64-source blocks per item of the FP32 path (1024 sources)
    int mixed_grid_mult = 0;    // blocks per CU of the persistent grid (0: occupancy limit)
    int plain_grid_mult = 0;
    // An instance whose own pair launch fills at least this many rounds of
    // resident blocks runs that launch instead of joining the persistent one
    // (0: every instance runs its own launch).
    double own_rounds = 1.0;
};
BatchTuning& batch_tuning();

// Device and pinned staging of the work lists (index 0: MFMA, 1: FP32) and
// of the batched staging and epilogue descriptors (2, 3).
struct BatchWorkspace {
    static constexpr int kSlots = 4;
    DevBuf dev[kSlots];
    void* host[kSlots] = {nullptr, nullptr, nullptr, nullptr};
    size_t host_bytes[kSlots] = {0, 0, 0, 0};
    hipEvent_t ev[kSlots] = {nullptr, nullptr, nullptr, nullptr};
    std::vector<hipEvent_t> sync;   // fork/join events between the call's stream and the handles' streams
    void release();
};

struct BatchStats {
    int64_t items[2] = {0, 0};
    int64_t grid[2] = {0, 0};
    int64_t instances[2] = {0, 0};
};

// All handles on the current device; each has its operators loaded.  Each
// entry's precision (baemm_set_precision) selects the launch it joins.
void apply_boundary_batch(std::vector<BatchEntry>& es, BatchWorkspace& bw, hipStream_t s, BatchStats* stats);

// One system of a batched GMRES solve (solvers_host.cpp): (P A) X = P B for
// the loaded boundary operator A of h; B, X device arrays, vertex_count x W
// complex double (ld W).
struct GmresBatchItem {
    Handle* h;
    const double2* B;
    double2* X;
    int64_t iterations = 0;
    int64_t cycles = 0;
    double rel_residual = 0.0;
    bool converged = false;
};
// Solves the systems together: every Arnoldi step applies the operators of
// all systems still iterating in one batched launch; the rest of a step runs
// on each handle's stream.  Synchronises s at the end.
void gmres_batch(std::vector<GmresBatchItem>& items, BatchWorkspace& bw, double tol, double cg_tol, hipStream_t s);

// One Gauss-Newton-step request of a fused batch (solvers_host.cpp): the
// forward solve and far field F (FarField), the Neumann data (Derivative_FF_
// Helper) and the adjoint derivative of g = F - F_obs (AdjointDerivative_FF).
// Host arrays: dirs wcs x 3, fobs m x W complex (may be null: zero data),
// F_out m x W complex, grad_out V x 3.
struct GnItem {
    Handle* h;
    const double* kappa;
    int64_t K, wcs;
    const double* dirs;
    int wave_type;
    const double* fobs;
    double* F_out;
    double* grad_out;
    double cg_tol, gmres_tol;
    int status = BAEMM_OK;      // BAEMM_ERR_NOCONV if a solve stopped above its tolerance
};
// Runs the steps of several requests phase by phase: each of the three
// boundary-integral solves of the batch is one batched GMRES (gmres_batch);
// the waves, far fields and gradients run per request on its handle's stream.
void gn_step_batch(std::vector<GnItem>& items, BatchWorkspace& bw, hipStream_t s);

}  // namespace baemm
