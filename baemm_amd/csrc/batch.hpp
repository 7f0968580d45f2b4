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
};
BatchTuning& batch_tuning();

// Device and pinned staging of the work lists (index 0: MFMA, 1: FP32).
struct BatchWorkspace {
    DevBuf dev[2];
    void* host[2] = {nullptr, nullptr};
    size_t host_bytes[2] = {0, 0};
    hipEvent_t ev[2] = {nullptr, nullptr};
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

}  // namespace baemm
