// ops.hpp -- the phases of a boundary-operator apply, shared by the
// single-instance path (ops.hip) and the batched persistent path (batch.hip).
#pragma once

#include <vector>

#include "pair_kernels.hpp"

namespace baemm {

// Arguments of the vertex epilogue: AvT of the summed pair partials (plus the
// single-layer diagonal), the mass term and alpha/beta.
struct EpiArgs {
    int nv, W, wcs, WCSP, K;
    int kernel_on, diag_on, mass_on;
    double ar, ai, br, bi;
    int S;
    int64_t seg_stride;
    int ldp, ldb;
    const int32_t* vt_ptr;
    const int32_t* vt_idx;
    const float* area;
    const float* sdiag;
    const float* part;
    const double2* tsum;   // non-null: the value of each triangle, ldb apart (tri_epilogue_kernel)
    const float2* b;
    const float* unscale;
    const int32_t* M_ptr;
    const int32_t* M_idx;
    const double* M_val;
    const void* B;
    int64_t ldB;
    void* C;
    int64_t ldC;
    float c1r[64], c1i[64], kap[64];
    float rotr[64], roti[64];   // phase per chunk multiplied into the summed partials
    double c0r[64], c0i[64];
};

// HIP events around the dominant pair-kernel launches of an apply, on the
// stream they run on (baemm_set_timing).
struct Timed {
    Handle& h;
    hipStream_t s;
    Timed(Handle& h_, hipStream_t s_) : h(h_), s(s_)
    {
        if (h.timing && !h.capturing) BAEMM_HIP(hipEventRecord(h.ev0, s));
    }
    void done()
    {
        if (!h.timing || h.capturing) return;
        BAEMM_HIP(hipEventRecord(h.ev1, s));
        BAEMM_HIP(hipEventSynchronize(h.ev1));
        float ms = 0.0f;
        BAEMM_HIP(hipEventElapsedTime(&ms, h.ev0, h.ev1));
        h.timed_ms += ms;
        h.timed_launches += 1;
    }
};

// Where an apply takes its device workspaces from: the handle's grow-only
// buffers (base == nullptr), or consecutive 256-byte-aligned pieces of a fixed
// device array of cap bytes (the no-allocation entry point).
struct WsSource {
    char* base = nullptr;
    size_t cap = 0, used = 0;
    void* take(DevBuf& buf, size_t n)
    {
        if (!base) {
            buf.reserve(n);
            return buf.ptr;
        }
        const size_t off = (used + 255) & ~(size_t)255;
        if (off + n > cap) throw ArgError{"internal: workspace larger than its planned size"};
        used = off + n;
        return base + off;
    }
};

// State of one apply between its phases.
// Staging launches of one boundary apply, recorded instead of launched when
// its plan is deferred (BdrPlan::defer): the batched apply runs those of all
// deferred instances of a call together (batch.hip).  b null: nothing staged.
struct StageDesc {
    // b = Av B (av_elem)
    const int4* tris = nullptr;
    const float* area = nullptr;
    int nt = 0, nt_pad = 0;
    const void* B = nullptr;
    int64_t ldB = 0;
    int wcs = 0, WCSP = 0, ldb = 0;
    float2* b = nullptr;
    // mixed precision: column maxima, scales and the hi/lo B operand
    int split = 0;
    int K = 0, nst = 0;
    unsigned* colmax = nullptr;
    float* scale = nullptr;
    float* unscale = nullptr;
    uint4* bop = nullptr;
    float sA[64] = {};
    // source normals scaled for the non-adjoint MFMA form (nrm_out null: none)
    const float4* nrm_in = nullptr;
    float4* nrm_out = nullptr;
    float nrm_s = 1.0f;
};

struct BdrPlan {
    bool kernel_on = false, mass_on = false, mixed = false, adj = false;
    // mixed: every chunk has c2 = 0 and c3 != 0 (single layer and adjoint double
    // layer), taken as the non-adjoint form divided by c3 with the target normals
    // in place of the source normals (adj is then false)
    bool tn = false;
    bool f32 = false;            // MFMA kernels in FP32 (BAEMM_PRECISION_FP32_MFMA); mixed is set too
    int WCSP = 0, ldb = 0, ldp = 0;
    int nst = 0;                 // 8-source steps (mixed) of the padded source range
    std::vector<float> sA;       // per-chunk power-of-two scale of A (mixed)
    std::vector<cfloat> rot;     // per-chunk factor the kernel values are divided by (mixed: c2, c3 with tn, or 1)
    const float4* nrm = nullptr; // normals of the pair kernel (scaled in the non-adjoint mixed form; target normals with tn)
    float2* b = nullptr;         // staged sources b = Av B
    const uint4* bop = nullptr;  // split MFMA operand (mixed)
    float* unscale = nullptr;    // per-column 1 / (A scale x b scale) of the mixed path
    // launch shape of the single-instance pair kernels (bdr_setup)
    bool planned = false;
    bool lean = false;           // set before bdr_setup: the fewest automatic segments (small partials)
    int S = 0;                   // source segments
    int seg_len = 0;             // plain kernels: sources per segment (multiple of 64)
    int steps = 0;               // MFMA kernels: 8-source steps per segment
    double rounds = 0.0;         // work-groups of one pair launch / resident work-groups of the device
    WsSource ws;
    EpiArgs ea{};
    // bdr_begin records its staging launches in sd instead of launching them
    // (mixed precision with the FP16 split, or FP32)
    bool defer = false;
    StageDesc sd{};
};

// Phase 0 (no launches): arithmetic, operand widths and the launch shape of
// the single-instance pair kernels for the loaded parameters.
void bdr_setup(Handle& h, BdrPlan& pl);
// Device workspace bytes of a single-instance apply planned by bdr_setup
// (an upper bound of what its phases take from pl.ws); need_rmin: the minimum
// centroid distance is not yet known and is computed in the workspace.
size_t bdr_workspace_bytes(const Handle& h, const BdrPlan& pl, bool need_rmin);
// Phase 1: parameters, Av gather and (mixed) the scaled hi/lo operand split.
void bdr_begin(Handle& h, const ApplyArgs& a, BdrPlan& pl, hipStream_t s);
// Reserves the partial buffer for S segment slices and records it in pl.ea.
float* bdr_reserve_part(Handle& h, BdrPlan& pl, int S);
// Phase 2 (single instance): the pair kernel launches.
void bdr_pairs(Handle& h, BdrPlan& pl, hipStream_t s);
// Phase 3: the vertex epilogue.
void bdr_end(Handle& h, const ApplyArgs& a, BdrPlan& pl, hipStream_t s);

int wcsp_plain(int wcs);
int wcsp_mixed(int wcs);
CoefPack make_pack(const Params& p, int q0, int kg, const float* scale_per_chunk, const cfloat* rot = nullptr);

}  // namespace baemm
