// batch.hip -- boundary-operator applies of several instances (meshes of
// uneven triangle counts, wave numbers and RHS widths) in ONE persistent
// launch per arithmetic mode.
//
// An instance large enough to fill the device with its own pair launch runs
// the single-instance path on its stream (Av gather, for the MFMA path the
// scaled hi/lo operand split, the pair launch, the vertex epilogue).  The
// others are staged together: each stage (Av gather with the scaled normals,
// column maxima, operand split, vertex epilogue) is one launch over all of
// them on the call's stream, from per-instance descriptors, computing what
// the single-instance kernels compute.  For their pair sums a work list of
// items (instance, target tile, source range, chunk group) is built on the
// host, ordered by decreasing cost, and walked by a persistent grid: every
// block takes the next item from a device counter (one vector atomic per
// item) until the list is exhausted, which every block reaches.  The partial
// sums of an item go to its instance's segment slice; the per-instance vertex
// epilogues follow.  Results are bit-identical to the single-instance path
// for the same segment boundaries, since every item computes exactly what the
// corresponding block of the single launch computes.
#include <algorithm>
#include <cstring>
#include <numeric>

#include "batch.hpp"
#include "ops.hpp"
#include "stage.hpp"

namespace baemm {

struct BatchInst {
    const float4* mid;
    const float4* nrm;
    const uint4* bop;     // mixed: split operand
    const float2* b;      // plain: staged sources
    float* part;
    int64_t seg_stride;
    int ldp, ldb;
    int nt, nt_pad, nst;
    int coef_off;         // first chunk coefficient of the instance in the coefficient table
    const int* skip;      // GMRES stop flag of the instance (null: none): its items return at once
};

struct BatchItem {
    int inst;
    int tile;
    int s0, s1;           // mixed: 8-source steps; plain: sources (multiples of 64)
    int q0;
    int variant;          // kernel variant id (see the switches below)
    int seg;
    int pad;
};

// ----------------------------------------------------------------- MFMA path
// Variants: (WCSP, KG) in {(16,1),(16,2),(32,1),(32,2),(64,1)} x (non-adjoint
// form, general adjoint form, non-adjoint form on the target normals).
static int mixed_variant(int WCSP, int KG, bool adj, bool tn)
{
    int v = -1;
    if (WCSP == 16) v = KG == 1 ? 0 : (KG == 2 ? 1 : -1);
    if (WCSP == 32) v = KG == 1 ? 2 : (KG == 2 ? 3 : -1);
    if (WCSP == 64 && KG == 1) v = 4;
    if (v < 0) throw ArgError{"internal: no batched mixed variant"};
    return tn ? 10 + v : 2 * v + (adj ? 1 : 0);
}

// Largest LDS need over the variants (all have 4 column blocks or fewer).
constexpr int kWsOpMax = WsLds<32, 2>::B_ELEMS;
constexpr int kWsAMax = WsLds<32, 2>::A_ELEMS;
static_assert(WsLds<64, 1>::B_ELEMS <= kWsOpMax && WsLds<64, 1>::A_ELEMS <= kWsAMax, "LDS bound");
static_assert(WsLds<16, 2>::B_ELEMS <= kWsOpMax && WsLds<16, 2>::A_ELEMS <= kWsAMax, "LDS bound");
static_assert(2 * MixedLds<16, 1>::OP_ELEMS <= kWsOpMax + kWsAMax, "LDS bound");

// Wave-specialised blocks of 8 waves (generation waves feed MFMA waves), as
// the single-instance MFMA kernel.
__global__ __launch_bounds__(512, 4) void bdr_mixed_batch_kernel(
    const BatchInst* __restrict__ insts, const BatchItem* __restrict__ items, const ChunkCoef* __restrict__ coefs,
    int n_items, unsigned* __restrict__ counter)
{
    __shared__ uint4 s_opa[kWsOpMax + kWsAMax];
    __shared__ int s_item;
    uint4* const s_op = s_opa;
    uint4* const s_a = s_opa + kWsOpMax;
    for (;;) {
        if (threadIdx.x == 0) s_item = (int)atomicAdd(counter, 1u);
        __syncthreads();
        const int it = s_item;
        __syncthreads();
        if (it >= n_items) break;
        const BatchItem w = items[it];
        const BatchInst in = insts[w.inst];
        if (in.skip && *in.skip) continue;   // uniform over the block: s_item was read by every thread
        const ChunkCoef* cf = coefs + in.coef_off + w.q0;
        float* out = in.part + (int64_t)w.seg * in.seg_stride;
#define BAEMM_MIXB(ID, WCSP_, KG_, ADJ_, TN_)                                                                     \
    case ID: {                                                                                                   \
        CoefPack cp;                                                                                             \
        for (int g = 0; g < KG_; ++g) cp.c[g] = cf[g];                                                           \
        mixed_ws_block<WCSP_, KG_, ADJ_, false, TN_>(in.mid, in.nt_pad, true, in.mid, in.nrm, in.nt, in.bop,      \
                                                     in.nst, cp, out,                                             \
                                         in.ldp, w.q0, w.tile,                                                     \
                                         w.s0, w.s1, s_op, s_a);                                                 \
        break;                                                                                                   \
    }
        // 16-column single chunks without the adjoint double layer: the
        // single-wave block body on two target tiles at
        // once (waves 0-3 tile 2t, waves 4-7 tile 2t+1): every wave generates
        // kernel values, which these generation-bound items need.
#define BAEMM_MIXB2(ID, KG_, ADJ_, TN_)                                                                          \
    case ID: {                                                                                                   \
        CoefPack cp;                                                                                             \
        for (int g = 0; g < KG_; ++g) cp.c[g] = cf[g];                                                           \
        using LL = MixedLds<16, KG_>;                                                                            \
        const int half = (threadIdx.x >> 6) >= 4;                                                                \
        const int t2 = 2 * w.tile + half;                                                                        \
        const bool real_tile = t2 < in.nt_pad / kMixTile;                                                        \
        mixed_block<16, KG_, ADJ_, 0, TN_>(in.mid, in.nt_pad, true, in.mid, in.nrm, in.nt, in.bop, in.nst, cp,     \
                                   real_tile ? out : nullptr, in.ldp,                                             \
                                   w.q0, real_tile ? t2 : 2 * w.tile, w.s0, w.s1, s_op + half * LL::OP_ELEMS);     \
        break;                                                                                                   \
    }
        switch (w.variant) {
            BAEMM_MIXB2(0, 1, false, false)
            BAEMM_MIXB(1, 16, 1, true, false)
            BAEMM_MIXB(2, 16, 2, false, false)
            BAEMM_MIXB(3, 16, 2, true, false)
            BAEMM_MIXB(4, 32, 1, false, false)
            BAEMM_MIXB(5, 32, 1, true, false)
            BAEMM_MIXB(6, 32, 2, false, false)
            BAEMM_MIXB(7, 32, 2, true, false)
            BAEMM_MIXB(8, 64, 1, false, false)
            BAEMM_MIXB(9, 64, 1, true, false)
            BAEMM_MIXB2(10, 1, false, true)
            BAEMM_MIXB(11, 16, 2, false, true)
            BAEMM_MIXB(12, 32, 1, false, true)
            BAEMM_MIXB(13, 32, 2, false, true)
            BAEMM_MIXB(14, 64, 1, false, true)
            default: break;
        }
#undef BAEMM_MIXB
#undef BAEMM_MIXB2
    }
}

// ----------------------------------------------------------------- FP32 path
// Variants: WCSP in {4, 8, 16, 32, 64} with KG = 1, x ADJ.
static int plain_variant(int WCSP, bool adj)
{
    int v = -1;
    switch (WCSP) {
        case 4: v = 0; break;
        case 8: v = 1; break;
        case 16: v = 2; break;
        case 32: v = 3; break;
        case 64: v = 4; break;
        default: break;
    }
    if (v < 0) throw ArgError{"internal: no batched plain variant"};
    return 2 * v + (adj ? 1 : 0);
}

__global__ __launch_bounds__(kTile) void bdr_plain_batch_kernel(const BatchInst* __restrict__ insts,
                                                                const BatchItem* __restrict__ items,
                                                                const ChunkCoef* __restrict__ coefs, int n_items,
                                                                unsigned* __restrict__ counter)
{
    __shared__ float4 s_y[kSrcBlock];
    __shared__ float4 s_n[kSrcBlock];
    __shared__ float2 s_b[kSrcBlock * 64];
    __shared__ int s_item;
    for (;;) {
        if (threadIdx.x == 0) s_item = (int)atomicAdd(counter, 1u);
        __syncthreads();
        const int it = s_item;
        __syncthreads();
        if (it >= n_items) break;
        const BatchItem w = items[it];
        const BatchInst in = insts[w.inst];
        if (in.skip && *in.skip) continue;
        CoefPack cp;
        cp.c[0] = coefs[in.coef_off + w.q0];
        float* out = in.part + (int64_t)w.seg * in.seg_stride;
#define BAEMM_PLB(ID, WCSP_, ADJ_)                                                                              \
    case ID:                                                                                                   \
        plain_block<OP_BDR, WCSP_, 1, ADJ_>(in.mid, in.nrm, in.nt, in.mid, in.nrm, in.nt, in.b, in.ldb, cp, out, \
                                            in.ldp, w.q0, w.tile, w.s0, w.s1, s_y, s_n, s_b);                  \
        break;
        switch (w.variant) {
            BAEMM_PLB(0, 4, false)
            BAEMM_PLB(1, 4, true)
            BAEMM_PLB(2, 8, false)
            BAEMM_PLB(3, 8, true)
            BAEMM_PLB(4, 16, false)
            BAEMM_PLB(5, 16, true)
            BAEMM_PLB(6, 32, false)
            BAEMM_PLB(7, 32, true)
            BAEMM_PLB(8, 64, false)
            BAEMM_PLB(9, 64, true)
            default: break;
        }
#undef BAEMM_PLB
    }
}

__global__ void batch_zero_kernel(unsigned* p) { p[threadIdx.x] = 0u; }

// ------------------------------------------------- batched staging, epilogue
// Instance k owns the flattened range [off[k], off[k+1]) of a launch; off has
// n + 1 entries, off[0] = 0.
__device__ __forceinline__ int owner_of(const int64_t* __restrict__ off, int n, int64_t idx)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= idx) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Per instance: b = Av B (nt_pad x ldb), the scaled source normals (nt_pad,
// non-adjoint MFMA form) and zeroed column maxima (ldb, MFMA path).
template <int BT>
__global__ void batch_stage_kernel(const StageDesc* __restrict__ sd, const int64_t* __restrict__ off, int n)
{
    const int64_t total = off[n];
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int k = owner_of(off, n, idx);
        const StageDesc& d = sd[k];
        int64_t e = idx - off[k];
        const int64_t nav = (int64_t)d.nt_pad * d.ldb;
        if (e < nav) {
            d.b[e] = av_elem<BT>(d.tris, d.area, d.nt, d.B, d.ldB, d.wcs, d.WCSP, d.ldb, e);
            continue;
        }
        e -= nav;
        const int64_t nn = d.nrm_out ? d.nt_pad : 0;
        if (e < nn) {
            const float4 v = d.nrm_in[e];
            d.nrm_out[e] = make_float4(d.nrm_s * v.x, d.nrm_s * v.y, d.nrm_s * v.z, v.w);
            continue;
        }
        d.colmax[e - nn] = 0u;
    }
}

constexpr int kColmaxRows = 256;   // rows per block of the batched column maxima

// Column maxima of |re|, |im| of b (rows < nt): block = 4 row lanes x 64
// columns over kColmaxRows rows of one instance; one atomicMax per column and
// block (positive floats order like their bit patterns).
__global__ __launch_bounds__(256) void batch_colmax_kernel(const StageDesc* __restrict__ sd,
                                                           const int64_t* __restrict__ boff, int n)
{
    __shared__ float sm[4][64];
    const int k = owner_of(boff, n, blockIdx.x);
    const StageDesc& d = sd[k];
    const int r0 = (int)(blockIdx.x - boff[k]) * kColmaxRows;
    const int r1 = min(r0 + kColmaxRows, d.nt);
    const int c = threadIdx.x & 63, rl = threadIdx.x >> 6;
    for (int c0 = 0; c0 < d.ldb; c0 += 64) {
        float m = 0.0f;
        if (c0 + c < d.ldb)
            for (int r = r0 + rl; r < r1; r += 4) {
                const float2 v = d.b[(int64_t)r * d.ldb + c0 + c];
                m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
            }
        sm[rl][c] = m;
        __syncthreads();
        if (rl == 0 && c0 + c < d.ldb) {
            m = fmaxf(fmaxf(sm[0][c], sm[1][c]), fmaxf(sm[2][c], sm[3][c]));
            atomicMax(&d.colmax[c0 + c], __float_as_uint(m));
        }
        __syncthreads();
    }
}

// Per MFMA instance: the hi/lo B operand (K x nst x NCB x 64 elements) with
// the column scales taken from the maxima, then the scales and unscales
// (ldb), as split_kernel computes them.
__global__ void batch_split_kernel(const StageDesc* __restrict__ sd, const int64_t* __restrict__ off, int n)
{
    const int64_t total = off[n];
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int k = owner_of(off, n, idx);
        const StageDesc& d = sd[k];
        const int64_t e = idx - off[k];
        const int64_t nb = (int64_t)d.K * d.nst * (2 * d.WCSP / 32) * 64;
        if (e < nb) {
            split_elem(d.b, d.ldb, d.nst, d.WCSP, e, [&](int pc) { return col_scale(d.colmax[pc]); }, d.bop);
        } else {
            const int pc = (int)(e - nb);
            const float sc = col_scale(d.colmax[pc]);
            d.scale[pc] = sc;
            d.unscale[pc] = 1.0f / (sc * d.sA[pc / d.WCSP]);
        }
    }
}

// Vertex epilogues of the instances: instance k owns blocks [eoff[k], eoff[k+1]).
template <int BT, int CT>
__global__ __launch_bounds__(256) void batch_epilogue_kernel(const EpiArgs* __restrict__ ea,
                                                             const int64_t* __restrict__ eoff, int n)
{
    const int k = owner_of(eoff, n, blockIdx.x);
    const EpiArgs& a = ea[k];
    const int lpv = epi_lanes(a.W);
    const int v = (int)(blockIdx.x - eoff[k]) * (256 / lpv) + threadIdx.x / lpv;
    if (v >= a.nv) return;
    epi_vertex<BT, CT>(a, v, threadIdx.x & (lpv - 1), lpv);
}

// ----------------------------------------------------------------- host

namespace {

struct Prepared {
    BdrPlan pl;
};

// Copies blob to the device buffer of slot through its pinned staging buffer
// (after the previous copy from it has completed); returns the device copy.
char* upload_blob(BatchWorkspace& bw, int slot, const std::vector<char>& blob, hipStream_t s)
{
    const size_t total = blob.size();
    bw.dev[slot].reserve(total);
    if (bw.host_bytes[slot] < total) {
        if (bw.host[slot]) {
            if (bw.ev[slot]) BAEMM_HIP(hipEventSynchronize(bw.ev[slot]));
            BAEMM_HIP(hipHostFree(bw.host[slot]));
            bw.host[slot] = nullptr;
        }
        BAEMM_HIP(hipHostMalloc(&bw.host[slot], total, hipHostMallocDefault));
        bw.host_bytes[slot] = total;
    }
    if (!bw.ev[slot]) BAEMM_HIP(hipEventCreateWithFlags(&bw.ev[slot], hipEventDisableTiming));
    BAEMM_HIP(hipEventSynchronize(bw.ev[slot]));
    std::memcpy(bw.host[slot], blob.data(), total);
    BAEMM_HIP(hipMemcpyAsync(bw.dev[slot].ptr, bw.host[slot], total, hipMemcpyHostToDevice, s));
    BAEMM_HIP(hipEventRecord(bw.ev[slot], s));
    return bw.dev[slot].as<char>();
}

template <class T>
size_t append(std::vector<char>& blob, const T* data, size_t count)
{
    const size_t at = (blob.size() + 255) / 256 * 256;
    blob.resize(at + sizeof(T) * count);
    std::memcpy(blob.data() + at, data, sizeof(T) * count);
    return at;
}

unsigned stride_grid(int64_t total)
{
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 8192));
}

// The staging launches the deferred instances recorded (bdr_begin with
// pl.defer), three launches for all of them on s.
void stage_deferred(std::vector<Prepared>& prep, const std::vector<char>& own, int Btype, BatchWorkspace& bw,
                    hipStream_t s)
{
    std::vector<StageDesc> sd;
    std::vector<int64_t> off(1, 0), boff(1, 0), soff(1, 0);
    for (size_t k = 0; k < prep.size(); ++k) {
        const StageDesc& d = prep[k].pl.sd;
        if (own[k] || !d.b) continue;
        sd.push_back(d);
        off.push_back(off.back() + (int64_t)d.nt_pad * d.ldb + (d.nrm_out ? d.nt_pad : 0) + (d.split ? d.ldb : 0));
        boff.push_back(boff.back() + (d.split ? (d.nt + kColmaxRows - 1) / kColmaxRows : 0));
        soff.push_back(soff.back() + (d.split ? (int64_t)d.K * d.nst * (2 * d.WCSP / 32) * 64 + d.ldb : 0));
    }
    if (sd.empty()) return;
    const int m = (int)sd.size();
    std::vector<char> blob;
    const size_t a_sd = append(blob, sd.data(), sd.size());
    const size_t a_off = append(blob, off.data(), off.size());
    const size_t a_boff = append(blob, boff.data(), boff.size());
    const size_t a_soff = append(blob, soff.data(), soff.size());
    char* dp = upload_blob(bw, 2, blob, s);
    const StageDesc* dsd = reinterpret_cast<const StageDesc*>(dp + a_sd);
    const int64_t* doff = reinterpret_cast<const int64_t*>(dp + a_off);
    if (Btype == kC128)
        hipLaunchKernelGGL(batch_stage_kernel<kC128>, dim3(stride_grid(off.back())), dim3(256), 0, s, dsd, doff, m);
    else
        hipLaunchKernelGGL(batch_stage_kernel<kC64>, dim3(stride_grid(off.back())), dim3(256), 0, s, dsd, doff, m);
    BAEMM_HIP(hipGetLastError());
    if (boff.back() > 0) {
        hipLaunchKernelGGL(batch_colmax_kernel, dim3((unsigned)boff.back()), dim3(256), 0, s, dsd,
                           reinterpret_cast<const int64_t*>(dp + a_boff), m);
        hipLaunchKernelGGL(batch_split_kernel, dim3(stride_grid(soff.back())), dim3(256), 0, s, dsd,
                           reinterpret_cast<const int64_t*>(dp + a_soff), m);
        BAEMM_HIP(hipGetLastError());
    }
}

template <int BT, int CT>
void launch_batch_epilogue(const EpiArgs* ea, const int64_t* eoff, int m, int64_t blocks, hipStream_t s)
{
    hipLaunchKernelGGL((batch_epilogue_kernel<BT, CT>), dim3((unsigned)blocks), dim3(256), 0, s, ea, eoff, m);
}

// The vertex epilogues of the deferred instances, one launch on s.
void epilogue_deferred(std::vector<Prepared>& prep, const std::vector<char>& own, int Btype, int Ctype,
                       BatchWorkspace& bw, hipStream_t s)
{
    std::vector<EpiArgs> ea;
    std::vector<int64_t> eoff(1, 0);
    for (size_t k = 0; k < prep.size(); ++k) {
        const BdrPlan& pl = prep[k].pl;
        if (own[k] || (!pl.kernel_on && !pl.mass_on)) continue;   // the reference leaves C_out untouched
        ea.push_back(pl.ea);
        const int vpb = 256 / epi_lanes(pl.ea.W);
        eoff.push_back(eoff.back() + (pl.ea.nv + vpb - 1) / vpb);
    }
    if (ea.empty()) return;
    const int m = (int)ea.size();
    std::vector<char> blob;
    const size_t a_ea = append(blob, ea.data(), ea.size());
    const size_t a_off = append(blob, eoff.data(), eoff.size());
    char* dp = upload_blob(bw, 3, blob, s);
    const EpiArgs* dea = reinterpret_cast<const EpiArgs*>(dp + a_ea);
    const int64_t* doff = reinterpret_cast<const int64_t*>(dp + a_off);
    const int64_t blocks = eoff.back();
    if (Btype == kC128 && Ctype == kC128) launch_batch_epilogue<kC128, kC128>(dea, doff, m, blocks, s);
    else if (Btype == kC128 && Ctype == kC64) launch_batch_epilogue<kC128, kC64>(dea, doff, m, blocks, s);
    else if (Btype == kC64 && Ctype == kC128) launch_batch_epilogue<kC64, kC128>(dea, doff, m, blocks, s);
    else launch_batch_epilogue<kC64, kC64>(dea, doff, m, blocks, s);
    BAEMM_HIP(hipGetLastError());
}

int persistent_grid(const void* fn, int block, int n_items, int grid_mult)
{
    int dev = 0, cus = 256, per_cu = 1;
    BAEMM_HIP(hipGetDevice(&dev));
    BAEMM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    BAEMM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block, 0));
    per_cu = std::max(1, per_cu);
    const int mult = grid_mult > 0 ? std::min(grid_mult, per_cu) : per_cu;
    return std::max(1, std::min(n_items, cus * mult));
}

}  // namespace

BatchTuning& batch_tuning()
{
    static BatchTuning t;
    return t;
}

void apply_boundary_batch(std::vector<BatchEntry>& es, BatchWorkspace& bw, hipStream_t s, BatchStats* stats)
{
    const BatchTuning tune = batch_tuning();
    const int n = (int)es.size();
    for (int k = 0; k < n; ++k)
        if (es[k].h->precision == BAEMM_PRECISION_FP32_MFMA)
            throw ArgError{"batched apply: instances must use BAEMM_PRECISION_FP32 or BAEMM_PRECISION_MIXED"};
    std::vector<Prepared> prep(n);
    // Events: [0] fork to the streams of the instances that run their own
    // launches, [1] end of the work on s, [2 + k] instance k: done on its
    // stream (own launch), or its stream's earlier work (staged on s).
    while ((int)bw.sync.size() < 2 + n) {
        hipEvent_t e = nullptr;
        BAEMM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        bw.sync.push_back(e);
    }
    BAEMM_HIP(hipEventRecord(bw.sync[0], s));
    // An instance whose own pair launch fills the device at least
    // tune.own_rounds times runs the single-instance path on its stream,
    // beside the others; the rest are staged together on s and share the
    // persistent launch, whose work list they fill.  The choice depends on the
    // instance and the tuning alone, so its result is the same in any mix.
    std::vector<char> own(n, 0);
    for (int k = 0; k < n; ++k) {
        Handle& h = *es[k].h;
        p0_invalidate(h);   // the batched apply leaves no P0 buffers
        BdrPlan& pl = prep[k].pl;
        bdr_setup(h, pl);
        if (pl.kernel_on && pl.rounds >= tune.own_rounds) {
            own[k] = 1;
            hipStream_t sk = h.stream;
            BAEMM_HIP(hipStreamWaitEvent(sk, bw.sync[0], 0));
            bdr_begin(h, es[k].a, pl, sk);
            bdr_pairs(h, pl, sk);
            bdr_end(h, es[k].a, pl, sk);
            BAEMM_HIP(hipEventRecord(bw.sync[2 + k], sk));
        } else {
            // staged on s: s waits for the instance stream's earlier work (its
            // inputs may come from there, and its workspaces are reused)
            BAEMM_HIP(hipEventRecord(bw.sync[2 + k], h.stream));
            BAEMM_HIP(hipStreamWaitEvent(s, bw.sync[2 + k], 0));
            pl.defer = true;
            bdr_begin(h, es[k].a, pl, s);   // records the staging launches in pl.sd
        }
    }
    stage_deferred(prep, own, es[0].a.Btype, bw, s);

    for (int mode = 0; mode < 2; ++mode) {     // 0: MFMA instances, 1: FP32 instances
        const bool mixed = (mode == 0);
        std::vector<BatchInst> insts;
        std::vector<ChunkCoef> coefs;
        std::vector<BatchItem> items;
        std::vector<double> cost;
        std::vector<int> owner;
        for (int k = 0; k < n; ++k) {
            BdrPlan& pl = prep[k].pl;
            if (!pl.kernel_on || pl.mixed != mixed || own[k]) continue;
            Handle& h = *es[k].h;
            const Params& p = h.prm;
            BatchInst in{};
            in.mid = h.d_mid;
            in.nrm = pl.nrm;
            in.bop = pl.bop;
            in.b = pl.b;
            in.ldp = pl.ldp;
            in.ldb = pl.ldb;
            in.nt = (int)h.nt;
            in.nt_pad = (int)h.nt_pad;
            in.nst = (int)(h.nt_pad / 8);
            in.coef_off = (int)coefs.size();
            in.skip = h.step_skip;
            for (int q = 0; q < p.K; ++q)
                coefs.push_back(make_pack(p, q, 1, mixed ? pl.sA.data() : nullptr, mixed ? pl.rot.data() : nullptr).c[0]);
            // source ranges of equal length across the batch
            int len, S, tiles, nsrc_units;
            if (mixed) {
                len = std::max(8, tune.mixed_steps);
                nsrc_units = in.nst;
                tiles = in.nt_pad / kMixTile;
            } else {
                len = std::max(1, tune.plain_blocks) * kSrcBlock;
                nsrc_units = in.nt_pad;
                tiles = in.nt_pad / kTile;
            }
            S = (nsrc_units + len - 1) / len;
            in.part = bdr_reserve_part(h, pl, S);
            in.seg_stride = pl.ea.seg_stride;
            const int inst_id = (int)insts.size();
            insts.push_back(in);
            owner.push_back(k);
            for (int q0 = 0; q0 < p.K;) {
                int kg = 1;
                if (mixed) {
                    const int maxkg = std::max(1, kMaxCols / pl.WCSP);
                    const int lim = std::min({maxkg, (int)p.K - q0, 2});
                    while (kg * 2 <= lim) kg *= 2;
                }
                const int var = mixed ? mixed_variant(pl.WCSP, kg, pl.adj, pl.tn) : plain_variant(pl.WCSP, pl.adj);
                // single 16-column chunks: items of two 128-target tiles
                const bool two = mixed && pl.WCSP == 16 && kg == 1 && !pl.adj;
                const int item_tiles = two ? (tiles + 1) / 2 : tiles;
                for (int seg = 0; seg < S; ++seg) {
                    const int s0 = seg * len, s1 = std::min(s0 + len, nsrc_units);
                    // cost model: kernel-value generation per chunk plus contraction per column
                    double per = mixed ? (double)(s1 - s0) * 8.0 * (kg * (1.0 + 0.35 * pl.WCSP / 32.0))
                                       : (double)(s1 - s0) * (kg * (1.0 + 0.12 * pl.WCSP));
                    if (two) per *= 2.0;
                    for (int t = 0; t < item_tiles; ++t) {
                        items.push_back(BatchItem{inst_id, t, s0, s1, q0, var, seg, 0});
                        cost.push_back(per);
                    }
                }
                q0 += kg;
            }
        }
        if (items.empty()) continue;
        // Longest items first (a greedy schedule for the dynamic queue); the
        // order inside equal costs keeps tiles of one source range together.
        std::vector<int> order(items.size());
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });
        std::vector<BatchItem> sorted(items.size());
        for (size_t k = 0; k < order.size(); ++k) sorted[k] = items[order[k]];

        const size_t b_insts = sizeof(BatchInst) * insts.size();
        const size_t b_coefs = sizeof(ChunkCoef) * coefs.size();
        const size_t b_items = sizeof(BatchItem) * sorted.size();
        const size_t off_coefs = (b_insts + 255) / 256 * 256;
        const size_t off_items = off_coefs + (b_coefs + 255) / 256 * 256;
        const size_t off_ctr = off_items + (b_items + 255) / 256 * 256;
        const size_t total = off_ctr + 256;
        bw.dev[mode].reserve(total);
        // pinned staging; wait until the previous upload from it has completed
        if (bw.host_bytes[mode] < total) {
            if (bw.host[mode]) {
                if (bw.ev[mode]) BAEMM_HIP(hipEventSynchronize(bw.ev[mode]));
                BAEMM_HIP(hipHostFree(bw.host[mode]));
                bw.host[mode] = nullptr;
            }
            BAEMM_HIP(hipHostMalloc(&bw.host[mode], total, hipHostMallocDefault));
            bw.host_bytes[mode] = total;
        }
        if (!bw.ev[mode]) BAEMM_HIP(hipEventCreateWithFlags(&bw.ev[mode], hipEventDisableTiming));
        BAEMM_HIP(hipEventSynchronize(bw.ev[mode]));
        char* hp = static_cast<char*>(bw.host[mode]);
        std::memcpy(hp, insts.data(), b_insts);
        std::memcpy(hp + off_coefs, coefs.data(), b_coefs);
        std::memcpy(hp + off_items, sorted.data(), b_items);
        char* dp = bw.dev[mode].as<char>();
        BAEMM_HIP(hipMemcpyAsync(dp, hp, off_ctr, hipMemcpyHostToDevice, s));
        BAEMM_HIP(hipEventRecord(bw.ev[mode], s));
        unsigned* ctr = reinterpret_cast<unsigned*>(dp + off_ctr);
        hipLaunchKernelGGL(batch_zero_kernel, dim3(1), dim3(64), 0, s, ctr);
        BAEMM_HIP(hipGetLastError());

        const void* fn = mixed ? (const void*)bdr_mixed_batch_kernel : (const void*)bdr_plain_batch_kernel;
        const int block = mixed ? 512 : kTile;
        const int grid = persistent_grid(fn, block, (int)sorted.size(), mixed ? tune.mixed_grid_mult : tune.plain_grid_mult);
        Handle& h0 = *es[owner[0]].h;
        Timed tm(h0, s);
        if (mixed)
            hipLaunchKernelGGL(bdr_mixed_batch_kernel, dim3(grid), dim3(block), 0, s,
                               reinterpret_cast<const BatchInst*>(dp), reinterpret_cast<const BatchItem*>(dp + off_items),
                               reinterpret_cast<const ChunkCoef*>(dp + off_coefs), (int)sorted.size(), ctr);
        else
            hipLaunchKernelGGL(bdr_plain_batch_kernel, dim3(grid), dim3(block), 0, s,
                               reinterpret_cast<const BatchInst*>(dp), reinterpret_cast<const BatchItem*>(dp + off_items),
                               reinterpret_cast<const ChunkCoef*>(dp + off_coefs), (int)sorted.size(), ctr);
        BAEMM_HIP(hipGetLastError());
        tm.done();
        if (stats) {
            stats->items[mode] = (int64_t)sorted.size();
            stats->grid[mode] = grid;
            stats->instances[mode] = (int64_t)insts.size();
        }
    }
    epilogue_deferred(prep, own, es[0].a.Btype, es[0].a.Ctype, bw, s);
    // results and workspaces of the instances staged on s are done there;
    // their streams wait for that before any later work on the handles
    BAEMM_HIP(hipEventRecord(bw.sync[1], s));
    for (int k = 0; k < n; ++k) {
        if (own[k]) BAEMM_HIP(hipStreamWaitEvent(s, bw.sync[2 + k], 0));
        else BAEMM_HIP(hipStreamWaitEvent(es[k].h->stream, bw.sync[1], 0));
    }
}

void BatchWorkspace::release()
{
    for (int m = 0; m < kSlots; ++m) {
        if (ev[m]) {
            (void)hipEventSynchronize(ev[m]);
            (void)hipEventDestroy(ev[m]);
            ev[m] = nullptr;
        }
        if (host[m]) (void)hipHostFree(host[m]);
        host[m] = nullptr;
        host_bytes[m] = 0;
        dev[m].release();
    }
    for (hipEvent_t e : sync) (void)hipEventDestroy(e);
    sync.clear();
}

}  // namespace baemm
