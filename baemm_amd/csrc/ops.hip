// ops.hip -- device-resident operator applies: staging, pair-kernel dispatch,
// epilogues (AvT, SL diagonal, mass, alpha/beta) and the incident waves.
#include <algorithm>
#include <numeric>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <utility>

#include "ops.hpp"
#include "stage.hpp"

namespace baemm {

// --------------------------------------------------------------- helpers

static inline int next_pow2(int x)
{
    int p = 1;
    while (p < x) p <<= 1;
    return p;
}

// Padded chunk width of the plain kernels (instantiated 4..64).
int wcsp_plain(int wcs) { return std::max(4, next_pow2(wcs)); }
// Padded chunk width of the MFMA kernel (16, 32 or 64: 2*WCSP real columns
// must fill 32-column MFMA blocks).
int wcsp_mixed(int wcs) { return std::max(16, next_pow2(wcs)); }

// --------------------------------------------------------------- staging

// b = Av B row by row (av_elem).
template <int BT>
__global__ void av_gather_kernel(const int4* __restrict__ tris, const float* __restrict__ area, int nt,
                                 int64_t total, const void* __restrict__ Bv, int64_t ldB, int wcs, int WCSP,
                                 int ldb, float2* __restrict__ b)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    b[idx] = av_elem<BT>(tris, area, nt, Bv, ldB, wcs, WCSP, ldb, idx);
}

// b[r][q*WCSP+kk] = (conj?) G[r][q*wcs+kk] for r < nrows, zero padding.
template <int BT>
__global__ void stage_rows_kernel(int nrows, int64_t total, const void* __restrict__ Gv, int64_t ldG, int wcs,
                                  int WCSP, int ldb, int conj, float2* __restrict__ b)
{
    using T = typename Elem<BT>::type;
    const T* G = static_cast<const T*>(Gv);
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int r = (int)(idx / ldb);
    const int pc = (int)(idx % ldb);
    const int q = pc / WCSP, kk = pc % WCSP;
    float2 out = make_float2(0.f, 0.f);
    if (r < nrows && kk < wcs) {
        const double2 v = to_d2(G[(int64_t)r * ldG + q * wcs + kk]);
        out = make_float2((float)v.x, (float)(conj ? -v.y : v.y));
    }
    b[idx] = out;
}

__global__ void fill_u32_kernel(unsigned* __restrict__ p, int n, unsigned v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Per-column max of |re|, |im| (positive floats order like their bit patterns).
// Block: 4 row lanes x 64 columns; blocks stride over rows; one atomicMax per
// column and block.  colmax must be zero on entry.
__global__ __launch_bounds__(256) void colmax_kernel(const float2* __restrict__ b, int nrows, int ldb,
                                                     unsigned* __restrict__ colmax)
{
    __shared__ float sm[4][64];
    const int c = threadIdx.x & 63, rl = threadIdx.x >> 6;
    for (int c0 = 0; c0 < ldb; c0 += 64) {
        float m = 0.0f;
        if (c0 + c < ldb)
            for (int r = blockIdx.x * 4 + rl; r < nrows; r += gridDim.x * 4) {
                const float2 v = b[(int64_t)r * ldb + c0 + c];
                m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
            }
        sm[rl][c] = m;
        __syncthreads();
        if (rl == 0 && c0 + c < ldb) {
            m = fmaxf(fmaxf(sm[0][c], sm[1][c]), fmaxf(sm[2][c], sm[3][c]));
            atomicMax(&colmax[c0 + c], __float_as_uint(m));
        }
        __syncthreads();
    }
}

// b = Av B (av_elem) and the maxima of |re|, |im| over each column of b
// (positive floats order like their bit patterns).  Threads stride over the
// elements by the grid's thread count, a multiple of ldb, so each thread stays
// in one column: its running maximum, then per block in LDS (ldb entries of
// dynamic shared memory), then one atomicMax per column and block.  colmax
// must be zero on entry.
// Largest grid of av_gather_max_kernel (more blocks: the separate gather and
// colmax_kernel).
constexpr int64_t kGatherMaxBlocks = 512;

template <int BT>
__global__ __launch_bounds__(256) void av_gather_max_kernel(const int4* __restrict__ tris,
                                                            const float* __restrict__ area, int nt, int64_t total,
                                                            const void* __restrict__ Bv, int64_t ldB, int wcs,
                                                            int WCSP, int ldb, float2* __restrict__ b,
                                                            unsigned* __restrict__ colmax)
{
    extern __shared__ unsigned s_max[];
    for (int c = threadIdx.x; c < ldb; c += blockDim.x) s_max[c] = 0u;
    __syncthreads();
    const int64_t first = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float m = 0.0f;
    for (int64_t idx = first; idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
        const float2 v = av_elem<BT>(tris, area, nt, Bv, ldB, wcs, WCSP, ldb, idx);
        b[idx] = v;
        m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
    }
    if (m > 0.0f) atomicMax(&s_max[(int)(first % ldb)], __float_as_uint(m));
    __syncthreads();
    for (int c = threadIdx.x; c < ldb; c += blockDim.x)
        if (s_max[c]) atomicMax(&colmax[c], s_max[c]);
}

// The power-of-two A scale sA_q of each chunk q.  The B scale s_pc of column
// pc is col_scale(colmax[pc]), a power of two that puts max |s b| in
// [2^13, 2^14); the epilogue multiplies by unscale[pc] = 1 / (sA_q * s_pc).
struct ChunkScales {
    float sA[64];
};

// B-operand of the MFMA kernel (split_elem) with the column scales
// col_scale(colmax[pc]); the first ncols threads also write
// unscale[pc] = 1 / (sA_q * s_pc) (ncols < the element count).
__global__ void split_kernel(const float2* __restrict__ b, int ldb, int nst, int WCSP, int K,
                             const unsigned* __restrict__ colmax, int ncols, ChunkScales cs,
                             float* __restrict__ unscale, uint4* __restrict__ bop)
{
    const int NCB = 2 * WCSP / 32;
    const int64_t total = (int64_t)K * nst * NCB * 64;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    if (idx < ncols) {
        const int pc = (int)idx;
        unscale[pc] = 1.0f / (col_scale(colmax[pc]) * cs.sA[pc / WCSP]);
    }
    split_elem(b, ldb, nst, WCSP, idx, [&](int pc) { return col_scale(colmax[pc]); }, bop);
}

// out[j] = s * nrm[j] for j < n: source normals carrying the common scale of
// the non-adjoint MFMA kernel values (out NULL: none); zero[j] = 0 for j < nzero.
__global__ void scale_nrm_kernel(const float4* __restrict__ nrm, int n, float s, float4* __restrict__ out,
                                 unsigned* __restrict__ zero, int nzero)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < nzero) zero[j] = 0u;
    if (j >= n || !out) return;
    const float4 v = nrm[j];
    out[j] = make_float4(s * v.x, s * v.y, s * v.z, v.w);
}

// B-operand of the FP32 MFMA kernel: for (chunk, step, col-block) and lane,
// the 8 FP32 values of MFMAs m = 0..7 (two 16-byte parts), in the order of the
// FP16 operand of split_kernel: rows (source 4 (lane / 32) + m / 2, re | im) of
// [[br, bi], [-bi, br]], unscaled.
__global__ void stage_f32_kernel(const float2* __restrict__ b, int ldb, int nst, int WCSP, int K,
                                 uint4* __restrict__ bop)
{
    const int NCB = 2 * WCSP / 32;
    const int64_t total = (int64_t)K * nst * NCB * 64;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int lane = (int)(idx % 64);
    int64_t rest = idx / 64;
    const int cb = (int)(rest % NCB);
    rest /= NCB;
    const int st = (int)(rest % nst);
    const int q = (int)(rest / nst);
    const int g = lane >> 5;
    const int c = cb * 32 + (lane & 31);
    const bool imcol = c >= WCSP;
    const int kk = imcol ? c - WCSP : c;
    const int pc = q * WCSP + kk;
    float e[8];
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
        const int j = 8 * st + 4 * g + jj;
        const float2 v = b[(int64_t)j * ldb + pc];
        e[2 * jj] = imcol ? v.y : v.x;
        e[2 * jj + 1] = imcol ? v.x : -v.y;
    }
    const int64_t base = ((((int64_t)q * nst + st) * NCB + cb) * 2) * 64 + lane;
    bop[base] = make_uint4(__float_as_uint(e[0]), __float_as_uint(e[1]), __float_as_uint(e[2]), __float_as_uint(e[3]));
    bop[base + 64] = make_uint4(__float_as_uint(e[4]), __float_as_uint(e[5]), __float_as_uint(e[6]), __float_as_uint(e[7]));
}

// Minimum squared distance between the targets x_i (i < ntgt) and the
// centroids y_j (j < nt), leaving out j = i when self (distinct centroids);
// sets the A scale of the mixed-precision MFMA path.  One target per thread,
// all centroids; the result (a non-negative float) is atomically min-ed as bits.
__global__ void rmin_kernel(const float4* __restrict__ tgt, int ntgt, const float4* __restrict__ mid, int nt,
                            int self, unsigned* __restrict__ out)
{
    __shared__ float4 s[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const float4 x = tgt[min(i, ntgt - 1)];
    float m = 3.0e38f;
    for (int jb = 0; jb < nt; jb += 256) {
        __syncthreads();
        if (jb + threadIdx.x < nt) s[threadIdx.x] = mid[jb + threadIdx.x];
        __syncthreads();
        const int jn = min(256, nt - jb);
        for (int jj = 0; jj < jn; ++jj) {
            const float4 y = s[jj];
            const float d0 = y.x - x.x, d1 = y.y - x.y, d2 = y.z - x.z;
            const float r2 = fmaf(d0, d0, fmaf(d1, d1, d2 * d2));
            if (!self || jb + jj != i) m = fminf(m, r2);
        }
    }
    if (i >= ntgt) m = 3.0e38f;
    for (int off = 32; off > 0; off >>= 1) m = fminf(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0 && i < ntgt) atomicMin(out, __float_as_uint(m));
}

// --------------------------------------------------------------- epilogues

// The boundary epilogue, one vertex per lpv lanes (epi_vertex).
template <int BT, int CT>
__global__ __launch_bounds__(256) void bdr_epilogue_kernel(EpiArgs a, int lpv)
{
    const int lane = threadIdx.x & (lpv - 1);
    const int v = blockIdx.x * (256 / lpv) + threadIdx.x / lpv;
    if (v >= a.nv) return;
    epi_vertex<BT, CT>(a, v, lane, lpv);
}

// Factors per wave-number chunk the MFMA near-field partials are multiplied by.
struct ChunkRot {
    float r[64], i[64];
};

// F[i][k] = alpha * rot_q * unscale_k * sum_s part_s[i][k] + beta * F[i][k]
// (far field, potential; unscale null and rot = 1 for the FP32 kernels).
template <int CT>
__global__ void rows_epilogue_kernel(int nrows, int W, int wcs, int WCSP, int S, int64_t seg_stride, int ldp,
                                     const float* __restrict__ part, double ar, double ai, double br, double bi,
                                     void* Cv, int64_t ldC, const float* __restrict__ unscale, ChunkRot rot)
{
    using TC = typename Elem<CT>::type;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)nrows * W) return;
    const int r = (int)(idx / W), k = (int)(idx % W);
    const int q = k / wcs, kk = k % wcs;
    const float* pp = part + (int64_t)r * ldp + q * 2 * WCSP + kk;
    double pr = 0.0, pi = 0.0;
    for (int s0 = 0; s0 < S; s0 += EPI_SEG) {
        float xr[EPI_SEG], xi[EPI_SEG];
#pragma unroll
        for (int j = 0; j < EPI_SEG; ++j) {
            const int64_t o = (int64_t)min(s0 + j, S - 1) * seg_stride;
            xr[j] = pp[o];
            xi[j] = pp[o + WCSP];
        }
#pragma unroll
        for (int j = 0; j < EPI_SEG; ++j)
            if (s0 + j < S) {
                pr += (double)xr[j];
                pi += (double)xi[j];
            }
    }
    if (unscale) {
        const double us = (double)unscale[q * WCSP + kk];
        pr *= us;
        pi *= us;
    }
    {
        const double rr = rot.r[q], ri = rot.i[q];
        const double tr = pr * rr - pi * ri;
        pi = pr * ri + pi * rr;
        pr = tr;
    }
    TC* C = static_cast<TC*>(Cv);
    TC& dst = C[(int64_t)r * ldC + k];
    double outr = ar * pr - ai * pi, outi = ar * pi + ai * pr;
    if (br != 0.0 || bi != 0.0) {
        const double2 c = to_d2(dst);
        outr += br * c.x - bi * c.y;
        outi += br * c.y + bi * c.x;
    }
    from_d2(dst, make_double2(outr, outi));
}

// Incident-wave values per triangle written as a single "segment" partial
// (IncidentWaveKernel_Plane.hpp:190-233, IncidentWaveKernel_Radial.hpp:5-57):
//   plane : (c1 + c2 i kappa (d.n)) exp(i kappa d.x)
//   radial: (c1 + c2 (-(n.delta)/R^2 + i kappa (n.delta)/R)) exp(i kappa R) / (4 pi R)
struct IncArgs {
    int nt, W, wcs, WCSP, ldp, radial;
    const float4* mid;
    const float4* nrm;
    float* part;
    float dirs[64][3];
    float kap[64], c1r[64], c1i[64], c2r[64], c2i[64];
};

__global__ void incident_kernel(IncArgs a)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)a.nt * a.W) return;
    const int t = (int)(idx / a.W), k = (int)(idx % a.W);
    const int q = k / a.wcs, j = k % a.wcs;
    const float4 x = a.mid[t], n = a.nrm[t];
    const double dx = a.dirs[j][0], dy = a.dirs[j][1], dz = a.dirs[j][2];
    const double kap = a.kap[q];
    double ur, ui;
    if (!a.radial) {
        const double ph = kap * (x.x * dx + x.y * dy + x.z * dz);
        const double dn = kap * (n.x * dx + n.y * dy + n.z * dz);
        const double fr = a.c1r[q] - a.c2i[q] * dn, fi = a.c1i[q] + a.c2r[q] * dn;
        const double cs = cos(ph), sn = sin(ph);
        ur = fr * cs - fi * sn;
        ui = fr * sn + fi * cs;
    } else {
        const double e0 = x.x - dx, e1 = x.y - dy, e2 = x.z - dz;
        const double R = sqrt(e0 * e0 + e1 * e1 + e2 * e2);
        const double nd = n.x * e0 + n.y * e1 + n.z * e2;
        const double A = 1.0 / (4.0 * M_PI * R);
        const double Br = -nd / (R * R), Bi = kap * nd / R;
        const double fr = A * (a.c1r[q] + a.c2r[q] * Br - a.c2i[q] * Bi);
        const double fi = A * (a.c1i[q] + a.c2r[q] * Bi + a.c2i[q] * Br);
        const double cs = cos(kap * R), sn = sin(kap * R);
        ur = fr * cs - fi * sn;
        ui = fr * sn + fi * cs;
    }
    float* out = a.part + (int64_t)t * a.ldp + q * 2 * a.WCSP + j;
    out[0] = (float)ur;
    out[a.WCSP] = (float)ui;
}

// --------------------------------------------------------------- dispatch

static void p0_record(Handle& h, const float2* b, int ldb, int WCSP, const EpiArgs* ea);

CoefPack make_pack(const Params& p, int q0, int kg, const float* scale_per_chunk, const cfloat* rot)
{
    CoefPack cp{};
    for (int g = 0; g < kg; ++g) {
        const float s = scale_per_chunk ? scale_per_chunk[q0 + g] : 1.0f;
        cp.c[g] = chunk_coefs(p, q0 + g, 1, s, rot)[0];
    }
    return cp;
}

#define PLAIN_CASES(OP, ADJ)                                                                             \
    if (WCSP == 4 && KG == 1) return &pair_plain_kernel<OP, 4, 1, ADJ>;                               \
    if (WCSP == 4 && KG == 2) return &pair_plain_kernel<OP, 4, 2, ADJ>;                               \
    if (WCSP == 8 && KG == 1) return &pair_plain_kernel<OP, 8, 1, ADJ>;                               \
    if (WCSP == 8 && KG == 2) return &pair_plain_kernel<OP, 8, 2, ADJ>;                               \
    if (WCSP == 16 && KG == 1) return &pair_plain_kernel<OP, 16, 1, ADJ>;                             \
    if (WCSP == 16 && KG == 2) return &pair_plain_kernel<OP, 16, 2, ADJ>;                             \
    if (WCSP == 32 && KG == 1) return &pair_plain_kernel<OP, 32, 1, ADJ>;                             \
    if (WCSP == 32 && KG == 2) return &pair_plain_kernel<OP, 32, 2, ADJ>;                             \
    if (WCSP == 64 && KG == 1) return &pair_plain_kernel<OP, 64, 1, ADJ>;

using PlainFn = void (*)(const float4*, const float4*, int, const float4*, const float4*, int, const float2*, int,
                         CoefPack, float*, int64_t, int, int, int, int, const int*);

static PlainFn plain_fn(int op, bool adj, int WCSP, int KG)
{
    if (op == OP_BDR && adj) { PLAIN_CASES(OP_BDR, true) }
    if (op == OP_BDR && !adj) { PLAIN_CASES(OP_BDR, false) }
    if (op == OP_NEAR) { PLAIN_CASES(OP_NEAR, false) }
    if (op == OP_FAR) { PLAIN_CASES(OP_FAR, false) }
    if (op == OP_HERG) { PLAIN_CASES(OP_HERG, false) }
    throw ArgError{"internal: no plain kernel instance"};
}

using MixedFn = void (*)(const float4*, int, int, const float4*, const float4*, int, const uint4*, int, CoefPack,
                         float*, int64_t, int, int, int, const int*);

using WsFn = void (*)(const float4*, int, int, const float4*, const float4*, int, const uint4*, int, CoefPack, float*,
                      int64_t, int, int, int, const int*);

static WsFn mixed_ws_fn(bool adj, int WCSP, int KG, bool f32, bool tn = false)
{
#define MIXWS(W_, K_)                                                                                    \
    if (WCSP == W_ && KG == K_) {                                                                        \
        if (tn) return f32 ? &bdr_mixed_ws_kernel<W_, K_, false, true, true> : &bdr_mixed_ws_kernel<W_, K_, false, false, true>; \
        if (f32) return adj ? &bdr_mixed_ws_kernel<W_, K_, true, true> : &bdr_mixed_ws_kernel<W_, K_, false, true>; \
        return adj ? &bdr_mixed_ws_kernel<W_, K_, true> : &bdr_mixed_ws_kernel<W_, K_, false>;           \
    }
    MIXWS(16, 1)
    MIXWS(16, 2)
    MIXWS(32, 1)
    MIXWS(32, 2)
    MIXWS(64, 1)
#undef MIXWS
    throw ArgError{"internal: no wave-specialised mixed kernel instance"};
}

// Coefficient form of chunk c for the single-wave kernel (its mixed_vals_q):
// 1 Re c1 = 0, 2 Im c1 = 0, else 0.
static int coef_form(const ChunkCoef& c) { return c.c1r == 0.0f ? 1 : (c.c1i == 0.0f ? 2 : 0); }

// The single-wave kernel for KG chunks with coefficients cp (NULL: the
// general form).  Without the adjoint term, one or two chunks take their own
// coefficient forms, four chunks a form they share; else the general form.
static MixedFn mixed_fn(bool adj, int WCSP, int KG, const CoefPack* cp = nullptr, bool tn = false)
{
    if (tn) {
        // the target-normal form: the general coefficient form, or every chunk
        // without Re c1 (the Dirichlet-to-Neumann rows (1/2, -i kappa, 0, 1))
        bool re0 = cp != nullptr;
        for (int q = 0; cp && q < KG; ++q) re0 = re0 && coef_form(cp->c[q]) == 1;
#define MIXTN(W_, K_, F_)                                                                      \
    if (WCSP == W_ && KG == K_)                                                                \
        return re0 ? &bdr_mixed_kernel<W_, K_, false, F_, true> : &bdr_mixed_kernel<W_, K_, false, 0, true>;
        MIXTN(16, 1, 1)
        MIXTN(16, 2, 4)
        MIXTN(16, 4, 40)
        MIXTN(32, 1, 1)
        MIXTN(32, 2, 4)
        MIXTN(64, 1, 1)
#undef MIXTN
        throw ArgError{"internal: no mixed kernel instance"};
    }
    int forms = 0;
    if (cp && !adj) {
        if (KG <= 2) {
            for (int q = KG - 1; q >= 0; --q) forms = 3 * forms + coef_form(cp->c[q]);
        } else {
            const int f = coef_form(cp->c[0]);
            bool same = true;
            for (int q = 1; q < KG; ++q) same = same && coef_form(cp->c[q]) == f;
            forms = same ? f * 40 : 0;   // 40 = 1 + 3 + 9 + 27: form f in every digit
        }
    }
#define MIXF1(W_)                                                     \
    if (WCSP == W_ && KG == 1) {                                      \
        if (adj) return &bdr_mixed_kernel<W_, 1, true>;               \
        if (forms == 1) return &bdr_mixed_kernel<W_, 1, false, 1>;    \
        if (forms == 2) return &bdr_mixed_kernel<W_, 1, false, 2>;    \
        return &bdr_mixed_kernel<W_, 1, false, 0>;                    \
    }
#define MIXF2(W_)                                                     \
    if (WCSP == W_ && KG == 2) {                                      \
        if (adj) return &bdr_mixed_kernel<W_, 2, true>;               \
        switch (forms) {                                              \
            case 1: return &bdr_mixed_kernel<W_, 2, false, 1>;        \
            case 2: return &bdr_mixed_kernel<W_, 2, false, 2>;        \
            case 3: return &bdr_mixed_kernel<W_, 2, false, 3>;        \
            case 4: return &bdr_mixed_kernel<W_, 2, false, 4>;        \
            case 5: return &bdr_mixed_kernel<W_, 2, false, 5>;        \
            case 6: return &bdr_mixed_kernel<W_, 2, false, 6>;        \
            case 7: return &bdr_mixed_kernel<W_, 2, false, 7>;        \
            case 8: return &bdr_mixed_kernel<W_, 2, false, 8>;        \
            default: return &bdr_mixed_kernel<W_, 2, false, 0>;       \
        }                                                             \
    }
    MIXF1(16)
    MIXF2(16)
    MIXF1(32)
    MIXF2(32)
    MIXF1(64)
#undef MIXF1
#undef MIXF2
    if (WCSP == 16 && KG == 4) {
        if (adj) return &bdr_mixed_kernel<16, 4, true>;
        if (forms == 40) return &bdr_mixed_kernel<16, 4, false, 40>;
        if (forms == 80) return &bdr_mixed_kernel<16, 4, false, 80>;
        return &bdr_mixed_kernel<16, 4, false, 0>;
    }
    throw ArgError{"internal: no mixed kernel instance"};
}

static int pick_kg(int WCSP, int remaining, bool mixed)
{
    const int maxkg = std::max(1, kMaxCols / WCSP);
    const int lim = mixed ? 4 : 2;
    int kg = 1;
    while (kg * 2 <= std::min({maxkg, remaining, lim})) kg *= 2;
    return kg;
}

// Number of source segments: enough work-groups to fill 256 CUs several times,
// and at most kMaxSegSources sources per segment, which bounds the length of
// each FP32 accumulation (the segment partials are summed in FP64).
constexpr int kMaxSegSources = 32768;
static int choose_segments(const Handle& h, int tiles, int max_seg, int64_t nsrc, int target = 2048)
{
    const int min_seg = (int)((nsrc + kMaxSegSources - 1) / kMaxSegSources);
    const int want = h.segments > 0 ? h.segments : (target + tiles - 1) / tiles;
    return std::max(1, std::min(std::max(want, min_seg), max_seg));
}

// Work-groups of kernel fn (bs threads each) resident on the current device at
// once: the occupancy per CU times the CU count, cached per device and kernel.
static int resident_blocks(const void* fn, int bs)
{
    static std::mutex mtx;
    static std::map<std::pair<int, const void*>, int> cache;
    int dev = 0;
    BAEMM_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mtx);
    auto it = cache.find({dev, fn});
    if (it != cache.end()) return it->second;
    int per_cu = 0, cus = 0;
    BAEMM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, bs, 0));
    BAEMM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int r = std::max(1, per_cu * cus);
    cache[{dev, fn}] = r;
    return r;
}

// Automatic segment count of a launch of tiles x S work-groups whose segments
// split nst source units (8-source steps or 64-source blocks): among the
// counts from S0 to 2 S0 (within max_seg), the one whose grid fills the last
// round of resident work-groups best (a part-empty last round costs a whole
// round of kernel time); the first count that fills it to 99.5 %, else the
// fullest (the smallest on ties).  Returns the count to split nst by.
static int balance_segments(int tiles, int nst, int S0, int max_seg, int resident)
{
    int best = S0;
    double best_fill = -1.0;
    for (int c = S0; c <= std::min(2 * S0, std::max(S0, max_seg)); ++c) {
        const int steps = (nst + c - 1) / c;
        const int S = (nst + steps - 1) / steps;
        const double rounds = (double)tiles * S / resident;
        const double fill = rounds / std::ceil(rounds);
        if (fill >= 0.995) return c;
        if (fill > best_fill + 1e-9) {
            best_fill = fill;
            best = c;
        }
    }
    return best;
}

// Which MFMA kernel a mixed plan runs: the wave-specialised one for the FP32
// contraction, when selected, and (automatic selection, the default) for
// operators with the adjoint double layer, whose single-wave kernel is built
// for 3 waves per SIMD; else the single-wave kernel.
static bool mixed_uses_ws(const Handle& h, const BdrPlan& pl)
{
    return pl.f32 || h.mixed_kernel == 1 || (h.mixed_kernel == 2 && pl.adj);
}

// Whether the vertex epilogue reads per-triangle values formed
// beforehand (tri_epilogue_kernel): with four or more segments, where re-reading
// every segment's partials for each of a triangle's three vertices costs more
// than one extra pass.
static bool use_tsum(const BdrPlan& pl) { return pl.kernel_on && pl.S >= 4; }

// Runs a plain pair kernel over all chunk groups: part has ntgt_pad rows.
static void run_plain(Handle& h, int op, bool adj, int WCSP, const float4* tx, const float4* tn, int ntgt,
                      int ntgt_pad, const float4* sy, const float4* sn, int nsrc, int nsrc_pad, const float2* b,
                      int ldb, float* part, int64_t seg_stride, int ldp, int S, int seg_len, hipStream_t s)
{
    const Params& p = h.prm;
    Timed tm(h, s);
    for (int q0 = 0; q0 < p.K;) {
        const int kg = pick_kg(WCSP, (int)p.K - q0, false);
        const CoefPack cp = make_pack(p, q0, kg, nullptr);
        PlainFn fn = plain_fn(op, adj, WCSP, kg);
        dim3 grid(ntgt_pad / kTile, S);
        hipLaunchKernelGGL(fn, grid, dim3(kTile), 0, s, tx, tn, ntgt, sy, sn, nsrc, b, ldb, cp, part, seg_stride, ldp,
                           q0, seg_len, nsrc_pad, op == OP_BDR ? h.step_skip : nullptr);
        BAEMM_HIP(hipGetLastError());
        q0 += kg;
    }
    tm.done();
}

// Minimum distance between the targets and the centroids, reduced in the
// 4-byte device word buf (synchronises s).
static float min_distance(Handle& h, const float4* tgt, int ntgt, bool self, unsigned* buf, hipStream_t s)
{
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, s, buf, 1, 0x7f7fffffu);
    hipLaunchKernelGGL(rmin_kernel, dim3((ntgt + 255) / 256), dim3(256), 0, s, tgt, ntgt, h.d_mid, (int)h.nt,
                       self ? 1 : 0, buf);
    BAEMM_HIP(hipGetLastError());
    unsigned out = 0;
    BAEMM_HIP(hipMemcpyAsync(&out, buf, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    BAEMM_HIP(hipStreamSynchronize(s));
    float r2;
    memcpy(&r2, &out, sizeof(float));
    return std::sqrt(std::max(r2, 1e-30f));
}

static float compute_rmin(Handle& h, WsSource& ws, hipStream_t s)
{
    if (h.rmin >= 0.0f) return h.rmin;
    h.rmin = min_distance(h, h.d_mid, (int)h.nt, true, static_cast<unsigned*>(ws.take(h.ws_misc, 256)), s);
    return h.rmin;
}

float compute_rmin(Handle& h, hipStream_t s)
{
    WsSource ws;
    return compute_rmin(h, ws, s);
}

template <int BT, int CT>
static void launch_bdr_epilogue(EpiArgs& ea, hipStream_t s)
{
    const int lpv = epi_lanes(ea.W);
    const int vpb = 256 / lpv;
    hipLaunchKernelGGL((bdr_epilogue_kernel<BT, CT>), dim3((ea.nv + vpb - 1) / vpb), dim3(256), 0, s, ea, lpv);
}

static void bdr_epilogue(Handle& h, EpiArgs& ea, int Btype, int Ctype, hipStream_t s)
{
    if (Btype == kC128 && Ctype == kC128) launch_bdr_epilogue<kC128, kC128>(ea, s);
    else if (Btype == kC128 && Ctype == kC64) launch_bdr_epilogue<kC128, kC64>(ea, s);
    else if (Btype == kC64 && Ctype == kC128) launch_bdr_epilogue<kC64, kC128>(ea, s);
    else launch_bdr_epilogue<kC64, kC64>(ea, s);
    BAEMM_HIP(hipGetLastError());
}

static EpiArgs base_epi(Handle& h, const ApplyArgs& a, int WCSP)
{
    const Params& p = h.prm;
    if (p.K > 64) throw ArgError{"at most 64 wave-number chunks are supported"};
    EpiArgs ea{};
    ea.nv = (int)h.nv;
    ea.W = (int)p.W;
    ea.wcs = (int)p.wcs;
    ea.WCSP = WCSP;
    ea.K = (int)p.K;
    ea.ar = a.alpha.real();
    ea.ai = a.alpha.imag();
    ea.br = a.beta.real();
    ea.bi = a.beta.imag();
    ea.vt_ptr = h.d_vt_ptr;
    ea.vt_idx = h.d_vt_idx;
    ea.area = h.d_area;
    ea.sdiag = h.d_sdiag;
    ea.M_ptr = h.d_M_ptr;
    ea.M_idx = h.d_M_idx;
    ea.M_val = h.d_M_val;
    ea.B = a.B;
    ea.ldB = a.ldB;
    ea.C = a.C;
    ea.ldC = a.ldC;
    for (int q = 0; q < p.K; ++q) {
        ea.rotr[q] = 1.0f;
        ea.roti[q] = 0.0f;
        ea.c0r[q] = p.c[4 * q].real();
        ea.c0i[q] = p.c[4 * q].imag();
        ea.c1r[q] = p.c[4 * q + 1].real();
        ea.c1i[q] = p.c[4 * q + 1].imag();
        ea.kap[q] = p.kappa[q];
    }
    return ea;
}

static void launch_av(Handle& h, const void* B, int64_t ldB, int Btype, int WCSP, float2* b, int ldb,
                      hipStream_t s)
{
    const int64_t total = h.nt_pad * (int64_t)ldb;
    const dim3 grid((unsigned)((total + 255) / 256));
    if (Btype == kC128)
        hipLaunchKernelGGL(av_gather_kernel<kC128>, grid, dim3(256), 0, s, h.d_tris, h.d_area, (int)h.nt, total, B,
                           ldB, (int)h.prm.wcs, WCSP, ldb, b);
    else
        hipLaunchKernelGGL(av_gather_kernel<kC64>, grid, dim3(256), 0, s, h.d_tris, h.d_area, (int)h.nt, total, B,
                           ldB, (int)h.prm.wcs, WCSP, ldb, b);
    BAEMM_HIP(hipGetLastError());
}

// Operands of the MFMA pair kernels for the staged sources pl.b: the B operand
// (hi/lo split with per-column scales, or FP32), the chunk factors pl.rot and
// scales pl.sA folded into the coefficients, the source normals pl.nrm the
// kernel reads, and pl.unscale (per column, null for FP32 MFMA).  rmin: lower
// bound of the target-centroid distances (mixed precision only).
static void stage_mfma(Handle& h, BdrPlan& pl, float rmin, hipStream_t s, const void* B, int64_t ldB, int Btype)
{
    const Params& p = h.prm;
    if (pl.f32) {
        if (pl.defer) throw ArgError{"internal: deferred staging of the FP32 MFMA operand"};
        launch_av(h, B, ldB, Btype, pl.WCSP, pl.b, pl.ldb, s);
        // FP32 MFMA: no FP16 range to fit, no scales; the non-adjoint form still
        // divides each chunk by c2 (the epilogue multiplies it back)
        pl.sA.assign(p.K, 1.0f);
        pl.rot.assign(p.K, cfloat(1.0f, 0.0f));
        if (!pl.adj)
            for (int q = 0; q < p.K; ++q) pl.rot[q] = cfloat(p.c[4 * q + (pl.tn ? 3 : 2)]);
        const int NCB = 2 * pl.WCSP / 32;
        const int64_t nbop = (int64_t)p.K * pl.nst * NCB * 64;
        uint4* bop = static_cast<uint4*>(pl.ws.take(h.ws_bop, sizeof(uint4) * nbop * 2));
        hipLaunchKernelGGL(stage_f32_kernel, dim3((unsigned)((nbop + 255) / 256)), dim3(256), 0, s, pl.b, pl.ldb,
                           pl.nst, pl.WCSP, (int)p.K, bop);
        BAEMM_HIP(hipGetLastError());
        pl.bop = bop;
        pl.unscale = nullptr;
        return;
    }
    // scales: A per chunk from a bound at the minimum target-centroid
    // distance rmin, b per column from its maximum.
    pl.sA.assign(p.K, 1.0f);
    pl.rot.assign(p.K, cfloat(1.0f, 0.0f));
    ChunkScales csc{};
    float nrm_sig = 0.0f;      // the common scale of the source normals (non-adjoint form)
    float4* nrm_out = nullptr;  // the scaled normals (non-adjoint form)
    if (!pl.adj) {
        // Every chunk has c2 != 0: the kernel values are taken as A' = s A / c2
        // with one scale s for all chunks, carried by the source normals
        // (Q' = c1/c2 s / r - s (z.n)/r^3 + i kappa s (z.n)/r^2); the epilogue
        // multiplies c2 / s back.
        // (with pl.tn: c3 in place of c2, the target normals in place of the
        // source normals)
        const int cn = pl.tn ? 3 : 2;
        double amax = 0.0;
        for (int q = 0; q < p.K; ++q) {
            const std::complex<double> c2(p.c[4 * q + cn]);
            const double rho = std::abs(std::complex<double>(p.c[4 * q + 1]) / c2);
            const double kr = (double)std::fabs(p.kappa[q]) * rmin;
            amax = std::max(amax, rho / rmin + (1.0 + kr) / ((double)rmin * rmin));
            pl.rot[q] = cfloat(p.c[4 * q + cn]);
        }
        int e = 0;
        std::frexp(std::max(amax, 1e-30), &e);
        const float sig = (float)std::ldexp(1.0, 14 - e);
        for (int q = 0; q < p.K; ++q) pl.sA[q] = csc.sA[q] = sig;
        float4* nrm = static_cast<float4*>(pl.ws.take(h.ws_nrm, sizeof(float4) * h.nt_pad));
        if (pl.defer) {
            pl.sd.nrm_in = h.d_nrm;
            pl.sd.nrm_out = nrm;
            pl.sd.nrm_s = sig;
        }
        nrm_sig = sig;
        nrm_out = nrm;
        pl.nrm = nrm;
    } else {
        for (int q = 0; q < p.K; ++q) {
            const double c1 = std::abs(std::complex<double>(p.c[4 * q + 1]));
            const double c23 =
                std::abs(std::complex<double>(p.c[4 * q + 2])) + std::abs(std::complex<double>(p.c[4 * q + 3]));
            const double kr = (double)std::fabs(p.kappa[q]) * rmin;
            const double amax = c1 / rmin + c23 * (1.0 + kr) / ((double)rmin * rmin);
            int e = 0;
            std::frexp(std::max(amax, 1e-30), &e);
            pl.sA[q] = (float)std::ldexp(1.0, 14 - e);
            csc.sA[q] = pl.sA[q];
        }
    }
    const int ncols = pl.ldb;
    unsigned* colmax = static_cast<unsigned*>(pl.ws.take(h.ws_scale, sizeof(float) * 3 * ncols));
    float* scale = reinterpret_cast<float*>(colmax + ncols);
    float* unscale = scale + ncols;
    const int NCB = 2 * pl.WCSP / 32;
    const int64_t nbop = (int64_t)p.K * pl.nst * NCB * 64;
    uint4* bop = static_cast<uint4*>(pl.ws.take(h.ws_bop, sizeof(uint4) * nbop * 2));
    pl.bop = bop;
    pl.unscale = unscale;
    if (pl.defer) {
        StageDesc& d = pl.sd;
        d.split = 1;
        d.K = (int)p.K;
        d.nst = pl.nst;
        d.colmax = colmax;
        d.scale = scale;
        d.unscale = unscale;
        d.bop = bop;
        for (int q = 0; q < p.K; ++q) d.sA[q] = csc.sA[q];
        return;
    }
    // the scaled source normals (non-adjoint form) with the column maxima
    // cleared; the Av gather and the column maxima (one launch for small
    // meshes, where the launches are the cost; two above, where the blocks'
    // atomics on the few maxima of the one launch cost more); the split operand
    // with the column scales and their inverses
    const int nn = pl.adj ? 0 : (int)h.nt_pad;
    hipLaunchKernelGGL(scale_nrm_kernel, dim3((unsigned)((std::max(nn, ncols) + 255) / 256)), dim3(256), 0, s,
                       h.d_nrm, nn, nrm_sig, nrm_out, colmax, ncols);
    const int64_t total = h.nt_pad * (int64_t)pl.ldb;
    if ((total + 255) / 256 <= kGatherMaxBlocks) {
        // a multiple of ldb / gcd(ldb, 256) blocks: the grid's thread count is
        // a multiple of ldb
        const int64_t unit = pl.ldb / std::gcd(pl.ldb, 256);
        const dim3 grid((unsigned)(((total + 255) / 256 + unit - 1) / unit * unit));
        const size_t lds = sizeof(unsigned) * (size_t)pl.ldb;
        if (Btype == kC128)
            hipLaunchKernelGGL(av_gather_max_kernel<kC128>, grid, dim3(256), lds, s, h.d_tris, h.d_area, (int)h.nt,
                               total, B, ldB, (int)p.wcs, pl.WCSP, pl.ldb, pl.b, colmax);
        else
            hipLaunchKernelGGL(av_gather_max_kernel<kC64>, grid, dim3(256), lds, s, h.d_tris, h.d_area, (int)h.nt,
                               total, B, ldB, (int)p.wcs, pl.WCSP, pl.ldb, pl.b, colmax);
    } else {
        launch_av(h, B, ldB, Btype, pl.WCSP, pl.b, pl.ldb, s);
        hipLaunchKernelGGL(colmax_kernel, dim3(256), dim3(256), 0, s, pl.b, (int)h.nt, pl.ldb, colmax);
    }
    hipLaunchKernelGGL(split_kernel, dim3((unsigned)((nbop + 255) / 256)), dim3(256), 0, s, pl.b, pl.ldb, pl.nst,
                       pl.WCSP, (int)p.K, colmax, ncols, csc, unscale, bop);
    BAEMM_HIP(hipGetLastError());
}

void bdr_setup(Handle& h, BdrPlan& pl)
{
    const Params& p = h.prm;
    if (!p.loaded) throw ArgError{"operators not loaded"};
    pl.planned = true;
    pl.kernel_on = p.kernel_on();
    pl.mass_on = p.mass_on();
    pl.mixed = (h.precision != BAEMM_PRECISION_FP32) && pl.kernel_on;
    pl.f32 = pl.mixed && h.precision == BAEMM_PRECISION_FP32_MFMA;
    pl.WCSP = pl.mixed ? wcsp_mixed((int)p.wcs) : wcsp_plain((int)p.wcs);
    pl.ldb = (int)p.K * pl.WCSP;
    pl.ldp = 2 * pl.ldb;
    pl.adj = p.any(3);
    // The MFMA kernels' non-adjoint form divides each chunk by its c2; chunks
    // without a double layer take the general (adjoint) form, unless every
    // chunk has the adjoint double layer and no double layer: the MFMA
    // kernels then take the non-adjoint form divided by c3, on the target
    // normals (the adjoint double layer has the double layer's values with the
    // target normal in place of the source normal).
    pl.tn = false;
    if (pl.mixed) {
        bool tn = true;
        for (int q = 0; q < p.K; ++q) {
            if (p.c[4 * q + 2] == cfloat(0.0f, 0.0f)) pl.adj = true;
            tn = tn && p.c[4 * q + 2] == cfloat(0.0f, 0.0f) && p.c[4 * q + 3] != cfloat(0.0f, 0.0f);
        }
        if (tn) {
            pl.tn = true;
            pl.adj = false;
        }
    }
    if (!pl.kernel_on) return;
    if (!pl.mixed) {
        const int tiles = (int)(h.nt_pad / kTile);
        const int max_seg = (int)(h.nt_pad / kSrcBlock);
        int S0 = choose_segments(h, tiles, max_seg, h.nt);
        const int res = resident_blocks((const void*)plain_fn(OP_BDR, pl.adj, pl.WCSP, pick_kg(pl.WCSP, (int)p.K, false)),
                                        kTile);
        if (h.segments == 0) S0 = balance_segments(tiles, max_seg, S0, max_seg, res);
        pl.seg_len = (int)(((h.nt_pad + S0 - 1) / S0 + kSrcBlock - 1) / kSrcBlock) * kSrcBlock;
        pl.S = (int)((h.nt_pad + pl.seg_len - 1) / pl.seg_len);
        pl.rounds = (double)tiles * pl.S / res;
        return;
    }
    pl.nst = (int)(h.nt_pad / 8);
    const int tiles = (int)(h.nt_pad / kMixTile);
    int S0 = choose_segments(h, tiles, pl.nst / 8, h.nt);
    int res;
    if (mixed_uses_ws(h, pl)) {
        const int kg0 = std::min(pick_kg(pl.WCSP, (int)p.K, true), 2);
        res = resident_blocks((const void*)mixed_ws_fn(pl.adj, pl.WCSP, kg0, pl.f32, pl.tn), 512);
    } else {
        res = resident_blocks((const void*)mixed_fn(pl.adj, pl.WCSP, pick_kg(pl.WCSP, (int)p.K, true), nullptr, pl.tn),
                              256);
        // The single-wave kernel runs 2-3 % faster with about four times the
        // work-groups (finer last rounds): up to 16 segments where the rule
        // above gives fewer, unless the plan must keep its partials small.
        if (h.segments == 0 && !pl.lean && S0 < 16)
            S0 = std::max(S0, std::min({16, (8192 + tiles - 1) / tiles, pl.nst / 8}));
    }
    if (h.segments == 0 && !pl.lean) S0 = balance_segments(tiles, pl.nst, S0, pl.nst / 8, res);
    pl.steps = (pl.nst + S0 - 1) / S0;
    pl.S = (pl.nst + pl.steps - 1) / pl.steps;
    pl.rounds = (double)tiles * pl.S / res;
}

// Per triangle t and column k of chunk q: the sum over the segments s (in
// ascending order, FP64) of part_s[t][q][k], real and imaginary parts, turned
// into the triangle's value (epi_tri_value) at tsum[t * ldb + q * WCSP + k % wcs]:
// the values the vertex epilogue forms, once per triangle instead of once per
// incident vertex.
__global__ __launch_bounds__(256) void tri_epilogue_kernel(EpiArgs a, int nt, double2* __restrict__ tsum)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)nt * a.W) return;
    const int t = (int)(idx / a.W), k = (int)(idx % a.W);
    const int q = k / a.wcs, kk = k % a.wcs;
    const int pc = q * a.WCSP + kk;
    const float* pp = a.part + (int64_t)t * a.ldp + q * 2 * a.WCSP + kk;
    double pr = 0.0, pi = 0.0;
    // EPI_SEG segments per batch, every load issued before the first is added
    for (int s0 = 0; s0 < a.S; s0 += EPI_SEG) {
        float xr[EPI_SEG], xi[EPI_SEG];
#pragma unroll
        for (int j = 0; j < EPI_SEG; ++j) {
            const int64_t o = (int64_t)min(s0 + j, a.S - 1) * a.seg_stride;
            xr[j] = pp[o];
            xi[j] = pp[o + a.WCSP];
        }
#pragma unroll
        for (int j = 0; j < EPI_SEG; ++j)
            if (s0 + j < a.S) {
                pr += (double)xr[j];
                pi += (double)xi[j];
            }
    }
    double qr, qi;
    epi_tri_value(a, t, q, pc, pr, pi, qr, qi);
    tsum[(int64_t)t * a.ldb + pc] = make_double2(qr, qi);
}

size_t bdr_workspace_bytes(const Handle& h, const BdrPlan& pl, bool need_rmin)
{
    if (!pl.kernel_on) return 0;
    const auto al = [](size_t n) { return (n + 255) & ~(size_t)255; };
    const Params& p = h.prm;
    size_t n = al(sizeof(float2) * h.nt_pad * pl.ldb);                       // staged sources b
    if (pl.mixed) {
        const int NCB = 2 * pl.WCSP / 32;
        n += al(sizeof(uint4) * 2 * (size_t)p.K * pl.nst * NCB * 64);       // MFMA B operand
        if (!pl.f32) {
            if (need_rmin) n += al(256);                                    // minimum distance
            if (!pl.adj) n += al(sizeof(float4) * h.nt_pad);                // scaled source normals
            n += al(sizeof(float) * 3 * pl.ldb);                             // column maxima and scales
        }
    }
    n += al(sizeof(float) * (size_t)pl.S * h.nt_pad * pl.ldp);              // segment partials
    if (use_tsum(pl)) n += al(sizeof(double2) * h.nt_pad * pl.ldb);         // their sums per triangle
    return n;
}

void bdr_begin(Handle& h, const ApplyArgs& a, BdrPlan& pl, hipStream_t s)
{
    const Params& p = h.prm;
    if (!pl.planned) bdr_setup(h, pl);
    if (!pl.kernel_on && !pl.mass_on) return;   // the reference leaves C_out untouched
    pl.nrm = h.d_nrm;
    EpiArgs& ea = pl.ea;
    ea = base_epi(h, a, pl.WCSP);
    ea.mass_on = pl.mass_on;
    ea.kernel_on = pl.kernel_on;
    ea.diag_on = pl.kernel_on && h.use_diagonal && p.any(1);
    if (!pl.kernel_on) return;
    pl.b = static_cast<float2*>(pl.ws.take(h.ws_b32, sizeof(float2) * h.nt_pad * pl.ldb));
    if (pl.defer) {
        StageDesc& d = pl.sd;
        d.tris = h.d_tris;
        d.area = h.d_area;
        d.nt = (int)h.nt;
        d.nt_pad = (int)h.nt_pad;
        d.B = a.B;
        d.ldB = a.ldB;
        d.wcs = (int)p.wcs;
        d.WCSP = pl.WCSP;
        d.ldb = pl.ldb;
        d.b = pl.b;
    } else if (!pl.mixed) {
        launch_av(h, a.B, a.ldB, a.Btype, pl.WCSP, pl.b, pl.ldb, s);
    }
    ea.b = pl.b;
    ea.ldp = pl.ldp;
    ea.ldb = pl.ldb;
    ea.unscale = nullptr;
    if (!pl.mixed) return;
    // (not deferred: the Av gather too)
    stage_mfma(h, pl, pl.f32 ? 0.0f : compute_rmin(h, pl.ws, s), s, a.B, a.ldB, a.Btype);
    ea.unscale = pl.unscale;
    for (int q = 0; q < p.K; ++q) {
        ea.rotr[q] = pl.rot[q].real();
        ea.roti[q] = pl.rot[q].imag();
    }
}

float* bdr_reserve_part(Handle& h, BdrPlan& pl, int S)
{
    const int64_t seg_stride = h.nt_pad * (int64_t)pl.ldp;
    float* part = static_cast<float*>(pl.ws.take(h.ws_part, sizeof(float) * seg_stride * S));
    pl.ea.S = S;
    pl.ea.seg_stride = seg_stride;
    pl.ea.part = part;
    return part;
}

void bdr_pairs(Handle& h, BdrPlan& pl, hipStream_t s)
{
    if (!pl.kernel_on) return;
    const Params& p = h.prm;
    float* part = bdr_reserve_part(h, pl, pl.S);
    if (!pl.mixed) {
        run_plain(h, OP_BDR, pl.adj, pl.WCSP, h.d_mid, h.d_nrm, (int)h.nt, (int)h.nt_pad, h.d_mid, h.d_nrm,
                  (int)h.nt, (int)h.nt_pad, pl.b, pl.ldb, part, pl.ea.seg_stride, pl.ldp, pl.S, pl.seg_len, s);
        return;
    }
    const int nst = pl.nst;
    const int tiles = (int)(h.nt_pad / kMixTile);
    // the FP32 MFMA contraction exists in the wave-specialised kernel only
    const bool ws = mixed_uses_ws(h, pl);
    Timed tm(h, s);
    for (int q0 = 0; q0 < p.K;) {
        int kg = pick_kg(pl.WCSP, (int)p.K - q0, true);
        if (ws) kg = std::min(kg, 2);
        const CoefPack cp = make_pack(p, q0, kg, pl.sA.data(), pl.rot.data());
        if (ws)
            hipLaunchKernelGGL(mixed_ws_fn(pl.adj, pl.WCSP, kg, pl.f32, pl.tn), dim3(tiles, pl.S), dim3(512), 0, s, h.d_mid,
                               (int)h.nt_pad, 1, h.d_mid, pl.nrm, (int)h.nt, pl.bop, nst, cp, part, pl.ea.seg_stride,
                               pl.ldp, q0, pl.steps, h.step_skip);
        else
            hipLaunchKernelGGL(mixed_fn(pl.adj, pl.WCSP, kg, &cp, pl.tn), dim3(tiles, pl.S), dim3(256), 0, s, h.d_mid,
                               (int)h.nt_pad, 1, h.d_mid, pl.nrm, (int)h.nt, pl.bop, nst, cp, part, pl.ea.seg_stride,
                               pl.ldp, q0, pl.steps, h.step_skip);
        BAEMM_HIP(hipGetLastError());
        q0 += kg;
    }
    tm.done();
}

void bdr_end(Handle& h, const ApplyArgs& a, BdrPlan& pl, hipStream_t s)
{
    if (!pl.kernel_on && !pl.mass_on) return;
    EpiArgs& ea = pl.ea;
    ea.tsum = nullptr;
    if (use_tsum(pl)) {
        // each triangle's value once, then the vertex gather reads them (three
        // vertices per triangle) instead of every segment's partials
        const int nt = (int)h.nt;
        double2* ts = static_cast<double2*>(pl.ws.take(h.ws_tsum, sizeof(double2) * h.nt_pad * pl.ldb));
        const int64_t total = (int64_t)nt * ea.W;
        hipLaunchKernelGGL(tri_epilogue_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, ea, nt, ts);
        BAEMM_HIP(hipGetLastError());
        ea.tsum = ts;
    }
    bdr_epilogue(h, ea, a.Btype, a.Ctype, s);
}

void apply_boundary_dev(Handle& h, const ApplyArgs& a, hipStream_t s, char* scratch, size_t scratch_bytes)
{
    BdrPlan pl;
    bdr_setup(h, pl);
    if (scratch && bdr_workspace_bytes(h, pl, h.rmin < 0.0f) > scratch_bytes) {
        pl = BdrPlan{};   // fewer segments, smaller partials
        pl.lean = true;
        bdr_setup(h, pl);
    }
    if (scratch) {
        const size_t need = bdr_workspace_bytes(h, pl, h.rmin < 0.0f);
        if (need > scratch_bytes)
            throw ArgError{"the apply needs " + std::to_string(need) + " bytes of workspace; the fixed scratch "
                           "array holds " + std::to_string(scratch_bytes)};
        pl.ws = WsSource{scratch, scratch_bytes, 0};
    }
    bdr_begin(h, a, pl, s);
    bdr_pairs(h, pl, s);
    bdr_end(h, a, pl, s);
    // the fixed scratch array is shared by every handle of the device: its
    // contents are no P0 buffer of this handle
    if (scratch) p0_invalidate(h);
    else if (pl.kernel_on) p0_record(h, pl.b, pl.ldb, pl.WCSP, &pl.ea);
}

static void rows_epilogue(int nrows, const Params& p, int WCSP, int S, int64_t seg_stride, int ldp, const float* part,
                          const ApplyArgs& a, hipStream_t s, const float* unscale = nullptr,
                          const std::vector<cfloat>* rotv = nullptr)
{
    ChunkRot rot;
    for (int q = 0; q < 64; ++q) {
        const cfloat f = (rotv && q < (int)rotv->size()) ? (*rotv)[q] : cfloat(1.0f, 0.0f);
        rot.r[q] = f.real();
        rot.i[q] = f.imag();
    }
    const int64_t total = (int64_t)nrows * p.W;
    const dim3 grid((unsigned)((total + 255) / 256));
    if (a.Ctype == kC128)
        hipLaunchKernelGGL(rows_epilogue_kernel<kC128>, grid, dim3(256), 0, s, nrows, (int)p.W, (int)p.wcs, WCSP, S,
                           seg_stride, ldp, part, a.alpha.real(), a.alpha.imag(), a.beta.real(), a.beta.imag(), a.C,
                           a.ldC, unscale, rot);
    else
        hipLaunchKernelGGL(rows_epilogue_kernel<kC64>, grid, dim3(256), 0, s, nrows, (int)p.W, (int)p.wcs, WCSP, S,
                           seg_stride, ldp, part, a.alpha.real(), a.alpha.imag(), a.beta.real(), a.beta.imag(), a.C,
                           a.ldC, unscale, rot);
    BAEMM_HIP(hipGetLastError());
}

// Far field / potential share the structure "targets x triangles", b = Av B.
static void apply_rows_op(Handle& h, int op, const ApplyArgs& a, const float4* tx, int ntgt, int ntgt_pad,
                          hipStream_t s)
{
    const Params& p = h.prm;
    if (!p.loaded) throw ArgError{"parameters not loaded"};
    if (!(p.any(1) || p.any(2))) return;   // ApplyFarFieldOperators.hpp:83
    if (p.K > 64) throw ArgError{"at most 64 wave-number chunks are supported"};
    const int WCSP = wcsp_plain((int)p.wcs);
    const int ldb = (int)p.K * WCSP, ldp = 2 * ldb;
    h.ws_b32.reserve(sizeof(float2) * h.nt_pad * ldb);
    float2* b = h.ws_b32.as<float2>();
    launch_av(h, a.B, a.ldB, a.Btype, WCSP, b, ldb, s);
    const int tiles = ntgt_pad / kTile;
    const int nsb = (int)(h.nt_pad / kSrcBlock);
    int S0 = choose_segments(h, tiles, nsb, h.nt);
    if (h.segments == 0)
        S0 = balance_segments(tiles, nsb, S0, nsb,
                              resident_blocks((const void*)plain_fn(op, false, WCSP, pick_kg(WCSP, (int)p.K, false)),
                                              kTile));
    const int seg_len = (int)(((h.nt_pad + S0 - 1) / S0 + kSrcBlock - 1) / kSrcBlock) * kSrcBlock;
    const int S = (int)((h.nt_pad + seg_len - 1) / seg_len);
    const int64_t seg_stride = (int64_t)ntgt_pad * ldp;
    h.ws_part.reserve(sizeof(float) * seg_stride * S);
    run_plain(h, op, false, WCSP, tx, nullptr, ntgt, ntgt_pad, h.d_mid, h.d_nrm, (int)h.nt, (int)h.nt_pad, b, ldb,
              h.ws_part.as<float>(), seg_stride, ldp, S, seg_len, s);
    rows_epilogue(ntgt, p, WCSP, S, seg_stride, ldp, h.ws_part.as<float>(), a, s);
    p0_record(h, b, ldb, WCSP, nullptr);
}

void apply_far_dev(Handle& h, const ApplyArgs& a, hipStream_t s)
{
    if (h.m < 1) throw ArgError{"no measurement directions"};
    apply_rows_op(h, OP_FAR, a, h.d_meas, (int)h.m, (int)h.m_pad, s);
}

// Potential at P points with the MFMA kernels (precision MIXED or FP32_MFMA):
// the wave-specialised boundary kernel with the points as targets and no pair
// left out (NearFieldOperatorKernel.cl uses the same single/double-layer
// kernel values), then the row epilogue with the scales and chunk factors.
static void apply_near_mfma(Handle& h, const ApplyArgs& a, const float4* d_pts, int P, hipStream_t s)
{
    const Params& p = h.prm;
    BdrPlan pl;
    pl.kernel_on = pl.mixed = true;
    pl.f32 = h.precision == BAEMM_PRECISION_FP32_MFMA;
    pl.WCSP = wcsp_mixed((int)p.wcs);
    pl.ldb = (int)p.K * pl.WCSP;
    pl.ldp = 2 * pl.ldb;
    // no adjoint double layer for points: the general form only for chunks without c2
    pl.adj = false;
    for (int q = 0; q < p.K; ++q)
        if (p.c[4 * q + 2] == cfloat(0.0f, 0.0f)) pl.adj = true;
    pl.nrm = h.d_nrm;
    pl.nst = (int)(h.nt_pad / 8);
    pl.b = static_cast<float2*>(pl.ws.take(h.ws_b32, sizeof(float2) * h.nt_pad * pl.ldb));
    stage_mfma(h, pl,
               pl.f32 ? 0.0f
                      : min_distance(h, d_pts, P, false, static_cast<unsigned*>(pl.ws.take(h.ws_misc, 256)), s),
               s, a.B, a.ldB, a.Btype);
    const int P_pad = (P + kMixTile - 1) / kMixTile * kMixTile;
    const int nst = pl.nst;
    const int tiles = P_pad / kMixTile;
    // the kernel and its segments as for the boundary operators (bdr_setup)
    const bool ws = mixed_uses_ws(h, pl);
    int S0 = choose_segments(h, tiles, nst / 8, h.nt);
    if (h.segments == 0) {
        int res;
        if (ws) {
            const int kg0 = std::min(pick_kg(pl.WCSP, (int)p.K, true), 2);
            res = resident_blocks((const void*)mixed_ws_fn(pl.adj, pl.WCSP, kg0, pl.f32), 512);
        } else {
            res = resident_blocks((const void*)mixed_fn(pl.adj, pl.WCSP, pick_kg(pl.WCSP, (int)p.K, true)), 256);
            if (S0 < 16) S0 = std::max(S0, std::min({16, (8192 + tiles - 1) / tiles, nst / 8}));
        }
        S0 = balance_segments(tiles, nst, S0, nst / 8, res);
    }
    const int steps = (nst + S0 - 1) / S0;
    const int S = (nst + steps - 1) / steps;
    const int64_t seg_stride = (int64_t)P_pad * pl.ldp;
    h.ws_part.reserve(sizeof(float) * seg_stride * S);
    float* part = h.ws_part.as<float>();
    // every pair of the source range is written: no clearing needed
    Timed tm(h, s);
    for (int q0 = 0; q0 < p.K;) {
        int kg = pick_kg(pl.WCSP, (int)p.K - q0, true);
        if (ws) kg = std::min(kg, 2);
        const CoefPack cp = make_pack(p, q0, kg, pl.sA.data(), pl.rot.data());
        if (ws)
            hipLaunchKernelGGL(mixed_ws_fn(pl.adj, pl.WCSP, kg, pl.f32), dim3(tiles, S), dim3(512), 0, s, d_pts, P, 0,
                               h.d_mid, pl.nrm, (int)h.nt, pl.bop, nst, cp, part, seg_stride, pl.ldp, q0, steps,
                               nullptr);
        else
            hipLaunchKernelGGL(mixed_fn(pl.adj, pl.WCSP, kg, &cp), dim3(tiles, S), dim3(256), 0, s, d_pts, P, 0, h.d_mid,
                               pl.nrm, (int)h.nt, pl.bop, nst, cp, part, seg_stride, pl.ldp, q0, steps, nullptr);
        BAEMM_HIP(hipGetLastError());
        q0 += kg;
    }
    tm.done();
    rows_epilogue(P, p, pl.WCSP, S, seg_stride, pl.ldp, part, a, s, pl.unscale, &pl.rot);
    p0_record(h, pl.b, pl.ldb, pl.WCSP, nullptr);
}

void apply_near_dev(Handle& h, const ApplyArgs& a, const float4* d_pts, int64_t P, hipStream_t s)
{
    const Params& p = h.prm;
    if (h.precision != BAEMM_PRECISION_FP32 && p.loaded && (p.any(1) || p.any(2)) && p.K <= 64 && P > 0) {
        apply_near_mfma(h, a, d_pts, (int)P, s);
        return;
    }
    const int P_pad = (int)(((P + kTile - 1) / kTile) * kTile);
    apply_rows_op(h, OP_NEAR, a, d_pts, (int)P, P_pad, s);
}

void herglotz_dev(Handle& h, const ApplyArgs& a, hipStream_t s)
{
    p0_invalidate(h);   // rewrites the workspaces; no P0 buffer of the reference kind
    const Params& p = h.prm;
    if (!p.loaded) throw ArgError{"parameters not loaded"};
    if (!(p.any(1) || p.any(2))) return;   // CreateHerglotzWave.hpp:200
    if (h.m < 1) throw ArgError{"no measurement directions"};
    if (p.K > 64) throw ArgError{"at most 64 wave-number chunks are supported"};
    const int WCSP = wcsp_plain((int)p.wcs);
    const int ldb = (int)p.K * WCSP, ldp = 2 * ldb;
    // g (m x W) -> conj(g) staged with m_pad rows (HerglotzWaveKernel.cl:143-144)
    h.ws_g.reserve(sizeof(float2) * h.m_pad * ldb);
    float2* gb = h.ws_g.as<float2>();
    const int64_t total = h.m_pad * (int64_t)ldb;
    const dim3 grid((unsigned)((total + 255) / 256));
    if (a.Btype == kC128)
        hipLaunchKernelGGL(stage_rows_kernel<kC128>, grid, dim3(256), 0, s, (int)h.m, total, a.B, a.ldB, (int)p.wcs,
                           WCSP, ldb, 1, gb);
    else
        hipLaunchKernelGGL(stage_rows_kernel<kC64>, grid, dim3(256), 0, s, (int)h.m, total, a.B, a.ldB, (int)p.wcs,
                           WCSP, ldb, 1, gb);
    BAEMM_HIP(hipGetLastError());
    const int tiles = (int)(h.nt_pad / kTile);
    const int nsb = (int)(h.m_pad / kSrcBlock);
    int S0 = choose_segments(h, tiles, nsb, h.m);
    if (h.segments == 0)
        S0 = balance_segments(tiles, nsb, S0, nsb,
                              resident_blocks((const void*)plain_fn(OP_HERG, false, WCSP,
                                                                    pick_kg(WCSP, (int)p.K, false)),
                                              kTile));
    const int seg_len = (int)(((h.m_pad + S0 - 1) / S0 + kSrcBlock - 1) / kSrcBlock) * kSrcBlock;
    const int S = (int)((h.m_pad + seg_len - 1) / seg_len);
    const int64_t seg_stride = h.nt_pad * (int64_t)ldp;
    h.ws_part.reserve(sizeof(float) * seg_stride * S);
    run_plain(h, OP_HERG, false, WCSP, h.d_mid, h.d_nrm, (int)h.nt, (int)h.nt_pad, h.d_meas, nullptr, (int)h.m,
              (int)h.m_pad, gb, ldb, h.ws_part.as<float>(), seg_stride, ldp, S, seg_len, s);
    EpiArgs ea = base_epi(h, a, WCSP);
    ea.kernel_on = 1;
    ea.diag_on = 0;
    ea.mass_on = 0;
    ea.S = S;
    ea.seg_stride = seg_stride;
    ea.ldp = ldp;
    ea.ldb = ldb;
    ea.part = h.ws_part.as<float>();
    ea.b = gb;
    ea.unscale = nullptr;
    bdr_epilogue(h, ea, kC64, a.Ctype, s);
}

void incident_dev(Handle& h, const ApplyArgs& a, const double* dirs, int64_t inc_count, int wave_type,
                  hipStream_t s)
{
    p0_invalidate(h);   // rewrites the workspaces; no P0 buffer of the reference kind
    const Params& p = h.prm;
    if (!p.loaded) throw ArgError{"parameters not loaded"};
    if (!(p.any(1) || p.any(2))) return;   // CreateIncidentWave.hpp:95
    if (!dirs) throw ArgError{"incident directions must not be NULL"};
    if (inc_count < p.wcs) throw ArgError{"inc_count must be >= wave_chunk_size"};
    if (p.wcs > 64 || p.K > 64) throw ArgError{"at most 64 directions and 64 chunks are supported"};
    if (wave_type != BAEMM_WAVE_PLANE && wave_type != BAEMM_WAVE_RADIAL) throw ArgError{"unknown wave type"};
    const int WCSP = (int)p.wcs;
    const int ldp = 2 * (int)p.K * WCSP;
    h.ws_part.reserve(sizeof(float) * h.nt_pad * ldp);
    IncArgs ia{};
    ia.nt = (int)h.nt;
    ia.W = (int)p.W;
    ia.wcs = (int)p.wcs;
    ia.WCSP = WCSP;
    ia.ldp = ldp;
    ia.radial = (wave_type == BAEMM_WAVE_RADIAL);
    ia.mid = h.d_mid;
    ia.nrm = h.d_nrm;
    ia.part = h.ws_part.as<float>();
    for (int j = 0; j < p.wcs; ++j)
        for (int d = 0; d < 3; ++d) ia.dirs[j][d] = (float)dirs[3 * j + d];
    for (int q = 0; q < p.K; ++q) {
        ia.kap[q] = p.kappa[q];
        ia.c1r[q] = p.c[4 * q + 1].real();
        ia.c1i[q] = p.c[4 * q + 1].imag();
        ia.c2r[q] = p.c[4 * q + 2].real();
        ia.c2i[q] = p.c[4 * q + 2].imag();
    }
    const int64_t total = h.nt * p.W;
    hipLaunchKernelGGL(incident_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, ia);
    BAEMM_HIP(hipGetLastError());
    EpiArgs ea = base_epi(h, a, WCSP);
    ea.kernel_on = 1;
    ea.diag_on = 0;
    ea.mass_on = 0;
    ea.S = 1;
    ea.seg_stride = 0;
    ea.ldp = ldp;
    ea.ldb = 0;
    ea.part = h.ws_part.as<float>();
    ea.b = nullptr;
    ea.unscale = nullptr;
    bdr_epilogue(h, ea, kC64, a.Ctype, s);
}


// --------------------------------------------------------------- P0 buffers

// The reference's P0 buffers (InputOutput.hpp) from the device data of the
// last apply, as complex float, simplex_count x W: which = 0: B, the staged
// sources b[t][k]; which = 1: C, the triangle values before AvT,
// rot_q * unscale_k * sum_s part_s[t][k] + D_t b[t][k] (the inner term of
// bdr_epilogue_kernel, summed in the same order in FP64).
__global__ void p0_kernel(P0Source src, int which, int nt, const float* __restrict__ sdiag,
                          float2* __restrict__ out)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)nt * src.W) return;
    const int t = (int)(idx / src.W), k = (int)(idx % src.W);
    const int q = k / src.wcs, kk = k % src.wcs;
    const int pc = q * src.WCSP + kk;
    const float2 bv = src.bptr[(int64_t)t * src.ldb + pc];
    if (which == 0) {
        out[idx] = bv;
        return;
    }
    double pr = 0.0, pi = 0.0;
    const float* pp = src.part + (int64_t)t * src.ldp + q * 2 * src.WCSP + kk;
    for (int sg = 0; sg < src.S; ++sg) {
        pr += (double)pp[(int64_t)sg * src.seg_stride];
        pi += (double)pp[(int64_t)sg * src.seg_stride + src.WCSP];
    }
    const double us = src.unscale ? (double)src.unscale[pc] : 1.0;
    double qr = pr * us, qi = pi * us;
    const double rr = src.rotr[q], ri = src.roti[q];
    const double tr = qr * rr - qi * ri;
    qi = qr * ri + qi * rr;
    qr = tr;
    if (src.diag) {
        const double fr = (double)src.c1r[q] * sdiag[t] - (double)src.c1i[q] * src.kap[q];
        const double fi = (double)src.c1i[q] * sdiag[t] + (double)src.c1r[q] * src.kap[q];
        qr += fr * bv.x - fi * bv.y;
        qi += fr * bv.y + fi * bv.x;
    }
    out[idx] = make_float2((float)qr, (float)qi);
}

// Records the device data of an apply as the source of the P0 buffers: b
// (B) and, for a boundary-operator apply (ea != null), the partials behind C.
static void p0_record(Handle& h, const float2* b, int ldb, int WCSP, const EpiArgs* ea)
{
    const Params& p = h.prm;
    P0Source& src = h.p0.src;
    src.b = b != nullptr;
    src.bptr = b;
    src.ldb = ldb;
    src.WCSP = WCSP;
    src.wcs = (int)p.wcs;
    src.K = (int)p.K;
    src.W = (int)p.W;
    h.p0.host_current[0] = false;
    h.p0.host_current[1] = false;
    src.c = ea != nullptr && ea->kernel_on;
    if (!src.c) return;
    src.part = ea->part;
    src.S = ea->S;
    src.ldp = ea->ldp;
    src.seg_stride = ea->seg_stride;
    src.unscale = ea->unscale;
    src.diag = ea->diag_on != 0;
    for (int q = 0; q < p.K; ++q) {
        src.rotr[q] = ea->rotr[q];
        src.roti[q] = ea->roti[q];
        src.c1r[q] = ea->c1r[q];
        src.c1i[q] = ea->c1i[q];
        src.kap[q] = ea->kap[q];
    }
}

void p0_invalidate(Handle& h)
{
    h.p0.src.b = h.p0.src.c = false;
    h.p0.host_current[0] = h.p0.host_current[1] = false;
}

bool p0_sync(Handle& h, int which, hipStream_t s)
{
    P0State& st = h.p0;
    if (st.host_current[which]) return true;
    const P0Source& src = st.src;
    if (!(which == 0 ? src.b : src.c)) return false;
    const int64_t n = h.nt * (int64_t)src.W;
    h.ws_p0.reserve(sizeof(float2) * std::max<int64_t>(n, 1));
    hipLaunchKernelGGL(p0_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, which, (int)h.nt, h.d_sdiag,
                       h.ws_p0.as<float2>());
    BAEMM_HIP(hipGetLastError());
    st.host[which].resize((size_t)n);
    BAEMM_HIP(hipMemcpyAsync(st.host[which].data(), h.ws_p0.ptr, sizeof(float2) * n, hipMemcpyDeviceToHost, s));
    BAEMM_HIP(hipStreamSynchronize(s));
    st.W[which] = src.W;
    st.host_current[which] = true;
    return true;
}

// --------------------------------------------------------------- diagnostics

__global__ void hw_sincos_kernel(const float* __restrict__ rev, int64_t n, float* __restrict__ s,
                                 float* __restrict__ c)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    s[i] = hw_sin_rev(rev[i]);
    c[i] = hw_cos_rev(rev[i]);
}

void hw_sincos_dev(const float* rev, int64_t n, float* sin_out, float* cos_out, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(hw_sincos_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, rev, n, sin_out,
                       cos_out);
    BAEMM_HIP(hipGetLastError());
}

}  // namespace baemm
