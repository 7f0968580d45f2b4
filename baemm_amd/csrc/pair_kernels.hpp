// pair_kernels.hpp -- triangle-pair interaction kernels for gfx950.
//
// Each kernel evaluates kernel values A(target i, source j) on the fly in FP32
// and contracts them with a block of complex right-hand sides:
//     part[seg][i][col] = sum_{j in seg} A_q(i, j) * b[j][col],  col in chunk q.
// The source range is split into segments (grid.y) so that a launch has
// enough work-groups for 256 CUs; the segment partials are summed in the
// epilogue kernels (deterministic, no atomics).
//
// Partial layout per target row (floats): for chunk q of padded width WCSP,
// the real parts at [q*2*WCSP, q*2*WCSP+WCSP) and the imaginary parts at
// [q*2*WCSP+WCSP, (q+1)*2*WCSP).  The MFMA kernel's 32-column output blocks map
// onto this layout directly.
#pragma once

#include <utility>

#include "internal.hpp"

namespace baemm {

enum PairOp { OP_BDR = 0, OP_NEAR = 1, OP_FAR = 2, OP_HERG = 3 };

struct CoefPack {
    ChunkCoef c[4];
};

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float float16_t __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float hw_sin_rev(float t) { return __builtin_amdgcn_sinf(t); }
__device__ __forceinline__ float hw_cos_rev(float t) { return __builtin_amdgcn_cosf(t); }

// kappa-independent geometry of one (target, source) pair for the boundary
// and potential kernels (BoundaryOperatorKernel.cl:104-112, 141, 151-155, 174-178).
struct BdrGeo {
    float rinv;   // 1/r, 0 where the pair is masked
    float r;      // r (0 where masked)
    float gj;     // (z . n_j) / r^2
    float hj;     // (z . n_j) / r^3
    float gi;     // -(z . n_i) / r^2  (adjoint double layer)
    float hi;     // -(z . n_i) / r^3
};

template <bool ADJ, bool MASK>
__device__ __forceinline__ BdrGeo bdr_geo(float4 x, float4 ni, float4 y, float4 nj, bool valid)
{
    const float z0 = y.x - x.x, z1 = y.y - x.y, z2 = y.z - x.z;
    const float r2 = fmaf(z0, z0, fmaf(z1, z1, z2 * z2));
    float rinv = __builtin_amdgcn_rsqf(r2);
    if constexpr (MASK) rinv = valid ? rinv : 0.0f;
    BdrGeo g;
    g.rinv = rinv;
    g.r = r2 * rinv;
    const float rinv2 = rinv * rinv;
    g.gj = fmaf(z0, nj.x, fmaf(z1, nj.y, z2 * nj.z)) * rinv2;
    g.hj = g.gj * rinv;
    if constexpr (ADJ) {
        g.gi = -fmaf(z0, ni.x, fmaf(z1, ni.y, z2 * ni.z)) * rinv2;
        g.hi = g.gi * rinv;
    } else {
        g.gi = g.hi = 0.0f;
    }
    return g;
}

// Formula A1, c1 E/r + (c2 (z.n_j) - c3 (z.n_i)) (i kappa r - 1) E / r^3,
// rearranged as A = E Q' with E = exp(i kappa r) and
//   Q' = c1/r - c2 (z.n_j)/r^3 - c3 gi/r + i kappa (c2 (z.n_j)/r^2 + c3 gi)
// (the kappa r factor of the double layer cancels one power of 1/r).
template <bool ADJ>
__device__ __forceinline__ void bdr_value(const BdrGeo& g, const ChunkCoef& c, float& Ar, float& Ai)
{
    const float t = g.r * c.kap_rev;
    const float cs = hw_cos_rev(t), sn = hw_sin_rev(t);
    float Qr = fmaf(-c.c2r, g.hj, fmaf(-c.kc2i, g.gj, c.c1r * g.rinv));
    float Qi = fmaf(-c.c2i, g.hj, fmaf(c.kc2r, g.gj, c.c1i * g.rinv));
    if constexpr (ADJ) {
        Qr = fmaf(-c.c3r, g.hi, fmaf(-c.kc3i, g.gi, Qr));
        Qi = fmaf(-c.c3i, g.hi, fmaf(c.kc3r, g.gi, Qi));
    }
    Ar = fmaf(cs, Qr, -sn * Qi);
    Ai = fmaf(cs, Qi, sn * Qr);
}

// Plane-wave kernels (formulas A2, A4): A = (c1 - i kappa (u.n) c2) exp(-i kappa u.y)
// with (u, y, n) = (xhat_i, y_j, n_j) for the far field and (d_j, x_t, n_t) for Herglotz.
__device__ __forceinline__ void plane_value(float uy, float un, const ChunkCoef& c, float& Ar, float& Ai)
{
    const float t = uy * c.kap_rev;
    const float cs = hw_cos_rev(t), sn = hw_sin_rev(t);
    const float kn = c.kap * un;
    const float wr = fmaf(kn, c.c2i, c.c1r);
    const float wi = fmaf(-kn, c.c2r, c.c1i);
    Ar = fmaf(wr, cs, wi * sn);
    Ai = fmaf(wi, cs, -wr * sn);
}

// ---------------------------------------------------------------------------
// Plain FP32 pair kernel: one target per thread, WCSP*KG complex accumulators,
// sources staged per 64 in LDS and broadcast to the whole block.
// ---------------------------------------------------------------------------
// One block of the plain kernel: targets [tile*256, tile*256+256) against the
// sources [s0, s1) (multiples of 64) of chunks q0..q0+KG-1; partial sums go to
// out[i*ldp + ...] (one segment slice).  s_y, s_n hold kSrcBlock float4 each,
// s_b kSrcBlock*WCSP*KG float2.  Ends with a barrier so the caller may reuse
// the LDS.
template <int OP, int WCSP, int KG, bool ADJ>
__device__ __forceinline__ void plain_block(const float4* __restrict__ tx, const float4* __restrict__ tn, int ntgt,
                                            const float4* __restrict__ sy, const float4* __restrict__ sn_, int nsrc,
                                            const float2* __restrict__ b, int ldb, const CoefPack& cp,
                                            float* __restrict__ outp, int ldp, int q0, int tile, int s0, int s1,
                                            float4* s_y, float4* s_n, float2* s_b)
{
    constexpr int NC = WCSP * KG;
    const int tid = threadIdx.x;
    const int i = tile * kTile + tid;
    const float4 x = tx[i];
    float4 ni = make_float4(0.f, 0.f, 0.f, 0.f);
    if constexpr (OP == OP_BDR || OP == OP_HERG) ni = tn[i];

    float accr[NC], acci[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) accr[k] = acci[k] = 0.0f;

    for (int jb = s0; jb < s1; jb += kSrcBlock) {
        __syncthreads();
        if (tid < kSrcBlock) s_y[tid] = sy[jb + tid];
        else if (tid < 2 * kSrcBlock) {
            if constexpr (OP != OP_HERG) s_n[tid - kSrcBlock] = sn_[jb + tid - kSrcBlock];
        }
        // b rows jb..jb+63, columns q0*WCSP .. q0*WCSP + NC (float4 = 2 complex)
        constexpr int V4 = kSrcBlock * NC / 2;
        for (int e = tid; e < V4; e += kTile) {
            const int row = e / (NC / 2 > 0 ? NC / 2 : 1);
            if constexpr (NC >= 2) {
                const int c2 = e % (NC / 2);
                const float4 v = *reinterpret_cast<const float4*>(b + (int64_t)(jb + row) * ldb + q0 * WCSP + 2 * c2);
                s_b[row * NC + 2 * c2] = make_float2(v.x, v.y);
                s_b[row * NC + 2 * c2 + 1] = make_float2(v.z, v.w);
            }
        }
        if constexpr (NC == 1) {
            if (tid < kSrcBlock) s_b[tid] = b[(int64_t)(jb + tid) * ldb + q0 * WCSP];
        }
        __syncthreads();
        const int jn = min(kSrcBlock, nsrc - jb);
        for (int jj = 0; jj < jn; ++jj) {
            const int j = jb + jj;
            float Ar[KG], Ai[KG];
            if constexpr (OP == OP_BDR || OP == OP_NEAR) {
                const bool valid = (OP == OP_NEAR) ? true : (j != i);
                const BdrGeo g = bdr_geo<ADJ, OP == OP_BDR>(x, ni, s_y[jj], s_n[jj], valid);
#pragma unroll
                for (int q = 0; q < KG; ++q) bdr_value<ADJ>(g, cp.c[q], Ar[q], Ai[q]);
            } else if constexpr (OP == OP_FAR) {
                const float4 y = s_y[jj], n = s_n[jj];
                const float uy = fmaf(x.x, y.x, fmaf(x.y, y.y, x.z * y.z));
                const float un = fmaf(x.x, n.x, fmaf(x.y, n.y, x.z * n.z));
#pragma unroll
                for (int q = 0; q < KG; ++q) plane_value(uy, un, cp.c[q], Ar[q], Ai[q]);
            } else {  // OP_HERG: sources are directions d_j
                const float4 d = s_y[jj];
                const float uy = fmaf(d.x, x.x, fmaf(d.y, x.y, d.z * x.z));
                const float un = fmaf(d.x, ni.x, fmaf(d.y, ni.y, d.z * ni.z));
#pragma unroll
                for (int q = 0; q < KG; ++q) plane_value(uy, un, cp.c[q], Ar[q], Ai[q]);
            }
            const float2* bj = s_b + jj * NC;
#pragma unroll
            for (int q = 0; q < KG; ++q) {
#pragma unroll
                for (int k = 0; k < WCSP; ++k) {
                    const float2 v = bj[q * WCSP + k];
                    accr[q * WCSP + k] = fmaf(Ar[q], v.x, fmaf(-Ai[q], v.y, accr[q * WCSP + k]));
                    acci[q * WCSP + k] = fmaf(Ar[q], v.y, fmaf(Ai[q], v.x, acci[q * WCSP + k]));
                }
            }
        }
    }
    if (i < ntgt) {
        float* out = outp + (int64_t)i * ldp + q0 * 2 * WCSP;
#pragma unroll
        for (int q = 0; q < KG; ++q) {
#pragma unroll
            for (int k = 0; k < WCSP; ++k) {
                out[q * 2 * WCSP + k] = accr[q * WCSP + k];
                out[q * 2 * WCSP + WCSP + k] = acci[q * WCSP + k];
            }
        }
    }
    __syncthreads();
}

template <int OP, int WCSP, int KG, bool ADJ>
__global__ __launch_bounds__(kTile) void pair_plain_kernel(
    const float4* __restrict__ tx, const float4* __restrict__ tn, int ntgt,
    const float4* __restrict__ sy, const float4* __restrict__ sn_, int nsrc,
    const float2* __restrict__ b, int ldb, CoefPack cp,
    float* __restrict__ part, int64_t seg_stride, int ldp, int q0, int seg_len, int nsrc_pad,
    const int* __restrict__ skip)
{
    if (skip && *skip) return;   // a converged solve: the step's result is not used
    __shared__ float4 s_y[kSrcBlock];
    __shared__ float4 s_n[kSrcBlock];
    __shared__ float2 s_b[kSrcBlock * WCSP * KG];
    const int s0 = blockIdx.y * seg_len;
    const int s1 = min(s0 + seg_len, nsrc_pad);
    plain_block<OP, WCSP, KG, ADJ>(tx, tn, ntgt, sy, sn_, nsrc, b, ldb, cp, part + (int64_t)blockIdx.y * seg_stride,
                                   ldp, q0, blockIdx.x, s0, s1, s_y, s_n, s_b);
}

// ---------------------------------------------------------------------------
// Mixed-precision boundary kernel on MFMA (v_mfma_f32_32x32x16_f16).
//
// Complex contraction as a real GEMM per chunk:
//   [Cr | Ci] = [Ar Ai] (interleaved along K) x [[br, bi], [-bi, br]]
// A-operand (generated in registers): lane l holds target i = l%32 and the
// 4 sources 4*(l/32)+0..3 of an 8-source step as (Ar, Ai) pairs.
// B-operand (precomputed by the split kernel, 16 bytes per lane): rows
// (j, re/im) x 32 real output columns.  Each value v is split
// v = hi + lo with hi = fp16(v) (round toward zero), lo = fp16(v - hi);
// the products hi*hi + hi*lo + lo*hi accumulate in FP32.  A carries a
// power-of-two scale folded into the coefficients, b a per-column
// power-of-two scale; the epilogue divides both out.
// Output lane l, register r: row (r/4)*8 + (l/32)*4 + r%4, column l%32.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t pack_rtz(float a, float b)
{
    auto h = __builtin_amdgcn_cvt_pkrtz(a, b);
    return __builtin_bit_cast(uint32_t, h);
}

// hi = fp16(a), fp16(b) rounded toward zero; lo = fp16 of the remainders
// a - hi, b - hi (exact in FP32: hi is a truncated to 11 significant bits).
// Each remainder is one v_fma_mix (-hi * 1.0 + a, the f16 operand read from
// the low or high half of hi) that rounds its result to f16 and writes it
// into the low (mixlo) or high (mixhi) half of lo, keeping the other half.
__device__ __forceinline__ void split2(float a, float b, uint32_t& hi, uint32_t& lo)
{
    hi = pack_rtz(a, b);
    uint32_t l;
    asm("v_fma_mixlo_f16 %0, -%1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(l) : "v"(hi), "v"(a));
    asm("v_fma_mixhi_f16 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "+v"(l) : "v"(hi), "v"(b));
    lo = l;
}

// FP16 hi/lo A operands of the 4 pairs of one lane for chunks q0..q0+KG-1
// (with F32: the FP32 values, ah = Re A, al = Im A).  Written stage by stage
// across the 4 sources so that the independent dependency chains interleave
// (the kernel is bound by VALU latency otherwise).
// The 4 pairs of one lane for one chunk (coefficients c): FP16 hi/lo A
// operands (with F32: ah = Re A, al = Im A).  Written stage by stage across
// the 4 sources so that the independent dependency chains interleave.
// FORM (no adjoint term only): which part of the chunk's single-layer
// coefficient c1 is zero and left out of the arithmetic, as the reference
// builds its kernels for the nonzero real and imaginary parts of its
// coefficients (LoadParameters.hpp:47-80): 0 none, 1 Re c1 = 0, 2 Im c1 = 0.
// The values are the same as with FORM 0 (a zero term adds nothing; at most
// the sign of a zero differs).
template <bool ADJ, bool F32 = false, int FORM = 0>
__device__ __forceinline__ void mixed_vals_q(const BdrGeo (&geo)[4], const ChunkCoef& c, uint32_t (&ah)[4],
                                             uint32_t (&al)[4])
{
    static_assert(FORM == 0 || !ADJ, "coefficient forms are for operators without the adjoint term");
    float cs[4], sn[4];
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
        const float t = geo[jj].r * c.kap_rev;
        cs[jj] = hw_cos_rev(t);
        sn[jj] = hw_sin_rev(t);
    }
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
        const BdrGeo& g = geo[jj];
        float Qr, Qi;
        if constexpr (ADJ) {
            Qr = fmaf(-c.c2r, g.hj, fmaf(-c.kc2i, g.gj, c.c1r * g.rinv));
            Qi = fmaf(-c.c2i, g.hj, fmaf(c.kc2r, g.gj, c.c1i * g.rinv));
            Qr = fmaf(-c.c3r, g.hi, fmaf(-c.kc3i, g.gi, Qr));
            Qi = fmaf(-c.c3i, g.hi, fmaf(c.kc3r, g.gi, Qi));
        } else {
            // The host divides each chunk's coefficients by c2 / s (s the
            // chunk scale; the epilogue multiplies c2 / s back) and the
            // source normals carry the factor s, so gj = s (z.n_j)/r^2,
            // hj = s (z.n_j)/r^3 and Q' = c1' / r - hj + i kappa gj.
            if constexpr (FORM == 1) Qr = -g.hj;
            else Qr = fmaf(c.c1r, g.rinv, -g.hj);
            if constexpr (FORM == 2) Qi = c.kap * g.gj;
            else Qi = fmaf(c.kap, g.gj, c.c1i * g.rinv);
        }
        const float Ar = fmaf(cs[jj], Qr, -sn[jj] * Qi);
        const float Ai = fmaf(cs[jj], Qi, sn[jj] * Qr);
        if constexpr (F32) {
            ah[jj] = __float_as_uint(Ar);
            al[jj] = __float_as_uint(Ai);
        } else {
            split2(Ar, Ai, ah[jj], al[jj]);
        }
    }
}

// FP16 hi/lo A operands of the 4 pairs of one lane for chunks q0..q0+KG-1.
template <int KG, bool ADJ, bool F32 = false>
__device__ __forceinline__ void mixed_vals(const BdrGeo (&geo)[4], const CoefPack& cp, uint32_t (&ah)[KG][4],
                                           uint32_t (&al)[KG][4])
{
#pragma unroll
    for (int q = 0; q < KG; ++q) mixed_vals_q<ADJ, F32>(geo, cp.c[q], ah[q], al[q]);
}

// Geometry of the 4 (target i, source jbase + jj) pairs of one lane when the
// sources are spread over the lanes instead of read by every lane: lane
// 16 r + jj (jj < 4) of each 16-lane row r holds source jj of its half-wave
// (centroid yv, normal nv); the other lanes hold copies.  Each difference
// z = y - x and each term of z . n_j takes the source value from lane jj of the
// row through a row_newbcast:jj DPP operand, so the broadcast costs no
// instruction: the same 24 VALU operations as bdr_geo, and the same rounding.
// The DPP source registers come from memory loads; s_nop 4 gives the wait
// states a DPP read needs after a VALU or EXEC write, in case the compiler
// moved a value just before.
// TN (no adjoint term): the normal products come from the target normal ni,
// dn = -(z . n_i), so that gj and hj hold the adjoint double layer's
// -(z . n_i)/r^2 and /r^3 (the gi, hi of the general form, same rounding); the
// source normals nv are not read.
template <bool ADJ, bool MASK, bool TN = false>
__device__ __forceinline__ void mixed_geo_dpp(const float4 x, const float4 ni, const float4 yv, const float4 nv,
                                              int jbase, int i, int nt, BdrGeo (&geo)[4])
{
    static_assert(!(ADJ && TN), "the target-normal form has no separate adjoint term");
    float z0[4], z1[4], z2[4], dn[4];
    asm("s_nop 4\n\t"
        "v_sub_f32_dpp %0, %12, %15 row_newbcast:0 row_mask:0xf bank_mask:0xf\n\t"
        "v_sub_f32_dpp %1, %13, %16 row_newbcast:0 row_mask:0xf bank_mask:0xf\n\t"
        "v_sub_f32_dpp %2, %14, %17 row_newbcast:0 row_mask:0xf bank_mask:0xf\n\t"
        "v_sub_f32_dpp %3, %12, %15 row_newbcast:1 row_mask:0xf bank_mask:0xf\n\t"
        "v_sub_f32_dpp %4, %13, %16 row_newbcast:1 row_mask:0xf bank_mask:0xf\n\t"
        "v_sub_f32_dpp %5, %14, %17 row_newbcast:1 row_mask:0xf bank_mask:0xf\n\t"
        "v_sub_f32_dpp %6, %12, %15 row_newbcast:2 row_mask:0xf bank_mask:0xf\n\t"
        "v_sub_f32_dpp %7, %13, %16 row_newbcast:2 row_mask:0xf bank_mask:0xf\n\t"
        "v_sub_f32_dpp %8, %14, %17 row_newbcast:2 row_mask:0xf bank_mask:0xf\n\t"
        "v_sub_f32_dpp %9, %12, %15 row_newbcast:3 row_mask:0xf bank_mask:0xf\n\t"
        "v_sub_f32_dpp %10, %13, %16 row_newbcast:3 row_mask:0xf bank_mask:0xf\n\t"
        "v_sub_f32_dpp %11, %14, %17 row_newbcast:3 row_mask:0xf bank_mask:0xf"
        : "=&v"(z0[0]), "=&v"(z1[0]), "=&v"(z2[0]), "=&v"(z0[1]), "=&v"(z1[1]), "=&v"(z2[1]), "=&v"(z0[2]),
          "=&v"(z1[2]), "=&v"(z2[2]), "=&v"(z0[3]), "=&v"(z1[3]), "=&v"(z2[3])
        : "v"(yv.x), "v"(yv.y), "v"(yv.z), "v"(x.x), "v"(x.y), "v"(x.z));
    if constexpr (TN) {
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) dn[jj] = -fmaf(z0[jj], ni.x, fmaf(z1[jj], ni.y, z2[jj] * ni.z));
    } else {
    // z . n_j = (z2 n_z + z1 n_y) + z0 n_x, in the order of bdr_geo's fmaf chain
    asm("s_nop 4\n\t"
        "v_mul_f32_dpp %0, %6, %9 row_newbcast:0 row_mask:0xf bank_mask:0xf\n\t"
        "v_mul_f32_dpp %1, %6, %12 row_newbcast:1 row_mask:0xf bank_mask:0xf\n\t"
        "v_mul_f32_dpp %2, %6, %15 row_newbcast:2 row_mask:0xf bank_mask:0xf\n\t"
        "v_mul_f32_dpp %3, %6, %18 row_newbcast:3 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %0, %5, %8 row_newbcast:0 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %1, %5, %11 row_newbcast:1 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %2, %5, %14 row_newbcast:2 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %3, %5, %17 row_newbcast:3 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %0, %4, %7 row_newbcast:0 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %1, %4, %10 row_newbcast:1 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %2, %4, %13 row_newbcast:2 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %3, %4, %16 row_newbcast:3 row_mask:0xf bank_mask:0xf"
        : "=&v"(dn[0]), "=&v"(dn[1]), "=&v"(dn[2]), "=&v"(dn[3])
        : "v"(nv.x), "v"(nv.y), "v"(nv.z), "v"(z0[0]), "v"(z1[0]), "v"(z2[0]), "v"(z0[1]), "v"(z1[1]),
          "v"(z2[1]), "v"(z0[2]), "v"(z1[2]), "v"(z2[2]), "v"(z0[3]), "v"(z1[3]), "v"(z2[3]));
    }
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
        const int j = jbase + jj;
        const float r2 = fmaf(z0[jj], z0[jj], fmaf(z1[jj], z1[jj], z2[jj] * z2[jj]));
        float rinv = __builtin_amdgcn_rsqf(r2);
        if constexpr (MASK) rinv = ((j != i) && (j < nt)) ? rinv : 0.0f;
        BdrGeo g;
        g.rinv = rinv;
        g.r = r2 * rinv;
        const float rinv2 = rinv * rinv;
        g.gj = dn[jj] * rinv2;
        g.hj = g.gj * rinv;
        if constexpr (ADJ) {
            g.gi = -fmaf(z0[jj], ni.x, fmaf(z1[jj], ni.y, z2[jj] * ni.z)) * rinv2;
            g.hi = g.gi * rinv;
        } else {
            g.gi = g.hi = 0.0f;
        }
        geo[jj] = g;
    }
}

// Buffer resource over the n float4 elements at p (raw, byte-addressed; word 3
// as the gfx9 family takes it for untyped dword loads).  Loads past n * 16
// bytes return zeros.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const float4* p, int n)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float4*>(p), (short)0, n * (int)sizeof(float4), 0x00020000);
}

// x, y, z of the float4 at byte offset voff + soff of r (w = 0).
__device__ __forceinline__ float4 load_xyz(__amdgpu_buffer_rsrc_t r, uint32_t voff, int soff)
{
    const auto v = __builtin_amdgcn_raw_buffer_load_b96(r, voff, soff, 0);
    return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]), 0.0f);
}

// Coefficient forms of the chunks of a single-wave launch, FORM of chunk q in
// base-3 digit q (see mixed_vals_q).
constexpr int chunk_form(int forms, int q) { return q == 0 ? forms % 3 : chunk_form(forms / 3, q - 1); }

template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, I>{}), ...);
}
// f(integral_constant<int, q>) for q = 0 .. N-1.
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f)
{
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

#ifndef BAEMM_MIX_MINB
#define BAEMM_MIX_MINB 4
#endif
// LDS footprint of one single-wave mixed block: the B operand tiles (uint4)
// of NS steps, double buffered.
template <int WCSP, int KG>
struct MixedLds {
    static constexpr int NCB = 2 * WCSP / 32;
    static constexpr int OPS = KG * NCB * 2;     // 1-KB operand tiles per step (chunk x col-block x hi/lo)
    static constexpr int NS = 2;                 // steps (8 sources each) per LDS stage
    static constexpr int TILES = NS * OPS;
    static constexpr int OP_ELEMS = 2 * TILES * 64;
};

// One block of the single-wave mixed kernel: targets [tile*128, tile*128+128)
// against the source steps [st0, st1) of chunks q0..q0+KG-1, partial sums
// written to out[row*ldp + ...] (one segment slice; out NULL: a tile computed
// only to keep the block's barriers in step).  The targets are tgt[0 .. ntgt)
// (reads clamp there): the centroids themselves (self, boundary operators:
// the pair i = j is left out, the target normals are nrm) or other points
// (potential: no pair left out, no target normals).  Each wave generates the FP16
// hi/lo A operands of its 32 targets in registers and contracts them itself,
// one wave-number chunk at a time, so that the MFMAs of a chunk run while the
// next chunk's values are generated.
// Each lane loads the one source (centroid, normal) of each step it holds for
// mixed_geo_dpp from global memory, one step ahead; only the B operands are
// staged in LDS (s_op: MixedLds<WCSP,KG>::OP_ELEMS elements), NS steps per
// stage and barrier.  Called by whole blocks of 256 threads (or 512 running
// two tiles side by side); ends with a barrier so the caller may reuse the LDS.
template <int WCSP, int KG, bool ADJ, int FORMS = 0, bool TN = false>
__device__ __forceinline__ void mixed_block(const float4* __restrict__ tgt, int ntgt, bool self,
                                            const float4* __restrict__ mid, const float4* __restrict__ nrm, int nt,
                                            const uint4* __restrict__ bop, int nst, const CoefPack& cp,
                                            float* __restrict__ out, int ldp, int q0, int tile, int st0, int st1,
                                            uint4* s_op)
{
    using L = MixedLds<WCSP, KG>;
    constexpr int NCB = L::NCB, OPS = L::OPS, NS = L::NS, TILES = L::TILES;
    const int lane = threadIdx.x & 63;
    // wave-uniform (scalar register): 4 waves per tile
    const int wave = __builtin_amdgcn_readfirstlane((threadIdx.x >> 6) & 3);
    const int g = lane >> 5;
    const int ibase = tile * kMixTile + wave * 32;
    const int i = ibase + (lane & 31);
    const float4 x = tgt[min(i, ntgt - 1)];
    float4 ni = make_float4(0.f, 0.f, 0.f, 0.f);
    if constexpr (ADJ || TN)
        if (self) ni = nrm[i];

    float16_t acc[KG][NCB];
#pragma unroll
    for (int q = 0; q < KG; ++q)
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[q][cb][r] = 0.0f;

    const int nt_pad = nst * 8;
    // steps to mask (wave-uniform): the 4 steps whose sources are this wave's
    // targets (self only), and steps reaching past nt
    const int st_diag = self ? __builtin_amdgcn_readfirstlane(ibase >> 3) : -(1 << 28);
    const int st_tail = nt >> 3;
    const int i_mask = self ? i : -1;   // the source index left out
    // the source of each step this lane holds for mixed_geo_dpp
    const int src_lane = 4 * g + (lane & 3);

    auto issue_stage = [&](int sbase, int buf) {
        for (int t = wave; t < TILES; t += 4) {
            const int st = sbase + t / OPS;
            if (st >= st1) break;
            const int o = t % OPS;
            const int q = o / (NCB * 2), cb = (o / 2) % NCB, pt = o % 2;
            const uint4* src = bop + ((((int64_t)(q0 + q) * nst + st) * NCB + cb) * 2 + pt) * 64 + lane;
            __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)&s_op[(buf * TILES + t) * 64],
                                             16, 0, 0);
        }
    };

    // The source of step st this lane holds is st * 8 + src_lane < nst * 8 =
    // nt_pad: a wave-uniform byte offset st * 128 plus this lane's fixed one,
    // read through buffer resources so that a step needs no address arithmetic.
    const __amdgpu_buffer_rsrc_t mid_rsrc = buffer_rsrc(mid, nt_pad);
    const __amdgpu_buffer_rsrc_t nrm_rsrc = buffer_rsrc(nrm, nt_pad);
    const uint32_t src_off = (uint32_t)src_lane * (uint32_t)sizeof(float4);
    float4 yv = make_float4(0.f, 0.f, 0.f, 0.f), nv = yv, yn = yv, nn = yv;
    if (st0 < st1) {
        yv = load_xyz(mid_rsrc, src_off, st0 * 128);
        nv = load_xyz(nrm_rsrc, src_off, st0 * 128);
        issue_stage(st0, 0);
    }

    // Step st (slot s of the stage in buffer buf) on the source (yc, nc); the
    // next step's source is loaded into (yx, nx).
    auto step = [&](int st, int s, int buf, const float4 yc, const float4 nc, float4& yx, float4& nx) {
        if (st + 1 < st1) {
            yx = load_xyz(mid_rsrc, src_off, (st + 1) * 128);
            nx = load_xyz(nrm_rsrc, src_off, (st + 1) * 128);
        }
        BdrGeo geo[4];
        const bool masked = (unsigned)(st - st_diag) < 4u || st >= st_tail;
        if (masked) mixed_geo_dpp<ADJ, true, TN>(x, ni, yc, nc, 8 * st + 4 * g, i_mask, nt, geo);
        else mixed_geo_dpp<ADJ, false, TN>(x, ni, yc, nc, 8 * st + 4 * g, i_mask, nt, geo);
        // chunk by chunk: the MFMAs of chunk q run while chunk q + 1 is generated
        static_for<KG>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            half8_t b_hi[NCB], b_lo[NCB];
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb) {
                const int o = (q * NCB + cb) * 2;
                b_hi[cb] = __builtin_bit_cast(half8_t, s_op[(buf * TILES + s * OPS + o) * 64 + lane]);
                b_lo[cb] = __builtin_bit_cast(half8_t, s_op[(buf * TILES + s * OPS + o + 1) * 64 + lane]);
            }
            uint32_t ahq[4], alq[4];
            mixed_vals_q<ADJ, false, chunk_form(FORMS, q)>(geo, cp.c[q], ahq, alq);
            const half8_t a_hi = __builtin_bit_cast(half8_t, make_uint4(ahq[0], ahq[1], ahq[2], ahq[3]));
            const half8_t a_lo = __builtin_bit_cast(half8_t, make_uint4(alq[0], alq[1], alq[2], alq[3]));
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb)
                acc[q][cb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a_hi, b_hi[cb], acc[q][cb], 0, 0, 0);
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb)
                acc[q][cb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a_hi, b_lo[cb], acc[q][cb], 0, 0, 0);
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb)
                acc[q][cb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a_lo, b_hi[cb], acc[q][cb], 0, 0, 0);
        });
    };

    // Two steps per stage: the first on (yv, nv) loading (yn, nn), the second
    // the other way round, so the source registers alternate without copies.
    // Only the last stage can hold a single step.
    static_assert(NS == 2, "the step loop alternates two source registers");
    int buf = 0;
    for (int sb = st0; sb < st1; sb += NS, buf ^= 1) {
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (sb + NS < st1) issue_stage(sb + NS, buf ^ 1);
        step(sb, 0, buf, yv, nv, yn, nn);
        if (sb + 1 < st1) step(sb + 1, 1, buf, yn, nn, yv, nv);
    }
    if (out) {
#pragma unroll
        for (int q = 0; q < KG; ++q)
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int row = ibase + (r >> 2) * 8 + g * 4 + (r & 3);
                    out[(int64_t)row * ldp + (q0 + q) * 2 * WCSP + cb * 32 + (lane & 31)] = acc[q][cb][r];
                }
    }
    __syncthreads();
}

// Waves per SIMD the single-wave kernel is built for: 4 (128 VGPRs) where the
// registers fit without spilling in the step loop, else 3.
template <int KG, bool ADJ>
constexpr int mixed_min_waves() { return (!ADJ && KG <= 2) ? BAEMM_MIX_MINB : 3; }

template <int WCSP, int KG, bool ADJ, int FORMS = 0, bool TN = false>
__global__ __launch_bounds__(256, (mixed_min_waves<KG, ADJ>())) void bdr_mixed_kernel(
    const float4* __restrict__ tgt, int ntgt, int self,
    const float4* __restrict__ mid, const float4* __restrict__ nrm, int nt,
    const uint4* __restrict__ bop, int nst, CoefPack cp,
    float* __restrict__ part, int64_t seg_stride, int ldp, int q0, int steps_per_seg, const int* __restrict__ skip)
{
    if (skip && *skip) return;   // a converged solve: the step's result is not used
    using L = MixedLds<WCSP, KG>;
    __shared__ uint4 s_op[L::OP_ELEMS];
    const int st0 = blockIdx.y * steps_per_seg;
    const int st1 = min(st0 + steps_per_seg, nst);
    mixed_block<WCSP, KG, ADJ, FORMS, TN>(tgt, ntgt, self != 0, mid, nrm, nt, bop, nst, cp,
                                      part + (int64_t)blockIdx.y * seg_stride, ldp, q0, blockIdx.x, st0, st1, s_op);
}

// ---------------------------------------------------------------------------
// Wave-specialised variant of the mixed kernel: 8 waves per block.  Waves 0-3
// (producers) generate the FP16 hi/lo A operands of their 32 targets exactly
// as bdr_mixed_kernel does and write them to LDS; waves 4-7 (consumers) read
// them with the staged B operands and issue the MFMAs, one 8-source step
// behind.  Kernel-value generation (VALU and transcendental units) and the
// contraction (matrix cores) then run in different waves of the same SIMD and
// overlap, instead of alternating inside one wave.
//
// Per step k (st0 + k), separated by block barriers:
//   producers: A(k) from geometry(k)            -> s_a[k % 2]
//   consumers: MFMA with A(k-1), B(k-1)          <- s_a[(k-1) % 2], s_op[(k-1) % 2]
//   loads:     B(k) -> s_op[k % 2] (consumer waves); each generation lane
//              loads the one source of step k+1 it holds for mixed_geo_dpp
//              into registers
// Every buffer is rewritten only after the barrier that ends its last read.
// ---------------------------------------------------------------------------
template <int WCSP, int KG>
struct WsLds {
    static constexpr int NCB = 2 * WCSP / 32;
    static constexpr int OPS = KG * NCB * 2;          // B tiles per step
    static constexpr int B_ELEMS = 2 * OPS * 64;      // uint4, double buffered
    static constexpr int A_ELEMS = 2 * 4 * KG * 2 * 64;
};

// One block (8 waves) of the wave-specialised kernel: targets
// [tile*128, tile*128+128) against source steps [st0, st1) of chunks
// q0..q0+KG-1, partial sums to out[row*ldp + ...].  The targets are tgt[0 ..
// ntgt) (reads clamp there): the centroids themselves (self, boundary
// operators: the pair i = j is left out, the target normals are nrm) or other
// points (potential: no pair left out, no target normals).  s_op, s_a: LDS of
// at least WsLds<WCSP,KG> B_ELEMS, A_ELEMS elements.  Ends with every wave past
// a common barrier, so the caller may reuse the LDS.
template <int WCSP, int KG, bool ADJ, bool F32 = false, bool TN = false>
__device__ __forceinline__ void mixed_ws_block(const float4* __restrict__ tgt, int ntgt, bool self,
                                               const float4* __restrict__ mid, const float4* __restrict__ nrm, int nt,
                                               const uint4* __restrict__ bop, int nst, const CoefPack& cp,
                                               float* __restrict__ out, int ldp, int q0, int tile, int st0, int st1,
                                               uint4* s_op, uint4* s_a)
{
    using L = WsLds<WCSP, KG>;
    constexpr int NCB = L::NCB, OPS = L::OPS;

    const int lane = threadIdx.x & 63;
    // wave-uniform role (scalar register), so the branches below are uniform
    // and every wave executes the same number of barriers
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const bool producer = wave < 4;
    const int w = wave & 3;
    const int g = lane >> 5;
    const int ibase = tile * kMixTile + w * 32;
    const int i = ibase + (lane & 31);
    const int T = max(0, st1 - st0);
    const int nt_pad = nst * 8;
    // Steps to mask (wave-uniform, kept in scalar registers): the 4 steps whose
    // sources are this wave's targets (self only), and steps reaching past nt.
    const int st_diag = self ? __builtin_amdgcn_readfirstlane(ibase >> 3) : -(1 << 28);
    const int st_tail = nt >> 3;
    const int i_mask = self ? i : -1;   // the source index left out

    auto issue_b = [&](int k) {
        const int st = st0 + k;
        for (int t = w; t < OPS; t += 4) {
            const int q = t / (NCB * 2), cb = (t / 2) % NCB, pt = t % 2;
            const uint4* src = bop + ((((int64_t)(q0 + q) * nst + st) * NCB + cb) * 2 + pt) * 64 + lane;
            __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)&s_op[((k & 1) * OPS + t) * 64],
                                             16, 0, 0);
        }
    };

    if (producer) {
        const float4 x = tgt[min(i, ntgt - 1)];
        float4 ni = make_float4(0.f, 0.f, 0.f, 0.f);
        if constexpr (ADJ || TN)
            if (self) ni = nrm[i];
        // the source of each step this lane holds for mixed_geo_dpp
        const int src_lane = 4 * g + (lane & 3);
        float4 yv = make_float4(0.f, 0.f, 0.f, 0.f), nv = yv;
        if (T > 0) {
            const int j = min(st0 * 8 + src_lane, nt_pad - 1);
            yv = mid[j];
            nv = nrm[j];
        }
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        for (int k = 0; k <= T; ++k) {
            float4 yn = yv, nn = nv;
            if (k + 1 < T) {
                const int j = min((st0 + k + 1) * 8 + src_lane, nt_pad - 1);
                yn = mid[j];
                nn = nrm[j];
            }
            if (k < T) {
                const int st = st0 + k;
                uint32_t ah[KG][4], al[KG][4];
                const bool masked = (unsigned)(st - st_diag) < 4u || st >= st_tail;
                BdrGeo geo[4];
                if (masked) mixed_geo_dpp<ADJ, true, TN>(x, ni, yv, nv, 8 * st + 4 * g, i_mask, nt, geo);
                else mixed_geo_dpp<ADJ, false, TN>(x, ni, yv, nv, 8 * st + 4 * g, i_mask, nt, geo);
                mixed_vals<KG, ADJ, F32>(geo, cp, ah, al);
#pragma unroll
                for (int q = 0; q < KG; ++q) {
                    const int base = (((k & 1) * 4 + w) * KG + q) * 2;
                    if constexpr (F32) {
                        // operands of the FP32 MFMAs m = 0..7: (Re, Im) of sources 0..3
                        s_a[base * 64 + lane] = make_uint4(ah[q][0], al[q][0], ah[q][1], al[q][1]);
                        s_a[(base + 1) * 64 + lane] = make_uint4(ah[q][2], al[q][2], ah[q][3], al[q][3]);
                    } else {
                        s_a[base * 64 + lane] = make_uint4(ah[q][0], ah[q][1], ah[q][2], ah[q][3]);
                        s_a[(base + 1) * 64 + lane] = make_uint4(al[q][0], al[q][1], al[q][2], al[q][3]);
                    }
                }
            }
            yv = yn;
            nv = nn;
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
        }
    } else {
        float16_t acc[KG][NCB];
#pragma unroll
        for (int q = 0; q < KG; ++q)
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[q][cb][r] = 0.0f;
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        for (int k = 0; k <= T; ++k) {
            if (k < T) issue_b(k);
            if constexpr (F32) {
                if (k >= 1) {
                    // FP32 MFMA m (K = 2): lane half h supplies K element h, i.e. the
                    // component (m & 1) of source 4 h + (m >> 1), for A and for B
                    const int kb = (k - 1) & 1;
                    float a[KG][8], bv[KG][NCB][8];
#pragma unroll
                    for (int q = 0; q < KG; ++q) {
                        const int base = ((kb * 4 + w) * KG + q) * 2;
                        const uint4 u0 = s_a[base * 64 + lane], u1 = s_a[(base + 1) * 64 + lane];
                        const uint32_t e[8] = {u0.x, u0.y, u0.z, u0.w, u1.x, u1.y, u1.z, u1.w};
#pragma unroll
                        for (int m = 0; m < 8; ++m) a[q][m] = __uint_as_float(e[m]);
#pragma unroll
                        for (int cb = 0; cb < NCB; ++cb) {
                            const int o = (q * NCB + cb) * 2;
                            const uint4 v0 = s_op[(kb * OPS + o) * 64 + lane], v1 = s_op[(kb * OPS + o + 1) * 64 + lane];
                            const uint32_t f[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
#pragma unroll
                            for (int m = 0; m < 8; ++m) bv[q][cb][m] = __uint_as_float(f[m]);
                        }
                    }
                    // m-major: consecutive MFMAs write different accumulators
#pragma unroll
                    for (int m = 0; m < 8; ++m)
#pragma unroll
                        for (int q = 0; q < KG; ++q)
#pragma unroll
                            for (int cb = 0; cb < NCB; ++cb)
                                acc[q][cb] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[q][m], bv[q][cb][m], acc[q][cb], 0, 0, 0);
                }
            } else if (k >= 1) {
                const int kb = (k - 1) & 1;
                half8_t a_hi[KG], a_lo[KG];
#pragma unroll
                for (int q = 0; q < KG; ++q) {
                    const int base = ((kb * 4 + w) * KG + q) * 2;
                    a_hi[q] = __builtin_bit_cast(half8_t, s_a[base * 64 + lane]);
                    a_lo[q] = __builtin_bit_cast(half8_t, s_a[(base + 1) * 64 + lane]);
                }
                half8_t b_hi[KG][NCB], b_lo[KG][NCB];
#pragma unroll
                for (int q = 0; q < KG; ++q)
#pragma unroll
                    for (int cb = 0; cb < NCB; ++cb) {
                        const int o = (q * NCB + cb) * 2;
                        b_hi[q][cb] = __builtin_bit_cast(half8_t, s_op[(kb * OPS + o) * 64 + lane]);
                        b_lo[q][cb] = __builtin_bit_cast(half8_t, s_op[(kb * OPS + o + 1) * 64 + lane]);
                    }
                // product-major: consecutive MFMAs write different accumulators
#pragma unroll
                for (int q = 0; q < KG; ++q)
#pragma unroll
                    for (int cb = 0; cb < NCB; ++cb)
                        acc[q][cb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a_hi[q], b_hi[q][cb], acc[q][cb], 0, 0, 0);
#pragma unroll
                for (int q = 0; q < KG; ++q)
#pragma unroll
                    for (int cb = 0; cb < NCB; ++cb)
                        acc[q][cb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a_hi[q], b_lo[q][cb], acc[q][cb], 0, 0, 0);
#pragma unroll
                for (int q = 0; q < KG; ++q)
#pragma unroll
                    for (int cb = 0; cb < NCB; ++cb)
                        acc[q][cb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a_lo[q], b_hi[q][cb], acc[q][cb], 0, 0, 0);
            }
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
        }
#pragma unroll
        for (int q = 0; q < KG; ++q)
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int row = ibase + (r >> 2) * 8 + g * 4 + (r & 3);
                    out[(int64_t)row * ldp + (q0 + q) * 2 * WCSP + cb * 32 + (lane & 31)] = acc[q][cb][r];
                }
    }
}

template <int WCSP, int KG, bool ADJ, bool F32 = false, bool TN = false>
__global__ __launch_bounds__(512, 4) void bdr_mixed_ws_kernel(
    const float4* __restrict__ tgt, int ntgt, int self,
    const float4* __restrict__ mid, const float4* __restrict__ nrm, int nt,
    const uint4* __restrict__ bop, int nst, CoefPack cp,
    float* __restrict__ part, int64_t seg_stride, int ldp, int q0, int steps_per_seg, const int* __restrict__ skip)
{
    if (skip && *skip) return;   // a converged solve: the step's result is not used
    using L = WsLds<WCSP, KG>;
    __shared__ uint4 s_op[L::B_ELEMS];
    __shared__ uint4 s_a[L::A_ELEMS];
    const int st0 = blockIdx.y * steps_per_seg;
    const int st1 = min(st0 + steps_per_seg, nst);
    mixed_ws_block<WCSP, KG, ADJ, F32, TN>(tgt, ntgt, self != 0, mid, nrm, nt, bop, nst, cp,
                                           part + (int64_t)blockIdx.y * seg_stride, ldp, q0, blockIdx.x, st0, st1,
                                           s_op, s_a);
}

}  // namespace baemm
