// stage.hpp -- per-element bodies of the boundary apply's staging and
// epilogue kernels, shared by the single-instance kernels (ops.hip) and their
// batched forms (batch.hip), so that both compute the same numbers.
#pragma once

#include "ops.hpp"

namespace baemm {

template <int T>
struct Elem;
template <>
struct Elem<kC128> {
    using type = double2;
};
template <>
struct Elem<kC64> {
    using type = float2;
};

__device__ __forceinline__ double2 to_d2(double2 v) { return v; }
__device__ __forceinline__ double2 to_d2(float2 v) { return make_double2(v.x, v.y); }
__device__ __forceinline__ void from_d2(double2& dst, double2 v) { dst = v; }
__device__ __forceinline__ void from_d2(float2& dst, double2 v) { dst = make_float2((float)v.x, (float)v.y); }

// Element idx (row t = idx / ldb, padded column pc = idx % ldb) of the staged
// sources b = Av B:  b[t][q*WCSP+kk] = (a_t/3) (B[i0][k] + B[i1][k] + B[i2][k]),
// k = q*wcs + kk (AvOp.Dot, ApplyBoundaryOperators.hpp:120-124; Av values
// Initialize.hpp:133-143).  Rows t >= nt and columns kk >= wcs are zero.
template <int BT>
__device__ __forceinline__ float2 av_elem(const int4* __restrict__ tris, const float* __restrict__ area, int nt,
                                          const void* __restrict__ Bv, int64_t ldB, int wcs, int WCSP, int ldb,
                                          int64_t idx)
{
    using T = typename Elem<BT>::type;
    const T* B = static_cast<const T*>(Bv);
    const int t = (int)(idx / ldb);
    const int pc = (int)(idx % ldb);
    const int q = pc / WCSP, kk = pc % WCSP;
    float2 out = make_float2(0.f, 0.f);
    if (t < nt && kk < wcs) {
        const int k = q * wcs + kk;
        const int4 tv = tris[t];
        const double2 v0 = to_d2(B[(int64_t)tv.x * ldB + k]);
        const double2 v1 = to_d2(B[(int64_t)tv.y * ldB + k]);
        const double2 v2 = to_d2(B[(int64_t)tv.z * ldB + k]);
        const double w = (double)area[t] / 3.0;
        out = make_float2((float)(w * (v0.x + v1.x + v2.x)), (float)(w * (v0.y + v1.y + v2.y)));
    }
    return out;
}

// Power-of-two scale of a column whose largest |re|, |im| has the bit pattern
// m (as the column maxima hold it): max |s b| lies in [2^13, 2^14).
__device__ __forceinline__ float col_scale(unsigned mbits)
{
    const float m = __uint_as_float(mbits);
    float s = 1.0f;
    if (m > 0.0f) {
        int e;
        frexpf(m, &e);          // m in [2^(e-1), 2^e)
        s = ldexpf(1.0f, 14 - e);
    }
    return s;
}

// Element idx of the B operand of the mixed MFMA kernels: 16 bytes hi and 16
// bytes lo per (chunk, step, col-block, lane), see bdr_mixed_kernel.  Values
// are scaled per column (scale_of(pc)), then split hi/lo (RN).
template <class ScaleOf>
__device__ __forceinline__ void split_elem(const float2* __restrict__ b, int ldb, int nst, int WCSP, int64_t idx,
                                           ScaleOf scale_of, uint4* __restrict__ bop)
{
    const int NCB = 2 * WCSP / 32;
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
    const float s = scale_of(pc);
    _Float16 hi[8], lo[8];
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
        const int j = 8 * st + 4 * g + jj;
        const float2 v = b[(int64_t)j * ldb + pc];
        const float br = v.x * s, bi = v.y * s;
        // rows (j, re) and (j, im) of [[br, bi], [-bi, br]]
        const float e0 = imcol ? bi : br;
        const float e1 = imcol ? br : -bi;
        const _Float16 h0 = (_Float16)e0, h1 = (_Float16)e1;
        hi[2 * jj] = h0;
        hi[2 * jj + 1] = h1;
        lo[2 * jj] = (_Float16)(e0 - (float)h0);
        lo[2 * jj + 1] = (_Float16)(e1 - (float)h1);
    }
    const int64_t base = ((((int64_t)q * nst + st) * NCB + cb) * 2) * 64 + lane;
    bop[base] = __builtin_bit_cast(uint4, hi);
    bop[base + 64] = __builtin_bit_cast(uint4, lo);
}

// Batch sizes of the epilogue's loads: incident triangles, and segment partials
// (also mass-matrix entries per row).
constexpr int EPI_TRI = 2;
constexpr int EPI_SEG = 8;

// Lanes per vertex of the epilogue: lanes run over the columns, so narrow
// right-hand sides leave no lanes idle.
__host__ __device__ inline int epi_lanes(int W) { return W <= 16 ? 16 : (W <= 32 ? 32 : 64); }

// Value of triangle t in column pc = q * WCSP + kk of the boundary epilogue,
// from its segment sums (pr, pi): unscaled, multiplied by the chunk's phase,
// plus the single-layer diagonal term D_t b[t][pc].  The vertex epilogue and
// tri_epilogue_kernel both form it here, so their values are the same bits.
__device__ __forceinline__ void epi_tri_value(const EpiArgs& a, int t, int q, int pc, double pr, double pi,
                                              double& qr, double& qi)
{
    const double us = a.unscale ? (double)a.unscale[pc] : 1.0;
    qr = pr * us;
    qi = pi * us;
    {
        const double rr = a.rotr[q], ri = a.roti[q];
        const double tr = qr * rr - qi * ri;
        qi = qr * ri + qi * rr;
        qr = tr;
    }
    if (a.diag_on) {
        const float2 bv = a.b[(int64_t)t * a.ldb + pc];
        const double fr = (double)a.c1r[q] * a.sdiag[t] - (double)a.c1i[q] * a.kap[q];
        const double fi = (double)a.c1i[q] * a.sdiag[t] + (double)a.c1r[q] * a.kap[q];
        qr += fr * bv.x - fi * bv.y;
        qi += fr * bv.y + fi * bv.x;
    }
}

// Incident triangles per batch of the vertex gather over triangle values
// (tri_epilogue_kernel): a vertex of a closed triangle mesh has about six.
constexpr int EPI_TRI_VAL = 8;

// Vertex v of the boundary epilogue, columns lane, lane + lpv, ...:
// C[v][k] = alpha * ( sum_{t ni v} a_t/3 (sum_s part_s[t][k] * unscale + D_t b[t][k])
//                     + c0_q (M B)[v][k] ) + beta * C[v][k]
// (AvOpTransp.Dot ApplyBoundaryOperators.hpp:136-140, ApplySingleLayerDiagonal.hpp:36-50,
//  MassOp.Dot :150-175).
template <int BT, int CT>
__device__ __forceinline__ void epi_vertex(const EpiArgs& a, int v, int lane, int lpv)
{
    using TB = typename Elem<BT>::type;
    using TC = typename Elem<CT>::type;
    for (int k = lane; k < a.W; k += lpv) {
        const int q = k / a.wcs, kk = k % a.wcs;
        const int pc = q * a.WCSP + kk;
        double sr = 0.0, si = 0.0;
        if (a.kernel_on && a.tsum) {
            // the triangles' values, formed once per triangle beforehand; all
            // loads of a batch are issued before the first is added, and the
            // sums run in the order of the loop below
            const int e0 = a.vt_ptr[v], e1 = a.vt_ptr[v + 1];
            for (int e = e0; e < e1; e += EPI_TRI_VAL) {
                int t[EPI_TRI_VAL];
                double2 tv[EPI_TRI_VAL];
                float ar[EPI_TRI_VAL];
#pragma unroll
                for (int u = 0; u < EPI_TRI_VAL; ++u) t[u] = a.vt_idx[min(e + u, e1 - 1)];
#pragma unroll
                for (int u = 0; u < EPI_TRI_VAL; ++u) {
                    tv[u] = a.tsum[(int64_t)t[u] * a.ldb + pc];
                    ar[u] = a.area[t[u]];
                }
#pragma unroll
                for (int u = 0; u < EPI_TRI_VAL; ++u) {
                    if (e + u >= e1) break;
                    const double w = (double)ar[u] / 3.0;
                    sr += w * tv[u].x;
                    si += w * tv[u].y;
                }
            }
        } else if (a.kernel_on) {
            const int e0 = a.vt_ptr[v], e1 = a.vt_ptr[v + 1];
            // EPI_TRI triangles and EPI_SEG segments per batch: every load of a
            // batch is issued before the first is added (indices clamped to the
            // valid range, so there is no branch between them), and the sums run
            // in the order of the plain loops (s ascending per triangle, then the
            // triangles in list order)
            for (int e = e0; e < e1; e += EPI_TRI) {
                int t[EPI_TRI];
                double pr[EPI_TRI], pi[EPI_TRI];
#pragma unroll
                for (int u = 0; u < EPI_TRI; ++u) {
                    t[u] = a.vt_idx[min(e + u, e1 - 1)];
                    pr[u] = pi[u] = 0.0;
                }
                for (int s0 = 0; s0 < a.S; s0 += EPI_SEG) {
                    float xr[EPI_TRI][EPI_SEG], xi[EPI_TRI][EPI_SEG];
#pragma unroll
                    for (int u = 0; u < EPI_TRI; ++u) {
                        const float* pp = a.part + (int64_t)t[u] * a.ldp + q * 2 * a.WCSP + kk;
#pragma unroll
                        for (int j = 0; j < EPI_SEG; ++j) {
                            const int64_t o = (int64_t)min(s0 + j, a.S - 1) * a.seg_stride;
                            xr[u][j] = pp[o];
                            xi[u][j] = pp[o + a.WCSP];
                        }
                    }
#pragma unroll
                    for (int u = 0; u < EPI_TRI; ++u)
#pragma unroll
                        for (int j = 0; j < EPI_SEG; ++j)
                            if (s0 + j < a.S) {
                                pr[u] += (double)xr[u][j];
                                pi[u] += (double)xi[u][j];
                            }
                }
#pragma unroll
                for (int u = 0; u < EPI_TRI; ++u) {
                    if (e + u >= e1) break;
                    double qr, qi;
                    epi_tri_value(a, t[u], q, pc, pr[u], pi[u], qr, qi);
                    const double w = (double)a.area[t[u]] / 3.0;
                    sr += w * qr;
                    si += w * qi;
                }
            }
        }
        if (a.mass_on) {
            const TB* B = static_cast<const TB*>(a.B);
            double mr = 0.0, mi = 0.0;
            const int e0 = a.M_ptr[v], e1 = a.M_ptr[v + 1];
            for (int e = e0; e < e1; e += EPI_SEG) {
                double mv[EPI_SEG];
                double2 bv[EPI_SEG];
#pragma unroll
                for (int j = 0; j < EPI_SEG; ++j) {
                    const int ej = min(e + j, e1 - 1);
                    mv[j] = a.M_val[ej];
                    bv[j] = to_d2(B[(int64_t)a.M_idx[ej] * a.ldB + k]);
                }
#pragma unroll
                for (int j = 0; j < EPI_SEG; ++j)
                    if (e + j < e1) {
                        mr += mv[j] * bv[j].x;
                        mi += mv[j] * bv[j].y;
                    }
            }
            sr += a.c0r[q] * mr - a.c0i[q] * mi;
            si += a.c0r[q] * mi + a.c0i[q] * mr;
        }
        TC* C = static_cast<TC*>(a.C);
        TC& dst = C[(int64_t)v * a.ldC + k];
        double outr = a.ar * sr - a.ai * si;
        double outi = a.ar * si + a.ai * sr;
        if (a.br != 0.0 || a.bi != 0.0) {
            const double2 c = to_d2(dst);
            outr += a.br * c.x - a.bi * c.y;
            outi += a.br * c.y + a.bi * c.x;
        }
        from_d2(dst, make_double2(outr, outi));
    }
}

}  // namespace baemm
