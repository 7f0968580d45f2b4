// tangent_point.hip -- the caller-side energy and metric of the Gauss-Newton
// example (Example_GaussNewton/main.cpp:301-346 takes them from Repulsor's
// TangentPointEnergy0 / TangentPointMetric0, which are not part of BAEMM),
// restated from the published discretisation (C. Yu, C. Brakensiek,
// H. Schumacher, K. Crane, "Repulsive Surfaces", ACM Trans. Graph. 40(6),
// 2021): centroid quadrature over all pairs of distinct triangles, P1 vertex
// functions.
//
//   energy  E = sum_{S != T} a_S a_T |n_S . (c_S - c_T)|^q |c_S - c_T|^-p
//   metric  <<u, v>> = sum_{S != T} w_ST <du_S - du_T, dv_S - dv_T>_F
//                    + sum_{S != T} w_ST k_ST <ubar_S - ubar_T, vbar_S - vbar_T>
//           w_ST = a_S a_T |c_S - c_T|^(-2 sigma - 2), sigma = (p - 2)/q - 1,
//           k_ST = (n_S . (c_S - c_T))^2 / |c_S - c_T|^4
//   (du_S the face gradient of a P1 vector function, ubar_S its face average).
//
// Like the boundary operators these are brute-force pair sums: one target
// triangle per thread, sources staged through LDS, the source range split
// into segments whose FP32 partials are summed in FP64 (deterministic).
#include <algorithm>
#include <cmath>
#include <vector>

#include "capi_util.hpp"
#include "solvers.hpp"

namespace baemm {

namespace {

constexpr int kTpSrc = 64;     // sources per LDS stage
constexpr int kTpMetric = 14;  // metric accumulators per target: W, G (9), Wt, H (3)
constexpr int kTpEnergy = 12;  // energy accumulators per target (see tp_energy_pair_kernel)

// Segments of a pair launch over nt targets: about 2048 work-groups.
int tp_segments(int64_t nt)
{
    const int64_t tiles = (nt + 255) / 256;
    const int64_t by_blocks = (2048 + tiles - 1) / tiles;
    const int64_t by_sources = (nt + kTpSrc - 1) / kTpSrc;
    return (int)std::max<int64_t>(1, std::min(by_blocks, by_sources));
}

// Metric pair sums of target S over the sources [j0, j1) of its segment:
// W = sum w_ST, G = sum w_ST du_T, Wt = sum wt_ST, H = sum wt_ST ubar_T with
// wt_ST = w_ST (k_ST + k_TS) (the metric's low-order form symmetrised).
// face: nt x 12 (du row-major component x space, then ubar).
__global__ __launch_bounds__(256) void tp_metric_pair_kernel(int nt, const float4* __restrict__ mid,
                                                             const float4* __restrict__ nrm,
                                                             const float* __restrict__ area,
                                                             const float* __restrict__ face, float expo, int seg_len,
                                                             float* __restrict__ part)
{
    __shared__ float4 s_c[kTpSrc], s_n[kTpSrc];
    __shared__ float s_a[kTpSrc];
    __shared__ float s_f[kTpSrc * 12];
    const int S = blockIdx.x * 256 + threadIdx.x;
    const int j0 = blockIdx.y * seg_len, j1 = min(j0 + seg_len, nt);
    const int Sc = min(S, nt - 1);
    const float4 cs = mid[Sc], ns = nrm[Sc];
    const float as = area[Sc];
    float acc[kTpMetric];
#pragma unroll
    for (int k = 0; k < kTpMetric; ++k) acc[k] = 0.0f;
    for (int jb = j0; jb < j1; jb += kTpSrc) {
        __syncthreads();
        const int jn = min(kTpSrc, j1 - jb);
        if (threadIdx.x < jn) {
            s_c[threadIdx.x] = mid[jb + threadIdx.x];
            s_n[threadIdx.x] = nrm[jb + threadIdx.x];
            s_a[threadIdx.x] = area[jb + threadIdx.x];
        }
        for (int e = threadIdx.x; e < jn * 12; e += 256) s_f[e] = face[(int64_t)jb * 12 + e];
        __syncthreads();
        for (int jj = 0; jj < jn; ++jj) {
            if (jb + jj == S) continue;
            const float4 ct = s_c[jj], nt4 = s_n[jj];
            const float z0 = ct.x - cs.x, z1 = ct.y - cs.y, z2 = ct.z - cs.z;
            const float r2 = fmaf(z0, z0, fmaf(z1, z1, z2 * z2));
            const float w = as * s_a[jj] * exp2f(expo * log2f(r2));
            const float rinv2 = 1.0f / r2;
            const float ds = fmaf(ns.x, z0, fmaf(ns.y, z1, ns.z * z2));
            const float dt = fmaf(nt4.x, z0, fmaf(nt4.y, z1, nt4.z * z2));
            const float wt = w * (ds * ds + dt * dt) * (rinv2 * rinv2);
            const float* f = s_f + jj * 12;
            acc[0] += w;
#pragma unroll
            for (int k = 0; k < 9; ++k) acc[1 + k] = fmaf(w, f[k], acc[1 + k]);
            acc[10] += wt;
#pragma unroll
            for (int k = 0; k < 3; ++k) acc[11 + k] = fmaf(wt, f[9 + k], acc[11 + k]);
        }
    }
    if (S < nt)
#pragma unroll
        for (int k = 0; k < kTpMetric; ++k) part[((int64_t)blockIdx.y * nt + S) * kTpMetric + k] = acc[k];
}

// Per target S, the segment sums in FP64 and the face terms of the metric:
// out = (g (9): 2 (W du_S - G), h (3): Wt ubar_S - H, W, Wt).
__global__ void tp_metric_face_kernel(int nt, int nseg, const float* __restrict__ part,
                                      const float* __restrict__ face, double* __restrict__ out)
{
    const int S = blockIdx.x * blockDim.x + threadIdx.x;
    if (S >= nt) return;
    double a[kTpMetric];
    for (int k = 0; k < kTpMetric; ++k) a[k] = 0.0;
    for (int g = 0; g < nseg; ++g)
        for (int k = 0; k < kTpMetric; ++k) a[k] += (double)part[((int64_t)g * nt + S) * kTpMetric + k];
    const float* f = face + (int64_t)S * 12;
    double* o = out + (int64_t)S * kTpMetric;
    for (int k = 0; k < 9; ++k) o[k] = 2.0 * (a[0] * (double)f[k] - a[1 + k]);
    for (int k = 0; k < 3; ++k) o[9 + k] = a[10] * (double)f[9 + k] - a[11 + k];
    o[12] = a[0];
    o[13] = a[10];
}

// Face inputs of the metric from a P1 vector function x (V x 3, FP64):
// du (component c, space k) = sum_j x[v_j][c] grad_j[k], ubar = mean of x[v_j].
__global__ void tp_face_input_kernel(int nt, const int4* __restrict__ tris, const float* __restrict__ grad,
                                     const double* __restrict__ x, float* __restrict__ face)
{
    const int S = blockIdx.x * blockDim.x + threadIdx.x;
    if (S >= nt) return;
    const int4 t = tris[S];
    const int v[3] = {t.x, t.y, t.z};
    const float* g = grad + (int64_t)S * 9;
    float* f = face + (int64_t)S * 12;
    for (int c = 0; c < 3; ++c) {
        double du[3] = {0.0, 0.0, 0.0}, ub = 0.0;
        for (int j = 0; j < 3; ++j) {
            const double xv = x[(int64_t)v[j] * 3 + c];
            for (int k = 0; k < 3; ++k) du[k] += xv * (double)g[j * 3 + k];
            ub += xv;
        }
        for (int k = 0; k < 3; ++k) f[c * 3 + k] = (float)du[k];
        f[9 + c] = (float)(ub / 3.0);
    }
}

// y[v][c] = scale sum_{S ni v} (g_S[c] . grad_{S, v} + h_S[c] / 3): the metric
// applied to x, gathered over the triangles around each vertex.
__global__ void tp_vertex_kernel(int nv, const int32_t* __restrict__ vt_ptr, const int32_t* __restrict__ vt_idx,
                                 const int4* __restrict__ tris, const float* __restrict__ grad,
                                 const double* __restrict__ fo, double scale, double* __restrict__ y)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int e = vt_ptr[v]; e < vt_ptr[v + 1]; ++e) {
        const int S = vt_idx[e];
        const int4 t = tris[S];
        const int k = (t.x == v) ? 0 : ((t.y == v) ? 1 : 2);
        const float* g = grad + (int64_t)S * 9 + k * 3;
        const double* o = fo + (int64_t)S * kTpMetric;
        for (int c = 0; c < 3; ++c)
            acc[c] += o[c * 3] * (double)g[0] + o[c * 3 + 1] * (double)g[1] + o[c * 3 + 2] * (double)g[2] +
                      o[9 + c] / 3.0;
    }
    for (int c = 0; c < 3; ++c) y[(int64_t)v * 3 + c] = scale * acc[c];
}

// The diagonal of the metric's matrix (equal for the three components of a
// vertex): d_v = sum_{S ni v} (2 W_S |grad_(S,v)|^2 + Wt_S / 9)
//              - sum_{S != T both at v} (2 w_ST grad_(S,v) . grad_(T,v) + wt_ST / 9),
// the second sum over the ordered pairs of triangles around v (FP64).
// fo: the face terms (W_S at 12, Wt_S at 13 of each row of 14).
__global__ void tp_diag_kernel(int nv, const int32_t* __restrict__ vt_ptr, const int32_t* __restrict__ vt_idx,
                               const int4* __restrict__ tris, const float* __restrict__ grad,
                               const float4* __restrict__ mid, const float4* __restrict__ nrm,
                               const float* __restrict__ area, const double* __restrict__ fo, double expo,
                               double* __restrict__ diag)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    double d = 0.0;
    for (int e = vt_ptr[v]; e < vt_ptr[v + 1]; ++e) {
        const int S = vt_idx[e];
        const int4 t = tris[S];
        const int k = (t.x == v) ? 0 : ((t.y == v) ? 1 : 2);
        const float* g = grad + (int64_t)S * 9 + k * 3;
        const double gg = (double)g[0] * g[0] + (double)g[1] * g[1] + (double)g[2] * g[2];
        d += 2.0 * fo[(int64_t)S * kTpMetric + 12] * gg + fo[(int64_t)S * kTpMetric + 13] / 9.0;
        const float4 cs = mid[S], ns = nrm[S];
        for (int e2 = vt_ptr[v]; e2 < vt_ptr[v + 1]; ++e2) {
            const int T = vt_idx[e2];
            if (T == S) continue;
            const int4 u = tris[T];
            const int kt = (u.x == v) ? 0 : ((u.y == v) ? 1 : 2);
            const float* h = grad + (int64_t)T * 9 + kt * 3;
            const float4 ct = mid[T], nt4 = nrm[T];
            const double z0 = (double)ct.x - cs.x, z1 = (double)ct.y - cs.y, z2 = (double)ct.z - cs.z;
            const double r2 = z0 * z0 + z1 * z1 + z2 * z2;
            const double w = (double)area[S] * area[T] * pow(r2, expo);
            const double ds = ns.x * z0 + ns.y * z1 + ns.z * z2, dt = nt4.x * z0 + nt4.y * z1 + nt4.z * z2;
            const double wt = w * (ds * ds + dt * dt) / (r2 * r2);
            const double gh = (double)g[0] * h[0] + (double)g[1] * h[1] + (double)g[2] * h[2];
            d -= 2.0 * w * gh + wt / 9.0;
        }
    }
    diag[v] = d;
}

// y = x / (scale d_v) row by row (0 where d_v = 0: a vertex of no triangle).
__global__ void tp_jacobi_kernel(int nv, const double* __restrict__ diag, double scale, const double* __restrict__ x,
                                 double* __restrict__ y)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const double d = scale * diag[v];
    for (int c = 0; c < 3; ++c) y[(int64_t)v * 3 + c] = d != 0.0 ? x[(int64_t)v * 3 + c] / d : 0.0;
}

// Energy pair sums of target F over the sources G of its segment (G != F),
// with d = c_F - c_G, R = |d|^-p, s1 = n_F . d, s2 = n_G . (c_G - c_F) and
// P = |s|^q, Pm = q |s|^(q-1) sgn s:
//   [0] sum a_G P1 R                         (the energy of F, times a_F)
//   [1] sum a_G R ((1 - q) P1 + P2)          (dE/dN_F along n_F)
//   [2..4] sum a_G R Pm1 d                   (dE/dN_F)
//   [5] sum a_G Pm1 R                        (dE/dc_F along n_F, times a_F)
//   [6..8] sum a_G p P1 R d / |d|^2          (dE/dc_F, times -a_F)
//   [9..11] sum a_G (Pm2 R n_G + p P2 R d / |d|^2)   (dE/dc_F, times -a_F)
__global__ __launch_bounds__(256) void tp_energy_pair_kernel(int nt, const float4* __restrict__ mid,
                                                             const float4* __restrict__ nrm,
                                                             const float* __restrict__ area, float q, float p,
                                                             int seg_len, float* __restrict__ part)
{
    __shared__ float4 s_c[kTpSrc], s_n[kTpSrc];
    __shared__ float s_a[kTpSrc];
    const int F = blockIdx.x * 256 + threadIdx.x;
    const int j0 = blockIdx.y * seg_len, j1 = min(j0 + seg_len, nt);
    const int Fc = min(F, nt - 1);
    const float4 cf = mid[Fc], nf = nrm[Fc];
    float acc[kTpEnergy];
#pragma unroll
    for (int k = 0; k < kTpEnergy; ++k) acc[k] = 0.0f;
    for (int jb = j0; jb < j1; jb += kTpSrc) {
        __syncthreads();
        const int jn = min(kTpSrc, j1 - jb);
        if (threadIdx.x < jn) {
            s_c[threadIdx.x] = mid[jb + threadIdx.x];
            s_n[threadIdx.x] = nrm[jb + threadIdx.x];
            s_a[threadIdx.x] = area[jb + threadIdx.x];
        }
        __syncthreads();
        for (int jj = 0; jj < jn; ++jj) {
            if (jb + jj == F) continue;
            const float4 cg = s_c[jj], ng = s_n[jj];
            const float ag = s_a[jj];
            const float d0 = cf.x - cg.x, d1 = cf.y - cg.y, d2 = cf.z - cg.z;
            const float r2 = fmaf(d0, d0, fmaf(d1, d1, d2 * d2));
            const float R = exp2f(-0.5f * p * log2f(r2));
            const float Rr = R / r2;
            const float s1 = fmaf(nf.x, d0, fmaf(nf.y, d1, nf.z * d2));
            const float s2 = -fmaf(ng.x, d0, fmaf(ng.y, d1, ng.z * d2));
            const float l1 = log2f(fabsf(s1)), l2 = log2f(fabsf(s2));
            const float P1 = exp2f(q * l1), P2 = exp2f(q * l2);
            const float Pm1 = q * copysignf(exp2f((q - 1.0f) * l1), s1);
            const float Pm2 = q * copysignf(exp2f((q - 1.0f) * l2), s2);
            acc[0] = fmaf(ag * P1, R, acc[0]);
            acc[1] = fmaf(ag * R, fmaf(1.0f - q, P1, P2), acc[1]);
            const float b = ag * R * Pm1;
            acc[2] = fmaf(b, d0, acc[2]);
            acc[3] = fmaf(b, d1, acc[3]);
            acc[4] = fmaf(b, d2, acc[4]);
            acc[5] = fmaf(ag * Pm1, R, acc[5]);
            const float dd = ag * p * P1 * Rr;
            acc[6] = fmaf(dd, d0, acc[6]);
            acc[7] = fmaf(dd, d1, acc[7]);
            acc[8] = fmaf(dd, d2, acc[8]);
            const float e1 = ag * Pm2 * R, e2 = ag * p * P2 * Rr;
            acc[9] = fmaf(e1, ng.x, fmaf(e2, d0, acc[9]));
            acc[10] = fmaf(e1, ng.y, fmaf(e2, d1, acc[10]));
            acc[11] = fmaf(e1, ng.z, fmaf(e2, d2, acc[11]));
        }
    }
    if (F < nt)
#pragma unroll
        for (int k = 0; k < kTpEnergy; ++k) part[((int64_t)blockIdx.y * nt + F) * kTpEnergy + k] = acc[k];
}

__global__ void tp_sum_segments_kernel(int nt, int nseg, int width, const float* __restrict__ part,
                                       double* __restrict__ out)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)nt * width) return;
    double a = 0.0;
    for (int g = 0; g < nseg; ++g) a += (double)part[(int64_t)g * nt * width + e];
    out[e] = a;
}

inline dim3 grid_of(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

// Hat-function gradients of every triangle (FP64 from the input coordinates,
// rounded to FP32), uploaded once per handle.
const float* tp_gradients(Handle& h, hipStream_t s)
{
    const size_t bytes = sizeof(float) * 9 * h.nt;
    if (h.ws_tp_grad.bytes >= bytes) return h.ws_tp_grad.as<float>();
    std::vector<float> g(9 * h.nt);
    for (int64_t t = 0; t < h.nt; ++t) {
        double x[3][3];
        for (int k = 0; k < 3; ++k)
            for (int d = 0; d < 3; ++d) x[k][d] = h.coords[3 * h.tris[3 * t + k] + d];
        double e1[3], e2[3], N[3];
        for (int d = 0; d < 3; ++d) {
            e1[d] = x[1][d] - x[0][d];
            e2[d] = x[2][d] - x[0][d];
        }
        N[0] = e1[1] * e2[2] - e1[2] * e2[1];
        N[1] = e1[2] * e2[0] - e1[0] * e2[2];
        N[2] = e1[0] * e2[1] - e1[1] * e2[0];
        const double nn = N[0] * N[0] + N[1] * N[1] + N[2] * N[2];   // (2 a)^2
        // grad lambda_k = N x (x_(k+2) - x_(k+1)) / |N|^2
        for (int k = 0; k < 3; ++k) {
            const int a = (k + 1) % 3, b = (k + 2) % 3;
            double e[3];
            for (int d = 0; d < 3; ++d) e[d] = x[b][d] - x[a][d];
            g[9 * t + 3 * k + 0] = (float)((N[1] * e[2] - N[2] * e[1]) / nn);
            g[9 * t + 3 * k + 1] = (float)((N[2] * e[0] - N[0] * e[2]) / nn);
            g[9 * t + 3 * k + 2] = (float)((N[0] * e[1] - N[1] * e[0]) / nn);
        }
    }
    h.ws_tp_grad.reserve(bytes);
    BAEMM_HIP(hipMemcpyAsync(h.ws_tp_grad.ptr, g.data(), bytes, hipMemcpyHostToDevice, s));
    BAEMM_HIP(hipStreamSynchronize(s));   // g is released on return
    return h.ws_tp_grad.as<float>();
}

void check_qp(double q, double p)
{
    if (!(q > 1.0) || !(p > 0.0) || !std::isfinite(q) || !std::isfinite(p))
        throw ArgError{"tangent-point exponents: q > 1 and p > 0 required"};
}

// The metric's face terms (ws_tp_out, nt x 14) for the P1 function x (device,
// V x 3) and, when (q, p) is new, the Jacobi weights.
void tp_metric_faces(Handle& h, double q, double p, const double* x, hipStream_t s)
{
    check_qp(q, p);
    const int nt = (int)h.nt;
    const float* grad = tp_gradients(h, s);
    h.ws_tp_face.reserve(sizeof(float) * 12 * nt);
    float* face = h.ws_tp_face.as<float>();
    hipLaunchKernelGGL(tp_face_input_kernel, grid_of(nt), dim3(256), 0, s, nt, h.d_tris, grad, x, face);
    const int S = tp_segments(nt);
    const int seg_len = (nt + S - 1) / S;
    h.ws_tp_part.reserve(sizeof(float) * (size_t)S * nt * kTpMetric);
    h.ws_tp_out.reserve(sizeof(double) * (size_t)nt * kTpMetric);
    const double sigma = (p - 2.0) / q - 1.0;
    hipLaunchKernelGGL(tp_metric_pair_kernel, dim3((unsigned)((nt + 255) / 256), (unsigned)S), dim3(256), 0, s, nt,
                       h.d_mid, h.d_nrm, h.d_area, face, (float)(-(sigma + 1.0)), seg_len, h.ws_tp_part.as<float>());
    hipLaunchKernelGGL(tp_metric_face_kernel, grid_of(nt), dim3(256), 0, s, nt, S, h.ws_tp_part.as<float>(), face,
                       h.ws_tp_out.as<double>());
    BAEMM_HIP(hipGetLastError());
}

// The diagonal of the metric for (q, p), kept until (q, p) changes; x is any
// device P1 function (the pair weights do not depend on it).
const double* tp_diagonal(Handle& h, double q, double p, const double* x, hipStream_t s)
{
    if (h.tp_diag_q == q && h.tp_diag_p == p) return h.ws_tp_diag.as<double>();
    tp_metric_faces(h, q, p, x, s);
    h.ws_tp_diag.reserve(sizeof(double) * h.nv);
    const double sigma = (p - 2.0) / q - 1.0;
    hipLaunchKernelGGL(tp_diag_kernel, grid_of(h.nv), dim3(256), 0, s, (int)h.nv, h.d_vt_ptr, h.d_vt_idx, h.d_tris,
                       h.ws_tp_grad.as<float>(), h.d_mid, h.d_nrm, h.d_area, h.ws_tp_out.as<double>(),
                       -(sigma + 1.0), h.ws_tp_diag.as<double>());
    BAEMM_HIP(hipGetLastError());
    h.tp_diag_q = q;
    h.tp_diag_p = p;
    return h.ws_tp_diag.as<double>();
}

void tp_metric_dev(Handle& h, double q, double p, double scale, const double* x, double* y, hipStream_t s)
{
    tp_metric_faces(h, q, p, x, s);
    hipLaunchKernelGGL(tp_vertex_kernel, grid_of(h.nv), dim3(256), 0, s, (int)h.nv, h.d_vt_ptr, h.d_vt_idx, h.d_tris,
                       h.ws_tp_grad.as<float>(), h.ws_tp_out.as<double>(), scale, y);
    BAEMM_HIP(hipGetLastError());
}

void tp_jacobi_dev(Handle& h, double q, double p, double scale, const double* x, double* y, hipStream_t s)
{
    const double* diag = tp_diagonal(h, q, p, x, s);
    hipLaunchKernelGGL(tp_jacobi_kernel, grid_of(h.nv), dim3(256), 0, s, (int)h.nv, diag, scale, x, y);
    BAEMM_HIP(hipGetLastError());
}

// A callback cannot return a status: on failure it leaves a message in the
// handle's last error and fills y with NaN (all bits set), so that the solve
// calling it fails instead of going on with stale values.
void tp_callback_failed(Handle& h, double* y, hipStream_t s)
{
    h.last_error = "tangent-point metric callback failed (q > 1 and p > 0 required; "
                   "baemm_tp_metric_prepare reports the cause)";
    (void)hipMemsetAsync(y, 0xff, sizeof(double) * 3 * h.nv, s);
}

}  // namespace
}  // namespace baemm

using namespace baemm;

extern "C" {

int baemm_tp_energy(baemm_handle h, double q, double p, double* energy, double* gradient)
{
    return guard(h, [&](Handle& hh) {
        check_qp(q, p);
        if (!energy && !gradient) throw ArgError{"energy and gradient are both NULL"};
        hipStream_t s = hh.stream;
        const int nt = (int)hh.nt;
        const int S = tp_segments(nt);
        const int seg_len = (nt + S - 1) / S;
        hh.ws_tp_part.reserve(sizeof(float) * (size_t)S * nt * kTpEnergy);
        hh.ws_tp_out.reserve(sizeof(double) * (size_t)nt * kTpEnergy);
        hipLaunchKernelGGL(tp_energy_pair_kernel, dim3((unsigned)((nt + 255) / 256), (unsigned)S), dim3(256), 0, s,
                           nt, hh.d_mid, hh.d_nrm, hh.d_area, (float)q, (float)p, seg_len,
                           hh.ws_tp_part.as<float>());
        hipLaunchKernelGGL(tp_sum_segments_kernel, grid_of((int64_t)nt * kTpEnergy), dim3(256), 0, s, nt, S, kTpEnergy,
                           hh.ws_tp_part.as<float>(), hh.ws_tp_out.as<double>());
        BAEMM_HIP(hipGetLastError());
        std::vector<double> fo((size_t)nt * kTpEnergy);
        BAEMM_HIP(hipMemcpyAsync(fo.data(), hh.ws_tp_out.ptr, sizeof(double) * fo.size(), hipMemcpyDeviceToHost, s));
        BAEMM_HIP(hipStreamSynchronize(s));
        // per triangle: dE/dN and dE/dc, then the chain rule to the vertices
        // (c = (x0 + x1 + x2) / 3, N = (x1 - x0) x (x2 - x0) / 2, so that
        // d(g . N)/dx_k = (x_(k+1) - x_(k+2)) x g / 2), in FP64 on the host
        double E = 0.0;
        if (gradient) std::fill(gradient, gradient + 3 * hh.nv, 0.0);
        for (int F = 0; F < nt; ++F) {
            const double* o = fo.data() + (size_t)F * kTpEnergy;
            double x[3][3];
            for (int k = 0; k < 3; ++k)
                for (int d = 0; d < 3; ++d) x[k][d] = hh.coords[3 * hh.tris[3 * F + k] + d];
            const double e1[3] = {x[1][0] - x[0][0], x[1][1] - x[0][1], x[1][2] - x[0][2]};
            const double e2[3] = {x[2][0] - x[0][0], x[2][1] - x[0][1], x[2][2] - x[0][2]};
            const double N[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                                 e1[0] * e2[1] - e1[1] * e2[0]};
            const double a2 = std::sqrt(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]);   // 2 a
            const double a = 0.5 * a2;
            const double n[3] = {N[0] / a2, N[1] / a2, N[2] / a2};
            E += a * o[0];
            if (!gradient) continue;
            double gN[3], gc[3];
            for (int d = 0; d < 3; ++d) {
                gN[d] = o[1] * n[d] + o[2 + d];
                gc[d] = a * (o[5] * n[d] - o[6 + d]) - a * o[9 + d];
            }
            for (int k = 0; k < 3; ++k) {
                const int i = (k + 1) % 3, j = (k + 2) % 3;
                const double e[3] = {x[i][0] - x[j][0], x[i][1] - x[j][1], x[i][2] - x[j][2]};
                const double cr[3] = {e[1] * gN[2] - e[2] * gN[1], e[2] * gN[0] - e[0] * gN[2], e[0] * gN[1] - e[1] * gN[0]};
                double* G = gradient + 3 * (int64_t)hh.tris[3 * F + k];
                for (int d = 0; d < 3; ++d) G[d] += gc[d] / 3.0 + 0.5 * cr[d];
            }
        }
        if (energy) *energy = E;
        return BAEMM_OK;
    });
}

int baemm_tp_metric_prepare(baemm_handle h, double q, double p)
{
    return guard(h, [&](Handle& hh) {
        check_qp(q, p);
        hipStream_t s = hh.stream;
        const size_t bytes = sizeof(double) * 3 * hh.nv;
        hh.io[4].reserve(bytes);
        BAEMM_HIP(hipMemsetAsync(hh.io[4].ptr, 0, bytes, s));
        tp_diagonal(hh, q, p, hh.io[4].as<double>(), s);
        BAEMM_HIP(hipStreamSynchronize(s));
        return BAEMM_OK;
    });
}

void baemm_tp_metric_apply(void* user, const double* x, double* y, void* stream)
{
    const auto* m = static_cast<const baemm_tp_metric*>(user);
    if (!m || !m->h || !x || !y) return;
    Handle& hh = *H(m->h);
    hipStream_t s = stream_of(hh, stream);
    try {
        tp_metric_dev(hh, m->q, m->p, m->scale, x, y, s);
    } catch (...) {
        tp_callback_failed(hh, y, s);
    }
}

void baemm_tp_metric_jacobi(void* user, const double* x, double* y, void* stream)
{
    const auto* m = static_cast<const baemm_tp_metric*>(user);
    if (!m || !m->h || !x || !y) return;
    Handle& hh = *H(m->h);
    hipStream_t s = stream_of(hh, stream);
    try {
        tp_jacobi_dev(hh, m->q, m->p, m->scale, x, y, s);
    } catch (...) {
        tp_callback_failed(hh, y, s);
    }
}

// The metric (jacobi false) or its Jacobi preconditioner (true) on host arrays.
static int tp_host(baemm_handle h, double q, double p, double scale, const double* x, double* y, bool jacobi)
{
    return guard(h, [&](Handle& hh) {
        if (!x || !y) throw ArgError{"x and y must not be NULL"};
        if (jacobi && !(scale != 0.0 && std::isfinite(scale))) throw ArgError{"scale must be finite and nonzero"};
        hipStream_t s = hh.stream;
        const size_t bytes = sizeof(double) * 3 * hh.nv;
        hh.io[4].reserve(bytes);
        hh.io[5].reserve(bytes);
        BAEMM_HIP(hipMemcpyAsync(hh.io[4].ptr, x, bytes, hipMemcpyHostToDevice, s));
        if (jacobi)
            tp_jacobi_dev(hh, q, p, scale, hh.io[4].as<double>(), hh.io[5].as<double>(), s);
        else
            tp_metric_dev(hh, q, p, scale, hh.io[4].as<double>(), hh.io[5].as<double>(), s);
        BAEMM_HIP(hipMemcpyAsync(y, hh.io[5].ptr, bytes, hipMemcpyDeviceToHost, s));
        BAEMM_HIP(hipStreamSynchronize(s));
        return BAEMM_OK;
    });
}

int baemm_tp_metric_apply_host(baemm_handle h, double q, double p, double scale, const double* x, double* y)
{
    return tp_host(h, q, p, scale, x, y, false);
}

int baemm_tp_metric_jacobi_host(baemm_handle h, double q, double p, double scale, const double* x, double* y)
{
    return tp_host(h, q, p, scale, x, y, true);
}

}  // extern "C"
