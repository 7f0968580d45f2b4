// solvers.hip -- device kernels of the Krylov solvers and the small
// per-vertex / per-triangle operators used by the solver compositions.
//
// All solver vectors are FP64, row-major (vertex-major, column-minor) with a
// leading dimension.  The mass-inverse PCG works on real columns: a complex
// column is two real ones (its real and imaginary parts), each solved as its
// own real CG with its own step lengths and stopping test.  The reference runs
// one complex CG per complex column whose operator and preconditioner view the
// column as two real ones (ApplyMassInverse.hpp:56-104), so its real and
// imaginary parts share the step lengths; both converge to M^{-1} X and differ
// only below the CG tolerance.
// Column reductions are two-pass (block partials, then a per-column sum) and
// deterministic.
#include <algorithm>
#include <mutex>
#include <set>
#include <vector>

#include "solvers.hpp"

namespace baemm {

// ---------------------------------------------------------------- sparse

// Y = alpha M X + beta Y, real columns (MassOp.Dot).  Thread per (row, col).
// skip (may be NULL): the kernel does nothing once *skip is set (device-side
// loop exit of the PCG, whose kernel sequence is fixed so that it can be
// captured in a graph).
__global__ void mass_spmm_kernel(int nv, const int32_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                                 const double* __restrict__ val, const double* __restrict__ X, int64_t ldX,
                                 double* __restrict__ Y, int64_t ldY, int ncols, double alpha, double beta,
                                 const int* __restrict__ skip)
{
    if (skip && *skip) return;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nv * ncols) return;
    const int v = (int)(t / ncols), c = (int)(t % ncols);
    double s = 0.0;
    for (int e = ptr[v]; e < ptr[v + 1]; ++e) s += val[e] * X[(int64_t)idx[e] * ldX + c];
    double* y = Y + (int64_t)v * ldY + c;
    *y = alpha * s + (beta != 0.0 ? beta * *y : 0.0);
}

// Y = diag(d) X (ApplyLumpedMassInverse, ApplyMassInverse.hpp:178-194).
__global__ void diag_scale_kernel(int nv, const double* __restrict__ d, const double* __restrict__ X, int64_t ldX,
                                  double* __restrict__ Y, int64_t ldY, int ncols, const int* __restrict__ skip)
{
    if (skip && *skip) return;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nv * ncols) return;
    const int v = (int)(t / ncols), c = (int)(t % ncols);
    Y[(int64_t)v * ldY + c] = d[v] * X[(int64_t)v * ldX + c];
}

// P0 <- P1: Y[t][c] = (a_t/3) sum_k X[tri_k][c]  (AvOp.Dot on real columns).
__global__ void av_real_kernel(int nt, const int4* __restrict__ tris, const float* __restrict__ area,
                               const double* __restrict__ X, int64_t ldX, double* __restrict__ Y, int64_t ldY,
                               int ncols)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nt * ncols) return;
    const int tr = (int)(t / ncols), c = (int)(t % ncols);
    const int4 v = tris[tr];
    Y[(int64_t)tr * ldY + c] =
        (double)area[tr] / 3.0 * (X[(int64_t)v.x * ldX + c] + X[(int64_t)v.y * ldX + c] + X[(int64_t)v.z * ldX + c]);
}

// P1 <- P0: Y[v][c] = alpha sum_{t ni v} (a_t/3) X[t][c] + beta Y  (AvOpTransp.Dot).
__global__ void avt_real_kernel(int nv, const int32_t* __restrict__ vt_ptr, const int32_t* __restrict__ vt_idx,
                                const float* __restrict__ area, const double* __restrict__ X, int64_t ldX,
                                double* __restrict__ Y, int64_t ldY, int ncols, double alpha, double beta)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nv * ncols) return;
    const int v = (int)(t / ncols), c = (int)(t % ncols);
    double s = 0.0;
    for (int e = vt_ptr[v]; e < vt_ptr[v + 1]; ++e) {
        const int tr = vt_idx[e];
        s += (double)area[tr] / 3.0 * X[(int64_t)tr * ldX + c];
    }
    double* y = Y + (int64_t)v * ldY + c;
    *y = alpha * s + (beta != 0.0 ? beta * *y : 0.0);
}

// DotWithNormals_PL pointwise part (FarField.hpp:990-1005): y_t = (n_t . X_t) / a_t, X: nt x 3.
__global__ void dot_normals_kernel(int nt, const float4* __restrict__ nrm, const float* __restrict__ area,
                                   const double* __restrict__ X, double* __restrict__ y)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const float4 n = nrm[t];
    y[t] = ((double)n.x * X[3 * t] + (double)n.y * X[3 * t + 1] + (double)n.z * X[3 * t + 2]) / (double)area[t];
}

// MultiplyWithNormals_PL pointwise part (FarField.hpp:924-935): Y_t = Re(x_t) / a_t n_t.
__global__ void mul_normals_kernel(int nt, const float4* __restrict__ nrm, const float* __restrict__ area,
                                   const double2* __restrict__ x, double* __restrict__ Y)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const float4 n = nrm[t];
    const double mul = x[t].x / (double)area[t];
    Y[3 * t] = mul * n.x;
    Y[3 * t + 1] = mul * n.y;
    Y[3 * t + 2] = mul * n.z;
}

// bdr[v][k] = -s[v] du[v][k]  (Derivative_FF, FarField.hpp:288-300).
__global__ void scale_rows_neg_kernel(int nv, int W, const double* __restrict__ s, const double2* __restrict__ du,
                                      double2* __restrict__ out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nv * W) return;
    const int v = (int)(t / W);
    const double2 d = du[t];
    out[t] = make_double2(-s[v] * d.x, -s[v] * d.y);
}

// p[v] = sum_k a[v][k] b[v][k] (unconjugated; HadamardProduct with ld_sum, HadamardProduct.hpp:18-25).
__global__ void row_product_sum_kernel(int nv, int W, const double2* __restrict__ a, const double2* __restrict__ b,
                                       double2* __restrict__ p)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    double sr = 0.0, si = 0.0;
    for (int k = 0; k < W; ++k) {
        const double2 x = a[(int64_t)v * W + k], y = b[(int64_t)v * W + k];
        sr += x.x * y.x - x.y * y.y;
        si += x.x * y.y + x.y * y.x;
    }
    p[v] = make_double2(sr, si);
}

// ---------------------------------------------------------------- reductions

// part[b][c] = sum over this block's rows of X[r][c] * Y[r][c]  (real columns).
__global__ __launch_bounds__(256) void coldot_part_kernel(int n, int ncols, const double* __restrict__ X,
                                                          int64_t ldX, const double* __restrict__ Y, int64_t ldY,
                                                          double* __restrict__ part, const int* __restrict__ skip)
{
    if (skip && *skip) return;
    __shared__ double sm[4][64];
    const int c = threadIdx.x & 63, rl = threadIdx.x >> 6;
    for (int c0 = 0; c0 < ncols; c0 += 64) {
        double s = 0.0;
        if (c0 + c < ncols)
            for (int r = blockIdx.x * 4 + rl; r < n; r += gridDim.x * 4)
                s += X[(int64_t)r * ldX + c0 + c] * Y[(int64_t)r * ldY + c0 + c];
        sm[rl][c] = s;
        __syncthreads();
        if (rl == 0 && c0 + c < ncols) part[(int64_t)blockIdx.x * ncols + c0 + c] = sm[0][c] + sm[1][c] + sm[2][c] + sm[3][c];
        __syncthreads();
    }
}

// Sum over b < nblocks of part[b * stride + e], by one wave: lane-strided
// partial sums, then a butterfly (a fixed order, so results are reproducible).
// The loads are issued eight at a time and then added in order, so the wave
// waits for one memory round trip per 512 partials, not one per 64.
// Every lane returns the sum.
__device__ __forceinline__ double wave_colsum(const double* __restrict__ part, int nblocks, int64_t stride, int64_t e)
{
    const int lane = threadIdx.x & 63;
    double s = 0.0;
    int b = lane;
    for (; b + 7 * 64 < nblocks; b += 8 * 64) {
        double x[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) x[j] = part[(int64_t)(b + 64 * j) * stride + e];
#pragma unroll
        for (int j = 0; j < 8; ++j) s += x[j];
    }
    for (; b < nblocks; b += 64) s += part[(int64_t)b * stride + e];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    return s;
}

// out[c] = sum_b part[b][c]; one 64-thread block per column c
__global__ __launch_bounds__(64) void colsum_kernel(int nblocks, int ncols, const double* __restrict__ part,
                                                    double* __restrict__ out, const int* __restrict__ skip)
{
    if (skip && *skip) return;
    const int c = blockIdx.x;
    const double v = wave_colsum(part, nblocks, ncols, c);
    if (threadIdx.x == 0) out[c] = v;
}

// Complex Gram-Schmidt dots: part[b][i][k] = sum_rows conj(V_i[r][k]) w[r][k], i < nb.
// V: basis vectors with stride vstride (elements) between vectors, ld W.
// GS_BATCH basis vectors per pass over the rows: each row's w and the batch's
// V_i are loaded together before they are used (w once per batch instead of
// once per vector); every dot still sums its rows in the same order.
constexpr int GS_BATCH = 8;

__global__ __launch_bounds__(256) void gs_dots_part_kernel(int n, int W, int nb, const double2* __restrict__ V,
                                                           int64_t vstride, const double2* __restrict__ w,
                                                           double2* __restrict__ part)
{
    __shared__ double2 sm[4][64];
    const int k = threadIdx.x & 63, rl = threadIdx.x >> 6;
    for (int i0 = 0; i0 < nb; i0 += GS_BATCH) {
        for (int k0 = 0; k0 < W; k0 += 64) {
            double sr[GS_BATCH], si[GS_BATCH];
#pragma unroll
            for (int j = 0; j < GS_BATCH; ++j) sr[j] = si[j] = 0.0;
            if (k0 + k < W)
                for (int r = blockIdx.x * 4 + rl; r < n; r += gridDim.x * 4) {
                    const double2 b = w[(int64_t)r * W + k0 + k];
                    double2 a[GS_BATCH];
#pragma unroll
                    for (int j = 0; j < GS_BATCH; ++j)   // vectors past nb read the last one (not used)
                        a[j] = V[(int64_t)min(i0 + j, nb - 1) * vstride + (int64_t)r * W + k0 + k];
#pragma unroll
                    for (int j = 0; j < GS_BATCH; ++j) {
                        sr[j] += a[j].x * b.x + a[j].y * b.y;
                        si[j] += a[j].x * b.y - a[j].y * b.x;
                    }
                }
#pragma unroll
            for (int j = 0; j < GS_BATCH; ++j) {
                if (i0 + j >= nb) break;
                sm[rl][k] = make_double2(sr[j], si[j]);
                __syncthreads();
                if (rl == 0 && k0 + k < W) {
                    double2 s = sm[0][k];
                    for (int q = 1; q < 4; ++q) {
                        s.x += sm[q][k].x;
                        s.y += sm[q][k].y;
                    }
                    part[((int64_t)blockIdx.x * nb + i0 + j) * W + k0 + k] = s;
                }
                __syncthreads();
            }
        }
    }
}

// out[e] = sum_b part[b][e]; one 64-thread block per element e
__global__ __launch_bounds__(64) void gs_dots_sum_kernel(int nblocks, int nbW, const double2* __restrict__ part,
                                                         double2* __restrict__ out)
{
    const int e = blockIdx.x;
    const double* pd = reinterpret_cast<const double*>(part);
    const double sr = wave_colsum(pd, nblocks, 2 * (int64_t)nbW, 2 * (int64_t)e);
    const double si = wave_colsum(pd, nblocks, 2 * (int64_t)nbW, 2 * (int64_t)e + 1);
    if (threadIdx.x == 0) out[e] = make_double2(sr, si);
}

// w[r][k] -= sum_{i<nb} h[i][k] V_i[r][k]   (or x += ... with sign = +1)
// nbp (may be null): the number of basis vectors is min(nb, *nbp).
__global__ void gs_update_kernel(int n, int W, int nb, const double2* __restrict__ V, int64_t vstride,
                                 const double2* __restrict__ h, double2* __restrict__ w, double sign,
                                 const int* __restrict__ nbp)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * W) return;
    if (nbp) nb = min(nb, *nbp);
    const int k = (int)(t % W);
    double sr = 0.0, si = 0.0;
    for (int i0 = 0; i0 < nb; i0 += GS_BATCH) {
        double2 a[GS_BATCH], c[GS_BATCH];
#pragma unroll
        for (int j = 0; j < GS_BATCH; ++j) {
            const int i = min(i0 + j, nb - 1);
            a[j] = V[(int64_t)i * vstride + t];
            c[j] = h[i * W + k];
        }
#pragma unroll
        for (int j = 0; j < GS_BATCH; ++j)
            if (i0 + j < nb) {
                sr += c[j].x * a[j].x - c[j].y * a[j].y;
                si += c[j].x * a[j].y + c[j].y * a[j].x;
            }
    }
    double2 o = w[t];
    o.x += sign * sr;
    o.y += sign * si;
    w[t] = o;
}

// dst[r][k] = src[r][k] * s[k]  (complex vector, real per-column scale)
__global__ void scale_cols_kernel(int n, int W, const double2* __restrict__ src, const double* __restrict__ s,
                                  double2* __restrict__ dst)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * W) return;
    const double f = s[t % W];
    const double2 v = src[t];
    dst[t] = make_double2(f * v.x, f * v.y);
}

// ---------------------------------------------------------------- CG steps (real columns)
// State per column: active (not yet converged), and one flag for the whole
// solve (all columns converged or the iteration limit reached).

// active = bb > 0; done = no column active; iterations = 0
// skip (may be null): a converged GMRES cycle; the PCG is then done at once.
__global__ void cg_init_kernel(int ncols, const double* __restrict__ bb, int* __restrict__ active,
                               int* __restrict__ done, int* __restrict__ iters, const int* __restrict__ skip)
{
    __shared__ int any;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    for (int c = threadIdx.x; c < ncols; c += blockDim.x) {
        const int a = bb[c] > 0.0;
        active[c] = a;
        if (a) atomicOr(&any, 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        *done = !any || (skip && *skip);
        *iters = 0;
        done[3] = 0;   // ticket and active count of cg_check_beta_kernel
        done[4] = 0;
    }
}

// dst (n x ncols, ld ncols) = src (ld lds); zero (n x ncols, ld ldz) = 0
__global__ void cg_start_kernel(int64_t n, int ncols, const double* __restrict__ src, int64_t lds,
                                double* __restrict__ dst, double* __restrict__ zero, int64_t ldz,
                                const int* __restrict__ skip)
{
    if (skip && *skip) return;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * ncols) return;
    const int64_t v = t / ncols;
    const int c = (int)(t % ncols);
    dst[t] = src[v * lds + c];
    zero[v * ldz + c] = 0.0;
}

// s[k] = 1 / sqrt(Re hn[k]) (0 for a zero norm): column scale of the next
// Krylov vector, from the norms computed on the device.
__global__ void inv_norm_kernel(int W, const double2* __restrict__ hn, double* __restrict__ s)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= W) return;
    const double v = hn[k].x;
    s[k] = v > 0.0 ? 1.0 / sqrt(v) : 0.0;
}

// ---------------------------------------------------------------- host wrappers

static inline dim3 grid_for(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

void dev_mass(Handle& h, double alpha, const double* X, int64_t ldX, double beta, double* Y, int64_t ldY, int ncols,
              hipStream_t s, const int* skip)
{
    hipLaunchKernelGGL(mass_spmm_kernel, grid_for(h.nv * ncols), dim3(256), 0, s, (int)h.nv, h.d_M_ptr, h.d_M_idx,
                       h.d_M_val, X, ldX, Y, ldY, ncols, alpha, beta, skip);
    BAEMM_HIP(hipGetLastError());
}

void dev_lumped(Handle& h, const double* X, int64_t ldX, double* Y, int64_t ldY, int ncols, hipStream_t s,
                const int* skip)
{
    hipLaunchKernelGGL(diag_scale_kernel, grid_for(h.nv * ncols), dim3(256), 0, s, (int)h.nv, h.d_lumped_inv, X, ldX,
                       Y, ldY, ncols, skip);
    BAEMM_HIP(hipGetLastError());
}

void dev_av(Handle& h, const double* X, int64_t ldX, double* Y, int64_t ldY, int ncols, hipStream_t s)
{
    hipLaunchKernelGGL(av_real_kernel, grid_for(h.nt * ncols), dim3(256), 0, s, (int)h.nt, h.d_tris, h.d_area, X, ldX,
                       Y, ldY, ncols);
    BAEMM_HIP(hipGetLastError());
}

void dev_avt(Handle& h, double alpha, const double* X, int64_t ldX, double beta, double* Y, int64_t ldY, int ncols,
             hipStream_t s)
{
    hipLaunchKernelGGL(avt_real_kernel, grid_for(h.nv * ncols), dim3(256), 0, s, (int)h.nv, h.d_vt_ptr, h.d_vt_idx,
                       h.d_area, X, ldX, Y, ldY, ncols, alpha, beta);
    BAEMM_HIP(hipGetLastError());
}

void dev_dot_normals(Handle& h, const double* X, double* y, hipStream_t s)
{
    hipLaunchKernelGGL(dot_normals_kernel, grid_for(h.nt), dim3(256), 0, s, (int)h.nt, h.d_nrm, h.d_area, X, y);
    BAEMM_HIP(hipGetLastError());
}

void dev_mul_normals(Handle& h, const double2* x, double* Y, hipStream_t s)
{
    hipLaunchKernelGGL(mul_normals_kernel, grid_for(h.nt), dim3(256), 0, s, (int)h.nt, h.d_nrm, h.d_area, x, Y);
    BAEMM_HIP(hipGetLastError());
}

void dev_scale_rows_neg(Handle& h, int W, const double* sv, const double2* du, double2* out, hipStream_t s)
{
    hipLaunchKernelGGL(scale_rows_neg_kernel, grid_for(h.nv * W), dim3(256), 0, s, (int)h.nv, W, sv, du, out);
    BAEMM_HIP(hipGetLastError());
}

void dev_row_product_sum(Handle& h, int W, const double2* a, const double2* b, double2* p, hipStream_t s)
{
    hipLaunchKernelGGL(row_product_sum_kernel, grid_for(h.nv), dim3(256), 0, s, (int)h.nv, W, a, b, p);
    BAEMM_HIP(hipGetLastError());
}

void dev_coldot(Workspace& ws, int64_t n, int ncols, const double* X, int64_t ldX, const double* Y, int64_t ldY,
                double* out, hipStream_t s, const int* skip)
{
    double* part = ws.red.as<double>();
    hipLaunchKernelGGL(coldot_part_kernel, dim3(kRedBlocks), dim3(256), 0, s, (int)n, ncols, X, ldX, Y, ldY, part,
                       skip);
    hipLaunchKernelGGL(colsum_kernel, dim3(ncols), dim3(64), 0, s, kRedBlocks, ncols, part, out, skip);
    BAEMM_HIP(hipGetLastError());
}

void dev_gs_dots(Workspace& ws, int64_t n, int W, int nb, const double2* V, int64_t vstride, const double2* w,
                 double2* out, hipStream_t s)
{
    double2* part = ws.red.as<double2>();
    hipLaunchKernelGGL(gs_dots_part_kernel, dim3(kRedBlocks), dim3(256), 0, s, (int)n, W, nb, V, vstride, w, part);
    hipLaunchKernelGGL(gs_dots_sum_kernel, dim3(nb * W), dim3(64), 0, s, kRedBlocks, nb * W, part, out);
    BAEMM_HIP(hipGetLastError());
}

void dev_gs_update(int64_t n, int W, int nb, const double2* V, int64_t vstride, const double2* hcoef, double2* w,
                   double sign, hipStream_t s, const int* nbp)
{
    hipLaunchKernelGGL(gs_update_kernel, grid_for(n * W), dim3(256), 0, s, (int)n, W, nb, V, vstride, hcoef, w, sign,
                       nbp);
    BAEMM_HIP(hipGetLastError());
}

void dev_inv_norm(int W, const double2* hn, double* sc, hipStream_t s)
{
    hipLaunchKernelGGL(inv_norm_kernel, grid_for(W), dim3(256), 0, s, W, hn, sc);
    BAEMM_HIP(hipGetLastError());
}

void dev_scale_cols(int64_t n, int W, const double2* src, const double* sc, double2* dst, hipStream_t s)
{
    hipLaunchKernelGGL(scale_cols_kernel, grid_for(n * W), dim3(256), 0, s, (int)n, W, src, sc, dst);
    BAEMM_HIP(hipGetLastError());
}

// ---------------------------------------------------------------- GMRES on the device
// The Hessenberg matrix, Givens rotations and least squares of the GMRES
// cycles, in FP64 on the device (one work-group, threads over the columns),
// so that an Arnoldi step needs no host round trip.  State per column k:
// H[k][i][j] ((m+1) x m, rotated in place), rotations cs[k][j], sn[k][j],
// rotated right-hand side g[k][0..m], initial norm bnorm[k], residual res[k].
// ctl[0]: stop flag (set once every column meets tol; the pair kernels and
// the PCG of later steps then return at once, and later Givens kernels do
// nothing), ctl[1]: completed steps of the cycle, ctl[2]: converged.
// worst[0]: the largest relative residual over the columns.

__device__ __forceinline__ double2 zmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 zadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 zsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

// Largest value over the work-group (256 threads); every thread gets it.
__device__ double block_max(double v)
{
    __shared__ double sm[4];
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
    __syncthreads();
    const double r = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
    __syncthreads();
    return r;
}

// Start of a cycle from the preconditioned residual r (its column norms^2 in
// hv[k].x): beta_k = |r_k|, bnorm_k = beta_k on the first cycle, g_k = beta_k e_1,
// sc_k = 1 / beta_k (the caller scales r to V_0 with it).
__global__ __launch_bounds__(256) void gm_start_kernel(int W, int m, const double2* __restrict__ hv, int first,
                                                       double tol, double* __restrict__ bnorm,
                                                       double* __restrict__ res, double2* __restrict__ g,
                                                       double* __restrict__ sc, int* __restrict__ ctl,
                                                       double* __restrict__ worst)
{
    double wmax = 0.0;
    for (int k = threadIdx.x; k < W; k += blockDim.x) {
        const double beta = sqrt(fmax(0.0, hv[k].x));
        if (first) bnorm[k] = beta;
        const double r = bnorm[k] > 0.0 ? beta / bnorm[k] : 0.0;
        res[k] = r;
        wmax = fmax(wmax, r);
        sc[k] = beta > 0.0 ? 1.0 / beta : 0.0;
        double2* gk = g + (int64_t)k * (m + 1);
        gk[0] = make_double2(beta, 0.0);
        for (int i = 1; i <= m; ++i) gk[i] = make_double2(0.0, 0.0);
    }
    wmax = block_max(wmax);
    if (threadIdx.x == 0) {
        worst[0] = wmax;
        ctl[1] = 0;
        ctl[2] = wmax <= tol ? 1 : 0;
        ctl[0] = ctl[2];
    }
}

// Step j of a cycle: Hessenberg column j from the Gram-Schmidt projections
// (h1 + h2, rows 0..j) and the new vector's norm^2 (h2[j+1].x), the previous
// rotations applied, a new rotation that zeroes H[j+1][j], g rotated.
__global__ __launch_bounds__(256) void gm_givens_kernel(int W, int m, int j, const double2* __restrict__ h1,
                                                        const double2* __restrict__ h2, double tol,
                                                        const double* __restrict__ bnorm, double2* __restrict__ H,
                                                        double2* __restrict__ cs, double2* __restrict__ sn,
                                                        double2* __restrict__ g, double* __restrict__ res,
                                                        int* __restrict__ ctl, double* __restrict__ worst)
{
    if (ctl[0]) return;   // converged at an earlier step of this cycle
    double wmax = 0.0;
    for (int k = threadIdx.x; k < W; k += blockDim.x) {
        double2* Hk = H + (int64_t)k * (m + 1) * m;
        double2* ck = cs + (int64_t)k * m;
        double2* sk = sn + (int64_t)k * m;
        double2* gk = g + (int64_t)k * (m + 1);
        for (int i = 0; i <= j; ++i) Hk[i * m + j] = zadd(h1[i * W + k], h2[i * W + k]);
        const double hn = sqrt(fmax(0.0, h2[(j + 1) * W + k].x));
        for (int i = 0; i < j; ++i) {
            const double2 c = ck[i], s = sk[i];
            const double2 a = Hk[i * m + j], b = Hk[(i + 1) * m + j];
            Hk[i * m + j] = zadd(zmul(c, a), zmul(s, b));
            Hk[(i + 1) * m + j] = zsub(zmul(c, b), zmul(zconj(s), a));
        }
        const double2 a = Hk[j * m + j];
        const double na = sqrt(a.x * a.x + a.y * a.y), nu = sqrt(na * na + hn * hn);
        double2 c, s;
        if (nu == 0.0) {
            c = make_double2(1.0, 0.0);
            s = make_double2(0.0, 0.0);
        } else if (na == 0.0) {
            c = make_double2(0.0, 0.0);
            s = make_double2(1.0, 0.0);
        } else {
            c = make_double2(na / nu, 0.0);
            s = make_double2(a.x / na * hn / nu, a.y / na * hn / nu);
        }
        ck[j] = c;
        sk[j] = s;
        Hk[j * m + j] = zadd(zmul(c, a), make_double2(s.x * hn, s.y * hn));
        Hk[(j + 1) * m + j] = make_double2(0.0, 0.0);
        const double2 gj = gk[j];
        gk[j] = zmul(c, gj);
        const double2 gn = zmul(zconj(s), gj);
        gk[j + 1] = make_double2(-gn.x, -gn.y);
        const double r = bnorm[k] > 0.0 ? sqrt(gk[j + 1].x * gk[j + 1].x + gk[j + 1].y * gk[j + 1].y) / bnorm[k]
                                        : 0.0;
        res[k] = r;
        wmax = fmax(wmax, r);
    }
    wmax = block_max(wmax);
    if (threadIdx.x == 0) {
        worst[0] = wmax;
        ctl[1] = j + 1;
        if (wmax <= tol) {
            ctl[2] = 1;
            ctl[0] = 1;
        }
    }
}

// End of a cycle: y_k = R_k^{-1} g_k over the ctl[1] completed steps (upper
// triangular back substitution), written as y[i * W + k] for the basis update.
__global__ __launch_bounds__(256) void gm_backsub_kernel(int W, int m, const double2* __restrict__ H,
                                                         const double2* __restrict__ g, const int* __restrict__ ctl,
                                                         double2* __restrict__ y)
{
    const int jd = ctl[1];
    for (int k = threadIdx.x; k < W; k += blockDim.x) {
        const double2* Hk = H + (int64_t)k * (m + 1) * m;
        const double2* gk = g + (int64_t)k * (m + 1);
        for (int i = jd - 1; i >= 0; --i) {
            double2 acc = gk[i];
            for (int l = i + 1; l < jd; ++l) acc = zsub(acc, zmul(Hk[i * m + l], y[l * W + k]));
            const double2 d = Hk[i * m + i];
            const double dd = d.x * d.x + d.y * d.y;
            y[i * W + k] = dd > 0.0 ? make_double2((acc.x * d.x + acc.y * d.y) / dd, (acc.y * d.x - acc.x * d.y) / dd)
                                    : make_double2(0.0, 0.0);
        }
    }
}

void dev_gm_start(int W, int m, const double2* hv, bool first, double tol, GmState& st, double* sc, hipStream_t s)
{
    hipLaunchKernelGGL(gm_start_kernel, dim3(1), dim3(256), 0, s, W, m, hv, first ? 1 : 0, tol, st.bnorm, st.res,
                       st.g, sc, st.ctl, st.worst);
    BAEMM_HIP(hipGetLastError());
}

void dev_gm_givens(int W, int m, int j, const double2* h1, const double2* h2, double tol, GmState& st,
                   hipStream_t s)
{
    hipLaunchKernelGGL(gm_givens_kernel, dim3(1), dim3(256), 0, s, W, m, j, h1, h2, tol, st.bnorm, st.H, st.cs,
                       st.sn, st.g, st.res, st.ctl, st.worst);
    BAEMM_HIP(hipGetLastError());
}

void dev_gm_backsub(int W, int m, GmState& st, double2* y, hipStream_t s)
{
    hipLaunchKernelGGL(gm_backsub_kernel, dim3(1), dim3(256), 0, s, W, m, st.H, st.g, st.ctl, y);
    BAEMM_HIP(hipGetLastError());
}

GmState dev_gm_state(Workspace& ws, int W, int m)
{
    const size_t nH = (size_t)W * (m + 1) * m, nr = (size_t)W * m, ng = (size_t)W * (m + 1);
    ws.gm.reserve(sizeof(double2) * (nH + 2 * nr + ng) + sizeof(double) * (2 * (size_t)W + 1) + 16 * sizeof(int));
    GmState st;
    st.H = ws.gm.as<double2>();
    st.cs = st.H + nH;
    st.sn = st.cs + nr;
    st.g = st.sn + nr;
    st.bnorm = reinterpret_cast<double*>(st.g + ng);
    st.res = st.bnorm + W;
    st.worst = st.res + W;
    st.ctl = reinterpret_cast<int*>(st.worst + 1);
    return st;
}

// ---------------------------------------------------------------- fused PCG steps
// Row-blocked (block b: rows b*CGW+rl, stride grid*CGW, for its CGW waves;
// 64 column lanes), deterministic two-pass reductions.  Block partials are
// stored column-major, part[c * nblocks + b], so the one-wave column sums read
// contiguous memory.

// Waves per block of the multi-block PCG passes, and rows per wave (the grid
// is about n / (CGW * CG_RPW) blocks); build-time settings.
#ifndef BAEMM_CG_WAVES
#define BAEMM_CG_WAVES 4
#endif
#ifndef BAEMM_CG_ROWS_PER_WAVE
#define BAEMM_CG_ROWS_PER_WAVE 4
#endif
constexpr int CGW = BAEMM_CG_WAVES;
constexpr int CG_RPW = BAEMM_CG_ROWS_PER_WAVE;

// Row entries of the mass matrix gathered per batch in the SpMV: the loads of a
// batch are issued together, then the products are added in entry order.
constexpr int kSpmvBatch = 8;

// The block's partial of x.y for column c0 + c (waves rl = 0..CGW-1 of the
// block hold the sums of their rows, added in wave order) goes to
// part[(c0 + c) * gridDim.x + blockIdx.x].
__device__ __forceinline__ void cg_store_partial(double (&sm)[CGW][64], double acc, int c0, int c, int rl, int ncols,
                                                 double* __restrict__ part)
{
    sm[rl][c] = acc;
    __syncthreads();
    if (rl == 0 && c0 + c < ncols) {
        double x = sm[0][c];
#pragma unroll
        for (int w = 1; w < CGW; ++w) x += sm[w][c];
        part[(int64_t)(c0 + c) * gridDim.x + blockIdx.x] = x;
    }
    __syncthreads();
}

// First step: q = M p and the block partials of p.q
__global__ __launch_bounds__(64 * CGW) void cg_spmv_dot_kernel(int n, int ncols, const int32_t* __restrict__ ptr,
                                                          const int32_t* __restrict__ idx,
                                                          const double* __restrict__ val, const double* __restrict__ p,
                                                          double* __restrict__ q, double* __restrict__ part,
                                                          const int* __restrict__ skip)
{
    if (*skip) return;
    __shared__ double sm[CGW][64];
    const int c = threadIdx.x & 63, rl = threadIdx.x >> 6;
    for (int c0 = 0; c0 < ncols; c0 += 64) {
        double acc = 0.0;
        if (c0 + c < ncols)
            for (int r = blockIdx.x * CGW + rl; r < n; r += gridDim.x * CGW) {
                const int e0 = ptr[r], e1 = ptr[r + 1];
                double sv = 0.0;
                for (int e = e0; e < e1; e += kSpmvBatch) {
                    double mv[kSpmvBatch], xv[kSpmvBatch];
#pragma unroll
                    for (int j = 0; j < kSpmvBatch; ++j) {
                        const int ej = min(e + j, e1 - 1);   // in the row: no branch between the loads
                        mv[j] = val[ej];
                        xv[j] = p[(int64_t)idx[ej] * ncols + c0 + c];
                    }
#pragma unroll
                    for (int j = 0; j < kSpmvBatch; ++j)
                        if (e + j < e1) sv += mv[j] * xv[j];
                }
                q[(int64_t)r * ncols + c0 + c] = sv;
                acc += p[(int64_t)r * ncols + c0 + c] * sv;
            }
        cg_store_partial(sm, acc, c0, c, rl, ncols, part);
    }
}

// Later steps: the new direction p = z + beta p_in with z = D r, formed where it
// is needed, and q = M p.  Each block writes p_out for its own rows and forms
// the entries of its neighbours' rows again (the same operations, so the same
// values), so the direction update needs no pass of its own.  p_in is not
// written in this launch (the directions are double-buffered).  Block partials
// of p_out.q.
__global__ __launch_bounds__(64 * CGW) void cg_spmv_p_dot_kernel(int n, int ncols, const int32_t* __restrict__ ptr,
                                                            const int32_t* __restrict__ idx,
                                                            const double* __restrict__ val,
                                                            const double* __restrict__ Dinv,
                                                            const double* __restrict__ r,
                                                            const double* __restrict__ beta,
                                                            const double* __restrict__ p_in,
                                                            double* __restrict__ p_out, double* __restrict__ q,
                                                            double* __restrict__ part, const int* __restrict__ skip)
{
    if (*skip) return;
    __shared__ double sm[CGW][64];
    const int c = threadIdx.x & 63, rl = threadIdx.x >> 6;
    for (int c0 = 0; c0 < ncols; c0 += 64) {
        double acc = 0.0;
        if (c0 + c < ncols) {
            const double bt = beta[c0 + c];
            for (int v = blockIdx.x * CGW + rl; v < n; v += gridDim.x * CGW) {
                const int e0 = ptr[v], e1 = ptr[v + 1];
                double sv = 0.0;
                for (int e = e0; e < e1; e += kSpmvBatch) {
                    double mv[kSpmvBatch], xv[kSpmvBatch];
#pragma unroll
                    for (int j = 0; j < kSpmvBatch; ++j) {
                        const int ej = min(e + j, e1 - 1);   // in the row: no branch between the loads
                        const int u = idx[ej];
                        const int64_t k = (int64_t)u * ncols + c0 + c;
                        mv[j] = val[ej];
                        xv[j] = fma(bt, p_in[k], Dinv[u] * r[k]);
                    }
#pragma unroll
                    for (int j = 0; j < kSpmvBatch; ++j)
                        if (e + j < e1) sv += mv[j] * xv[j];
                }
                const int64_t k = (int64_t)v * ncols + c0 + c;
                const double pv = fma(bt, p_in[k], Dinv[v] * r[k]);
                p_out[k] = pv;
                q[k] = sv;
                acc += pv * sv;
            }
        }
        cg_store_partial(sm, acc, c0, c, rl, ncols, part);
    }
}

// pq = sum_b part[c][b]; alpha = active ? rz / pq : 0; one 64-thread block per column
__global__ __launch_bounds__(64) void cg_alpha_sum_kernel(int nblocks, int ncols, const double* __restrict__ part,
                                                          const double* __restrict__ rz,
                                                          const int* __restrict__ active, double* __restrict__ alpha,
                                                          const int* __restrict__ skip)
{
    if (*skip) return;
    const int c = blockIdx.x;
    const double pq = wave_colsum(part, nblocks, 1, (int64_t)c * nblocks);
    if (threadIdx.x == 0) alpha[c] = (active[c] && pq != 0.0) ? rz[c] / pq : 0.0;
}

// Y += alpha p; r -= alpha q; block partials part1 = r.r and part2 = r.z with
// z = D r (not stored: the next step forms it again)
__global__ __launch_bounds__(64 * CGW) void cg_update_dots_kernel(int n, int ncols, const double* __restrict__ alpha,
                                                             const double* __restrict__ p,
                                                             const double* __restrict__ q,
                                                             const double* __restrict__ Dinv, double* __restrict__ Y,
                                                             int64_t ldY, double* __restrict__ r,
                                                             double* __restrict__ part1, double* __restrict__ part2,
                                                             const int* __restrict__ skip)
{
    if (*skip) return;
    __shared__ double sm1[CGW][64], sm2[CGW][64];
    const int c = threadIdx.x & 63, rl = threadIdx.x >> 6;
    for (int c0 = 0; c0 < ncols; c0 += 64) {
        double a1 = 0.0, a2 = 0.0;
        if (c0 + c < ncols) {
            const double al = alpha[c0 + c];
            for (int v = blockIdx.x * CGW + rl; v < n; v += gridDim.x * CGW) {
                const int64_t k = (int64_t)v * ncols + c0 + c;
                Y[(int64_t)v * ldY + c0 + c] += al * p[k];
                const double rv = r[k] - al * q[k];
                r[k] = rv;
                const double zv = Dinv[v] * rv;
                a1 += rv * rv;
                a2 += rv * zv;
            }
        }
        cg_store_partial(sm1, a1, c0, c, rl, ncols, part1);
        cg_store_partial(sm2, a2, c0, c, rl, ncols, part2);
    }
}

// Column c = blockIdx.x (64 threads): rr = sum part1, rz_new = sum part2;
// freeze the column once rr <= tol2 bb; beta = rz_new / rz.  The last block to
// finish (ticket ctr[3]) reads the count of still-active columns (ctr[4]),
// advances the iteration counters and sets the done flag, then clears ctr[3..4]
// for the next iteration.  ctr: done, iterations, total, ticket, active count.
__global__ __launch_bounds__(64) void cg_check_beta_kernel(int nblocks, int ncols, const double* __restrict__ part1,
                                                           const double* __restrict__ part2,
                                                           const double* __restrict__ bb, double tol2, int max_iter,
                                                           int* __restrict__ active, double* __restrict__ rz,
                                                           double* __restrict__ beta, int* __restrict__ ctr)
{
    if (ctr[0]) return;
    const int c = blockIdx.x;
    const double rr = wave_colsum(part1, nblocks, 1, (int64_t)c * nblocks);
    const double rzn = wave_colsum(part2, nblocks, 1, (int64_t)c * nblocks);
    if (threadIdx.x != 0) return;
    int still = 0;
    if (active[c]) {
        if (rr <= tol2 * bb[c]) active[c] = 0;
        else still = 1;
    }
    beta[c] = (rz[c] != 0.0) ? rzn / rz[c] : 0.0;
    rz[c] = rzn;
    if (still) atomicAdd(&ctr[4], 1);
    __threadfence();
    if (atomicAdd(&ctr[3], 1) == (int)gridDim.x - 1) {
        __threadfence();
        const int any = atomicAdd(&ctr[4], 0);
        const int it = ctr[1] + 1;
        ctr[1] = it;
        ctr[2] += 1;
        if (!any || it >= max_iter) ctr[0] = 1;
        atomicExch(&ctr[4], 0);
        atomicExch(&ctr[3], 0);
    }
}

// ---------------------------------------------------------------- LDS block PCG
// The block PCG of small meshes with its working set on chip: the mass
// matrix and the search directions p live in LDS (copied there once), and
// each thread keeps r, q and Y of its R rows (tid + k * 256) in registers.
// The loop touches no global memory; Y is written once at the end.  Same
// algorithm and stopping rule as the multi-block PCG below; ctr[1] receives
// the largest iteration count over the work-groups (zeroed before the launch)
// and pcg_total_kernel adds it to ctr[2].
// The matrix is held in one of two forms:
//  * CSR (TRI = false): val (nnz doubles), ptr (n + 1 ints), idx (nnz ints);
//  * triangles (TRI = true), for meshes whose CSR does not fit: per triangle
//    its weight a/12 (double, from the same FP32 area as the CSR values) and
//    its vertices (16-bit), and the vertex-to-triangle incidence (16-bit), so
//    that (M p)_i = sum over the triangles t of i of w_t (p_a + p_b + p_c + p_i)
//    (a/6 on the diagonal, a/12 off it, as the CSR assembles it).
// pcg_lds_bytes gives the dynamic LDS size of either form.
constexpr int kPcgLdsThreads = 256;
constexpr size_t kPcgLdsMaxBytes = 150u * 1024u;
// One column per work-group: FarField solves with 32 plane waves measured
// 25-40 % faster on MI355X than with four columns per work-group (more
// work-groups, fewer rows x columns per thread, a smaller LDS footprint).

struct PcgMatrix {
    int n, nnz, nt, nvt;    // nvt: vertex-to-triangle incidences
    const int32_t* ptr;     // CSR
    const int32_t* idx;
    const double* val;
    const int4* tris;       // triangle form
    const float* area;
    const int32_t* vt_ptr;
    const int32_t* vt_idx;
};

static size_t pcg_lds_bytes(int64_t n, int64_t nnz, int64_t nt, int64_t nvt, int CGC, bool tri)
{
    if (!tri) return (size_t)(8 * nnz + 8 * n * CGC + 4 * (n + 1) + 4 * nnz);
    return (size_t)(8 * nt + 8 * n * CGC + 8 * nt + 2 * (n + 1) + 2 * nvt);
}

template <int CGC, int R, bool TRI>
__global__ __launch_bounds__(kPcgLdsThreads) void pcg_lds_kernel(PcgMatrix A, const double* __restrict__ Dinv,
                                                                 const double* __restrict__ X, int64_t ldX,
                                                                 double* __restrict__ Y, int64_t ldY, int ncols,
                                                                 double tol2, int max_iter, int* __restrict__ ctr,
                                                                 const int* __restrict__ skip)
{
    if (skip && *skip) return;   // a converged GMRES cycle: the step's result is not used
    constexpr int NT = kPcgLdsThreads, NW = NT / 64;
    const int n = A.n, nnz = A.nnz, nt = A.nt;
    extern __shared__ double pcg_lds[];
    // CSR form
    double* s_val = pcg_lds;
    double* s_p = TRI ? pcg_lds + nt : s_val + nnz;
    int32_t* s_ptr = reinterpret_cast<int32_t*>(s_p + (int64_t)n * CGC);
    int32_t* s_idx = s_ptr + (n + 1);
    // triangle form
    double* s_tw = pcg_lds;
    ushort4* s_tv = reinterpret_cast<ushort4*>(s_p + (int64_t)n * CGC);
    uint16_t* s_vptr = reinterpret_cast<uint16_t*>(s_tv + nt);
    uint16_t* s_vidx = s_vptr + (n + 1);
    __shared__ double red[2][CGC][NW];
    __shared__ double s_rz[CGC], s_bb[CGC];
    __shared__ int s_active[CGC];
    const int c0 = blockIdx.x * CGC;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    bool okc[CGC];
#pragma unroll
    for (int c = 0; c < CGC; ++c) okc[c] = c0 + c < ncols;

    // block sums of per-thread values v[k][c] for k < K (waves reduce by
    // shuffles, then LDS in wave order)
    auto block_sum = [&](double (&v)[2][CGC], int K) {
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            if (k >= K) break;
#pragma unroll
            for (int c = 0; c < CGC; ++c) {
                double x = v[k][c];
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
                if (lane == 0) red[k][c][wv] = x;
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            if (k >= K) break;
#pragma unroll
            for (int c = 0; c < CGC; ++c) {
                double x = red[k][c][0];
#pragma unroll
                for (int w = 1; w < NW; ++w) x += red[k][c][w];
                v[k][c] = x;
            }
        }
        __syncthreads();
    };

    if constexpr (TRI) {
        for (int t = tid; t < nt; t += NT) {
            const int4 tv = A.tris[t];
            s_tv[t] = make_ushort4((unsigned short)tv.x, (unsigned short)tv.y, (unsigned short)tv.z, 0);
            s_tw[t] = (double)A.area[t] / 12.0;
        }
        for (int v = tid; v <= n; v += NT) s_vptr[v] = (uint16_t)A.vt_ptr[v];
        for (int e = tid; e < A.nvt; e += NT) s_vidx[e] = (uint16_t)A.vt_idx[e];
    } else {
        for (int e = tid; e < nnz; e += NT) {
            s_val[e] = A.val[e];
            s_idx[e] = A.idx[e];
        }
        for (int v = tid; v <= n; v += NT) s_ptr[v] = A.ptr[v];
    }

    // r = X, Y = 0, p = D r; bb = r.r, rz = r.(D r)
    double r[R][CGC], q[R][CGC], y[R][CGC], d[R];
    double acc[2][CGC];
#pragma unroll
    for (int c = 0; c < CGC; ++c) acc[0][c] = acc[1][c] = 0.0;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int v = tid + k * NT;
        d[k] = v < n ? Dinv[v] : 0.0;
#pragma unroll
        for (int c = 0; c < CGC; ++c) {
            const double x = (v < n && okc[c]) ? X[(int64_t)v * ldX + c0 + c] : 0.0;
            r[k][c] = x;
            q[k][c] = 0.0;
            y[k][c] = 0.0;
            if (v < n) s_p[(int64_t)v * CGC + c] = d[k] * x;
            acc[0][c] += x * x;
            acc[1][c] += x * d[k] * x;
        }
    }
    block_sum(acc, 2);   // its barriers also publish the LDS copies
    if (tid == 0) {
#pragma unroll
        for (int c = 0; c < CGC; ++c) {
            s_bb[c] = acc[0][c];
            s_rz[c] = acc[1][c];
            s_active[c] = okc[c] && acc[0][c] > 0.0;
        }
    }
    __syncthreads();
    int it = 0;
    for (;;) {
        bool any = false;
#pragma unroll
        for (int c = 0; c < CGC; ++c) any = any || s_active[c];
        if (!any || it >= max_iter) break;
        ++it;
        // q = M p, pq
#pragma unroll
        for (int c = 0; c < CGC; ++c) acc[0][c] = 0.0;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const int v = tid + k * NT;
            if (v < n) {
                double sv[CGC];
#pragma unroll
                for (int c = 0; c < CGC; ++c) sv[c] = 0.0;
                if constexpr (TRI) {
                    for (int e = s_vptr[v]; e < s_vptr[v + 1]; ++e) {
                        const int t = s_vidx[e];
                        const ushort4 tv = s_tv[t];
                        const double w = s_tw[t];
#pragma unroll
                        for (int c = 0; c < CGC; ++c)
                            sv[c] += w * (s_p[tv.x * CGC + c] + s_p[tv.y * CGC + c] + s_p[tv.z * CGC + c] +
                                          s_p[(int64_t)v * CGC + c]);
                    }
                } else {
                    for (int e = s_ptr[v]; e < s_ptr[v + 1]; ++e) {
                        const double m = s_val[e];
                        const int64_t u = (int64_t)s_idx[e] * CGC;
#pragma unroll
                        for (int c = 0; c < CGC; ++c) sv[c] += m * s_p[u + c];
                    }
                }
#pragma unroll
                for (int c = 0; c < CGC; ++c) {
                    q[k][c] = sv[c];
                    acc[0][c] += s_p[(int64_t)v * CGC + c] * sv[c];
                }
            }
        }
        block_sum(acc, 1);
        double alpha[CGC];
#pragma unroll
        for (int c = 0; c < CGC; ++c) alpha[c] = (s_active[c] && acc[0][c] != 0.0) ? s_rz[c] / acc[0][c] : 0.0;
        // Y += alpha p, r -= alpha q; rr and r.(D r)
#pragma unroll
        for (int c = 0; c < CGC; ++c) acc[0][c] = acc[1][c] = 0.0;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const int v = tid + k * NT;
            if (v < n) {
#pragma unroll
                for (int c = 0; c < CGC; ++c) {
                    y[k][c] += alpha[c] * s_p[(int64_t)v * CGC + c];
                    r[k][c] -= alpha[c] * q[k][c];
                    acc[0][c] += r[k][c] * r[k][c];
                    acc[1][c] += r[k][c] * d[k] * r[k][c];
                }
            }
        }
        block_sum(acc, 2);
        double beta[CGC];
#pragma unroll
        for (int c = 0; c < CGC; ++c) beta[c] = (s_rz[c] != 0.0) ? acc[1][c] / s_rz[c] : 0.0;
        __syncthreads();   // every thread has read s_rz and s_active
        if (tid == 0) {
#pragma unroll
            for (int c = 0; c < CGC; ++c) {
                if (s_active[c] && acc[0][c] <= tol2 * s_bb[c]) s_active[c] = 0;
                s_rz[c] = acc[1][c];
            }
        }
        // p = z + beta p on the own rows (the SpMV reads of p ended at the
        // barriers above; the next SpMV reads after the barrier below)
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const int v = tid + k * NT;
            if (v < n) {
#pragma unroll
                for (int c = 0; c < CGC; ++c) {
                    const int64_t kk = (int64_t)v * CGC + c;
                    s_p[kk] = d[k] * r[k][c] + beta[c] * s_p[kk];
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int v = tid + k * NT;
        if (v < n) {
#pragma unroll
            for (int c = 0; c < CGC; ++c)
                if (okc[c]) Y[(int64_t)v * ldY + c0 + c] = y[k][c];
        }
    }
    if (tid == 0) atomicMax(&ctr[1], it);
}

// Launches pcg_lds_kernel<CGC, R, TRI> on nb work-groups with lds bytes of
// dynamic LDS, first allowing that much on this device (once per device and
// kernel: a host-side attribute, no stream work).
template <int CGC, int R, bool TRI>
static void launch_pcg_lds(Handle& h, int nb, size_t lds, const double* X, int64_t ldX, double* Y, int64_t ldY,
                           int ncols, double tol, int max_iter, int* ctr, hipStream_t s)
{
    static std::mutex mtx;
    static std::set<int> allowed;
    {
        std::lock_guard<std::mutex> lk(mtx);
        if (!allowed.count(h.device)) {
            BAEMM_HIP(hipFuncSetAttribute((const void*)pcg_lds_kernel<CGC, R, TRI>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)kPcgLdsMaxBytes));
            allowed.insert(h.device);
        }
    }
    const PcgMatrix A{(int)h.nv, (int)h.h_M_idx.size(), (int)h.nt, (int)h.h_vt_idx.size(), h.d_M_ptr, h.d_M_idx,
                      h.d_M_val, h.d_tris, h.d_area, h.d_vt_ptr, h.d_vt_idx};
    hipLaunchKernelGGL((pcg_lds_kernel<CGC, R, TRI>), dim3(nb), dim3(kPcgLdsThreads), lds, s, A, h.d_lumped_inv, X,
                       ldX, Y, ldY, ncols, tol * tol, max_iter, ctr, h.step_skip);
}

// The LDS block PCG for CGC columns per work-group and R rows per thread.
template <bool TRI>
static void launch_pcg_lds_any(int CGC, int R, Handle& h, int nb, size_t lds, const double* X, int64_t ldX,
                               double* Y, int64_t ldY, int ncols, double tol, int max_iter, int* ctr, hipStream_t s)
{
#define BAEMM_PCGL(C_, R_)                                                                        \
    if (CGC == C_ && R == R_) {                                                                   \
        launch_pcg_lds<C_, R_, TRI>(h, nb, lds, X, ldX, Y, ldY, ncols, tol, max_iter, ctr, s);     \
        return;                                                                                   \
    }
    BAEMM_PCGL(1, 1) BAEMM_PCGL(1, 2) BAEMM_PCGL(1, 4) BAEMM_PCGL(1, 8) BAEMM_PCGL(1, 16)
#undef BAEMM_PCGL
    throw ArgError{"internal: no LDS PCG variant"};
}

// ctr[2] += ctr[1] (iterations of the block PCG just run); ctr[1] = 0 before it.
__global__ void pcg_reset_kernel(int* __restrict__ ctr) { ctr[1] = 0; }
__global__ void pcg_total_kernel(int* __restrict__ ctr) { ctr[2] += ctr[1]; }

// PCG for M Y = X on real columns (ApplyMassInverse.hpp:6-156): zero initial
// guess, lumped-diagonal preconditioner, at most max_iter iterations, per
// column relative residual ||r|| <= tol ||X||.  X: n x ncols (ldX), Y: (ldY).
// Convergence is decided on the device: the sequence of launches is the same
// for every call (no host synchronisation, so it can be captured in a graph);
// once every column has converged the remaining launches return at once, and
// a frozen column is no longer updated, so the result equals stopping there.
// Counters (ws.cgctr, int[8]): done flag, iterations of the last call, total
// iterations since the buffer was created, then the ticket and active-column
// count of cg_check_beta_kernel.
void dev_mass_cg(Handle& h, Workspace& ws, const double* X, int64_t ldX, double* Y, int64_t ldY, int ncols, double tol,
                 int max_iter, hipStream_t s)
{
    const int64_t n = h.nv;
    const int64_t N = n * ncols;
    // LDS block PCG: the CSR form when it fits, else the triangle form (16-bit
    // vertex and triangle indices)
    const int64_t nnz = (int64_t)h.h_M_idx.size(), nvt = (int64_t)h.h_vt_idx.size();
    const size_t b_csr = pcg_lds_bytes(n, nnz, h.nt, nvt, 1, false);
    const size_t b_tri = pcg_lds_bytes(n, nnz, h.nt, nvt, 1, true);
    const bool fit_csr = n <= 8 * kPcgLdsThreads && b_csr <= kPcgLdsMaxBytes;
    const bool fit_tri = !fit_csr && n <= 16 * kPcgLdsThreads && h.nt <= 65535 && nvt <= 65535 &&
                         b_tri <= kPcgLdsMaxBytes;
    if (fit_csr || fit_tri) {
        // working set on chip: one launch with no global traffic in the loop
        int* ctr = dev_cg_counters(ws, s);
        ws.scal.reserve(sizeof(double) * 8 * ncols + sizeof(int) * ncols);
        hipLaunchKernelGGL(pcg_reset_kernel, dim3(1), dim3(1), 0, s, ctr);
        int R = 1;
        while (R * kPcgLdsThreads < n) R *= 2;
        if (fit_csr)
            launch_pcg_lds_any<false>(1, R, h, ncols, b_csr, X, ldX, Y, ldY, ncols, tol, max_iter, ctr, s);
        else
            launch_pcg_lds_any<true>(1, R, h, ncols, b_tri, X, ldX, Y, ldY, ncols, tol, max_iter, ctr, s);
        hipLaunchKernelGGL(pcg_total_kernel, dim3(1), dim3(1), 0, s, ctr);
        BAEMM_HIP(hipGetLastError());
        return;
    }
    ws.cg.reserve(sizeof(double) * (4 * N));
    // partial-sum blocks of the SpMV and update passes: about 16 rows per block
    // (4 per wave), so the row loops stay short; the column sums read them with
    // one wave per column
    const int nred = (int)std::min<int64_t>(4096, std::max<int64_t>(kRedBlocks, (n + CG_RPW * CGW - 1) / (CG_RPW * CGW)));
    ws.red.reserve(sizeof(double) * 2 * std::max(nred, kRedBlocks) * ncols);
    ws.scal.reserve(sizeof(double) * 8 * ncols + sizeof(int) * ncols);
    int* ctr = dev_cg_counters(ws, s);
    // the search directions are double-buffered: step it uses pa (even it) or pb
    double* r = ws.cg.as<double>();
    double* pa = r + N;
    double* pb = pa + N;
    double* q = pb + N;
    double* rz = ws.scal.as<double>();
    double* rz_new = rz + ncols;
    double* pq = rz_new + ncols;
    double* alpha = pq + ncols;
    double* beta = alpha + ncols;
    double* rr = beta + ncols;
    double* bb = rr + ncols;
    int* active = reinterpret_cast<int*>(bb + 2 * ncols);
    int* done = ctr;
    int* iters = ctr + 1;
    // r = X (contiguous), Y = 0
    hipLaunchKernelGGL(cg_start_kernel, grid_for(N), dim3(256), 0, s, n, ncols, X, ldX, r, Y, ldY, h.step_skip);
    dev_coldot(ws, n, ncols, r, ncols, r, ncols, bb, s, h.step_skip);
    hipLaunchKernelGGL(cg_init_kernel, dim3(1), dim3(256), 0, s, ncols, bb, active, done, iters, h.step_skip);
    dev_lumped(h, r, ncols, pa, ncols, ncols, s, done);      // p = z = D r
    dev_coldot(ws, n, ncols, r, ncols, pa, ncols, rz, s, done);
    // four launches per iteration: p = z + beta p (after the first) and q = M p
    // with p.q partials; alpha; updates of Y and r with r.r and r.z partials;
    // convergence and beta
    double* part1 = ws.red.as<double>();
    double* part2 = part1 + (int64_t)nred * ncols;
    for (int it = 0; it < max_iter; ++it) {
        double* p = (it & 1) ? pb : pa;
        if (it == 0)
            hipLaunchKernelGGL(cg_spmv_dot_kernel, dim3(nred), dim3(64 * CGW), 0, s, (int)n, ncols, h.d_M_ptr, h.d_M_idx,
                               h.d_M_val, p, q, part1, done);
        else
            hipLaunchKernelGGL(cg_spmv_p_dot_kernel, dim3(nred), dim3(64 * CGW), 0, s, (int)n, ncols, h.d_M_ptr,
                               h.d_M_idx, h.d_M_val, h.d_lumped_inv, r, beta, (it & 1) ? pa : pb, p, q, part1,
                               done);
        hipLaunchKernelGGL(cg_alpha_sum_kernel, dim3(ncols), dim3(64), 0, s, nred, ncols, part1, rz, active, alpha,
                           done);
        hipLaunchKernelGGL(cg_update_dots_kernel, dim3(nred), dim3(64 * CGW), 0, s, (int)n, ncols, alpha, p, q,
                           h.d_lumped_inv, Y, ldY, r, part1, part2, done);
        hipLaunchKernelGGL(cg_check_beta_kernel, dim3(ncols), dim3(64), 0, s, nred, ncols, part1, part2, bb,
                           tol * tol, max_iter, active, rz, beta, ctr);
        BAEMM_HIP(hipGetLastError());
    }
    BAEMM_HIP(hipGetLastError());
}

// The PCG counters, created zeroed on first use (outside any graph capture:
// the first PCG of a handle always runs eagerly).
int* dev_cg_counters(Workspace& ws, hipStream_t s)
{
    if (ws.cgctr.bytes < 8 * sizeof(int)) {
        ws.cgctr.reserve(8 * sizeof(int));
        BAEMM_HIP(hipMemsetAsync(ws.cgctr.ptr, 0, 8 * sizeof(int), s));
    }
    return ws.cgctr.as<int>();
}

// Iterations of the last dev_mass_cg (the largest over its columns;
// synchronises the stream).
int dev_mass_cg_result(Workspace& ws, hipStream_t s)
{
    int ctr[4] = {0, 0, 0, 0};
    BAEMM_HIP(hipMemcpyAsync(ctr, ws.cgctr.ptr, sizeof(ctr), hipMemcpyDeviceToHost, s));
    BAEMM_HIP(hipStreamSynchronize(s));
    return ctr[1];
}

// Total PCG iterations of the handle so far (synchronises the stream).
int64_t dev_cg_total(Workspace& ws, hipStream_t s)
{
    const int* ctr = dev_cg_counters(ws, s);
    int h[4];
    BAEMM_HIP(hipMemcpyAsync(h, ctr, sizeof(h), hipMemcpyDeviceToHost, s));
    BAEMM_HIP(hipStreamSynchronize(s));
    return h[2];
}

}  // namespace baemm

namespace baemm {

// r = b - t  (complex vectors of N elements)
__global__ void sub_c_kernel(int64_t N, const double2* __restrict__ b, const double2* __restrict__ t,
                             double2* __restrict__ r)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    r[i] = make_double2(b[i].x - t[i].x, b[i].y - t[i].y);
}

void dev_sub_c(int64_t N, const double2* b, const double2* t, double2* r, hipStream_t s)
{
    hipLaunchKernelGGL(sub_c_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, N, b, t, r);
    BAEMM_HIP(hipGetLastError());
}

}  // namespace baemm

// ---------------------------------------------------------------------------
// Real vectors of the device-side Gauss-Newton solve (baemm_gauss_newton_solve_dev):
// the outer GMRES over vec(X) (3 V doubles) keeps its Krylov basis in HBM.
// ---------------------------------------------------------------------------
namespace baemm {

constexpr int kRdotThreads = 256;

// out[b] = sum_i V[b vs + i] w[i], b < gridDim.x: one block per vector, lane-
// strided partial sums then a fixed tree, so results are reproducible.
__global__ __launch_bounds__(kRdotThreads) void rdots_kernel(int64_t n, const double* __restrict__ V, int64_t vs,
                                                             const double* __restrict__ w, double* __restrict__ out)
{
    const double* v = V + (int64_t)blockIdx.x * vs;
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += kRdotThreads) acc = fma(v[i], w[i], acc);
    __shared__ double red[kRdotThreads];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = kRdotThreads / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// w[i] += sign sum_{b < nb} c[b] V[b vs + i]   (c on the device)
__global__ void rcombine_kernel(int64_t n, int nb, const double* __restrict__ V, int64_t vs,
                                const double* __restrict__ c, double sign, double* __restrict__ w)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc = w[i];
    for (int b = 0; b < nb; ++b) acc = fma(sign * c[b], V[b * vs + i], acc);
    w[i] = acc;
}

// y = a x + b y  (b == 0: y is not read)
__global__ void raxpby_kernel(int64_t n, double a, const double* __restrict__ x, double b, double* __restrict__ y)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    y[i] = (b != 0.0) ? fma(a, x[i], b * y[i]) : a * x[i];
}

// Y (nv x 3, ld ldY) = a X (nv x 3, ld ldX) + b Y  (b == 0: Y is not read)
__global__ void rows3_axpby_kernel(int64_t nv, double a, const double* __restrict__ X, int64_t ldX, double b,
                                   double* __restrict__ Y, int64_t ldY)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 3 * nv) return;
    const int64_t v = e / 3, d = e % 3;
    const double x = X[v * ldX + d];
    Y[v * ldY + d] = (b != 0.0) ? fma(a, x, b * Y[v * ldY + d]) : a * x;
}

// y = x / (c + d m_v) per vertex v (3 columns), m_v the mass-matrix row sum
// 3 / lumped_inv[v] (the row sum of the P1 mass matrix is a third of the
// areas around v; lumped_inv holds 9 over that area sum).
__global__ void metric_jacobi_kernel(int64_t nv, double c, double d, const double* __restrict__ lumped_inv,
                                     const double* __restrict__ x, double* __restrict__ y)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 3 * nv) return;
    const double li = lumped_inv[e / 3];
    const double mv = li > 0.0 ? 3.0 / li : 0.0;
    y[e] = x[e] / (c + d * mv);
}

void dev_rdots(int64_t n, int nb, const double* V, int64_t vs, const double* w, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(rdots_kernel, dim3((unsigned)nb), dim3(kRdotThreads), 0, s, n, V, vs, w, out);
    BAEMM_HIP(hipGetLastError());
}

void dev_rcombine(int64_t n, int nb, const double* V, int64_t vs, const double* c, double sign, double* w,
                  hipStream_t s)
{
    hipLaunchKernelGGL(rcombine_kernel, grid_for(n), dim3(256), 0, s, n, nb, V, vs, c, sign, w);
    BAEMM_HIP(hipGetLastError());
}

void dev_raxpby(int64_t n, double a, const double* x, double b, double* y, hipStream_t s)
{
    hipLaunchKernelGGL(raxpby_kernel, grid_for(n), dim3(256), 0, s, n, a, x, b, y);
    BAEMM_HIP(hipGetLastError());
}

void dev_rows3_axpby(int64_t nv, double a, const double* X, int64_t ldX, double b, double* Y, int64_t ldY,
                     hipStream_t s)
{
    hipLaunchKernelGGL(rows3_axpby_kernel, grid_for(3 * nv), dim3(256), 0, s, nv, a, X, ldX, b, Y, ldY);
    BAEMM_HIP(hipGetLastError());
}

void dev_metric_jacobi(Handle& h, double c, double d, const double* x, double* y, hipStream_t s)
{
    hipLaunchKernelGGL(metric_jacobi_kernel, grid_for(3 * h.nv), dim3(256), 0, s, h.nv, c, d, h.d_lumped_inv, x, y);
    BAEMM_HIP(hipGetLastError());
}

}  // namespace baemm
