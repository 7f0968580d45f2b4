// solvers.hpp -- device primitives of the Krylov solvers (solvers.hip) and the
// host drivers / compositions built on them (solvers_host.cpp).
#pragma once

#include <algorithm>
#include <vector>

#include "internal.hpp"

namespace baemm {

constexpr int kRedBlocks = 240;   // blocks of the first pass of column reductions

void dev_mass(Handle& h, double alpha, const double* X, int64_t ldX, double beta, double* Y, int64_t ldY, int ncols,
              hipStream_t s, const int* skip = nullptr);
void dev_lumped(Handle& h, const double* X, int64_t ldX, double* Y, int64_t ldY, int ncols, hipStream_t s,
                const int* skip = nullptr);
void dev_av(Handle& h, const double* X, int64_t ldX, double* Y, int64_t ldY, int ncols, hipStream_t s);
void dev_avt(Handle& h, double alpha, const double* X, int64_t ldX, double beta, double* Y, int64_t ldY, int ncols,
             hipStream_t s);
void dev_dot_normals(Handle& h, const double* X, double* y, hipStream_t s);
void dev_mul_normals(Handle& h, const double2* x, double* Y, hipStream_t s);
void dev_scale_rows_neg(Handle& h, int W, const double* sv, const double2* du, double2* out, hipStream_t s);
void dev_row_product_sum(Handle& h, int W, const double2* a, const double2* b, double2* p, hipStream_t s);
void dev_coldot(Workspace& ws, int64_t n, int ncols, const double* X, int64_t ldX, const double* Y, int64_t ldY,
                double* out, hipStream_t s, const int* skip = nullptr);
void dev_gs_dots(Workspace& ws, int64_t n, int W, int nb, const double2* V, int64_t vstride, const double2* w,
                 double2* out, hipStream_t s);
void dev_gs_update(int64_t n, int W, int nb, const double2* V, int64_t vstride, const double2* hcoef, double2* w,
                   double sign, hipStream_t s, const int* nbp = nullptr);
void dev_scale_cols(int64_t n, int W, const double2* src, const double* sc, double2* dst, hipStream_t s);
void dev_inv_norm(int W, const double2* hn, double* sc, hipStream_t s);
void dev_sub_c(int64_t N, const double2* b, const double2* t, double2* r, hipStream_t s);
void dev_mass_cg(Handle& h, Workspace& ws, const double* X, int64_t ldX, double* Y, int64_t ldY, int ncols, double tol,
                 int max_iter, hipStream_t s);
int dev_mass_cg_result(Workspace& ws, hipStream_t s);

// Real vectors of the device-side Gauss-Newton GMRES.
// out[b] = V_b . w for b < nb (V_b = V + b vs), one reproducible sum per vector.
void dev_rdots(int64_t n, int nb, const double* V, int64_t vs, const double* w, double* out, hipStream_t s);
// w += sign sum_b c[b] V_b (c on the device)
void dev_rcombine(int64_t n, int nb, const double* V, int64_t vs, const double* c, double sign, double* w,
                  hipStream_t s);
// y = a x + b y (b == 0: y not read)
void dev_raxpby(int64_t n, double a, const double* x, double b, double* y, hipStream_t s);
// Y = a X + b Y on nv x 3 arrays with leading dimensions (b == 0: Y not read)
void dev_rows3_axpby(int64_t nv, double a, const double* X, int64_t ldX, double b, double* Y, int64_t ldY,
                     hipStream_t s);
// y = x / (c + d m_v) row by row (m_v: row sum of the mass matrix), nv x 3
void dev_metric_jacobi(Handle& h, double c, double d, const double* x, double* y, hipStream_t s);

// GMRES cycle state on the device (Hessenberg, rotations, least squares).
struct GmState {
    double2* H = nullptr;   // W x (m+1) x m
    double2* cs = nullptr;  // W x m
    double2* sn = nullptr;  // W x m
    double2* g = nullptr;   // W x (m+1)
    double* bnorm = nullptr;
    double* res = nullptr;
    double* worst = nullptr;
    int* ctl = nullptr;     // stop flag, completed steps, converged
};
GmState dev_gm_state(Workspace& ws, int W, int m);
void dev_gm_start(int W, int m, const double2* hv, bool first, double tol, GmState& st, double* sc, hipStream_t s);
void dev_gm_givens(int W, int m, int j, const double2* h1, const double2* h2, double tol, GmState& st,
                   hipStream_t s);
void dev_gm_backsub(int W, int m, GmState& st, double2* y, hipStream_t s);
int* dev_cg_counters(Workspace& ws, hipStream_t s);
int64_t dev_cg_total(Workspace& ws, hipStream_t s);

}  // namespace baemm
