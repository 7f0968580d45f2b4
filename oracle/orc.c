/*
 * orc.c -- CPU oracle for the brute-force boundary-element hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (baemm_amd/, include/)
 * links, loads or calls this file.  It is used by tests/, by
 * __graft_entry__.smoke() as the checker, and by bench.py's cpu_baseline leg.
 *
 * Two families of routines live here:
 *
 *   *_f64  : FP64 brute-force restatement of the operator mathematics
 *            (the "oracle" proper).  Pinned by the analytic checks in
 *            tests/test_oracle_*.py (sphere spectra, symmetry identities,
 *            Gauss law, Funk-Hecke, polar quadrature of the self term ...).
 *
 *   *_f32  : FP32 restatement that follows the reference OpenCL kernels'
 *            arithmetic and loop order (64-source blocks staged per
 *            block, a private row of 64 kernel values, then the
 *            contraction; per-wave-number-chunk outer loop; self term
 *            masked).  This is the stand-in for "the reference's own
 *            operator apply" because the reference cannot run here
 *            (no OpenCL device, Tensors/Tools submodules absent).
 *
 * Complex arrays are interleaved (re, im) doubles or floats, row-major,
 * with the leading dimension given in complex elements.
 *
 * Reference files cited relative to the BAEMM repository root.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_BLOCK 64 /* block_size default, MemberVariables.hpp:55 */

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_get_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Geometry                                                                  */
/* ------------------------------------------------------------------------ */

/* Self term of the single layer: (1/a) * int_T 1/|c - y| dy, c the centroid.
 * Edge formula of Initialize.hpp:249-302: for the sub-triangle (c, x_a, x_b)
 * with h the distance from c to the edge line and alpha, beta the angles at c
 * between the foot of the perpendicular and x_a, x_b:
 *     int 1/rho = h * (atanh(sin alpha) + atanh(sin beta)).               */
static double self_term_f64(const double x[3][3], double c[3], double area)
{
    double u[3][3];
    for (int k = 0; k < 3; ++k)
        for (int d = 0; d < 3; ++d) u[k][d] = x[k][d] - c[d];
    double sum = 0.0;
    for (int e = 0; e < 3; ++e) {
        /* edge (a, b) = (1,2), (2,0), (0,1) as in Initialize.hpp:269-300 */
        const int a = (e + 1) % 3, b = (e + 2) % 3;
        const double uaua = u[a][0]*u[a][0] + u[a][1]*u[a][1] + u[a][2]*u[a][2];
        const double ubub = u[b][0]*u[b][0] + u[b][1]*u[b][1] + u[b][2]*u[b][2];
        const double uaub = u[a][0]*u[b][0] + u[a][1]*u[b][1] + u[a][2]*u[b][2];
        const double La = sqrt(uaua), Lb = sqrt(ubub);
        const double p = uaua - uaub, q = ubub - uaub;
        const double vinv = 1.0 / sqrt(fabs(p + q));
        const double sa = p * vinv / La, sb = q * vinv / Lb;
        sum += La * sqrt(1.0 - sa * sa) * (atanh(sa) + atanh(sb));
    }
    return sum / area;
}

/* FP64 geometry: centroid, unit normal, area, single-layer self term.
 * Initialize.hpp:95-130 and :249-302.  T holds 0-based vertex indices. */
void orc_geometry_f64(int64_t nt, const double *V, const int64_t *T,
                      double *mid, double *nrm, double *area, double *sdiag)
{
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < nt; ++t) {
        double x[3][3], c[3];
        for (int k = 0; k < 3; ++k)
            for (int d = 0; d < 3; ++d) x[k][d] = V[3 * T[3 * t + k] + d];
        for (int d = 0; d < 3; ++d) c[d] = (x[0][d] + x[1][d] + x[2][d]) / 3.0;
        double e0[3], e1[3], an[3];
        for (int d = 0; d < 3; ++d) { e0[d] = x[1][d] - x[0][d]; e1[d] = x[2][d] - x[0][d]; }
        an[0] = 0.5 * (e0[1] * e1[2] - e0[2] * e1[1]);
        an[1] = 0.5 * (e0[2] * e1[0] - e0[0] * e1[2]);
        an[2] = 0.5 * (e0[0] * e1[1] - e0[1] * e1[0]);
        const double a = sqrt(an[0] * an[0] + an[1] * an[1] + an[2] * an[2]);
        for (int d = 0; d < 3; ++d) { mid[3 * t + d] = c[d]; nrm[3 * t + d] = an[d] / a; }
        area[t] = a;
        sdiag[t] = self_term_f64(x, c, a);
    }
}

/* FP32 geometry in the reference's arithmetic order: coordinates are cast to
 * float first (Helmholtz_OpenCL.hpp:187), then Initialize.hpp:95-130 and
 * :249-302 run in float. Outputs float4-padded midpoints/normals. */
void orc_geometry_f32(int64_t nt, const double *V, const int64_t *T,
                      float *mid4, float *nrm4, float *area, float *sdiag)
{
    const float third = 1.0f / 3.0f, half = 0.5f;
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < nt; ++t) {
        float x[3][3], c[3], u[3][3];
        for (int k = 0; k < 3; ++k)
            for (int d = 0; d < 3; ++d) x[k][d] = (float)V[3 * T[3 * t + k] + d];
        for (int d = 0; d < 3; ++d) c[d] = third * (x[0][d] + x[1][d] + x[2][d]);
        for (int k = 0; k < 3; ++k)
            for (int d = 0; d < 3; ++d) u[k][d] = x[k][d] - c[d];
        for (int d = 0; d < 3; ++d) mid4[4 * t + d] = c[d];
        mid4[4 * t + 3] = 0.0f;
        float df0[3], df1[3], an[3];
        for (int d = 0; d < 3; ++d) { df0[d] = x[1][d] - x[0][d]; df1[d] = x[2][d] - x[0][d]; }
        an[0] = half * (df0[1] * df1[2] - df0[2] * df1[1]);
        an[1] = half * (df0[2] * df1[0] - df0[0] * df1[2]);
        an[2] = half * (df0[0] * df1[1] - df0[1] * df1[0]);
        const float a = sqrtf(an[0] * an[0] + an[1] * an[1] + an[2] * an[2]);
        area[t] = a;
        const float ainv = 1.0f / a;
        for (int d = 0; d < 3; ++d) nrm4[4 * t + d] = ainv * an[d];
        nrm4[4 * t + 3] = 0.0f;
        float dot[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                dot[i][j] = u[i][0] * u[j][0] + u[i][1] * u[j][1] + u[i][2] * u[j][2];
        const float L[3] = {sqrtf(dot[0][0]), sqrtf(dot[1][1]), sqrtf(dot[2][2])};
        const float Linv[3] = {1.0f / L[0], 1.0f / L[1], 1.0f / L[2]};
        float sum = 0.0f;
        for (int e = 0; e < 3; ++e) {
            const int ia = (e + 1) % 3, ib = (e + 2) % 3;
            const float p = dot[ia][ia] - dot[ia][ib];
            const float q = dot[ib][ib] - dot[ia][ib];
            const float vinv = 1.0f / sqrtf(fabsf(p + q));
            const float sa = Linv[ia] * vinv * p;
            const float sb = Linv[ib] * vinv * q;
            sum += L[ia] * sqrtf(1.0f - sa * sa) * (atanhf(sa) + atanhf(sb));
        }
        sdiag[t] = sum / a;
    }
}

/* ------------------------------------------------------------------------ */
/* Pair kernels, FP64                                                        */
/* ------------------------------------------------------------------------ */

/* Boundary operator, off-diagonal part (formula A1; BoundaryOperatorKernel.cl
 * :104-193): for target i and source j != i, z = y_j - x_i, r = |z|,
 * E = exp(i kappa r):
 *   A_ij = c1 E / r + (c2 (z.n_j) - c3 (z.n_i)) (i kappa r - 1) E / r^3
 * C[row][k] = sum_j A_ij b[j][k] for k in chunk q = k / wcs.
 * coef holds K rows of 4 complex (c0..c3), already multiplied by 1/(4 pi)
 * (LoadParameters.hpp:65-79); c0 is unused here.
 * rows: target indices to evaluate (nrows of them); C is nrows x ldc.       */
void orc_bdr_f64(int64_t nt, const double *mid, const double *nrm,
                 int64_t W, int64_t wcs, const double *kappa, const double *coef,
                 const double *b, int64_t ldb, double *C, int64_t ldc,
                 const int64_t *rows, int64_t nrows)
{
    const int64_t K = W / wcs;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t r = 0; r < nrows; ++r) {
        const int64_t i = rows[r];
        double *Ci = C + 2 * ldc * r;
        for (int64_t k = 0; k < 2 * W; ++k) Ci[k] = 0.0;
        const double *xi = mid + 3 * i, *ni = nrm + 3 * i;
        for (int64_t j = 0; j < nt; ++j) {
            if (j == i) continue;
            const double *yj = mid + 3 * j, *nj = nrm + 3 * j;
            const double z0 = yj[0] - xi[0], z1 = yj[1] - xi[1], z2 = yj[2] - xi[2];
            const double r2 = z0 * z0 + z1 * z1 + z2 * z2;
            const double rr = sqrt(r2), rinv = 1.0 / rr;
            const double znj = z0 * nj[0] + z1 * nj[1] + z2 * nj[2];
            const double zni = z0 * ni[0] + z1 * ni[1] + z2 * ni[2];
            for (int64_t q = 0; q < K; ++q) {
                const double kap = kappa[q];
                const double *c = coef + 8 * q;
                const double cs = cos(kap * rr), sn = sin(kap * rr);
                const double er = cs * rinv, ei = sn * rinv;          /* E/r */
                /* (i kappa r - 1) E / r^3 */
                const double fr = (-er - kap * rr * ei) * rinv * rinv;
                const double fi = (kap * rr * er - ei) * rinv * rinv;
                /* A = c1 (er + i ei) + (c2 znj - c3 zni) (fr + i fi) */
                const double dr = c[4] * znj - c[6] * zni;
                const double di = c[5] * znj - c[7] * zni;
                const double Ar = c[2] * er - c[3] * ei + dr * fr - di * fi;
                const double Ai = c[2] * ei + c[3] * er + dr * fi + di * fr;
                const double *bj = b + 2 * (ldb * j + q * wcs);
                double *Cq = Ci + 2 * q * wcs;
                for (int64_t k = 0; k < wcs; ++k) {
                    const double br = bj[2 * k], bi = bj[2 * k + 1];
                    Cq[2 * k] += Ar * br - Ai * bi;
                    Cq[2 * k + 1] += Ar * bi + Ai * br;
                }
            }
        }
    }
}

/* Far-field map (formula A2; FarFieldOperatorKernel.cl:91-133):
 *   F[i][k] = sum_j (c1 - i kappa (xhat_i . n_j) c2) exp(-i kappa xhat_i . y_j) b[j][k]
 * coef as above (c1 = coef[2..3], c2 = coef[4..5]).                          */
void orc_far_f64(int64_t nt, const double *mid, const double *nrm,
                 int64_t m, const double *dirs,
                 int64_t W, int64_t wcs, const double *kappa, const double *coef,
                 const double *b, int64_t ldb, double *F, int64_t ldf)
{
    const int64_t K = W / wcs;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t i = 0; i < m; ++i) {
        double *Fi = F + 2 * ldf * i;
        for (int64_t k = 0; k < 2 * W; ++k) Fi[k] = 0.0;
        const double *x = dirs + 3 * i;
        for (int64_t j = 0; j < nt; ++j) {
            const double *y = mid + 3 * j, *n = nrm + 3 * j;
            const double xy = x[0] * y[0] + x[1] * y[1] + x[2] * y[2];
            const double xn = x[0] * n[0] + x[1] * n[1] + x[2] * n[2];
            for (int64_t q = 0; q < K; ++q) {
                const double kap = kappa[q];
                const double *c = coef + 8 * q;
                const double cs = cos(kap * xy), sn = -sin(kap * xy); /* exp(-i kappa x.y) */
                /* w = c1 - i kappa xn c2 */
                const double wr = c[2] + kap * xn * c[5];
                const double wi = c[3] - kap * xn * c[4];
                const double Ar = wr * cs - wi * sn, Ai = wr * sn + wi * cs;
                const double *bj = b + 2 * (ldb * j + q * wcs);
                double *Fq = Fi + 2 * q * wcs;
                for (int64_t k = 0; k < wcs; ++k) {
                    const double br = bj[2 * k], bi = bj[2 * k + 1];
                    Fq[2 * k] += Ar * br - Ai * bi;
                    Fq[2 * k + 1] += Ar * bi + Ai * br;
                }
            }
        }
    }
}

/* Potential at arbitrary points (formula A3; NearFieldOperatorKernel.cl:92-149):
 *   U[p][k] = sum_j (c1 E/r + c2 (z.n_j)(i kappa r - 1) E/r^3) b[j][k],  z = y_j - x_p */
void orc_near_f64(int64_t nt, const double *mid, const double *nrm,
                  int64_t P, const double *pts,
                  int64_t W, int64_t wcs, const double *kappa, const double *coef,
                  const double *b, int64_t ldb, double *U, int64_t ldu)
{
    const int64_t K = W / wcs;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t p = 0; p < P; ++p) {
        double *Up = U + 2 * ldu * p;
        for (int64_t k = 0; k < 2 * W; ++k) Up[k] = 0.0;
        const double *x = pts + 3 * p;
        for (int64_t j = 0; j < nt; ++j) {
            const double *y = mid + 3 * j, *n = nrm + 3 * j;
            const double z0 = y[0] - x[0], z1 = y[1] - x[1], z2 = y[2] - x[2];
            const double rr = sqrt(z0 * z0 + z1 * z1 + z2 * z2), rinv = 1.0 / rr;
            const double zn = z0 * n[0] + z1 * n[1] + z2 * n[2];
            for (int64_t q = 0; q < K; ++q) {
                const double kap = kappa[q];
                const double *c = coef + 8 * q;
                const double cs = cos(kap * rr), sn = sin(kap * rr);
                const double er = cs * rinv, ei = sn * rinv;
                const double fr = (-er - kap * rr * ei) * rinv * rinv;
                const double fi = (kap * rr * er - ei) * rinv * rinv;
                const double dr = c[4] * zn, di = c[5] * zn;
                const double Ar = c[2] * er - c[3] * ei + dr * fr - di * fi;
                const double Ai = c[2] * ei + c[3] * er + dr * fi + di * fr;
                const double *bj = b + 2 * (ldb * j + q * wcs);
                double *Uq = Up + 2 * q * wcs;
                for (int64_t k = 0; k < wcs; ++k) {
                    const double br = bj[2 * k], bi = bj[2 * k + 1];
                    Uq[2 * k] += Ar * br - Ai * bi;
                    Uq[2 * k + 1] += Ar * bi + Ai * br;
                }
            }
        }
    }
}

/* Herglotz kernel (formula A4; HerglotzWaveKernel.cl:91-144), P0 output:
 *   H[t][k] = sum_j (c1 - i kappa (n_t . d_j) c2) exp(-i kappa d_j . x_t) conj(g[j][k])
 * coef already carries the 4 pi / m factor (CreateHerglotzWave.hpp:40).       */
void orc_herglotz_f64(int64_t nt, const double *mid, const double *nrm,
                      int64_t m, const double *dirs,
                      int64_t W, int64_t wcs, const double *kappa, const double *coef,
                      const double *g, int64_t ldg, double *H, int64_t ldh)
{
    const int64_t K = W / wcs;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t t = 0; t < nt; ++t) {
        double *Ht = H + 2 * ldh * t;
        for (int64_t k = 0; k < 2 * W; ++k) Ht[k] = 0.0;
        const double *x = mid + 3 * t, *n = nrm + 3 * t;
        for (int64_t j = 0; j < m; ++j) {
            const double *d = dirs + 3 * j;
            const double dx = d[0] * x[0] + d[1] * x[1] + d[2] * x[2];
            const double dn = d[0] * n[0] + d[1] * n[1] + d[2] * n[2];
            for (int64_t q = 0; q < K; ++q) {
                const double kap = kappa[q];
                const double *c = coef + 8 * q;
                const double cs = cos(kap * dx), sn = -sin(kap * dx);
                const double wr = c[2] + kap * dn * c[5];
                const double wi = c[3] - kap * dn * c[4];
                const double Ar = wr * cs - wi * sn, Ai = wr * sn + wi * cs;
                const double *gj = g + 2 * (ldg * j + q * wcs);
                double *Hq = Ht + 2 * q * wcs;
                for (int64_t k = 0; k < wcs; ++k) {
                    const double gr = gj[2 * k], gi = -gj[2 * k + 1];
                    Hq[2 * k] += Ar * gr - Ai * gi;
                    Hq[2 * k + 1] += Ar * gi + Ai * gr;
                }
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Pair kernels, FP32 restatement of the reference OpenCL loop structure     */
/* ------------------------------------------------------------------------ */

/* BoundaryOperatorKernel.cl:16-238 restated on the CPU, one target per
 * iteration of the outer loop ("work-item").  flags bit0..5 =
 * Re/Im single, Re/Im double, Re/Im adjoint double (StringManipulation.hpp:
 * 7-17); kappa/coef in float as after LoadParameters.hpp:65-79.
 * mid4/nrm4: float4 padded.  b: rows >= nt are never read (the kernel masks
 * them through delta; the host zero-pads).  C: nt x W complex float.        */
void orc_bdr_cl_f32(int64_t nt, const float *mid4, const float *nrm4,
                    int64_t W, int64_t wcs, const float *kappa, const float *coef,
                    int flags, const float *b, float *C,
                    int64_t row_begin, int64_t row_end)
{
    const int64_t kcc = W / wcs;
    const int64_t nblk = (nt + ORC_BLOCK - 1) / ORC_BLOCK;
    const int reS = flags & 1, imS = (flags >> 1) & 1;
    const int reD = (flags >> 2) & 1, imD = (flags >> 3) & 1;
    const int reA = (flags >> 4) & 1, imA = (flags >> 5) & 1;
#pragma omp parallel
    {
        float *Ar = (float *)malloc(sizeof(float) * 2 * ORC_BLOCK);
        float *Ci = (float *)malloc(sizeof(float) * 2 * (size_t)wcs);
#pragma omp for schedule(dynamic, 16)
        for (int64_t i = row_begin; i < row_end; ++i) {
            const float x0 = mid4[4 * i], x1 = mid4[4 * i + 1], x2 = mid4[4 * i + 2];
            const float n0 = nrm4[4 * i], n1 = nrm4[4 * i + 1], n2 = nrm4[4 * i + 2];
            for (int64_t q = 0; q < kcc; ++q) {
                const float kap = kappa[q];
                const float *c = coef + 8 * q;
                for (int64_t k = 0; k < 2 * wcs; ++k) Ci[k] = 0.0f;
                for (int64_t jb = 0; jb < nblk; ++jb) {
                    for (int jl = 0; jl < ORC_BLOCK; ++jl) {
                        const int64_t j = ORC_BLOCK * jb + jl;
                        float ar = 0.0f, ai = 0.0f;
                        if (j < nt) {
                            const float delta = (float)(i == j);
                            const float z0 = mid4[4 * j] - x0, z1 = mid4[4 * j + 1] - x1,
                                        z2 = mid4[4 * j + 2] - x2;
                            const float r = sqrtf(z0 * z0 + z1 * z1 + z2 * z2);
                            const float rinv = (1.0f - delta) / (r + delta);
                            const float kr = kap * r;
                            const float cs = cosf(kr), sn = sinf(kr);
                            if (reS && imS) {
                                ar += (c[2] * cs - c[3] * sn) * rinv;
                                ai += (c[2] * sn + c[3] * cs) * rinv;
                            } else if (reS) {
                                ar += (c[2] * cs) * rinv;
                                ai += (c[2] * sn) * rinv;
                            } else if (imS) {
                                ar += (-c[3] * sn) * rinv;
                                ai += (c[3] * cs) * rinv;
                            }
                            if (reD || imD || reA || imA) {
                                const float r3 = rinv * rinv * rinv;
                                const float krc = kr * cs, krs = kr * sn;
                                const float a0 = -(krs + cs) * r3;
                                const float a1 = (krc - sn) * r3;
                                if (reD || imD) {
                                    const float f = z0 * nrm4[4 * j] + z1 * nrm4[4 * j + 1] +
                                                    z2 * nrm4[4 * j + 2];
                                    const float b0 = f * a0, b1 = f * a1;
                                    if (reD) { ar += b0 * c[4]; ai += b1 * c[4]; }
                                    if (imD) { ar -= b1 * c[5]; ai += b0 * c[5]; }
                                }
                                if (reA || imA) {
                                    const float f = -(z0 * n0 + z1 * n1 + z2 * n2);
                                    const float b0 = f * a0, b1 = f * a1;
                                    if (reA) { ar += b0 * c[6]; ai += b1 * c[6]; }
                                    if (imA) { ar -= b1 * c[7]; ai += b0 * c[7]; }
                                }
                            }
                        }
                        Ar[2 * jl] = ar;
                        Ar[2 * jl + 1] = ai;
                    }
                    const int64_t jmax = (nt - ORC_BLOCK * jb) < ORC_BLOCK ? (nt - ORC_BLOCK * jb) : ORC_BLOCK;
                    for (int jl = 0; jl < jmax; ++jl) {
                        const int64_t j = ORC_BLOCK * jb + jl;
                        const float *bj = b + 2 * (W * j + q * wcs);
                        const float pr = Ar[2 * jl], pi = Ar[2 * jl + 1];
                        for (int64_t k = 0; k < wcs; ++k) {
                            Ci[2 * k] += pr * bj[2 * k] - pi * bj[2 * k + 1];
                            Ci[2 * k + 1] += pr * bj[2 * k + 1] + pi * bj[2 * k];
                        }
                    }
                }
                float *Cq = C + 2 * (W * i + q * wcs);
                for (int64_t k = 0; k < 2 * wcs; ++k) Cq[k] = Ci[k];
            }
        }
        free(Ar);
        free(Ci);
    }
}

/* FarFieldOperatorKernel.cl:18-177 restated (FP32).  flags bits 0..3 =
 * Re/Im single, Re/Im double.  dirs4: float4 padded directions; F: m x W.   */
void orc_far_cl_f32(int64_t nt, const float *mid4, const float *nrm4,
                    int64_t m, const float *dirs4,
                    int64_t W, int64_t wcs, const float *kappa, const float *coef,
                    int flags, const float *b, float *F)
{
    const int64_t kcc = W / wcs;
    const int reS = flags & 1, imS = (flags >> 1) & 1;
    const int reD = (flags >> 2) & 1, imD = (flags >> 3) & 1;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t i = 0; i < m; ++i) {
        const float x0 = dirs4[4 * i], x1 = dirs4[4 * i + 1], x2 = dirs4[4 * i + 2];
        for (int64_t q = 0; q < kcc; ++q) {
            const float kap = kappa[q];
            const float *c = coef + 8 * q;
            float *Fq = F + 2 * (W * i + q * wcs);
            for (int64_t k = 0; k < 2 * wcs; ++k) Fq[k] = 0.0f;
            for (int64_t j = 0; j < nt; ++j) {
                const float kr = kap * (mid4[4 * j] * x0 + mid4[4 * j + 1] * x1 + mid4[4 * j + 2] * x2);
                const float cs = cosf(kr), sn = sinf(kr);
                float ar = 0.0f, ai = 0.0f;
                if (reS && imS) { ar += c[2] * cs + c[3] * sn; ai += -c[2] * sn + c[3] * cs; }
                else if (reS) { ar += c[2] * cs; ai -= c[2] * sn; }
                else if (imS) { ar += c[3] * sn; ai += c[3] * cs; }
                if (reD || imD) {
                    const float dk = kap * (nrm4[4 * j] * x0 + nrm4[4 * j + 1] * x1 + nrm4[4 * j + 2] * x2);
                    const float dkc = dk * cs, dks = dk * sn;
                    if (reD) { ar -= dks * c[4]; ai -= dkc * c[4]; }
                    if (imD) { ar += dkc * c[5]; ai -= dks * c[5]; }
                }
                const float *bj = b + 2 * (W * j + q * wcs);
                for (int64_t k = 0; k < wcs; ++k) {
                    Fq[2 * k] += ar * bj[2 * k] - ai * bj[2 * k + 1];
                    Fq[2 * k + 1] += ar * bj[2 * k + 1] + ai * bj[2 * k];
                }
            }
        }
    }
}
