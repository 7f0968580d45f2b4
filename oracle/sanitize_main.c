/* Driver of the AddressSanitizer / UndefinedBehaviorSanitizer build of the
 * oracle (test infrastructure only; built by `make -C oracle asan` into
 * oracle/_asan/).  Runs every oracle routine on the icosahedron (12 vertices,
 * 20 triangles) with uneven column counts, wave-number chunks, padded leading
 * dimensions and a row subset, so that an out-of-bounds index or undefined
 * arithmetic in orc.c stops the program.  Exit 0 when every output is finite. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

void orc_geometry_f64(int64_t nt, const double *V, const int64_t *T, double *mid, double *nrm, double *area,
                      double *sdiag);
void orc_geometry_f32(int64_t nt, const double *V, const int64_t *T, float *mid4, float *nrm4, float *area,
                      float *sdiag);
void orc_bdr_f64(int64_t nt, const double *mid, const double *nrm, int64_t W, int64_t wcs, const double *kappa,
                 const double *coef, const double *b, int64_t ldb, double *C, int64_t ldc, const int64_t *rows,
                 int64_t nrows);
void orc_far_f64(int64_t nt, const double *mid, const double *nrm, int64_t m, const double *dirs, int64_t W,
                 int64_t wcs, const double *kappa, const double *coef, const double *b, int64_t ldb, double *F,
                 int64_t ldf);
void orc_near_f64(int64_t nt, const double *mid, const double *nrm, int64_t P, const double *pts, int64_t W,
                  int64_t wcs, const double *kappa, const double *coef, const double *b, int64_t ldb, double *U,
                  int64_t ldu);
void orc_herglotz_f64(int64_t nt, const double *mid, const double *nrm, int64_t m, const double *dirs, int64_t W,
                      int64_t wcs, const double *kappa, const double *coef, const double *g, int64_t ldg, double *H,
                      int64_t ldh);
void orc_bdr_cl_f32(int64_t nt, const float *mid4, const float *nrm4, int64_t W, int64_t wcs, const float *kappa,
                    const float *coef, int flags, const float *b, float *C, int64_t row_begin, int64_t row_end);
void orc_far_cl_f32(int64_t nt, const float *mid4, const float *nrm4, int64_t m, const float *dirs4, int64_t W,
                    int64_t wcs, const float *kappa, const float *coef, int flags, const float *b, float *F);

static int finite_all(const double *a, size_t n)
{
    for (size_t k = 0; k < n; ++k)
        if (!isfinite(a[k])) return 0;
    return 1;
}

static int finite_all_f(const float *a, size_t n)
{
    for (size_t k = 0; k < n; ++k)
        if (!isfinite(a[k])) return 0;
    return 1;
}

int main(void)
{
    const double p = (1.0 + sqrt(5.0)) / 2.0;
    double V[36] = {-1, p, 0, 1, p, 0, -1, -p, 0, 1, -p, 0, 0, -1, p, 0, 1, p,
                    0, -1, -p, 0, 1, -p, p, 0, -1, p, 0, 1, -p, 0, -1, -p, 0, 1};
    const int64_t T[60] = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
                           3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    const int64_t nt = 20, W = 6, wcs = 3, K = 2, ld = 7, m = 5, P = 4;
    double mid[60], nrm[60], area[20], sdiag[20];
    float mid4[80], nrm4[80], areaf[20], sdiagf[20];
    orc_geometry_f64(nt, V, T, mid, nrm, area, sdiag);
    orc_geometry_f32(nt, V, T, mid4, nrm4, areaf, sdiagf);
    const double kappa[2] = {2.0, 3.5};
    double coef[16];
    for (int k = 0; k < 16; ++k) coef[k] = 0.1 * (k + 1) * ((k & 1) ? -1.0 : 1.0);
    double *b = malloc(sizeof(double) * 2 * nt * ld);
    for (int64_t k = 0; k < 2 * nt * ld; ++k) b[k] = cos(0.37 * (double)k);
    double *C = calloc(2 * nt * ld, sizeof(double));
    const int64_t rows[3] = {19, 0, 7};
    orc_bdr_f64(nt, mid, nrm, W, wcs, kappa, coef, b, ld, C, ld, rows, 3);
    double dirs[3 * 5], pts[3 * 4];
    for (int k = 0; k < 3 * m; ++k) dirs[k] = sin(1.1 * k);
    for (int i = 0; i < m; ++i) {
        const double r = sqrt(dirs[3 * i] * dirs[3 * i] + dirs[3 * i + 1] * dirs[3 * i + 1] +
                              dirs[3 * i + 2] * dirs[3 * i + 2]);
        for (int d = 0; d < 3; ++d) dirs[3 * i + d] /= r;
    }
    for (int k = 0; k < 3 * P; ++k) pts[k] = 3.0 * cos(0.9 * k + 0.3);
    double *F = calloc(2 * m * ld, sizeof(double));
    double *U = calloc(2 * P * ld, sizeof(double));
    double *H = calloc(2 * nt * ld, sizeof(double));
    orc_far_f64(nt, mid, nrm, m, dirs, W, wcs, kappa, coef, b, ld, F, ld);
    orc_near_f64(nt, mid, nrm, P, pts, W, wcs, kappa, coef, b, ld, U, ld);
    orc_herglotz_f64(nt, mid, nrm, m, dirs, W, wcs, kappa, coef, F, ld, H, ld);
    /* FP32 restatements: b and C are nt x W complex float, dense */
    float *bf = malloc(sizeof(float) * 2 * nt * W), *Cf = calloc(2 * nt * W, sizeof(float));
    float *Ff = calloc(2 * m * W, sizeof(float)), dirs4[4 * 5], kf[2] = {2.0f, 3.5f}, cf[16];
    for (int64_t k = 0; k < 2 * nt * W; ++k) bf[k] = (float)cos(0.21 * (double)k);
    for (int k = 0; k < 16; ++k) cf[k] = (float)coef[k];
    for (int i = 0; i < m; ++i) {
        for (int d = 0; d < 3; ++d) dirs4[4 * i + d] = (float)dirs[3 * i + d];
        dirs4[4 * i + 3] = 0.0f;
    }
    orc_bdr_cl_f32(nt, mid4, nrm4, W, wcs, kf, cf, 63, bf, Cf, 2, nt);
    orc_far_cl_f32(nt, mid4, nrm4, m, dirs4, W, wcs, kf, cf, 15, bf, Ff);
    int ok = finite_all(mid, 60) && finite_all(area, 20) && finite_all(sdiag, 20) && finite_all(C, 2 * nt * ld) &&
             finite_all(F, 2 * m * ld) && finite_all(U, 2 * P * ld) && finite_all(H, 2 * nt * ld) &&
             finite_all_f(mid4, 80) && finite_all_f(Cf, 2 * nt * W) && finite_all_f(Ff, 2 * m * W);
    (void)K;
    free(b);
    free(C);
    free(F);
    free(U);
    free(H);
    free(bf);
    free(Cf);
    free(Ff);
    printf("oracle sanitizer run: %s\n", ok ? "OK" : "NON-FINITE OUTPUT");
    return ok ? 0 : 1;
}
