/*
 * baemm.h -- C ABI of libbaemm_hip, the MI355X (gfx950) implementation of the
 * BAEMM brute-force boundary-element hot path.
 *
 * Each entry point replaces a method of the reference class
 * BAEMM::Helmholtz_OpenCL (header-only C++, OpenCL backend); the replaced
 * interface is cited as file:line relative to the BAEMM repository root.
 * The C++ drop-in class in Helmholtz_HIP.hpp restores the reference's
 * method names, templates and argument order on top of these functions.
 *
 * Conventions
 *  - Every function returns an int status (BAEMM_OK == 0).  On failure the
 *    handle keeps a message readable with baemm_last_error().  The reference
 *    prints and returns instead (eprint/wprint); callers that ignored those
 *    prints can ignore the status.
 *  - Complex arrays are interleaved (re, im) doubles ("complex<double>"),
 *    row-major (node-major, wave-minor) with an explicit leading dimension
 *    counted in complex elements, exactly as the reference's C_ext arrays.
 *  - alpha/beta are complex scalars passed as double[2].  Results are
 *    C = alpha * Op(B) + beta * C.  beta == 0 does not read C.
 *  - Host-pointer functions are synchronous.  *_dev functions take device
 *    pointers and a hipStream_t (passed as void*, NULL = the handle's stream)
 *    and return without synchronising.
 *  - A handle is bound to one GPU (device index) and serialises its calls.
 */
#ifndef BAEMM_H
#define BAEMM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct baemm_handle_s* baemm_handle;

enum {
    BAEMM_OK = 0,
    BAEMM_ERR_ARG = 1,      /* invalid argument (sizes, null pointers, W % wcs != 0, ...) */
    BAEMM_ERR_HIP = 2,      /* HIP runtime error */
    BAEMM_ERR_STATE = 3,    /* call order (e.g. apply before load) */
    BAEMM_ERR_NOCONV = 4,   /* an iterative solver stopped before its tolerance */
    BAEMM_ERR_INTERNAL = 5
};

/* Element types of complex device arrays. */
enum { BAEMM_C128 = 0, BAEMM_C64 = 1 };

/* Arithmetic of the triangle-pair contraction.  FP32: kernel values and
 * contraction in FP32 on the vector ALU.  MIXED: kernel values in FP32, the
 * contraction on MFMA as FP16 hi+lo split products with FP32 accumulation. */
enum { BAEMM_PRECISION_FP32 = 0, BAEMM_PRECISION_MIXED = 1 };

/* WaveType, BAEMM_Common.hpp:7-11 */
enum { BAEMM_WAVE_PLANE = 201, BAEMM_WAVE_RADIAL = 202 };

/* Library version (major * 10000 + minor * 100 + patch). */
int baemm_version(void);

/* ---------------------------------------------------------------- handle */

/* Helmholtz_OpenCL(coords, V, tris, T, meas, m, device_num, threads),
 * Helmholtz_OpenCL.hpp:175-249 (geometry: Initialize.hpp:3-356).
 * coords: V x 3, triangles: T x 3 zero-based, meas_directions: m x 3. */
int baemm_create(const double* coords, int64_t vertex_count,
                 const int64_t* triangles, int64_t simplex_count,
                 const double* meas_directions, int64_t meas_count,
                 int device, baemm_handle* out);

/* Same with 32-bit triangle indices (I_ext = int32). */
int baemm_create_i32(const double* coords, int64_t vertex_count,
                     const int32_t* triangles, int64_t simplex_count,
                     const double* meas_directions, int64_t meas_count,
                     int device, baemm_handle* out);

/* ~Helmholtz_OpenCL, Helmholtz_OpenCL.hpp:251-282. */
int baemm_destroy(baemm_handle h);

/* Message of the last failed call on h (empty if none). */
const char* baemm_last_error(baemm_handle h);

/* ------------------------------------------------------ getters / setters */

/* VertexCount / SimplexCount / GetMeasCount, GetSetters.hpp:8-16,69-72. */
int64_t baemm_vertex_count(baemm_handle h);
int64_t baemm_simplex_count(baemm_handle h);
int64_t baemm_meas_count(baemm_handle h);

/* UseDiagonal, GetSetters.hpp:115-118. */
int baemm_use_diagonal(baemm_handle h, int use_it);

/* Get/SetBlockSize, GetSetters.hpp:64-93.  Kept for interface parity; the HIP
 * kernels choose their own tiles, the value only changes what the getter returns. */
int baemm_set_block_size(baemm_handle h, int block_size);
int baemm_get_block_size(baemm_handle h);

/* Get/SetWaveChunkSize, GetSetters.hpp:47-57. */
int baemm_set_wave_chunk_size(baemm_handle h, int wave_chunk_size);
int baemm_get_wave_chunk_size(baemm_handle h);

/* Pair-contraction arithmetic (no reference counterpart: the reference has
 * only FP32).  Default BAEMM_PRECISION_FP32. */
int baemm_set_precision(baemm_handle h, int precision);
int baemm_get_precision(baemm_handle h);

/* Number of source segments the pair kernels split the sources into (tuning
 * knob, no reference counterpart); 0 selects automatically. */
int baemm_set_segments(baemm_handle h, int segments);

/* gmres_max_iter / gmres_max_restarts public members, MemberVariables.hpp:64-65. */
int baemm_set_gmres_limits(baemm_handle h, int max_iter, int max_restarts);

/* TriangleMidpoints / TriangleNormals (float4 padded, 4*T floats),
 * TriangleAreas, SingleLayerDiagonal (T floats), GetSetters.hpp:95-113.
 * Any pointer may be NULL. */
int baemm_get_triangle_data(baemm_handle h, float* midpoints4, float* normals4,
                            float* areas, float* single_diag);

/* MassMatrix(), GetSetters.hpp:18-21: CSR of the P1 mass matrix
 * (V x V, nnz = baemm_mass_nnz).  Any pointer may be NULL. */
int64_t baemm_mass_nnz(baemm_handle h);
int baemm_get_mass_matrix(baemm_handle h, int32_t* rowptr, int32_t* colidx, double* values);

/* MassMatrix().Dot<NRHS>(alpha, X, ldX, beta, Y, ldY, nrhs) on real columns,
 * as called at Example_GaussNewton/main.cpp:395-399 and FarField.hpp:567. */
int baemm_mass_dot(baemm_handle h, double alpha, const double* X, int64_t ldX,
                   double beta, double* Y, int64_t ldY, int64_t nrhs);

/* ----------------------------------------------------------- parameters */

/* LoadParameters(kappa, coeff, W, wcs, factor), LoadParameters.hpp:198-259:
 * kappa[K], coeff[K x 4] complex, K = ceil(W / wcs); c_j (j >= 1) are scaled
 * by factor / (4 pi).  W must be a multiple of wcs (the reference kernels
 * silently skip a partial last chunk, BoundaryOperatorKernel.cl:17). */
int baemm_load_parameters(baemm_handle h, const double* kappa, const double* coeff,
                          int64_t wave_count, int64_t wave_chunk_size, double factor);

/* GetWaveNumbers / GetCoefficients / GetWaveCount, GetSetters.hpp:35-82:
 * kappa[K] and the scaled coefficients c[K x 4] (complex float interleaved).
 * Any pointer may be NULL. */
int baemm_get_parameters(baemm_handle h, float* kappa, float* coeff,
                         int64_t* wave_count, int64_t* wave_chunk_size, int64_t* chunk_count);

/* LoadBoundaryOperators_PL / UnloadBoundaryOperators_PL,
 * ApplyBoundaryOperators.hpp:48-71 (factor 1). */
int baemm_load_boundary_operators(baemm_handle h, const double* kappa, const double* coeff,
                                  int64_t wave_count, int64_t wave_chunk_size);
int baemm_unload_boundary_operators(baemm_handle h);

/* --------------------------------------------------------------- applies */

/* ApplyBoundaryOperators_PL(alpha, B, ldB, beta, C, ldC) with loaded
 * parameters, ApplyBoundaryOperators.hpp:77-178:
 * C = alpha * [sum_q c_q0 M + AvT (K_q + D_q) Av] B + beta * C   (V x W). */
int baemm_apply_boundary(baemm_handle h, const double* alpha, const double* B, int64_t ldB,
                         const double* beta, double* C, int64_t ldC);

/* ApplyFarFieldOperators_PL with loaded parameters, ApplyFarFieldOperators.hpp:49-104:
 * F = alpha * FF(Av B) + beta * F  (m x W). */
int baemm_apply_far_field(baemm_handle h, const double* alpha, const double* B, int64_t ldB,
                          const double* beta, double* F, int64_t ldF);

/* ApplyNearFieldOperators_PL with loaded parameters, ApplyNearFieldOperators.hpp:160-226:
 * U = alpha * NF(Av B)(pts) + beta * U  (P x W); pts: P x 3.
 * ldU is honoured (the reference uses its internal ldC there, Q5). */
int baemm_apply_near_field(baemm_handle h, const double* alpha, const double* B, int64_t ldB,
                           const double* beta, double* U, int64_t ldU,
                           const double* pts, int64_t P);

/* CreateHerglotzWave_PL with loaded parameters (load with factor 4 pi / m),
 * CreateHerglotzWave.hpp:177-222: C = alpha * AvT Herg(conj g) + beta * C. */
int baemm_create_herglotz_wave(baemm_handle h, const double* alpha, const double* g, int64_t ldg,
                               const double* beta, double* C, int64_t ldC);

/* CreateIncidentWave_PL with loaded parameters (load with factor 4 pi),
 * CreateIncidentWave.hpp:55-127: C = alpha * AvT U + beta * C;
 * dirs: inc_count x 3 (>= wave_chunk_size rows used), wave_type 201/202. */
int baemm_create_incident_wave(baemm_handle h, const double* alpha, const double* dirs,
                               int64_t inc_count, const double* beta, double* C, int64_t ldC,
                               int wave_type);

/* Device-pointer variants.  B_type / C_type: BAEMM_C128 or BAEMM_C64. */
int baemm_apply_boundary_dev(baemm_handle h, const double* alpha, const void* B, int64_t ldB, int B_type,
                             const double* beta, void* C, int64_t ldC, int C_type, void* stream);
int baemm_apply_far_field_dev(baemm_handle h, const double* alpha, const void* B, int64_t ldB, int B_type,
                              const double* beta, void* F, int64_t ldF, int F_type, void* stream);

/* ------------------------------------------------------------- timing */

/* Records HIP events around every launch of the dominant pair kernel on the
 * stream it runs on; baemm_get_timing returns the summed milliseconds and the
 * number of launches since the last reset (no reference counterpart). */
int baemm_set_timing(baemm_handle h, int on);
int baemm_get_timing(baemm_handle h, double* total_ms, int64_t* launches);

/* Device memory helpers for the *_dev entry points. */
int baemm_dev_alloc(void** ptr, size_t bytes);
int baemm_dev_free(void* ptr);
int baemm_dev_upload(void* dst, const void* src, size_t bytes);
int baemm_dev_download(void* dst, const void* src, size_t bytes);
int baemm_dev_synchronize(void);

#ifdef __cplusplus
}
#endif

#endif /* BAEMM_H */
