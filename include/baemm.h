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
 *  - A handle is bound to one GPU (device index) and serialises its calls.  A
 *    call on a handle from inside a callback that a call on the same handle is
 *    running (the M and P of baemm_gauss_newton_solve) returns BAEMM_ERR_STATE
 *    (message in baemm_last_error(NULL)); such callbacks use another handle.
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
enum { BAEMM_PRECISION_FP32 = 0, BAEMM_PRECISION_MIXED = 1, BAEMM_PRECISION_FP32_MFMA = 2 };

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

/* Message of the last failed call on h (empty if none).  With h == NULL: the
 * message of the last failed baemm_create* on the calling thread. */
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
 * only FP32).  Kernel values are generated in FP32 in every mode; the
 * contraction with the staged sources runs
 *   BAEMM_PRECISION_FP32:      in FP32 on the vector ALUs (the default),
 *   BAEMM_PRECISION_MIXED:     on MFMA as FP16 hi+lo products with FP32
 *                              accumulation (within 1e-6 of FP32),
 *   BAEMM_PRECISION_FP32_MFMA: on MFMA in FP32 (v_mfma_f32_32x32x2_f32) with
 *                              FP32 accumulation.
 * The MFMA modes apply to the boundary operators and the potential
 * (baemm_apply_near_field); the far-field and Herglotz kernels stay FP32.
 * baemm_apply_boundary_batch takes FP32 and MIXED instances only. */
int baemm_set_precision(baemm_handle h, int precision);
int baemm_get_precision(baemm_handle h);

/* Number of source segments the pair kernels split the sources into (tuning
 * knob, no reference counterpart); 0 selects automatically. */
int baemm_set_segments(baemm_handle h, int segments);

/* No reference counterpart: the MFMA kernel variant of the mixed precision
 * for the boundary operators.
 * 0 = one wave generates its kernel values and issues their MFMAs;
 * 1 = wave-specialised blocks (generation waves feed MFMA waves through LDS);
 * 2 = automatic, the default: the single-wave kernel, the wave-specialised one
 * for operators with both the double and the adjoint double layer.
 * With the same source segments (baemm_set_segments) the variants give
 * identical results; their automatic segment counts differ (results within
 * rounding).  Operators whose chunks all have the adjoint double layer and no
 * double layer (the Dirichlet-to-Neumann rows) take a faster form of their own
 * on every MFMA kernel.
 * The FP32 MFMA precision always uses the wave-specialised kernel; the
 * potential (baemm_apply_near_field) follows the same choice. */
int baemm_set_mixed_kernel(baemm_handle h, int variant);

/* gmres_max_iter / gmres_max_restarts public members, MemberVariables.hpp:64-65. */
int baemm_set_gmres_limits(baemm_handle h, int max_iter, int max_restarts);

/* No reference counterpart: capture each GMRES Arnoldi step (operator apply,
 * PCG preconditioner, Gram-Schmidt, normalisation) in a hipGraph and replay
 * it (on != 0), instead of launching its ~100 kernels one by one.  A graph is
 * kept per Krylov index and re-captured when the loaded operator, precision
 * or workspaces change.  Results are identical either way.  Default on. */
int baemm_set_graphs(baemm_handle h, int on);
int baemm_get_graph_stats(baemm_handle h, int64_t* captures, int64_t* launches);

/* TriangleMidpoints / TriangleNormals (float4 padded, 4*T floats),
 * TriangleAreas, SingleLayerDiagonal (T floats), GetSetters.hpp:95-113.
 * Any pointer may be NULL. */
int baemm_get_triangle_data(baemm_handle h, float* midpoints4, float* normals4,
                            float* areas, float* single_diag);

/* VertexCoordinates / TriangleCoordinates, Helmholtz_OpenCL.hpp:333-341 (FP32):
 * vertex_coords V x 3; triangle_coords T x 12 (the three vertices of each
 * triangle as x, y, z, 0, Initialize.hpp:82-93).  Any pointer may be NULL. */
int baemm_get_coordinates(baemm_handle h, float* vertex_coords, float* triangle_coords);

/* DeviceInfo, Helmholtz_OpenCL.hpp:343-391: vendor, name, architecture,
 * driver and memory sizes of the handle's GPU as text (NUL-terminated,
 * truncated to size bytes). */
int baemm_device_info(baemm_handle h, char* buf, int64_t size);

/* PrintBooleans, LoadParameters.hpp:85-95: the eight flags of the loaded
 * parameters, flags[2j] = Re and flags[2j+1] = Im of coefficient column j
 * (mass matrix, single layer, double layer, adjoint double layer). */
int baemm_get_flags(baemm_handle h, int* flags);

/* ReadB / WriteB / ReadC / WriteC, InputOutput.hpp:30-89: the reference's
 * P0 buffers, simplex_count x wave_count complex float (interleaved, the
 * reference's internal Complex), rows = triangles.  B holds the triangle
 * values b = Av B_in of the last operator apply (boundary, far field or
 * potential), C the triangle values (K + D) b of the last boundary-operator
 * apply before AvT (ApplyBoundaryOperators.hpp:120-140), also when the apply
 * ran inside a solver.  Read* overwrite the handle's copy (ld_input >=
 * wave_count) until the next apply replaces it, as in the reference.
 * Write* return BAEMM_ERR_STATE while the buffer has no value: B before the
 * first apply; C before the first boundary apply or after a Herglotz wave,
 * incident wave, batched or no-allocation apply rewrote the workspaces. */
int baemm_read_b(baemm_handle h, const float* input, int64_t ld_input, int64_t wave_count);
int baemm_write_b(baemm_handle h, float* output, int64_t ld_output);
int baemm_read_c(baemm_handle h, const float* input, int64_t ld_input, int64_t wave_count);
int baemm_write_c(baemm_handle h, float* output, int64_t ld_output);

/* B(i,k) / C(i,k), InputOutput.hpp:5-25: the handle's host copy of buffer
 * `which` (0 = B, 1 = C), brought up to date with the device, as *data
 * (complex float interleaved, element (i, k) at 2 (ld i + k)); valid until
 * the next call on h.  Same states as baemm_write_b/_c. */
int baemm_p0_buffer(baemm_handle h, int which, float** data, int64_t* ld);

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

/* No reference counterpart: autotuner of the single-instance apply.  Times the
 * loaded boundary operator (current precision) with source-segment counts
 * giving launches of about 512..8192 blocks, on synthetic data in the handle's
 * own buffers, and keeps the fastest (baemm_set_segments); later applies and
 * GMRES solves with this handle use it. */
int baemm_autotune_segments(baemm_handle h, int repeats, int* best_segments, double* best_ms);

/* Device-pointer variants.  B_type / C_type: BAEMM_C128 or BAEMM_C64. */
int baemm_apply_boundary_dev(baemm_handle h, const double* alpha, const void* B, int64_t ldB, int B_type,
                             const double* beta, void* C, int64_t ldC, int C_type, void* stream);
int baemm_apply_far_field_dev(baemm_handle h, const double* alpha, const void* B, int64_t ldB, int B_type,
                              const double* beta, void* F, int64_t ldF, int F_type, void* stream);

/* No reference counterpart: baemm_apply_boundary_dev for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and results; it calls no
 * hipMalloc/hipFree: every workspace of the apply lives in one fixed device
 * array of baemm_noalloc_scratch_bytes() bytes (256 MiB) that each GPU holds
 * from the first call on it until the process ends.  An apply that needs more
 * (mixed precision at 64 columns: about 1.5 KB per triangle plus 0.5 KB
 * per triangle and source segment, so meshes up to roughly 50,000 triangles)
 * fails with BAEMM_ERR_ARG before launching anything.  All calls on one GPU share that
 * array, so calls that overlap on one GPU (other streams or threads) race;
 * the array keeps the last call's contents.  The first mixed-precision call
 * of a handle synchronises its stream once (minimum centroid distance). */
int baemm_apply_boundary_dev_noalloc(baemm_handle h, const double* alpha, const void* B, int64_t ldB, int B_type,
                                     const double* beta, void* C, int64_t ldC, int C_type, void* stream);
/* Size of the fixed scratch array of baemm_apply_boundary_dev_noalloc. */
int64_t baemm_noalloc_scratch_bytes(void);

/* ------------------------------------------------------------ batching
 * No reference counterpart.  Applies the loaded boundary operators of n
 * handles (different meshes, wave numbers and W) in one persistent launch per
 * arithmetic mode: instances whose precision is BAEMM_PRECISION_MIXED share
 * one MFMA launch, FP32 instances one FP32 launch.  Device pointers;
 * alpha/beta are 2n doubles (re, im per instance) or NULL (1 and 0).  Each
 * handle may appear once; all on one device; the call is asynchronous on
 * `stream` (NULL: the first handle's stream).  An instance's result does not
 * depend on the other instances of the batch. */
int baemm_apply_boundary_batch(int n, const baemm_handle* handles, const double* alpha, const void* const* B,
                               const int64_t* ldB, int B_type, const double* beta, void* const* C,
                               const int64_t* ldC, int C_type, void* stream);

/* Tile shapes of the persistent launches: source range per work item (MFMA:
 * 8-source steps, >= 8; FP32: 64-source blocks, >= 1) and blocks per CU of
 * the persistent grid (0: occupancy limit).  Process-wide. */
int baemm_batch_set_tuning(int mixed_steps, int plain_blocks, int mixed_grid_mult, int plain_grid_mult);
int baemm_batch_get_tuning(int* mixed_steps, int* plain_blocks, int* mixed_grid_mult, int* plain_grid_mult);

/* Which instances skip the persistent launch (no reference counterpart): an
 * instance whose own pair launch fills at least min_rounds rounds of resident
 * work-groups runs that launch on its handle's stream, beside the others
 * (default 1; 0: every instance, i.e. the schedule of separate applies).  The
 * autotuner sets it too.  An instance's result depends on it only through the
 * source segments its partial sums are split into (rounding). */
int baemm_batch_set_own_launch(double min_rounds);
int baemm_batch_get_own_launch(double* min_rounds);

/* Work items and persistent grid size of the last batched call. */
int baemm_batch_get_stats(int64_t* items_mixed, int64_t* items_fp32, int64_t* grid_mixed, int64_t* grid_fp32);

/* Autotuner: times the candidate tile shapes on this batch (C is overwritten)
 * and keeps the fastest as the process-wide tuning; *best_ms receives its time. */
int baemm_batch_autotune(int n, const baemm_handle* handles, const void* const* B, const int64_t* ldB, int B_type,
                         void* const* C, const int64_t* ldC, int C_type, int repeats, double* best_ms);

/* ----------------------------------------------------------- mass inverse */

/* ApplyMassInverse<NRHS>(X, ldX, Y, ldY, cg_tol, nrhs), ApplyMassInverse.hpp:6-156:
 * Y = M^{-1} X by preconditioned CG (lumped-diagonal preconditioner,
 * at most 20 iterations, relative residual cg_tol per column); real columns. */
int baemm_apply_mass_inverse(baemm_handle h, const double* X, int64_t ldX, double* Y, int64_t ldY,
                             int64_t nrhs, double cg_tol);

/* Same for complex columns (re and im solved as 2*nrhs real columns,
 * ApplyMassInverse.hpp:71-89); ldX, ldY in complex elements. */
int baemm_apply_mass_inverse_complex(baemm_handle h, const double* X, int64_t ldX, double* Y,
                                     int64_t ldY, int64_t nrhs, double cg_tol);

/* ApplyLumpedMassInverse, ApplyMassInverse.hpp:161-197 (real columns). */
int baemm_apply_lumped_mass_inverse(baemm_handle h, const double* X, int64_t ldX, double* Y,
                                    int64_t ldY, int64_t nrhs);

/* ----------------------------------------------------------------- solvers
 * kappa[K] wave numbers, wcs = wave_chunk_size, W = K * wcs columns; vertex
 * arrays are V x W complex (ld = W), far-field arrays m x W complex.  The
 * boundary integral equations are solved by left-preconditioned restarted
 * GMRES on the device (restart and cycle limits: baemm_set_gmres_limits,
 * default 30 and 3, MemberVariables.hpp:64-65) with the mass inverse (PCG,
 * cg_tol) as preconditioner.  A solve that stops above its tolerance still
 * writes its result and returns BAEMM_ERR_NOCONV (the reference discards the
 * GMRES flag, FarField.hpp:764). */

/* BoundaryPotential / BoundaryPotential_parameters, FarField.hpp:626-771:
 * (1/2 M - i eta S + K) phi = wave.  eta may be NULL (eta = kappa); coeff
 * (K x 4 complex, may be NULL) receives the coefficient rows. */
int baemm_boundary_potential(baemm_handle h, const double* kappa, int64_t K, const double* eta,
                             double* coeff, const double* wave, double* phi, int64_t wcs,
                             double cg_tol, double gmres_tol);

/* DirichletToNeumann, FarField.hpp:790-879: (1/2 M - i kappa S + K') psi = wave
 * (wave in weak form, psi strong). */
int baemm_dirichlet_to_neumann(baemm_handle h, const double* kappa, int64_t K, const double* wave,
                               double* neumann_trace, int64_t wcs, double cg_tol, double gmres_tol);

/* FarField<WC>, FarField.hpp:28-70, and FarField_parameters (:90-144, eta may
 * be NULL).  dirs: wcs x 3 incident directions (plane) or point sources
 * (radial); Y: m x W far field. */
int baemm_far_field(baemm_handle h, const double* kappa, int64_t K, const double* dirs, int64_t wcs,
                    double* Y, int wave_type, double cg_tol, double gmres_tol);
int baemm_far_field_parameters(baemm_handle h, const double* kappa, int64_t K, const double* dirs,
                               int64_t wcs, double* Y, int wave_type, const double* eta,
                               double cg_tol, double gmres_tol);

/* Derivative_FF_Helper, FarField.hpp:148-207: du_dn (V x W) = Neumann data
 * of the total field. */
int baemm_derivative_ff_helper(baemm_handle h, const double* kappa, int64_t K, const double* dirs,
                               int64_t wcs, double* du_dn, int wave_type, double cg_tol,
                               double gmres_tol);

/* Derivative_FF, FarField.hpp:227-321: Y_out (m x W) = DF X_in, X_in V x 3.
 * du_dn is the reference's C_ext*& argument: if *du_dn is NULL it is
 * allocated with calloc (free it with free()) and filled by the helper. */
int baemm_derivative_ff(baemm_handle h, const double* kappa, int64_t K, const double* dirs,
                        int64_t wcs, const double* X_in, double* Y_out, double** du_dn,
                        int wave_type, double cg_tol, double gmres_tol);

/* AdjointDerivative_FF, FarField.hpp:342-428: Y_out (V x 3) = DF* g, g m x W. */
int baemm_adjoint_derivative_ff(baemm_handle h, const double* kappa, int64_t K, const double* dirs,
                                int64_t wcs, const double* g, double* Y_out, double** du_dn,
                                int wave_type, double cg_tol, double gmres_tol);

/* Caller-supplied real operator y = Op(x) on V x 3 arrays (ld 3), the M and P
 * function handles of GaussNewtonSolve. */
typedef void (*baemm_real_op)(void* user, const double* x, double* y);

/* GaussNewtonSolve, FarField.hpp:457-604: X = alpha (M + Mass DF* DF)^{-1} B
 * + beta X by left-preconditioned real GMRES (preconditioner P) with tolerance
 * gmres_tol_outer; the inner solves use gmres_tol_inner.  *succeeded receives
 * the outer GMRES flag (the reference's return value). */
int baemm_gauss_newton_solve(baemm_handle h, const double* kappa, int64_t K, const double* dirs,
                             int64_t wcs, baemm_real_op M, void* M_user, baemm_real_op P, void* P_user,
                             double alpha, const double* B, int64_t ldB, double beta, double* X,
                             int64_t ldX, double** du_dn, int wave_type, double cg_tol,
                             double gmres_tol_inner, double gmres_tol_outer, int* succeeded);

/* Device-side GaussNewtonSolve (no reference counterpart: it replaces the
 * host callbacks of Example_GaussNewton/main.cpp:326-346).  The same solve as
 * baemm_gauss_newton_solve, with every vector in HBM: the outer GMRES keeps its
 * Krylov basis on the device, and M and P are called with device arrays
 * (V x 3 doubles, ld 3) and the stream the solve runs on, so a metric written
 * for the GPU runs without host copies.  B (ld ldB >= 3), X (ld ldX >= 3):
 * device arrays.  du_dn: device array V x W complex double; when du_dn_ready
 * is 0 it is filled (Derivative_FF_Helper), else used as is; NULL: computed
 * into a buffer of the handle.  stream NULL: the handle's stream.
 * Synchronous. */
typedef void (*baemm_dev_real_op)(void* user, const double* x, double* y, void* stream);
int baemm_gauss_newton_solve_dev(baemm_handle h, const double* kappa, int64_t K, const double* dirs,
                                 int64_t wcs, baemm_dev_real_op M, void* M_user, baemm_dev_real_op P,
                                 void* P_user, double alpha, const double* B, int64_t ldB, double beta,
                                 double* X, int64_t ldX, void* du_dn, int du_dn_ready, int wave_type,
                                 double cg_tol, double gmres_tol_inner, double gmres_tol_outer,
                                 int* succeeded, void* stream);

/* A metric for baemm_gauss_newton_solve_dev that runs on the device (no
 * reference counterpart; Repulsor's TangentPointMetric0, the metric of
 * Example_GaussNewton/main.cpp:301-346, is not part of BAEMM):
 *   M x = c x + d (Mass (x) I_3) x,   P x = x / (c + d m_v) row by row,
 * with m_v the row sums of the P1 mass matrix (a Jacobi preconditioner).
 * user points to a baemm_mass_metric whose handle is the one being solved on;
 * the functions are meant to be passed to baemm_gauss_newton_solve_dev, which
 * calls them with that handle's lock held. */
typedef struct {
    baemm_handle h;
    double c, d;
} baemm_mass_metric;
void baemm_mass_metric_apply(void* user, const double* x, double* y, void* stream);
void baemm_mass_metric_jacobi(void* user, const double* x, double* y, void* stream);

/* The caller-side energy and metric of the Gauss-Newton example
 * (Example_GaussNewton/main.cpp:301-346 builds them with Repulsor's
 * TangentPointEnergy0 and TangentPointMetric0, which are not part of BAEMM),
 * on the handle's mesh, restated from Yu, Brakensiek, Schumacher & Crane,
 * "Repulsive Surfaces" (2021) with centroid quadrature over all pairs of
 * distinct triangles S != T:
 *   E = sum a_S a_T |n_S . (c_S - c_T)|^q / |c_S - c_T|^p          (q > 1, p > 0)
 *   <<u, v>> = sum w_ST <du_S - du_T, dv_S - dv_T>_F
 *            + sum w_ST k_ST <ubar_S - ubar_T, vbar_S - vbar_T>,
 *   w_ST = a_S a_T |c_S - c_T|^(-2 sigma - 2), sigma = (p - 2)/q - 1,
 *   k_ST = (n_S . (c_S - c_T))^2 / |c_S - c_T|^4,
 * du_S the face gradient and ubar_S the face average of a P1 vector function.
 * Pair terms in FP32, sums in FP64.
 *
 * baemm_tp_energy: E and (gradient non-NULL) dE/dx, vertex_count x 3 doubles,
 * on the host.  Synchronous. */
int baemm_tp_energy(baemm_handle h, double q, double p, double* energy, double* gradient);

/* The metric as a device callback for baemm_gauss_newton_solve_dev:
 * M x = scale * (matrix of <<., .>>) x on V x 3 device arrays, and the Jacobi
 * preconditioner P x = x / (scale * d_v) row by row, d_v the diagonal of that
 * matrix (equal for the three components of vertex v).
 * baemm_tp_metric_prepare allocates the callbacks' buffers and computes the
 * diagonal for (q, p); call it first (a callback that fails
 * cannot return a status: it fills y with NaN and leaves a message in
 * baemm_last_error).  Calls on one handle must not overlap, as for the
 * mass metric above.  baemm_tp_metric_apply_host applies M to host arrays. */
typedef struct {
    baemm_handle h;
    double q, p, scale;
} baemm_tp_metric;
int baemm_tp_metric_prepare(baemm_handle h, double q, double p);
void baemm_tp_metric_apply(void* user, const double* x, double* y, void* stream);
void baemm_tp_metric_jacobi(void* user, const double* x, double* y, void* stream);
int baemm_tp_metric_apply_host(baemm_handle h, double q, double p, double scale, const double* x, double* y);
/* The Jacobi preconditioner on host arrays: y = x / (scale * d_v) row by row
 * (scale finite and nonzero). */
int baemm_tp_metric_jacobi_host(baemm_handle h, double q, double p, double scale, const double* x, double* y);

/* No reference counterpart: the GMRES of the solvers above on the loaded
 * boundary operator A: solves (P A) X = P B (P = mass inverse by PCG, cg_tol)
 * to gmres_tol.  B, X: device arrays, vertex_count x W complex double (ld W).
 * *iterations (may be NULL) receives the Arnoldi steps.  Synchronous. */
int baemm_gmres_dev(baemm_handle h, const void* B, void* X, double cg_tol, double gmres_tol, int64_t* iterations,
                    void* stream);

/* No reference counterpart: baemm_gmres_dev for n handles solved together
 * (different meshes, wave numbers and W; precision FP32 or MIXED; equal
 * gmres limits).  Each Arnoldi step applies the operators of every system
 * still iterating in one batched persistent launch (as
 * baemm_apply_boundary_batch); the preconditioner, Gram-Schmidt and Givens
 * steps of the systems run on their handles' streams and overlap.
 * iterations[k], converged[k] (either may be NULL) per system; returns
 * BAEMM_ERR_NOCONV if any system stopped above gmres_tol.  Synchronous. */
int baemm_gmres_batch(int n, const baemm_handle* handles, const void* const* B, void* const* X, double cg_tol,
                      double gmres_tol, int64_t* iterations, int* converged, void* stream);

/* Counters of the solvers since the last reset (no reference counterpart). */
int baemm_get_solver_stats(baemm_handle h, int64_t* gmres_iterations, int64_t* gmres_cycles,
                           double* last_rel_residual, int64_t* cg_iterations,
                           int64_t* operator_applies);
int baemm_reset_solver_stats(baemm_handle h);

/* ------------------------------------------------------------- serving
 * No reference counterpart.  A request loop for Gauss-Newton-step requests:
 * forward solve + far field (FarField), Neumann data (Derivative_FF_Helper)
 * and the adjoint derivative of the misfit g = F - F_obs
 * (AdjointDerivative_FF).  `slots` worker threads run requests concurrently,
 * each on its own handle and HIP stream; a queued request starts when a slot
 * is free and its estimated device memory fits into `hbm_budget_bytes` next
 * to what is resident.  Geometry and solver workspaces of meshes with a
 * mesh_id >= 0 stay resident for later requests on the same mesh (evicted
 * least-recently-used when memory is needed).  A request's results do not
 * depend on the other requests in flight. */
typedef struct baemm_server_s* baemm_server;

typedef struct {
    const double* coords;          /* vertex_count x 3 */
    int64_t vertex_count;
    const int64_t* triangles;      /* simplex_count x 3 */
    int64_t simplex_count;
    int64_t mesh_id;               /* >= 0: geometry may be cached under this id */
    const double* meas_directions; /* meas_count x 3 */
    int64_t meas_count;
    const double* kappa;           /* kappa_count wave numbers */
    int64_t kappa_count;
    const double* inc_directions;  /* wave_chunk_size x 3 */
    int64_t wave_chunk_size;       /* W = kappa_count * wave_chunk_size <= 64 */
    const double* far_field_data;  /* meas_count x W complex, or NULL (zero data) */
    int wave_type;                 /* BAEMM_WAVE_PLANE (0 means plane) or BAEMM_WAVE_RADIAL */
    int precision;                 /* BAEMM_PRECISION_FP32 or BAEMM_PRECISION_MIXED */
    double cg_tol;                 /* <= 0: 1e-10 */
    double gmres_tol;              /* <= 0: 1e-6 */
} baemm_gn_request;

typedef struct {
    int64_t completed, failed, queued, running, max_in_flight;
    int64_t resident_bytes, peak_bytes, gmres_iterations;
    double p50_ms, p99_ms, mean_ms;   /* latency from submission to completion */
    double solves_per_s;              /* completed / (last completion - first submission) */
} baemm_server_stats_t;

int baemm_server_create(int device, int slots, int64_t hbm_budget_bytes, baemm_server* out);
/* Admission order (no reference counterpart).  BAEMM_SERVE_FIFO (the default):
 * the oldest queued request; one that does not fit holds back the later ones.
 * BAEMM_SERVE_SHORTEST_FIRST: the queued request with the least estimated work
 * (N_t^2 * W) that fits, except that a request passed over max_skips times
 * (>= 0) goes first and holds back the others until it fits, so none waits
 * forever.  Applies to admissions after the call. */
enum { BAEMM_SERVE_FIFO = 0, BAEMM_SERVE_SHORTEST_FIRST = 1 };
int baemm_server_set_policy(baemm_server srv, int policy, int max_skips);
/* Fused mode (on != 0; off by default): instead of running each request on
 * its own worker, the loop gathers the queued requests that fit (up to
 * `slots`) into a batch and runs them phase by phase, each of the batch's
 * three boundary-integral solves one batched GMRES (baemm_gmres_batch: every
 * Arnoldi step applies the operators of all requests still iterating in one
 * launch).  Requests that arrive meanwhile join the next batch.  A request's
 * result does not depend on the other requests of its batch.  Applies to
 * admissions after the call. */
int baemm_server_set_fusion(baemm_server srv, int on);
/* Waits for every submitted request, then releases the cached handles. */
int baemm_server_destroy(baemm_server srv);
/* Inputs are copied; far_field_out (meas_count x W complex) and gradient_out
 * (vertex_count x 3) must stay valid until baemm_server_wait returns. */
int baemm_server_submit_gn(baemm_server srv, const baemm_gn_request* request, double* far_field_out,
                           double* gradient_out, int64_t* request_id);
/* Blocks until the request is done; returns its status (BAEMM_ERR_NOCONV:
 * results written, a solve stopped above its tolerance).  Each id is waited
 * for once. */
int baemm_server_wait(baemm_server srv, int64_t request_id, double* latency_ms, int64_t* gmres_iterations);
int baemm_server_stats(baemm_server srv, baemm_server_stats_t* stats);
/* The device-memory estimate admission uses for a request on a fresh handle
 * (bytes; -1 for an invalid request). */
int64_t baemm_server_request_bytes(const baemm_gn_request* request);

/* ------------------------------------------------------------- timing */

/* Records HIP events around every launch of the dominant pair kernel on the
 * stream it runs on; baemm_get_timing returns the summed milliseconds and the
 * number of launches since the last reset (no reference counterpart). */
int baemm_set_timing(baemm_handle h, int on);
int baemm_get_timing(baemm_handle h, double* total_ms, int64_t* launches);

/* Diagnostic, no reference counterpart: the hardware sine and cosine
 * (v_sin_f32 / v_cos_f32, argument in revolutions: sin(2 pi rev)) that every
 * pair kernel uses for e^{i kappa r}, evaluated on `device` for rev[0..n).
 * The pair kernels accept phases kappa r / (2 pi) up to 65,536 revolutions
 * (baemm_load_parameters rejects kappa * mesh extent beyond that); this entry
 * point lets a test confirm the hardware's accuracy over that range. */
int baemm_hw_sincos(int device, const float* rev, int64_t n, float* sin_out, float* cos_out);

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
