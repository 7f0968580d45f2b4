// solvers_host.cpp -- host drivers of the device Krylov solvers and the solver
// compositions of FarField.hpp (FarField, BoundaryPotential, DirichletToNeumann,
// Derivative_FF, AdjointDerivative_FF, GaussNewtonSolve), with their C ABI.
//
// Vectors live on the device for the whole composition; only the caller's
// inputs and outputs cross PCIe.  GMRES: left-preconditioned, restarted
// (restart length gmres_max_iter, at most gmres_max_restarts cycles), one
// Krylov space per column advanced together, classical Gram-Schmidt with
// re-orthogonalisation on the device, and the Hessenberg matrix, Givens
// rotations and least squares in FP64 on the device as well.  The preconditioner is the mass
// inverse by PCG, as in the reference (lumped_mass_as_prec_for_intopsQ false,
// MemberVariables.hpp:95).
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "batch.hpp"
#include "capi_util.hpp"
#include "solvers.hpp"

namespace baemm {

namespace {

struct GmresResult {
    int iterations = 0;
    int cycles = 0;
    double rel_residual = 0.0;
    bool converged = false;
};

void apply_op(Handle& h, const double2* in, double2* out, hipStream_t s)
{
    ApplyArgs a{cdouble(1.0, 0.0), cdouble(0.0, 0.0), in, h.prm.W, kC128, out, h.prm.W, kC128};
    apply_boundary_dev(h, a, s);
    if (!h.capturing) h.stats.operator_applies++;
}

void apply_prec(Handle& h, const double2* in, double2* out, double cg_tol, hipStream_t s)
{
    const int W = (int)h.prm.W;
    dev_mass_cg(h, h.sws, reinterpret_cast<const double*>(in), 2 * W, reinterpret_cast<double*>(out), 2 * W, 2 * W,
                cg_tol, 20, s);
}

// Everything an Arnoldi step's captured graph depends on: the device
// buffers it touches, the loaded operator (wave numbers and coefficients are
// baked into the kernel arguments), precision and shape.  A graph is
// re-captured when any of it changes; solves that reload the same operator
// (every DtN solve at the same wave numbers, for instance) replay it.
struct GraphKeyHead {
    const void* ptr[20];
    int64_t val[12];
    double tol, gm_tol;
};

std::vector<char> graph_key(Handle& h, int j, double cg_tol, double gm_tol)
{
    GraphKeyHead k;
    std::memset(&k, 0, sizeof(k));
    const Workspace& ws = h.sws;
    const void* p[] = {ws.V.ptr, ws.w.ptr, ws.t.ptr, ws.hc.ptr, ws.sc.ptr, ws.red.ptr, ws.cg.ptr, ws.scal.ptr,
                       ws.cgctr.ptr, h.ws_b32.ptr, h.ws_bop.ptr, h.ws_part.ptr, h.ws_scale.ptr, h.ws_misc.ptr,
                       h.ws_nrm.ptr, ws.gm.ptr, h.ws_tsum.ptr};
    static_assert(sizeof(p) / sizeof(p[0]) <= 20, "key size");
    for (size_t i = 0; i < sizeof(p) / sizeof(p[0]); ++i) k.ptr[i] = p[i];
    int rmin_bits = 0;
    std::memcpy(&rmin_bits, &h.rmin, sizeof(float));
    const int64_t v[] = {j, h.nv, h.prm.W, h.prm.wcs, h.prm.K, h.precision, h.use_diagonal ? 1 : 0, h.segments,
                         h.gmres_max_iter, rmin_bits, h.mixed_kernel};
    for (size_t i = 0; i < sizeof(v) / sizeof(v[0]); ++i) k.val[i] = v[i];
    k.tol = cg_tol;
    k.gm_tol = gm_tol;
    std::vector<char> key(reinterpret_cast<const char*>(&k), reinterpret_cast<const char*>(&k) + sizeof(k));
    const auto add = [&](const void* d, size_t n) {
        key.insert(key.end(), static_cast<const char*>(d), static_cast<const char*>(d) + n);
    };
    add(h.prm.kappa.data(), sizeof(float) * h.prm.kappa.size());
    add(h.prm.c.data(), sizeof(cfloat) * h.prm.c.size());
    for (int q = 0; q < 4; ++q) {
        const char f[2] = {(char)h.prm.re[q], (char)h.prm.im[q]};
        add(f, 2);
    }
    return key;
}

template <class F>
void run_graph(Handle& h, int j, double cg_tol, double gm_tol, hipStream_t s, F&& enqueue)
{
    if ((int)h.graphs.size() <= j) h.graphs.resize(j + 1);
    GraphSlot& g = h.graphs[j];
    std::vector<char> key = graph_key(h, j, cg_tol, gm_tol);
    if (!g.exec || g.key != key) {
        if (g.exec) {
            // the previous launch of this graph may still run on s
            BAEMM_HIP(hipStreamSynchronize(s));
            BAEMM_HIP(hipGraphExecDestroy(g.exec));
            g.exec = nullptr;
        }
        hipGraph_t graph = nullptr;
        BAEMM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        h.capturing = true;
        try {
            enqueue();
        } catch (...) {
            h.capturing = false;
            (void)hipStreamEndCapture(s, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        h.capturing = false;
        BAEMM_HIP(hipStreamEndCapture(s, &graph));
        const hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        BAEMM_HIP(e);
        g.key.swap(key);
        h.graph_captures++;
    }
    BAEMM_HIP(hipGraphLaunch(g.exec, s));
    h.graph_launches++;
}

// Pinned host words the GMRES flags are copied to (two slots, one per step in
// flight: ctl[0..3] as ints, then the largest residual as a double), and the
// events that mark each step's copy.
void gm_host_state(Handle& h)
{
    if (!h.pin_ctl) BAEMM_HIP(hipHostMalloc((void**)&h.pin_ctl, 64, hipHostMallocDefault));
    for (hipEvent_t& e : h.ev_step)
        if (!e) BAEMM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
}

// Copies the GMRES flags and the largest residual into slot `slot` of the
// pinned words, after the work enqueued so far, and records the slot's event.
void gm_snapshot(Handle& h, const GmState& st, int slot, hipStream_t s)
{
    int* dst = h.pin_ctl + 8 * slot;
    BAEMM_HIP(hipMemcpyAsync(dst, st.ctl, 4 * sizeof(int), hipMemcpyDeviceToHost, s));
    BAEMM_HIP(hipMemcpyAsync(dst + 4, st.worst, sizeof(double), hipMemcpyDeviceToHost, s));
    BAEMM_HIP(hipEventRecord(h.ev_step[slot], s));
}

// Solves (P A) X = P B for the loaded boundary operator A on W columns.
//
// Left-preconditioned restarted GMRES, one Krylov space per column advanced
// together.  An Arnoldi step (operator apply, PCG preconditioner, classical
// Gram-Schmidt twice, normalisation, then the Hessenberg column, the Givens
// rotations and the residual estimate of every column) runs on the device
// with no host round trip; the host enqueues step j + 1 before it waits for
// the flags of step j, so the device never idles between steps.  Once every
// column meets tol, a device flag makes the pair kernels and the PCG of the
// step already in flight return at once, and its Givens kernel does nothing;
// the cycle's least squares (back substitution) uses the steps completed
// before the flag.  The host synchronises once per cycle start.
GmresResult gmres_boundary(Handle& h, const double2* B, double2* X, double tol, double cg_tol, hipStream_t s)
{
    Workspace& ws = h.sws;
    const int W = (int)h.prm.W;
    const int64_t n = h.nv, NW = n * W;
    const int m = h.gmres_max_iter;
    const int cycles = h.gmres_max_restarts;
    ws.V.reserve(sizeof(double2) * NW * (m + 1));
    ws.w.reserve(sizeof(double2) * NW);
    ws.t.reserve(sizeof(double2) * NW);
    ws.hc.reserve(sizeof(double2) * (2 * (m + 2) * W));
    ws.sc.reserve(sizeof(double) * W);
    ws.red.reserve(sizeof(double2) * kRedBlocks * (m + 2) * W);
    double2* V = ws.V.as<double2>();
    double2* w = ws.w.as<double2>();
    double2* t = ws.t.as<double2>();
    double2* hc1 = ws.hc.as<double2>();
    double2* hc2 = hc1 + (m + 2) * W;
    double* sc = ws.sc.as<double>();
    GmState st = dev_gm_state(ws, W, m);
    gm_host_state(h);
    const int64_t cg0 = dev_cg_total(ws, s);

    BAEMM_HIP(hipMemsetAsync(X, 0, sizeof(double2) * NW, s));

    // Iterative refinement around GMRES(m): every restart cycle starts from the
    // true preconditioned residual P (B - A X), accumulated in FP64 from the
    // FP64 iterate X; with the mixed-precision operator the residual is taken
    // with the FP32 operator, so the cycles correct towards the FP32
    // operator's solution while their Krylov steps use the faster one.  A solve converges
    // when the Givens estimate of a cycle, or the residual a restart measures,
    // meets tol.
    GmresResult out;
    for (int cyc = 0; cyc < cycles; ++cyc) {
        if (cyc == 0) {
            apply_prec(h, B, V, cg_tol, s);
        } else {
            const int prec = h.precision;
            if (prec == BAEMM_PRECISION_MIXED) h.precision = BAEMM_PRECISION_FP32;
            try {
                apply_op(h, X, t, s);
            } catch (...) {
                h.precision = prec;
                throw;
            }
            h.precision = prec;
            dev_sub_c(NW, B, t, w, s);
            apply_prec(h, w, V, cg_tol, s);
        }
        dev_gs_dots(ws, n, W, 1, V, NW, V, hc1, s);
        dev_gm_start(W, m, hc1, cyc == 0, tol, st, sc, s);
        gm_snapshot(h, st, 0, s);
        BAEMM_HIP(hipEventSynchronize(h.ev_step[0]));
        out.rel_residual = *reinterpret_cast<const double*>(h.pin_ctl + 4);
        if (h.pin_ctl[2]) {
            out.converged = true;
            break;
        }
        out.cycles = cyc + 1;
        dev_scale_cols(n, W, V, sc, V, s);
        // One Arnoldi step: V_{j+1} = normalised (P A V_j minus its
        // projections), projections in hc1 + hc2, norm^2 in hc2[j+1], then the
        // Givens update.  The first step of a solve runs eagerly (it sizes
        // every workspace); later steps replay a captured graph when enabled.
        auto enqueue_step = [&](int j) {
            auto step = [&]() {
                double2* Vj = V + (int64_t)j * NW;
                double2* Vn = V + (int64_t)(j + 1) * NW;
                apply_op(h, Vj, t, s);
                apply_prec(h, t, w, cg_tol, s);
                // classical Gram-Schmidt, twice
                dev_gs_dots(ws, n, W, j + 1, V, NW, w, hc1, s);
                dev_gs_update(n, W, j + 1, V, NW, hc1, w, -1.0, s);
                dev_gs_dots(ws, n, W, j + 1, V, NW, w, hc2, s);
                dev_gs_update(n, W, j + 1, V, NW, hc2, w, -1.0, s);
                dev_gs_dots(ws, n, W, 1, w, NW, w, hc2 + (j + 1) * W, s);
                dev_inv_norm(W, hc2 + (j + 1) * W, sc, s);
                dev_scale_cols(n, W, w, sc, Vn, s);
                dev_gm_givens(W, m, j, hc1, hc2, tol, st, s);
            };
            h.step_skip = st.ctl;
            try {
                if (h.use_graphs && !(cyc == 0 && j == 0)) {
                    run_graph(h, j, cg_tol, tol, s, step);
                    h.stats.operator_applies++;
                } else {
                    step();
                }
            } catch (...) {
                h.step_skip = nullptr;
                throw;
            }
            h.step_skip = nullptr;
            gm_snapshot(h, st, j & 1, s);
        };
        enqueue_step(0);
        for (int j = 0; j < m; ++j) {
            if (j + 1 < m) enqueue_step(j + 1);
            // flags after step j (step j + 1 runs meanwhile)
            BAEMM_HIP(hipEventSynchronize(h.ev_step[j & 1]));
            const int* c = h.pin_ctl + 8 * (j & 1);
            out.rel_residual = *reinterpret_cast<const double*>(c + 4);
            if (c[1] == j + 1) {
                out.iterations++;
                h.stats.gmres_iterations++;
            }
            if (c[0]) break;
        }
        // y = R^{-1} g per column over the completed steps; X += V y
        dev_gm_backsub(W, m, st, hc1, s);
        dev_gs_update(n, W, m, V, NW, hc1, X, 1.0, s, st.ctl + 1);
        if (out.rel_residual <= tol) {
            out.converged = true;
            BAEMM_HIP(hipStreamSynchronize(s));
            break;
        }
    }
    h.stats.cg_iterations += dev_cg_total(ws, s) - cg0;
    h.stats.gmres_cycles += out.cycles;
    h.stats.gmres_rel_residual = out.rel_residual;
    h.stats.converged = h.stats.converged && out.converged;
    return out;
}

std::vector<double> rows_coeff(int64_t K, const std::vector<cdouble>& row_of)
{
    std::vector<double> c(8 * K);
    for (int64_t q = 0; q < K; ++q)
        for (int j = 0; j < 4; ++j) {
            c[8 * q + 2 * j] = row_of[4 * q + j].real();
            c[8 * q + 2 * j + 1] = row_of[4 * q + j].imag();
        }
    return c;
}

// BoundaryPotential_parameters (FarField.hpp:678-771): (1/2 M - i eta S + K) phi = wave.
GmresResult boundary_potential_dev(Handle& h, const double* kappa, const double* eta, int64_t K, int64_t wcs,
                                   const double2* wave, double2* phi, double cg_tol, double gmres_tol, hipStream_t s,
                                   double* coeff_out = nullptr)
{
    std::vector<cdouble> rows(4 * K);
    for (int64_t q = 0; q < K; ++q) {
        rows[4 * q] = {0.5, 0.0};
        rows[4 * q + 1] = {0.0, -eta[q]};
        rows[4 * q + 2] = {1.0, 0.0};
        rows[4 * q + 3] = {0.0, 0.0};
    }
    const std::vector<double> c = rows_coeff(K, rows);
    if (coeff_out) std::memcpy(coeff_out, c.data(), sizeof(double) * c.size());
    load_params(h, kappa, c.data(), K * wcs, wcs, 1.0);
    return gmres_boundary(h, wave, phi, gmres_tol, cg_tol, s);
}

// DirichletToNeumann (FarField.hpp:790-879): (1/2 M - i kappa S + K') psi = wave.
GmresResult dtn_dev(Handle& h, const double* kappa, int64_t K, int64_t wcs, const double2* wave, double2* psi,
                    double cg_tol, double gmres_tol, hipStream_t s)
{
    std::vector<cdouble> rows(4 * K);
    for (int64_t q = 0; q < K; ++q) {
        rows[4 * q] = {0.5, 0.0};
        rows[4 * q + 1] = {0.0, -kappa[q]};
        rows[4 * q + 2] = {0.0, 0.0};
        rows[4 * q + 3] = {1.0, 0.0};
    }
    const std::vector<double> c = rows_coeff(K, rows);
    load_params(h, kappa, c.data(), K * wcs, wcs, 1.0);
    return gmres_boundary(h, wave, psi, gmres_tol, cg_tol, s);
}

// CreateIncidentWave_PL with coefficient rows (0, c1_q, c2_q, 0), factor 4 pi (alpha 1, beta 0).
void incident_rows_dev(Handle& h, const double* kappa, int64_t K, int64_t wcs, const double* dirs, int wave_type,
                       const std::vector<cdouble>& c12, double2* out, hipStream_t s)
{
    std::vector<cdouble> rows(4 * K);
    for (int64_t q = 0; q < K; ++q) {
        rows[4 * q] = 0.0;
        rows[4 * q + 1] = c12[2 * q];
        rows[4 * q + 2] = c12[2 * q + 1];
        rows[4 * q + 3] = 0.0;
    }
    const std::vector<double> c = rows_coeff(K, rows);
    load_params(h, kappa, c.data(), K * wcs, wcs, 4.0 * M_PI);
    ApplyArgs a{cdouble(1.0, 0.0), cdouble(0.0, 0.0), nullptr, 0, kC128, out, K * wcs, kC128};
    incident_dev(h, a, dirs, wcs, wave_type, s);
}

// ApplyFarFieldOperators_PL with coefficient rows `c` (FarField.hpp:139-143): F (device, m x W).
void far_field_dev(Handle& h, const double* kappa, int64_t K, int64_t wcs, const std::vector<double>& c,
                   const double2* phi, double2* F, hipStream_t s)
{
    load_params(h, kappa, c.data(), K * wcs, wcs, 1.0);
    ApplyArgs a{cdouble(1.0, 0.0), cdouble(0.0, 0.0), phi, K * wcs, kC128, F, K * wcs, kC128};
    apply_far_dev(h, a, s);
}

// Derivative_FF_Helper (FarField.hpp:148-207): du_dn = DtN(du_i/dn - i kappa u_i).
void du_dn_dev(Handle& h, const double* kappa, int64_t K, int64_t wcs, const double* dirs, int wave_type,
               double2* du, double cg_tol, double gmres_tol, hipStream_t s)
{
    const int64_t W = K * wcs;
    h.sws.a.reserve(sizeof(double2) * h.nv * W);
    double2* inc = h.sws.a.as<double2>();
    std::vector<cdouble> c12(2 * K);
    for (int64_t q = 0; q < K; ++q) {
        c12[2 * q] = {0.0, -kappa[q]};
        c12[2 * q + 1] = {1.0, 0.0};
    }
    incident_rows_dev(h, kappa, K, wcs, dirs, wave_type, c12, inc, s);
    dtn_dev(h, kappa, K, wcs, inc, du, cg_tol, gmres_tol, s);
}

// PCG mass solve for n_cols real columns of a device array (ApplyMassInverse).
void mass_inverse_dev(Handle& h, const double* X, int64_t ldX, double* Y, int64_t ldY, int ncols, double tol,
                      hipStream_t s)
{
    dev_mass_cg(h, h.sws, X, ldX, Y, ldY, ncols, tol, 20, s);
    h.stats.cg_iterations += dev_mass_cg_result(h.sws, s);
}

// DotWithNormals_PL (FarField.hpp:956-1022): Xn (device, V) = M^{-1} AvT ((n . Av X)/a).
void dot_with_normals_dev(Handle& h, const double* X, double* Xn, double cg_tol, hipStream_t s)
{
    h.sws.b.reserve(sizeof(double) * (4 * h.nt + h.nv));
    double* Xt = h.sws.b.as<double>();
    double* yt = Xt + 3 * h.nt;
    double* yw = yt + h.nt;
    dev_av(h, X, 3, Xt, 3, 3, s);
    dev_dot_normals(h, Xt, yt, s);
    dev_avt(h, 1.0, yt, 1, 0.0, yw, 1, 1, s);
    mass_inverse_dev(h, yw, 1, Xn, 1, 1, cg_tol, s);
}

// MultiplyWithNormals_PL (FarField.hpp:886-952): Y (device, V x 3) = M^{-1} (factor AvT (Re(Av x)/a n)).
void multiply_with_normals_dev(Handle& h, const double2* x, double* Y, double factor, double cg_tol, hipStream_t s)
{
    h.sws.b.reserve(sizeof(double) * (2 * h.nt + 3 * h.nt + 3 * h.nv));
    double* xt = h.sws.b.as<double>();               // nt complex (as 2 real columns)
    double* Yt = xt + 2 * h.nt;
    double* Yw = Yt + 3 * h.nt;
    dev_av(h, reinterpret_cast<const double*>(x), 2, xt, 2, 2, s);
    dev_mul_normals(h, reinterpret_cast<const double2*>(xt), Yt, s);
    dev_avt(h, factor, Yt, 3, 0.0, Yw, 3, 3, s);
    mass_inverse_dev(h, Yw, 3, Y, 3, 3, cg_tol, s);
}

// Derivative_FF core (FarField.hpp:285-318) with du_dn on the device: F = DF X.
void derivative_dev(Handle& h, const double* kappa, int64_t K, int64_t wcs, const double* Xdev, const double2* du,
                    double2* F, double cg_tol, double gmres_tol, hipStream_t s)
{
    const int64_t W = K * wcs, NW = h.nv * W;
    h.sws.c.reserve(sizeof(double2) * 2 * NW + sizeof(double) * h.nv);
    double2* bdr = h.sws.c.as<double2>();
    double2* phi = bdr + NW;
    double* xn = reinterpret_cast<double*>(phi + NW);
    dot_with_normals_dev(h, Xdev, xn, cg_tol, s);
    dev_scale_rows_neg(h, (int)W, xn, du, bdr, s);
    h.sws.a.reserve(sizeof(double2) * NW);
    double2* weak = h.sws.a.as<double2>();
    dev_mass(h, 1.0, reinterpret_cast<const double*>(bdr), 2 * W, 0.0, reinterpret_cast<double*>(weak), 2 * W,
             (int)(2 * W), s);
    std::vector<double> coeff(8 * K);
    boundary_potential_dev(h, kappa, kappa, K, wcs, weak, phi, cg_tol, gmres_tol, s, coeff.data());
    far_field_dev(h, kappa, K, wcs, coeff, phi, F, s);
}

// AdjointDerivative_FF core (FarField.hpp:394-425) with du_dn on the device: Y (V x 3) = DF* g.
void adjoint_dev(Handle& h, const double* kappa, int64_t K, int64_t wcs, const double2* g, const double2* du,
                 double* Y, double cg_tol, double gmres_tol, hipStream_t s)
{
    const int64_t W = K * wcs, NW = h.nv * W;
    h.sws.d.reserve(sizeof(double2) * (2 * NW + h.nv));
    double2* herg = h.sws.d.as<double2>();
    double2* dv = herg + NW;
    double2* prod = dv + NW;
    std::vector<cdouble> rows(4 * K);
    for (int64_t q = 0; q < K; ++q) {
        rows[4 * q] = 0.0;
        rows[4 * q + 1] = {0.0, -kappa[q]};
        rows[4 * q + 2] = 1.0;
        rows[4 * q + 3] = 0.0;
    }
    const std::vector<double> c = rows_coeff(K, rows);
    load_params(h, kappa, c.data(), W, wcs, 4.0 * M_PI / (double)h.m);
    ApplyArgs a{cdouble(1.0, 0.0), cdouble(0.0, 0.0), g, W, kC128, herg, W, kC128};
    herglotz_dev(h, a, s);
    dtn_dev(h, kappa, K, wcs, herg, dv, cg_tol, gmres_tol, s);
    dev_row_product_sum(h, (int)W, du, dv, prod, s);
    multiply_with_normals_dev(h, prod, Y, -1.0 / (double)W, cg_tol, s);
}

void check_common(Handle& h, const double* kappa, int64_t K, int64_t wcs)
{
    if (!kappa) throw ArgError{"kappa must not be NULL"};
    if (K < 1 || wcs < 1) throw ArgError{"wave_chunk_count and wave_chunk_size must be >= 1"};
    if (K > 64) throw ArgError{"at most 64 wave-number chunks are supported"};
}

// Device array of an entry point, held in the handle's grow-only I/O slots:
// no allocation (and no device-wide synchronisation of hipFree) per call.
struct DevArray {
    void* p = nullptr;
    DevArray(Handle& h, int slot, size_t bytes)
    {
        h.io[slot].reserve(bytes > 0 ? bytes : 16);
        p = h.io[slot].ptr;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

void upload(void* dst, const void* src, size_t bytes, hipStream_t s)
{
    BAEMM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
}

void download(void* dst, const void* src, size_t bytes, hipStream_t s)
{
    BAEMM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    BAEMM_HIP(hipStreamSynchronize(s));
}

// du_dn of the reference's C_ext*& argument: allocated with calloc when NULL
// (FarField.hpp:273-283), filled by Derivative_FF_Helper.
double2* ensure_du_dn(Handle& h, double** du_dn, const double* kappa, int64_t K, int64_t wcs, const double* dirs,
                      int wave_type, double cg_tol, double gmres_tol, DevArray& ddu, hipStream_t s)
{
    const int64_t W = K * wcs, NW = h.nv * W;
    if (!du_dn) throw ArgError{"du_dn must point to a pointer (which may be NULL)"};
    if (*du_dn == nullptr) {
        *du_dn = static_cast<double*>(std::calloc((size_t)NW * 2, sizeof(double)));
        if (!*du_dn) throw std::bad_alloc();
        du_dn_dev(h, kappa, K, wcs, dirs, wave_type, ddu.as<double2>(), cg_tol, gmres_tol, s);
        download(*du_dn, ddu.p, sizeof(double2) * NW, s);
    } else {
        upload(ddu.p, *du_dn, sizeof(double2) * NW, s);
    }
    return ddu.as<double2>();
}

}  // namespace

// Batched GMRES: the cycles of gmres_boundary for several systems in lockstep.
// Per Arnoldi step, one batched launch applies the operators of every system
// still iterating (apply_boundary_batch; a system that converged at the step
// before has its items return at once through its stop flag), then each
// system's PCG, Gram-Schmidt and Givens run on its handle's stream, so those
// of different systems overlap.  As in gmres_boundary, step j + 1 is enqueued
// before the flags of step j are read, restart residuals of the mixed
// precision use the FP32 operator, and each system's result is the one its
// own GMRES reaches (the batched apply's segment split is the batch's own, so
// results differ from gmres_boundary's in the last bits).
void gmres_batch(std::vector<GmresBatchItem>& items, BatchWorkspace& bw, double tol, double cg_tol, hipStream_t s)
{
    struct Sys {
        Handle* h;
        GmresBatchItem* it;
        int W;
        int64_t n, NW;
        double2 *V, *w, *t, *hc1, *hc2;
        double* sc;
        GmState st;
        int64_t cg0;
        bool done = false;      // converged (or out of cycles)
        bool iterating = false; // in the current cycle, not yet stopped
    };
    const int nsys = (int)items.size();
    if (nsys == 0) return;
    int m = items[0].h->gmres_max_iter, cycles = items[0].h->gmres_max_restarts;
    std::vector<Sys> sy(nsys);
    for (int k = 0; k < nsys; ++k) {
        Sys& q = sy[k];
        q.h = items[k].h;
        q.it = &items[k];
        Handle& h = *q.h;
        if (h.gmres_max_iter != m || h.gmres_max_restarts != cycles)
            throw ArgError{"batched GMRES: all handles need the same gmres limits"};
        Workspace& ws = h.sws;
        q.W = (int)h.prm.W;
        q.n = h.nv;
        q.NW = q.n * q.W;
        ws.V.reserve(sizeof(double2) * q.NW * (m + 1));
        ws.w.reserve(sizeof(double2) * q.NW);
        ws.t.reserve(sizeof(double2) * q.NW);
        ws.hc.reserve(sizeof(double2) * (2 * (m + 2) * q.W));
        ws.sc.reserve(sizeof(double) * q.W);
        ws.red.reserve(sizeof(double2) * kRedBlocks * (m + 2) * q.W);
        q.V = ws.V.as<double2>();
        q.w = ws.w.as<double2>();
        q.t = ws.t.as<double2>();
        q.hc1 = ws.hc.as<double2>();
        q.hc2 = q.hc1 + (m + 2) * q.W;
        q.sc = ws.sc.as<double>();
        q.st = dev_gm_state(ws, q.W, m);
        gm_host_state(h);
        q.cg0 = dev_cg_total(ws, h.stream);
        BAEMM_HIP(hipMemsetAsync(q.it->X, 0, sizeof(double2) * q.NW, h.stream));
    }
    // the operator applies of the listed systems, in: in_k, out: t_k
    auto batch_apply = [&](const std::vector<int>& ks, auto&& in_of) {
        std::vector<BatchEntry> es;
        for (int k : ks) {
            Sys& q = sy[k];
            es.push_back(BatchEntry{q.h, ApplyArgs{cdouble(1.0, 0.0), cdouble(0.0, 0.0), in_of(k), q.W, kC128, q.t,
                                                   q.W, kC128}});
            q.h->stats.operator_applies++;
        }
        apply_boundary_batch(es, bw, s, nullptr);
    };
    auto enqueue_step = [&](int j, const std::vector<int>& ks) {
        for (int k : ks) sy[k].h->step_skip = sy[k].st.ctl;
        try {
            batch_apply(ks, [&](int k) { return (const void*)(sy[k].V + (int64_t)j * sy[k].NW); });
            for (int k : ks) {
                Sys& q = sy[k];
                Handle& h = *q.h;
                Workspace& ws = h.sws;
                hipStream_t sk = h.stream;
                double2* Vn = q.V + (int64_t)(j + 1) * q.NW;
                apply_prec(h, q.t, q.w, cg_tol, sk);
                dev_gs_dots(ws, q.n, q.W, j + 1, q.V, q.NW, q.w, q.hc1, sk);
                dev_gs_update(q.n, q.W, j + 1, q.V, q.NW, q.hc1, q.w, -1.0, sk);
                dev_gs_dots(ws, q.n, q.W, j + 1, q.V, q.NW, q.w, q.hc2, sk);
                dev_gs_update(q.n, q.W, j + 1, q.V, q.NW, q.hc2, q.w, -1.0, sk);
                dev_gs_dots(ws, q.n, q.W, 1, q.w, q.NW, q.w, q.hc2 + (j + 1) * q.W, sk);
                dev_inv_norm(q.W, q.hc2 + (j + 1) * q.W, q.sc, sk);
                dev_scale_cols(q.n, q.W, q.w, q.sc, Vn, sk);
                dev_gm_givens(q.W, m, j, q.hc1, q.hc2, tol, q.st, sk);
                gm_snapshot(h, q.st, j & 1, sk);
            }
        } catch (...) {
            for (Sys& q : sy) q.h->step_skip = nullptr;
            throw;
        }
        for (int k : ks) sy[k].h->step_skip = nullptr;
    };
    for (int cyc = 0; cyc < cycles; ++cyc) {
        std::vector<int> act;
        for (int k = 0; k < nsys; ++k)
            if (!sy[k].done) act.push_back(k);
        if (act.empty()) break;
        if (cyc == 0) {
            for (int k : act) apply_prec(*sy[k].h, sy[k].it->B, sy[k].V, cg_tol, sy[k].h->stream);
        } else {
            // restart residuals P (B - A X), the mixed precision's with the FP32 operator
            std::vector<int> prec(nsys);
            for (int k : act) {
                prec[k] = sy[k].h->precision;
                if (prec[k] == BAEMM_PRECISION_MIXED) sy[k].h->precision = BAEMM_PRECISION_FP32;
            }
            try {
                batch_apply(act, [&](int k) { return (const void*)sy[k].it->X; });
            } catch (...) {
                for (int k : act) sy[k].h->precision = prec[k];
                throw;
            }
            for (int k : act) {
                Sys& q = sy[k];
                q.h->precision = prec[k];
                dev_sub_c(q.NW, q.it->B, q.t, q.w, q.h->stream);
                apply_prec(*q.h, q.w, q.V, cg_tol, q.h->stream);
            }
        }
        for (int k : act) {
            Sys& q = sy[k];
            dev_gs_dots(q.h->sws, q.n, q.W, 1, q.V, q.NW, q.V, q.hc1, q.h->stream);
            dev_gm_start(q.W, m, q.hc1, cyc == 0, tol, q.st, q.sc, q.h->stream);
            gm_snapshot(*q.h, q.st, 0, q.h->stream);
        }
        std::vector<int> run;
        for (int k : act) {
            Sys& q = sy[k];
            BAEMM_HIP(hipEventSynchronize(q.h->ev_step[0]));
            q.it->rel_residual = *reinterpret_cast<const double*>(q.h->pin_ctl + 4);
            if (q.h->pin_ctl[2]) {
                q.done = true;
                q.it->converged = true;
                continue;
            }
            q.it->cycles++;
            q.iterating = true;
            dev_scale_cols(q.n, q.W, q.V, q.sc, q.V, q.h->stream);
            run.push_back(k);
        }
        if (run.empty()) break;
        const std::vector<int> cycle_set = run;
        enqueue_step(0, run);
        std::vector<int> inflight = run;   // systems whose step j is enqueued
        for (int j = 0; j < m && !inflight.empty(); ++j) {
            std::vector<int> next;
            if (j + 1 < m && !run.empty()) {
                next = run;
                enqueue_step(j + 1, next);
            }
            for (int k : inflight) {
                Sys& q = sy[k];
                BAEMM_HIP(hipEventSynchronize(q.h->ev_step[j & 1]));
                const int* c = q.h->pin_ctl + 8 * (j & 1);
                q.it->rel_residual = *reinterpret_cast<const double*>(c + 4);
                if (c[1] == j + 1) {
                    q.it->iterations++;
                    q.h->stats.gmres_iterations++;
                }
                if (c[0]) q.iterating = false;
            }
            run.clear();
            for (int k : cycle_set)
                if (sy[k].iterating) run.push_back(k);
            // systems that stopped at step j are in step j + 1 already (their
            // work there returns at once); step j + 2 leaves them out
            inflight = next;
        }
        for (int k : cycle_set) {
            Sys& q = sy[k];
            dev_gm_backsub(q.W, m, q.st, q.hc1, q.h->stream);
            dev_gs_update(q.n, q.W, m, q.V, q.NW, q.hc1, q.it->X, 1.0, q.h->stream, q.st.ctl + 1);
            if (q.it->rel_residual <= tol) {
                q.done = true;
                q.it->converged = true;
            }
        }
    }
    for (Sys& q : sy) {
        BAEMM_HIP(hipStreamSynchronize(q.h->stream));
        Handle& h = *q.h;
        h.stats.cg_iterations += dev_cg_total(h.sws, h.stream) - q.cg0;
        h.stats.gmres_cycles += q.it->cycles;
        h.stats.gmres_rel_residual = q.it->rel_residual;
        h.stats.converged = h.stats.converged && q.it->converged;
    }
    BAEMM_HIP(hipStreamSynchronize(s));
}

// Fused Gauss-Newton steps: the per-request pipeline of the serving loop
// (FarField, Derivative_FF_Helper, AdjointDerivative_FF) with each of its three
// solves batched over the requests.  Device arrays per request live in the
// handle's I/O slots: [0] right-hand side, [1] potential, [2] far field,
// [3] Neumann data du, [4] Herglotz solve dv, [5] misfit, product and gradient.
void gn_step_batch(std::vector<GnItem>& items, BatchWorkspace& bw, hipStream_t s)
{
    const int n = (int)items.size();
    if (n == 0) return;
    struct Buf {
        double2 *rhs, *phi, *F, *du, *dv, *g, *prod;
        double* Y;
        std::vector<double> coeff;
        int64_t W, NW;
    };
    std::vector<Buf> b(n);
    for (int k = 0; k < n; ++k) {
        GnItem& it = items[k];
        Handle& h = *it.h;
        check_common(h, it.kappa, it.K, it.wcs);
        if (!it.dirs || !it.F_out || !it.grad_out) throw ArgError{"request arrays must not be NULL"};
        if (h.m < 1) throw ArgError{"no measurement directions"};
        Buf& q = b[k];
        q.W = it.K * it.wcs;
        q.NW = h.nv * q.W;
        q.rhs = DevArray(h, 0, sizeof(double2) * q.NW).as<double2>();
        q.phi = DevArray(h, 1, sizeof(double2) * q.NW).as<double2>();
        q.F = DevArray(h, 2, sizeof(double2) * h.m * q.W).as<double2>();
        q.du = DevArray(h, 3, sizeof(double2) * q.NW).as<double2>();
        q.dv = DevArray(h, 4, sizeof(double2) * q.NW).as<double2>();
        char* p5 = DevArray(h, 5, sizeof(double2) * (h.m * q.W + h.nv) + sizeof(double) * 3 * h.nv).as<char>();
        q.g = reinterpret_cast<double2*>(p5);
        q.prod = q.g + h.m * q.W;
        q.Y = reinterpret_cast<double*>(q.prod + h.nv);
        h.stats.converged = true;
    }
    auto solve = [&](auto&& rhs_of, auto&& x_of, double tol_of_first) {
        (void)tol_of_first;
        std::vector<GmresBatchItem> g(n);
        for (int k = 0; k < n; ++k) {
            g[k].h = items[k].h;
            g[k].B = rhs_of(k);
            g[k].X = x_of(k);
        }
        // one tolerance per batch: the requests of a fused batch share theirs
        gmres_batch(g, bw, items[0].gmres_tol, items[0].cg_tol, s);
        for (int k = 0; k < n; ++k)
            if (!g[k].converged) items[k].status = BAEMM_ERR_NOCONV;
    };
    for (int k = 1; k < n; ++k)
        if (items[k].gmres_tol != items[0].gmres_tol || items[k].cg_tol != items[0].cg_tol)
            throw ArgError{"a fused batch needs one GMRES and one CG tolerance"};
    // 1. incident wave (0, -1, 0, 0) and the potential operator (1/2, -i kappa, 1, 0)
    for (int k = 0; k < n; ++k) {
        GnItem& it = items[k];
        Handle& h = *it.h;
        std::vector<cdouble> c12(2 * it.K, cdouble(0.0, 0.0));
        for (int64_t q = 0; q < it.K; ++q) c12[2 * q] = -1.0;
        incident_rows_dev(h, it.kappa, it.K, it.wcs, it.dirs, it.wave_type, c12, b[k].rhs, h.stream);
        std::vector<cdouble> rows(4 * it.K);
        for (int64_t q = 0; q < it.K; ++q) {
            rows[4 * q] = {0.5, 0.0};
            rows[4 * q + 1] = {0.0, -it.kappa[q]};
            rows[4 * q + 2] = {1.0, 0.0};
            rows[4 * q + 3] = {0.0, 0.0};
        }
        b[k].coeff = rows_coeff(it.K, rows);
        load_params(h, it.kappa, b[k].coeff.data(), b[k].W, it.wcs, 1.0);
    }
    solve([&](int k) { return (const double2*)b[k].rhs; }, [&](int k) { return b[k].phi; }, 0.0);
    // 2. far field; incident wave (0, -i kappa, 1, 0) and the DtN operator (1/2, -i kappa, 0, 1)
    for (int k = 0; k < n; ++k) {
        GnItem& it = items[k];
        Handle& h = *it.h;
        far_field_dev(h, it.kappa, it.K, it.wcs, b[k].coeff, b[k].phi, b[k].F, h.stream);
        BAEMM_HIP(hipMemcpyAsync(it.F_out, b[k].F, sizeof(double2) * h.m * b[k].W, hipMemcpyDeviceToHost, h.stream));
        std::vector<cdouble> c12(2 * it.K);
        for (int64_t q = 0; q < it.K; ++q) {
            c12[2 * q] = {0.0, -it.kappa[q]};
            c12[2 * q + 1] = {1.0, 0.0};
        }
        incident_rows_dev(h, it.kappa, it.K, it.wcs, it.dirs, it.wave_type, c12, b[k].rhs, h.stream);
        std::vector<cdouble> rows(4 * it.K);
        for (int64_t q = 0; q < it.K; ++q) {
            rows[4 * q] = {0.5, 0.0};
            rows[4 * q + 1] = {0.0, -it.kappa[q]};
            rows[4 * q + 2] = {0.0, 0.0};
            rows[4 * q + 3] = {1.0, 0.0};
        }
        const std::vector<double> c = rows_coeff(it.K, rows);
        load_params(h, it.kappa, c.data(), b[k].W, it.wcs, 1.0);
    }
    solve([&](int k) { return (const double2*)b[k].rhs; }, [&](int k) { return b[k].du; }, 0.0);
    // 3. misfit g = F - F_obs, its Herglotz wave (0, -i kappa, 1, 0) with factor 4 pi / m, DtN again
    for (int k = 0; k < n; ++k) {
        GnItem& it = items[k];
        Handle& h = *it.h;
        BAEMM_HIP(hipStreamSynchronize(h.stream));     // F_out is on the host
        const int64_t mW = h.m * b[k].W;
        std::vector<double> g(it.F_out, it.F_out + 2 * mW);
        if (it.fobs)
            for (int64_t e = 0; e < 2 * mW; ++e) g[e] -= it.fobs[e];
        BAEMM_HIP(hipMemcpyAsync(b[k].g, g.data(), sizeof(double2) * mW, hipMemcpyHostToDevice, h.stream));
        std::vector<cdouble> rows(4 * it.K);
        for (int64_t q = 0; q < it.K; ++q) {
            rows[4 * q] = 0.0;
            rows[4 * q + 1] = {0.0, -it.kappa[q]};
            rows[4 * q + 2] = 1.0;
            rows[4 * q + 3] = 0.0;
        }
        const std::vector<double> c = rows_coeff(it.K, rows);
        load_params(h, it.kappa, c.data(), b[k].W, it.wcs, 4.0 * M_PI / (double)h.m);
        ApplyArgs a{cdouble(1.0, 0.0), cdouble(0.0, 0.0), b[k].g, b[k].W, kC128, b[k].rhs, b[k].W, kC128};
        herglotz_dev(h, a, h.stream);
        BAEMM_HIP(hipStreamSynchronize(h.stream));     // the host copy of g is released here
        std::vector<cdouble> dr(4 * it.K);
        for (int64_t q = 0; q < it.K; ++q) {
            dr[4 * q] = {0.5, 0.0};
            dr[4 * q + 1] = {0.0, -it.kappa[q]};
            dr[4 * q + 2] = {0.0, 0.0};
            dr[4 * q + 3] = {1.0, 0.0};
        }
        const std::vector<double> cd = rows_coeff(it.K, dr);
        load_params(h, it.kappa, cd.data(), b[k].W, it.wcs, 1.0);
    }
    solve([&](int k) { return (const double2*)b[k].rhs; }, [&](int k) { return b[k].dv; }, 0.0);
    // 4. gradient: M^{-1} AvT ((-1/W) Re(sum_k du dv) averaged to P0 / a n)
    for (int k = 0; k < n; ++k) {
        GnItem& it = items[k];
        Handle& h = *it.h;
        dev_row_product_sum(h, (int)b[k].W, b[k].du, b[k].dv, b[k].prod, h.stream);
        multiply_with_normals_dev(h, b[k].prod, b[k].Y, -1.0 / (double)b[k].W, it.cg_tol, h.stream);
        BAEMM_HIP(hipMemcpyAsync(it.grad_out, b[k].Y, sizeof(double) * 3 * h.nv, hipMemcpyDeviceToHost, h.stream));
    }
    for (int k = 0; k < n; ++k) BAEMM_HIP(hipStreamSynchronize(items[k].h->stream));
}

}  // namespace baemm

using namespace baemm;

extern "C" {

int baemm_apply_mass_inverse(baemm_handle h, const double* X, int64_t ldX, double* Y, int64_t ldY, int64_t nrhs,
                             double cg_tol)
{
    return guard(h, [&](Handle& hh) {
        if (!X || !Y || nrhs < 1 || ldX < nrhs || ldY < nrhs) throw ArgError{"invalid mass-inverse arguments"};
        hipStream_t s = hh.stream;
        DevArray dx(hh, 0, sizeof(double) * hh.nv * nrhs), dy(hh, 1, sizeof(double) * hh.nv * nrhs);
        BAEMM_HIP(copy_rows_async(dx.p, sizeof(double) * nrhs, X, sizeof(double) * ldX, sizeof(double) * nrhs, hh.nv,
                                   hipMemcpyHostToDevice, s));
        dev_mass_cg(hh, hh.sws, dx.as<double>(), nrhs, dy.as<double>(), nrhs, (int)nrhs, cg_tol, 20, s);
        hh.stats.cg_iterations += dev_mass_cg_result(hh.sws, s);
        BAEMM_HIP(copy_rows_async(Y, sizeof(double) * ldY, dy.p, sizeof(double) * nrhs, sizeof(double) * nrhs, hh.nv,
                                   hipMemcpyDeviceToHost, s));
        BAEMM_HIP(hipStreamSynchronize(s));
        return BAEMM_OK;
    });
}

int baemm_apply_mass_inverse_complex(baemm_handle h, const double* X, int64_t ldX, double* Y, int64_t ldY,
                                     int64_t nrhs, double cg_tol)
{
    if (nrhs < 1 || ldX < nrhs || ldY < nrhs) return BAEMM_ERR_ARG;
    return baemm_apply_mass_inverse(h, X, 2 * ldX, Y, 2 * ldY, 2 * nrhs, cg_tol);
}

int baemm_apply_lumped_mass_inverse(baemm_handle h, const double* X, int64_t ldX, double* Y, int64_t ldY,
                                    int64_t nrhs)
{
    return guard(h, [&](Handle& hh) {
        if (!X || !Y || nrhs < 1 || ldX < nrhs || ldY < nrhs) throw ArgError{"invalid arguments"};
        for (int64_t v = 0; v < hh.nv; ++v)
            for (int64_t k = 0; k < nrhs; ++k) Y[v * ldY + k] = hh.h_lumped_inv[v] * X[v * ldX + k];
        return BAEMM_OK;
    });
}

int baemm_boundary_potential(baemm_handle h, const double* kappa, int64_t K, const double* eta, double* coeff,
                             const double* wave, double* phi, int64_t wcs, double cg_tol, double gmres_tol)
{
    return guard(h, [&](Handle& hh) {
        check_common(hh, kappa, K, wcs);
        if (!wave || !phi) throw ArgError{"wave and phi must not be NULL"};
        const int64_t NW = hh.nv * K * wcs;
        hipStream_t s = hh.stream;
        DevArray dw(hh, 0, sizeof(double2) * NW), dp(hh, 1, sizeof(double2) * NW);
        upload(dw.p, wave, sizeof(double2) * NW, s);
        const GmresResult r = boundary_potential_dev(hh, kappa, eta ? eta : kappa, K, wcs, dw.as<double2>(),
                                                     dp.as<double2>(), cg_tol, gmres_tol, s, coeff);
        download(phi, dp.p, sizeof(double2) * NW, s);
        return r.converged ? BAEMM_OK : BAEMM_ERR_NOCONV;
    });
}

int baemm_gmres_dev(baemm_handle h, const void* B, void* X, double cg_tol, double gmres_tol, int64_t* iterations,
                    void* stream)
{
    return guard(h, [&](Handle& hh) {
        if (!hh.prm.loaded) throw StateError{"operators are not loaded"};
        if (!B || !X) throw ArgError{"B and X must not be NULL"};
        hipStream_t s = stream_of(hh, stream);
        const GmresResult r = gmres_boundary(hh, static_cast<const double2*>(B), static_cast<double2*>(X), gmres_tol,
                                             cg_tol, s);
        BAEMM_HIP(hipStreamSynchronize(s));
        if (iterations) *iterations = r.iterations;
        return r.converged ? BAEMM_OK : BAEMM_ERR_NOCONV;
    });
}

int baemm_dirichlet_to_neumann(baemm_handle h, const double* kappa, int64_t K, const double* wave, double* neumann,
                               int64_t wcs, double cg_tol, double gmres_tol)
{
    return guard(h, [&](Handle& hh) {
        check_common(hh, kappa, K, wcs);
        if (!wave || !neumann) throw ArgError{"wave and neumann_trace must not be NULL"};
        const int64_t NW = hh.nv * K * wcs;
        hipStream_t s = hh.stream;
        DevArray dw(hh, 0, sizeof(double2) * NW), dp(hh, 1, sizeof(double2) * NW);
        upload(dw.p, wave, sizeof(double2) * NW, s);
        const GmresResult r = dtn_dev(hh, kappa, K, wcs, dw.as<double2>(), dp.as<double2>(), cg_tol, gmres_tol, s);
        download(neumann, dp.p, sizeof(double2) * NW, s);
        return r.converged ? BAEMM_OK : BAEMM_ERR_NOCONV;
    });
}

int baemm_far_field_parameters(baemm_handle h, const double* kappa, int64_t K, const double* dirs, int64_t wcs,
                               double* Y, int wave_type, const double* eta, double cg_tol, double gmres_tol)
{
    return guard(h, [&](Handle& hh) {
        check_common(hh, kappa, K, wcs);
        if (!dirs || !Y) throw ArgError{"directions and output must not be NULL"};
        if (hh.m < 1) throw ArgError{"no measurement directions"};
        const int64_t W = K * wcs, NW = hh.nv * W;
        hipStream_t s = hh.stream;
        DevArray dwave(hh, 0, sizeof(double2) * NW);
        DevArray dphi(hh, 1, sizeof(double2) * NW);
        DevArray dF(hh, 2, sizeof(double2) * hh.m * W);
        std::vector<cdouble> c12(2 * K, cdouble(0.0, 0.0));
        for (int64_t q = 0; q < K; ++q) c12[2 * q] = -1.0;     // (0, -1, 0, 0): FarField.hpp:120-126
        incident_rows_dev(hh, kappa, K, wcs, dirs, wave_type, c12, dwave.as<double2>(), s);
        std::vector<double> coeff(8 * K);
        const GmresResult r = boundary_potential_dev(hh, kappa, eta ? eta : kappa, K, wcs, dwave.as<double2>(),
                                                     dphi.as<double2>(), cg_tol, gmres_tol, s, coeff.data());
        far_field_dev(hh, kappa, K, wcs, coeff, dphi.as<double2>(), dF.as<double2>(), s);
        download(Y, dF.p, sizeof(double2) * hh.m * W, s);
        return r.converged ? BAEMM_OK : BAEMM_ERR_NOCONV;
    });
}

int baemm_far_field(baemm_handle h, const double* kappa, int64_t K, const double* dirs, int64_t wcs, double* Y,
                    int wave_type, double cg_tol, double gmres_tol)
{
    return baemm_far_field_parameters(h, kappa, K, dirs, wcs, Y, wave_type, kappa, cg_tol, gmres_tol);
}

int baemm_derivative_ff_helper(baemm_handle h, const double* kappa, int64_t K, const double* dirs, int64_t wcs,
                               double* du_dn, int wave_type, double cg_tol, double gmres_tol)
{
    return guard(h, [&](Handle& hh) {
        check_common(hh, kappa, K, wcs);
        if (!dirs || !du_dn) throw ArgError{"directions and du_dn must not be NULL"};
        const int64_t NW = hh.nv * K * wcs;
        hipStream_t s = hh.stream;
        DevArray ddu(hh, 0, sizeof(double2) * NW);
        hh.stats.converged = true;
        du_dn_dev(hh, kappa, K, wcs, dirs, wave_type, ddu.as<double2>(), cg_tol, gmres_tol, s);
        download(du_dn, ddu.p, sizeof(double2) * NW, s);
        return hh.stats.converged ? BAEMM_OK : BAEMM_ERR_NOCONV;
    });
}

int baemm_derivative_ff(baemm_handle h, const double* kappa, int64_t K, const double* dirs, int64_t wcs,
                        const double* X_in, double* Y_out, double** du_dn, int wave_type, double cg_tol,
                        double gmres_tol)
{
    return guard(h, [&](Handle& hh) {
        check_common(hh, kappa, K, wcs);
        if (!dirs || !X_in || !Y_out) throw ArgError{"directions, X_in and Y_out must not be NULL"};
        if (hh.m < 1) throw ArgError{"no measurement directions"};
        const int64_t W = K * wcs, NW = hh.nv * W;
        hipStream_t s = hh.stream;
        hh.stats.converged = true;
        DevArray ddu(hh, 0, sizeof(double2) * NW);
        DevArray dX(hh, 1, sizeof(double) * 3 * hh.nv);
        DevArray dF(hh, 2, sizeof(double2) * hh.m * W);
        const double2* du = ensure_du_dn(hh, du_dn, kappa, K, wcs, dirs, wave_type, cg_tol, gmres_tol, ddu, s);
        upload(dX.p, X_in, sizeof(double) * 3 * hh.nv, s);
        derivative_dev(hh, kappa, K, wcs, dX.as<double>(), du, dF.as<double2>(), cg_tol, gmres_tol, s);
        download(Y_out, dF.p, sizeof(double2) * hh.m * W, s);
        return hh.stats.converged ? BAEMM_OK : BAEMM_ERR_NOCONV;
    });
}

int baemm_adjoint_derivative_ff(baemm_handle h, const double* kappa, int64_t K, const double* dirs, int64_t wcs,
                                const double* g, double* Y_out, double** du_dn, int wave_type, double cg_tol,
                                double gmres_tol)
{
    return guard(h, [&](Handle& hh) {
        check_common(hh, kappa, K, wcs);
        if (!dirs || !g || !Y_out) throw ArgError{"directions, g and Y_out must not be NULL"};
        if (hh.m < 1) throw ArgError{"no measurement directions"};
        const int64_t W = K * wcs, NW = hh.nv * W;
        hipStream_t s = hh.stream;
        hh.stats.converged = true;
        DevArray ddu(hh, 0, sizeof(double2) * NW);
        DevArray dg(hh, 1, sizeof(double2) * hh.m * W);
        DevArray dY(hh, 2, sizeof(double) * 3 * hh.nv);
        const double2* du = ensure_du_dn(hh, du_dn, kappa, K, wcs, dirs, wave_type, cg_tol, gmres_tol, ddu, s);
        upload(dg.p, g, sizeof(double2) * hh.m * W, s);
        adjoint_dev(hh, kappa, K, wcs, dg.as<double2>(), du, dY.as<double>(), cg_tol, gmres_tol, s);
        download(Y_out, dY.p, sizeof(double) * 3 * hh.nv, s);
        return hh.stats.converged ? BAEMM_OK : BAEMM_ERR_NOCONV;
    });
}

int baemm_gauss_newton_solve(baemm_handle h, const double* kappa, int64_t K, const double* dirs, int64_t wcs,
                             baemm_real_op M, void* M_user, baemm_real_op P, void* P_user, double alpha,
                             const double* B, int64_t ldB, double beta, double* X, int64_t ldX, double** du_dn,
                             int wave_type, double cg_tol, double gmres_tol_inner, double gmres_tol_outer,
                             int* succeeded)
{
    return guard(h, [&](Handle& hh) {
        check_common(hh, kappa, K, wcs);
        if (!dirs || !M || !P || !B || !X || ldB < 3 || ldX < 3) throw ArgError{"invalid Gauss-Newton arguments"};
        if (hh.m < 1) throw ArgError{"no measurement directions"};
        const int64_t W = K * wcs, NW = hh.nv * W, n = hh.nv, N = 3 * n;
        hipStream_t s = hh.stream;
        hh.stats.converged = true;
        DevArray ddu(hh, 0, sizeof(double2) * NW);
        DevArray dX(hh, 1, sizeof(double) * N);
        DevArray dF(hh, 2, sizeof(double2) * hh.m * W);
        DevArray dY(hh, 3, sizeof(double) * N);
        const double2* du = ensure_du_dn(hh, du_dn, kappa, K, wcs, dirs, wave_type, cg_tol, gmres_tol_inner, ddu, s);
        std::vector<double> tmp(N), ys(N);
        // A x = M x + Mass DF* DF x  (FarField.hpp:551-572)
        auto A = [&](const double* x, double* y) {
            M(M_user, x, y);
            upload(dX.p, x, sizeof(double) * N, s);
            derivative_dev(hh, kappa, K, wcs, dX.as<double>(), du, dF.as<double2>(), cg_tol, gmres_tol_inner, s);
            adjoint_dev(hh, kappa, K, wcs, dF.as<double2>(), du, dY.as<double>(), cg_tol, gmres_tol_inner, s);
            download(ys.data(), dY.p, sizeof(double) * N, s);
            for (int64_t v = 0; v < n; ++v)
                for (int d = 0; d < 3; ++d) {
                    double acc = 0.0;
                    for (int32_t e = hh.h_M_ptr[v]; e < hh.h_M_ptr[v + 1]; ++e)
                        acc += hh.h_M_val[e] * ys[3 * (int64_t)hh.h_M_idx[e] + d];
                    y[3 * v + d] += acc;
                }
        };
        // Left-preconditioned restarted GMRES on the host (one real column of length 3V).
        std::vector<double> b(N), xs(N, 0.0), r(N), z(N);
        for (int64_t v = 0; v < n; ++v)
            for (int d = 0; d < 3; ++d) b[3 * v + d] = B[v * ldB + d];
        const int m = hh.gmres_max_iter;
        std::vector<double> Vb((size_t)(m + 1) * N), Hm((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), yv(m);
        auto dot = [&](const double* a, const double* c) {
            double acc = 0.0;
            for (int64_t i = 0; i < N; ++i) acc += a[i] * c[i];
            return acc;
        };
        bool ok = false;
        double bnorm = 0.0;
        for (int cyc = 0; cyc < hh.gmres_max_restarts && !ok; ++cyc) {
            if (cyc == 0) {
                P(P_user, b.data(), z.data());
            } else {
                A(xs.data(), tmp.data());
                for (int64_t i = 0; i < N; ++i) r[i] = b[i] - tmp[i];
                P(P_user, r.data(), z.data());
            }
            const double bt = std::sqrt(dot(z.data(), z.data()));
            if (cyc == 0) bnorm = bt;
            if (bnorm == 0.0 || bt <= gmres_tol_outer * bnorm) {
                ok = true;
                break;
            }
            for (int64_t i = 0; i < N; ++i) Vb[i] = z[i] / bt;
            std::fill(g.begin(), g.end(), 0.0);
            g[0] = bt;
            int jd = 0;
            for (int j = 0; j < m; ++j) {
                A(&Vb[(size_t)j * N], tmp.data());
                P(P_user, tmp.data(), z.data());
                for (int i = 0; i <= j; ++i) {   // modified Gram-Schmidt
                    const double hij = dot(&Vb[(size_t)i * N], z.data());
                    Hm[(size_t)i * m + j] = hij;
                    for (int64_t l = 0; l < N; ++l) z[l] -= hij * Vb[(size_t)i * N + l];
                }
                const double hn = std::sqrt(dot(z.data(), z.data()));
                for (int64_t l = 0; l < N; ++l) Vb[(size_t)(j + 1) * N + l] = hn > 0.0 ? z[l] / hn : 0.0;
                for (int i = 0; i < j; ++i) {
                    const double a0 = Hm[(size_t)i * m + j], a1 = Hm[(size_t)(i + 1) * m + j];
                    Hm[(size_t)i * m + j] = cs[i] * a0 + sn[i] * a1;
                    Hm[(size_t)(i + 1) * m + j] = -sn[i] * a0 + cs[i] * a1;
                }
                const double a0 = Hm[(size_t)j * m + j];
                const double nu = std::sqrt(a0 * a0 + hn * hn);
                cs[j] = nu > 0.0 ? a0 / nu : 1.0;
                sn[j] = nu > 0.0 ? hn / nu : 0.0;
                Hm[(size_t)j * m + j] = nu;
                g[j + 1] = -sn[j] * g[j];
                g[j] = cs[j] * g[j];
                jd = j + 1;
                hh.stats.gmres_iterations++;
                if (std::fabs(g[j + 1]) <= gmres_tol_outer * bnorm || hn == 0.0) break;
            }
            for (int i = jd - 1; i >= 0; --i) {
                double acc = g[i];
                for (int l = i + 1; l < jd; ++l) acc -= Hm[(size_t)i * m + l] * yv[l];
                yv[i] = Hm[(size_t)i * m + i] != 0.0 ? acc / Hm[(size_t)i * m + i] : 0.0;
            }
            for (int i = 0; i < jd; ++i)
                for (int64_t l = 0; l < N; ++l) xs[l] += yv[i] * Vb[(size_t)i * N + l];
            hh.stats.gmres_rel_residual = std::fabs(g[jd]) / bnorm;
            if (std::fabs(g[jd]) <= gmres_tol_outer * bnorm) ok = true;
        }
        // X = alpha (M + Mass DF* DF)^{-1} B + beta X   (FarField.hpp:433-435, 581-599)
        for (int64_t v = 0; v < n; ++v)
            for (int d = 0; d < 3; ++d)
                X[v * ldX + d] = alpha * xs[3 * v + d] + (beta != 0.0 ? beta * X[v * ldX + d] : 0.0);
        if (succeeded) *succeeded = ok ? 1 : 0;
        return BAEMM_OK;
    });
}

int baemm_gauss_newton_solve_dev(baemm_handle h, const double* kappa, int64_t K, const double* dirs, int64_t wcs,
                                 baemm_dev_real_op M, void* M_user, baemm_dev_real_op P, void* P_user, double alpha,
                                 const double* B, int64_t ldB, double beta, double* X, int64_t ldX, void* du_dn,
                                 int du_dn_ready, int wave_type, double cg_tol, double gmres_tol_inner,
                                 double gmres_tol_outer, int* succeeded, void* stream)
{
    return guard(h, [&](Handle& hh) {
        check_common(hh, kappa, K, wcs);
        if (!dirs || !M || !P || !B || !X || ldB < 3 || ldX < 3) throw ArgError{"invalid Gauss-Newton arguments"};
        if (hh.m < 1) throw ArgError{"no measurement directions"};
        const int64_t W = K * wcs, NW = hh.nv * W, n = hh.nv, N = 3 * n;
        const int m = hh.gmres_max_iter;
        hipStream_t s = stream_of(hh, stream);
        void* sv = static_cast<void*>(s);
        hh.stats.converged = true;
        DevArray ddu(hh, 0, du_dn ? 16 : sizeof(double2) * NW);
        double2* du = du_dn ? static_cast<double2*>(du_dn) : ddu.as<double2>();
        if (!du_dn || !du_dn_ready) du_dn_dev(hh, kappa, K, wcs, dirs, wave_type, du, cg_tol, gmres_tol_inner, s);
        // Krylov basis (m + 1) x N, then b, x, r, z, t and DF* DF x (N each), then
        // two passes of Gram-Schmidt coefficients (m + 1 each)
        DevArray dv(hh, 1, sizeof(double) * ((size_t)(m + 1) * N + 6 * (size_t)N + 2 * (size_t)(m + 1)));
        double* Vb = dv.as<double>();
        double* b = Vb + (size_t)(m + 1) * N;
        double* xs = b + N;
        double* r = xs + N;
        double* z = r + N;
        double* t = z + N;
        double* dY = t + N;
        double* hd = dY + N;
        DevArray dF(hh, 2, sizeof(double2) * hh.m * W);
        // A x = M x + Mass DF* DF x  (FarField.hpp:551-572), all on the device
        auto A = [&](const double* x, double* y) {
            M(M_user, x, y, sv);
            derivative_dev(hh, kappa, K, wcs, x, du, dF.as<double2>(), cg_tol, gmres_tol_inner, s);
            adjoint_dev(hh, kappa, K, wcs, dF.as<double2>(), du, dY, cg_tol, gmres_tol_inner, s);
            dev_mass(hh, 1.0, dY, 3, 1.0, y, 3, 3, s);
        };
        auto norm = [&](const double* x) {
            dev_rdots(N, 1, x, N, x, hd, s);
            double v = 0.0;
            download(&v, hd, sizeof(double), s);
            return std::sqrt(v);
        };
        dev_rows3_axpby(n, 1.0, B, ldB, 0.0, b, 3, s);
        BAEMM_HIP(hipMemsetAsync(xs, 0, sizeof(double) * N, s));
        // Left-preconditioned restarted GMRES over vec(X): vectors in HBM, two
        // passes of classical Gram-Schmidt per step, the (m + 1) x m Hessenberg
        // matrix and its Givens rotations on the host (j + 1 numbers per step).
        std::vector<double> Hm((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), yv(m), h1(m + 1), h2(m + 1);
        bool ok = false;
        double bnorm = 0.0;
        for (int cyc = 0; cyc < hh.gmres_max_restarts && !ok; ++cyc) {
            if (cyc == 0) {
                P(P_user, b, z, sv);
            } else {
                A(xs, t);
                dev_raxpby(N, 1.0, b, 0.0, r, s);
                dev_raxpby(N, -1.0, t, 1.0, r, s);
                P(P_user, r, z, sv);
            }
            const double bt = norm(z);
            if (cyc == 0) bnorm = bt;
            if (bnorm == 0.0 || bt <= gmres_tol_outer * bnorm) {
                ok = true;
                break;
            }
            dev_raxpby(N, 1.0 / bt, z, 0.0, Vb, s);
            std::fill(g.begin(), g.end(), 0.0);
            g[0] = bt;
            int jd = 0;
            for (int j = 0; j < m; ++j) {
                A(Vb + (size_t)j * N, t);
                P(P_user, t, z, sv);
                dev_rdots(N, j + 1, Vb, N, z, hd, s);
                dev_rcombine(N, j + 1, Vb, N, hd, -1.0, z, s);
                dev_rdots(N, j + 1, Vb, N, z, hd + (m + 1), s);
                dev_rcombine(N, j + 1, Vb, N, hd + (m + 1), -1.0, z, s);
                download(h1.data(), hd, sizeof(double) * (j + 1), s);
                download(h2.data(), hd + (m + 1), sizeof(double) * (j + 1), s);
                for (int i = 0; i <= j; ++i) Hm[(size_t)i * m + j] = h1[i] + h2[i];
                const double hn = norm(z);
                dev_raxpby(N, hn > 0.0 ? 1.0 / hn : 0.0, z, 0.0, Vb + (size_t)(j + 1) * N, s);
                for (int i = 0; i < j; ++i) {
                    const double a0 = Hm[(size_t)i * m + j], a1 = Hm[(size_t)(i + 1) * m + j];
                    Hm[(size_t)i * m + j] = cs[i] * a0 + sn[i] * a1;
                    Hm[(size_t)(i + 1) * m + j] = -sn[i] * a0 + cs[i] * a1;
                }
                const double a0 = Hm[(size_t)j * m + j];
                const double nu = std::sqrt(a0 * a0 + hn * hn);
                cs[j] = nu > 0.0 ? a0 / nu : 1.0;
                sn[j] = nu > 0.0 ? hn / nu : 0.0;
                Hm[(size_t)j * m + j] = nu;
                g[j + 1] = -sn[j] * g[j];
                g[j] = cs[j] * g[j];
                jd = j + 1;
                hh.stats.gmres_iterations++;
                if (std::fabs(g[j + 1]) <= gmres_tol_outer * bnorm || hn == 0.0) break;
            }
            for (int i = jd - 1; i >= 0; --i) {
                double acc = g[i];
                for (int l = i + 1; l < jd; ++l) acc -= Hm[(size_t)i * m + l] * yv[l];
                yv[i] = Hm[(size_t)i * m + i] != 0.0 ? acc / Hm[(size_t)i * m + i] : 0.0;
            }
            // x += sum_i y_i V_i (the coefficients through the device)
            upload(hd, yv.data(), sizeof(double) * jd, s);
            dev_rcombine(N, jd, Vb, N, hd, 1.0, xs, s);
            BAEMM_HIP(hipStreamSynchronize(s));   // yv is reused by the next cycle
            hh.stats.gmres_rel_residual = std::fabs(g[jd]) / bnorm;
            if (std::fabs(g[jd]) <= gmres_tol_outer * bnorm) ok = true;
        }
        // X = alpha (M + Mass DF* DF)^{-1} B + beta X
        dev_rows3_axpby(n, alpha, xs, 3, beta, X, ldX, s);
        BAEMM_HIP(hipStreamSynchronize(s));
        if (succeeded) *succeeded = ok ? 1 : 0;
        return BAEMM_OK;
    });
}

void baemm_mass_metric_apply(void* user, const double* x, double* y, void* stream)
{
    const auto* mm = static_cast<const baemm_mass_metric*>(user);
    if (!mm || !mm->h || !x || !y) return;
    Handle& hh = *H(mm->h);
    hipStream_t s = stream_of(hh, stream);
    dev_mass(hh, mm->d, x, 3, 0.0, y, 3, 3, s);
    dev_raxpby(3 * hh.nv, mm->c, x, 1.0, y, s);
}

void baemm_mass_metric_jacobi(void* user, const double* x, double* y, void* stream)
{
    const auto* mm = static_cast<const baemm_mass_metric*>(user);
    if (!mm || !mm->h || !x || !y) return;
    Handle& hh = *H(mm->h);
    dev_metric_jacobi(hh, mm->c, mm->d, x, y, stream_of(hh, stream));
}

int baemm_get_solver_stats(baemm_handle h, int64_t* gmres_iterations, int64_t* gmres_cycles, double* rel_residual,
                           int64_t* cg_iterations, int64_t* operator_applies)
{
    return guard(h, [&](Handle& hh) {
        if (gmres_iterations) *gmres_iterations = hh.stats.gmres_iterations;
        if (gmres_cycles) *gmres_cycles = hh.stats.gmres_cycles;
        if (rel_residual) *rel_residual = hh.stats.gmres_rel_residual;
        if (cg_iterations) *cg_iterations = hh.stats.cg_iterations;
        if (operator_applies) *operator_applies = hh.stats.operator_applies;
        return BAEMM_OK;
    });
}

int baemm_reset_solver_stats(baemm_handle h)
{
    return guard(h, [&](Handle& hh) {
        hh.stats = SolverStats{};
        return BAEMM_OK;
    });
}

}  // extern "C"
