// sanitize_host.cpp -- host-only driver of the AddressSanitizer +
// UndefinedBehaviorSanitizer build of the library (make asan).  Needs no GPU:
// it runs the paths that do their work before the first device call and the
// host-side setup routines directly:
//  * every extern "C" entry point with value-initialised arguments (NULL
//    handles, NULL pointers, zero sizes): each must return an error code,
//    never dereference what it was handed;
//  * baemm_create / baemm_create_i32 with malformed meshes (NULL arrays,
//    empty meshes, out-of-range and negative triangle indices, bad
//    measurement counts) and, on a valid mesh, the clean-up path taken when no
//    device is present;
//  * the host geometry (centroids, normals, self terms, vertex-triangle
//    incidence, P1 mass matrix in CSR, lumped inverse areas) of an
//    icosahedron, checked against closed forms, and the parameter loading with
//    its argument and phase-range checks;
//  * the device-memory estimate of the request loop.
// Exit status 0 when every check passed; the sanitizers abort on a finding.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <type_traits>
#include <vector>

#include "internal.hpp"

using namespace baemm;

namespace {

int g_fail = 0;

void check(bool ok, const char* what)
{
    if (!ok) {
        std::printf("FAILED: %s\n", what);
        ++g_fail;
    }
}

template <class T>
T zero_value()
{
    return T{};
}

// Calls f with every argument value-initialised.
template <class R, class... A>
R call_zero(R (*f)(A...))
{
    return f(zero_value<typename std::remove_cv<A>::type>()...);
}

template <class R, class... A>
void expect_error(const char* name, R (*f)(A...))
{
    const R r = call_zero(f);
    bool ok;
    if constexpr (std::is_pointer<R>::value)
        ok = true;   // baemm_last_error(NULL): any string
    else if constexpr (std::is_same<R, int64_t>::value)
        ok = r < 0;            // counts and estimates: -1
    else
        ok = r != BAEMM_OK;    // status codes (getters: -1)
    if (!ok) std::printf("  %s returned %lld\n", name, (long long)r);
    check(ok, name);
}

void icosahedron(std::vector<double>& V, std::vector<int64_t>& T)
{
    const double p = (1.0 + std::sqrt(5.0)) / 2.0;
    V = {-1, p, 0, 1, p, 0, -1, -p, 0, 1, -p, 0, 0, -1, p, 0, 1, p, 0, -1, -p, 0, 1, -p, p, 0, -1, p, 0, 1, -p, 0, -1,
         -p, 0, 1};
    T = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
         3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
}

#define NULL_CALL(f) expect_error(#f, &f)

void null_arguments()
{
    NULL_CALL(baemm_create);
    NULL_CALL(baemm_create_i32);
    NULL_CALL(baemm_destroy);
    NULL_CALL(baemm_last_error);
    NULL_CALL(baemm_vertex_count);
    NULL_CALL(baemm_simplex_count);
    NULL_CALL(baemm_meas_count);
    NULL_CALL(baemm_use_diagonal);
    NULL_CALL(baemm_set_block_size);
    NULL_CALL(baemm_get_block_size);
    NULL_CALL(baemm_set_wave_chunk_size);
    NULL_CALL(baemm_get_wave_chunk_size);
    NULL_CALL(baemm_set_precision);
    NULL_CALL(baemm_get_precision);
    NULL_CALL(baemm_set_segments);
    NULL_CALL(baemm_set_mixed_kernel);
    NULL_CALL(baemm_set_gmres_limits);
    NULL_CALL(baemm_set_graphs);
    NULL_CALL(baemm_get_graph_stats);
    NULL_CALL(baemm_get_triangle_data);
    NULL_CALL(baemm_get_coordinates);
    NULL_CALL(baemm_device_info);
    NULL_CALL(baemm_get_flags);
    NULL_CALL(baemm_read_b);
    NULL_CALL(baemm_write_b);
    NULL_CALL(baemm_read_c);
    NULL_CALL(baemm_write_c);
    NULL_CALL(baemm_p0_buffer);
    NULL_CALL(baemm_mass_nnz);
    NULL_CALL(baemm_get_mass_matrix);
    NULL_CALL(baemm_mass_dot);
    NULL_CALL(baemm_load_parameters);
    NULL_CALL(baemm_get_parameters);
    NULL_CALL(baemm_load_boundary_operators);
    NULL_CALL(baemm_unload_boundary_operators);
    NULL_CALL(baemm_apply_boundary);
    NULL_CALL(baemm_apply_far_field);
    NULL_CALL(baemm_apply_near_field);
    NULL_CALL(baemm_create_herglotz_wave);
    NULL_CALL(baemm_create_incident_wave);
    NULL_CALL(baemm_autotune_segments);
    NULL_CALL(baemm_apply_boundary_dev);
    NULL_CALL(baemm_apply_far_field_dev);
    NULL_CALL(baemm_apply_boundary_dev_noalloc);
    NULL_CALL(baemm_apply_boundary_batch);
    NULL_CALL(baemm_batch_autotune);
    NULL_CALL(baemm_apply_mass_inverse);
    NULL_CALL(baemm_apply_mass_inverse_complex);
    NULL_CALL(baemm_apply_lumped_mass_inverse);
    NULL_CALL(baemm_boundary_potential);
    NULL_CALL(baemm_dirichlet_to_neumann);
    NULL_CALL(baemm_far_field);
    NULL_CALL(baemm_far_field_parameters);
    NULL_CALL(baemm_derivative_ff_helper);
    NULL_CALL(baemm_derivative_ff);
    NULL_CALL(baemm_adjoint_derivative_ff);
    NULL_CALL(baemm_gauss_newton_solve);
    NULL_CALL(baemm_gauss_newton_solve_dev);
    NULL_CALL(baemm_tp_energy);
    NULL_CALL(baemm_tp_metric_prepare);
    NULL_CALL(baemm_tp_metric_apply_host);
    NULL_CALL(baemm_tp_metric_jacobi_host);
    NULL_CALL(baemm_gmres_dev);
    NULL_CALL(baemm_gmres_batch);
    NULL_CALL(baemm_get_solver_stats);
    NULL_CALL(baemm_reset_solver_stats);
    NULL_CALL(baemm_server_create);
    NULL_CALL(baemm_server_set_policy);
    NULL_CALL(baemm_server_set_fusion);
    NULL_CALL(baemm_server_destroy);
    NULL_CALL(baemm_server_submit_gn);
    NULL_CALL(baemm_server_wait);
    NULL_CALL(baemm_server_stats);
    NULL_CALL(baemm_server_request_bytes);
    NULL_CALL(baemm_set_timing);
    NULL_CALL(baemm_get_timing);
    NULL_CALL(baemm_dev_alloc);
    NULL_CALL(baemm_batch_set_tuning);
    NULL_CALL(baemm_batch_get_own_launch);
    check(baemm_batch_set_own_launch(-1.0) == BAEMM_ERR_ARG, "baemm_batch_set_own_launch(-1)");
    check(baemm_batch_get_tuning(nullptr, nullptr, nullptr, nullptr) == BAEMM_OK ||
              baemm_batch_get_tuning(nullptr, nullptr, nullptr, nullptr) == BAEMM_ERR_ARG,
          "baemm_batch_get_tuning(NULL...)");
    int64_t st4[4];
    check(baemm_batch_get_stats(nullptr, nullptr, nullptr, nullptr) == BAEMM_OK &&
              baemm_batch_get_stats(st4, st4 + 1, st4 + 2, st4 + 3) == BAEMM_OK,
          "baemm_batch_get_stats");
    baemm_mass_metric_apply(nullptr, nullptr, nullptr, nullptr);   // returns at once
    baemm_mass_metric_jacobi(nullptr, nullptr, nullptr, nullptr);
    baemm_tp_metric_apply(nullptr, nullptr, nullptr, nullptr);
    baemm_tp_metric_jacobi(nullptr, nullptr, nullptr, nullptr);
    check(baemm_version() > 0, "baemm_version");
    // n < 0 is refused before the device is looked at
    check(baemm_hw_sincos(0, nullptr, -1, nullptr, nullptr) == BAEMM_ERR_ARG, "baemm_hw_sincos(n < 0)");
}

void malformed_meshes(bool have_device)
{
    std::vector<double> V;
    std::vector<int64_t> T;
    icosahedron(V, T);
    std::vector<int32_t> T32(T.begin(), T.end());
    const double dirs[6] = {1, 0, 0, 0, 0, 1};
    baemm_handle h = nullptr;
    struct Case {
        const char* what;
        const double* v;
        int64_t nv;
        const int64_t* t;
        int64_t nt;
        const double* m;
        int64_t nm;
    } cases[] = {
        {"NULL coordinates", nullptr, 12, T.data(), 20, nullptr, 0},
        {"NULL triangles", V.data(), 12, nullptr, 20, nullptr, 0},
        {"no triangles", V.data(), 12, T.data(), 0, nullptr, 0},
        {"two vertices", V.data(), 2, T.data(), 20, nullptr, 0},
        {"index past the vertex count", V.data(), 11, T.data(), 20, nullptr, 0},
        {"negative measurement count", V.data(), 12, T.data(), 20, dirs, -1},
        {"NULL measurement directions", V.data(), 12, T.data(), 20, nullptr, 2},
        {"more than 2^28 triangles", V.data(), 12, T.data(), (int64_t)1 << 29, nullptr, 0},
    };
    for (const Case& c : cases) {
        h = reinterpret_cast<baemm_handle>(0x1);
        const int st = baemm_create(c.v, c.nv, c.t, c.nt, c.m, c.nm, 0, &h);
        const char* msg = baemm_last_error(nullptr);
        std::printf("  create, %s: %d '%s'\n", c.what, st, msg ? msg : "");
        check(st == BAEMM_ERR_ARG && h == nullptr && msg && msg[0], c.what);
    }
    std::vector<int64_t> Tneg = T;
    Tneg[17] = -3;
    check(baemm_create(V.data(), 12, Tneg.data(), 20, nullptr, 0, 0, &h) == BAEMM_ERR_ARG, "negative index");
    std::vector<int32_t> T32bad = T32;
    T32bad[59] = 12;
    check(baemm_create_i32(V.data(), 12, T32bad.data(), 20, nullptr, 0, 0, &h) == BAEMM_ERR_ARG,
          "int32 index past the vertex count");
    check(baemm_create(V.data(), 12, T.data(), 20, dirs, 2, 0, nullptr) == BAEMM_ERR_ARG, "NULL out");
    // A valid mesh: without a device the call fails in its first HIP call and
    // releases what it had built.
    const int st = baemm_create(V.data(), 12, T.data(), 20, dirs, 2, 0, &h);
    std::printf("  create, valid mesh: %d '%s'\n", st, baemm_last_error(nullptr));
    if (have_device) {
        check(st == BAEMM_OK && h, "valid mesh");
        if (h) baemm_destroy(h);
    } else {
        check(st == BAEMM_ERR_HIP && h == nullptr, "valid mesh without a device");
    }
    baemm_server srv = nullptr;
    check(baemm_server_create(0, 0, 1 << 20, &srv) == BAEMM_ERR_ARG, "server with no slots");
    check(baemm_server_create(0, 2, 0, &srv) == BAEMM_ERR_ARG, "server with no memory budget");
}

void host_geometry()
{
    std::vector<double> V;
    std::vector<int64_t> T;
    icosahedron(V, T);
    Handle h;
    h.nv = 12;
    h.nt = 20;
    h.m = 3;
    h.coords = V;
    h.tris.assign(T.begin(), T.end());
    build_host_geometry(h);
    // vertices (0, +-1, +-phi) cyclically: edge 2, area sqrt(3), centroid
    // distance phi^2 / sqrt(3), bounding box [-phi, phi]^3 (diagonal 2 sqrt(3) phi)
    const double phi = (1.0 + std::sqrt(5.0)) / 2.0;
    double area = 0.0, worst_nrm = 0.0, worst_mid = 0.0;
    for (int t = 0; t < 20; ++t) {
        area += h.h_area[t];
        const float* c = &h.h_mid4[4 * t];
        const float* n = &h.h_nrm4[4 * t];
        const double r = std::sqrt((double)c[0] * c[0] + (double)c[1] * c[1] + (double)c[2] * c[2]);
        worst_mid = std::max(worst_mid, std::fabs(r - phi * phi / std::sqrt(3.0)));
        // outward unit normal parallel to the centroid
        worst_nrm = std::max(worst_nrm, std::fabs((c[0] * n[0] + c[1] * n[1] + c[2] * n[2]) / r - 1.0));
        check(h.h_sdiag[t] > 0.0f && std::isfinite(h.h_sdiag[t]), "self term finite and positive");
    }
    check(std::fabs(area - 20.0 * std::sqrt(3.0)) < 1e-5, "total area 20 sqrt(3)");
    check(worst_mid < 1e-6 && worst_nrm < 1e-6, "centroids and outward normals");
    check(std::fabs(h.extent - 2.0 * std::sqrt(3.0) * phi) < 1e-12, "extent");
    check(h.nt_pad % kTile == 0 && h.nt_pad >= 20 && (int64_t)h.h_mid4.size() == 4 * h.nt_pad, "padding");
    // every vertex has 5 incident triangles; the mass matrix has 12 + 60 entries
    bool five = true;
    for (int v = 0; v < 12; ++v) five = five && (h.h_vt_ptr[v + 1] - h.h_vt_ptr[v] == 5);
    check(five, "vertex-triangle incidence");
    check(h.h_M_ptr[12] == 72, "mass matrix nonzeros");
    double msum = 0.0;
    for (double x : h.h_M_val) msum += x;
    check(std::fabs(msum - area) < 1e-9 * area, "mass matrix sums to the area");
    for (int v = 0; v < 12; ++v) check(std::fabs(h.h_lumped_inv[v] * 5.0 * std::sqrt(3.0) / 9.0 - 1.0) < 1e-6,
                                       "lumped inverse areas");
}

void parameters()
{
    std::vector<double> V;
    std::vector<int64_t> T;
    icosahedron(V, T);
    Handle h;
    h.nv = 12;
    h.nt = 20;
    h.coords = V;
    h.tris.assign(T.begin(), T.end());
    build_host_geometry(h);
    const double kappa[4] = {1.0, 2.0, 3.0, 4.0};
    double coeff[32];
    for (int k = 0; k < 32; ++k) coeff[k] = (k % 3) - 1.0;
    auto refused = [&](const double* kp, const double* c, int64_t W, int64_t wcs) {
        try {
            load_params(h, kp, c, W, wcs, 1.0);
        } catch (const ArgError&) {
            return true;
        }
        return false;
    };
    check(refused(kappa, coeff, 0, 1), "W = 0");
    check(refused(kappa, coeff, 6, 4), "W not a multiple of wcs");
    check(refused(nullptr, coeff, 4, 1), "NULL kappa");
    check(refused(kappa, nullptr, 4, 1), "NULL coeff");
    const double huge[1] = {1e9};
    check(refused(huge, coeff, 1, 1), "phase range");
    check(!refused(kappa, coeff, 64, 16), "valid parameters");
    check(h.prm.K == 4 && h.prm.W == 64 && h.prm.loaded, "parameter shape");
    const std::vector<ChunkCoef> cc = chunk_coefs(h.prm, 1, 3, 0.5f, nullptr);
    check(cc.size() == 3 && cc[2].kap == 4.0f, "chunk coefficients");
}

void request_estimate()
{
    std::vector<double> V;
    std::vector<int64_t> T;
    icosahedron(V, T);
    const double kappa[2] = {1.0, 2.0}, inc[3] = {0, 0, 1};
    baemm_gn_request r{};
    r.coords = V.data();
    r.vertex_count = 12;
    r.triangles = T.data();
    r.simplex_count = 20;
    r.mesh_id = -1;
    r.meas_directions = inc;
    r.meas_count = 1;
    r.kappa = kappa;
    r.kappa_count = 2;
    r.inc_directions = inc;
    r.wave_chunk_size = 1;
    const int64_t b = baemm_server_request_bytes(&r);
    check(b > 0, "request estimate");
    r.wave_chunk_size = 0;
    check(baemm_server_request_bytes(&r) == -1, "request estimate of an invalid request");
}

}  // namespace

int main()
{
    int ndev = 0;
    const bool have_device = hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0;
    std::printf("devices: %d\n", have_device ? ndev : 0);
    null_arguments();
    malformed_meshes(have_device);
    host_geometry();
    parameters();
    request_estimate();
    std::printf("host sanitizer run: %s (%d failed checks)\n", g_fail ? "FAILED" : "OK", g_fail);
    return g_fail ? 1 : 0;
}
