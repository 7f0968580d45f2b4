// capi.cpp -- extern "C" entry points of libbaemm_hip (declared in include/baemm.h).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <new>
#include <string>

#include "capi_util.hpp"

using namespace baemm;

namespace {

// Message of the last failed call on this thread that has no single handle
// to carry it (baemm_create*, the batched entry points); returned by
// baemm_last_error(NULL).
thread_local std::string g_create_error;

template <class I>
int create_impl(const double* coords, int64_t nv, const I* tris, int64_t nt, const double* meas, int64_t m,
                int device, baemm_handle* out)
{
    g_create_error.clear();
    auto fail = [](const char* msg) {
        g_create_error = msg;
        return BAEMM_ERR_ARG;
    };
    if (!out) return fail("out is NULL");
    *out = nullptr;
    if (!coords || !tris || nv < 3 || nt < 1) return fail("empty or NULL mesh");
    if (m < 0 || (m > 0 && !meas)) return fail("invalid measurement directions");
    if (nt > (int64_t)1 << 28 || nv > (int64_t)1 << 28) return fail("mesh larger than 2^28 vertices or triangles");
    Handle* h = new (std::nothrow) Handle();
    if (!h) {
        g_create_error = "host allocation failed";
        return BAEMM_ERR_INTERNAL;
    }
    int st = BAEMM_OK;
    try {
        // Arguments are checked before the first HIP call.
        h->tris.resize(3 * nt);
        for (int64_t k = 0; k < 3 * nt; ++k) {
            const int64_t v = (int64_t)tris[k];
            if (v < 0 || v >= nv) throw ArgError{"triangle index out of range"};
            h->tris[k] = (int32_t)v;
        }
        int ndev = 0;
        BAEMM_HIP(hipGetDeviceCount(&ndev));
        // Helmholtz_OpenCL.hpp:216-223: an out-of-range device index falls back to device 0.
        h->device = (device >= 0 && device < ndev) ? device : 0;
        BAEMM_HIP(hipSetDevice(h->device));
        h->nv = nv;
        h->nt = nt;
        h->m = m;
        h->coords.assign(coords, coords + 3 * nv);
        if (m > 0) h->meas.assign(meas, meas + 3 * m);
        build_host_geometry(*h);
        // Non-blocking: the handle's work never waits for the legacy null stream.
        BAEMM_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        upload_geometry(*h);
        BAEMM_HIP(hipEventCreate(&h->ev0));
        BAEMM_HIP(hipEventCreate(&h->ev1));
    } catch (const ArgError& e) {
        g_create_error = e.msg;
        st = BAEMM_ERR_ARG;
    } catch (const HipError& e) {
        g_create_error = std::string("HIP error '") + hipGetErrorString(e.err) + "' in " + e.what;
        st = BAEMM_ERR_HIP;
    } catch (const std::bad_alloc&) {
        g_create_error = "host allocation failed";
        st = BAEMM_ERR_INTERNAL;
    } catch (const std::exception& e) {
        g_create_error = e.what();
        st = BAEMM_ERR_INTERNAL;
    }
    if (st != BAEMM_OK) {
        free_device(*h);
        if (h->ev0) (void)hipEventDestroy(h->ev0);
        if (h->ev1) (void)hipEventDestroy(h->ev1);
        if (h->stream) (void)hipStreamDestroy(h->stream);
        delete h;
        return st;
    }
    *out = reinterpret_cast<baemm_handle>(h);
    return BAEMM_OK;
}

void check_ld(int64_t ld, int64_t W, const char* what)
{
    if (ld < W) throw ArgError{std::string(what) + ": leading dimension smaller than wave_count"};
}

// Host-pointer apply: copy B (rows x W) and, when beta != 0, C in; run; copy C out.
template <class Dev>
void host_apply(Handle& h, const double* alpha, const double* B, int64_t ldB, int64_t brows, const double* beta,
                double* C, int64_t ldC, int64_t crows, int64_t ncols_in, int64_t ncols_out, Dev&& dev)
{
    const Params& p = h.prm;
    if (!p.loaded) throw StateError{"parameters are not loaded"};
    if (!B || !C) throw ArgError{"B and C must not be NULL"};
    check_ld(ldB, ncols_in, "B");
    check_ld(ldC, ncols_out, "C");
    hipStream_t s = h.stream;
    h.ws_B.reserve(sizeof(cdouble) * brows * ncols_in);
    h.ws_out.reserve(sizeof(cdouble) * crows * ncols_out);
    BAEMM_HIP(copy_rows_async(h.ws_B.ptr, sizeof(cdouble) * ncols_in, B, sizeof(cdouble) * ldB,
                               sizeof(cdouble) * ncols_in, brows, hipMemcpyHostToDevice, s));
    const cdouble be = cplx(beta, 0.0);
    if (be != cdouble(0.0, 0.0))
        BAEMM_HIP(copy_rows_async(h.ws_out.ptr, sizeof(cdouble) * ncols_out, C, sizeof(cdouble) * ldC,
                                   sizeof(cdouble) * ncols_out, crows, hipMemcpyHostToDevice, s));
    ApplyArgs a{cplx(alpha, 1.0), be, h.ws_B.ptr, ncols_in, kC128, h.ws_out.ptr, ncols_out, kC128};
    const bool touched = dev(a, s);
    if (touched)
        BAEMM_HIP(copy_rows_async(C, sizeof(cdouble) * ldC, h.ws_out.ptr, sizeof(cdouble) * ncols_out,
                                   sizeof(cdouble) * ncols_out, crows, hipMemcpyDeviceToHost, s));
    BAEMM_HIP(hipStreamSynchronize(s));
}

}  // namespace

namespace baemm {
void set_thread_error(const std::string& msg) { g_create_error = msg; }
}  // namespace baemm

extern "C" {

int baemm_version(void) { return 10000; }

int baemm_create(const double* coords, int64_t vertex_count, const int64_t* triangles, int64_t simplex_count,
                 const double* meas_directions, int64_t meas_count, int device, baemm_handle* out)
{
    return create_impl(coords, vertex_count, triangles, simplex_count, meas_directions, meas_count, device, out);
}

int baemm_create_i32(const double* coords, int64_t vertex_count, const int32_t* triangles, int64_t simplex_count,
                     const double* meas_directions, int64_t meas_count, int device, baemm_handle* out)
{
    return create_impl(coords, vertex_count, triangles, simplex_count, meas_directions, meas_count, device, out);
}

int baemm_destroy(baemm_handle h)
{
    if (!h) return BAEMM_ERR_ARG;
    Handle* hh = H(h);
    {
        std::lock_guard<std::mutex> lk(hh->mtx);
        (void)hipSetDevice(hh->device);
        if (hh->stream) (void)hipStreamSynchronize(hh->stream);
        free_device(*hh);
        if (hh->ev0) (void)hipEventDestroy(hh->ev0);
        if (hh->ev1) (void)hipEventDestroy(hh->ev1);
        for (hipEvent_t e : hh->ev_step)
            if (e) (void)hipEventDestroy(e);
        if (hh->pin_ctl) (void)hipHostFree(hh->pin_ctl);
        if (hh->stream) (void)hipStreamDestroy(hh->stream);
    }
    delete hh;
    return BAEMM_OK;
}

const char* baemm_last_error(baemm_handle h)
{
    if (h) return H(h)->last_error.c_str();
    return g_create_error.empty() ? "null handle" : g_create_error.c_str();
}

int64_t baemm_vertex_count(baemm_handle h) { return h ? H(h)->nv : -1; }
int64_t baemm_simplex_count(baemm_handle h) { return h ? H(h)->nt : -1; }
int64_t baemm_meas_count(baemm_handle h) { return h ? H(h)->m : -1; }

int baemm_use_diagonal(baemm_handle h, int use_it)
{
    return guard(h, [&](Handle& hh) {
        hh.use_diagonal = use_it != 0;
        return BAEMM_OK;
    });
}

int baemm_set_block_size(baemm_handle h, int block_size)
{
    return guard(h, [&](Handle& hh) {
        if (block_size < 1) throw ArgError{"block_size must be >= 1"};
        hh.block_size = block_size;
        return BAEMM_OK;
    });
}

int baemm_get_block_size(baemm_handle h) { return h ? H(h)->block_size : -1; }

int baemm_set_wave_chunk_size(baemm_handle h, int wcs)
{
    return guard(h, [&](Handle& hh) {
        if (wcs < 1) throw ArgError{"wave_chunk_size must be >= 1"};
        hh.wave_chunk_size = wcs;
        return BAEMM_OK;
    });
}

int baemm_get_wave_chunk_size(baemm_handle h) { return h ? H(h)->wave_chunk_size : -1; }

int baemm_set_precision(baemm_handle h, int precision)
{
    return guard(h, [&](Handle& hh) {
        if (precision != BAEMM_PRECISION_FP32 && precision != BAEMM_PRECISION_MIXED &&
            precision != BAEMM_PRECISION_FP32_MFMA)
            throw ArgError{"unknown precision mode"};
        hh.precision = precision;
        return BAEMM_OK;
    });
}

int baemm_get_precision(baemm_handle h) { return h ? H(h)->precision : -1; }

int baemm_set_segments(baemm_handle h, int segments)
{
    return guard(h, [&](Handle& hh) {
        if (segments < 0) throw ArgError{"segments must be >= 0"};
        hh.segments = segments;
        return BAEMM_OK;
    });
}

int baemm_set_mixed_kernel(baemm_handle h, int variant)
{
    return guard(h, [&](Handle& hh) {
        if (variant < 0 || variant > 2) throw ArgError{"mixed kernel variant must be 0, 1 or 2"};
        hh.mixed_kernel = variant;
        return BAEMM_OK;
    });
}

int baemm_set_gmres_limits(baemm_handle h, int max_iter, int max_restarts)
{
    return guard(h, [&](Handle& hh) {
        if (max_iter < 1 || max_restarts < 1) throw ArgError{"GMRES limits must be >= 1"};
        hh.gmres_max_iter = max_iter;
        hh.gmres_max_restarts = max_restarts;
        return BAEMM_OK;
    });
}

int baemm_set_graphs(baemm_handle h, int on)
{
    return guard(h, [&](Handle& hh) {
        hh.use_graphs = on != 0;
        return BAEMM_OK;
    });
}

int baemm_get_graph_stats(baemm_handle h, int64_t* captures, int64_t* launches)
{
    return guard(h, [&](Handle& hh) {
        if (captures) *captures = hh.graph_captures;
        if (launches) *launches = hh.graph_launches;
        return BAEMM_OK;
    });
}

int baemm_get_triangle_data(baemm_handle h, float* mid4, float* nrm4, float* areas, float* sdiag)
{
    return guard(h, [&](Handle& hh) {
        if (mid4) std::memcpy(mid4, hh.h_mid4.data(), sizeof(float) * 4 * hh.nt);
        if (nrm4) std::memcpy(nrm4, hh.h_nrm4.data(), sizeof(float) * 4 * hh.nt);
        if (areas) std::memcpy(areas, hh.h_area.data(), sizeof(float) * hh.nt);
        if (sdiag) std::memcpy(sdiag, hh.h_sdiag.data(), sizeof(float) * hh.nt);
        return BAEMM_OK;
    });
}

int64_t baemm_mass_nnz(baemm_handle h) { return h ? (int64_t)H(h)->h_M_idx.size() : -1; }

int baemm_get_mass_matrix(baemm_handle h, int32_t* rowptr, int32_t* colidx, double* values)
{
    return guard(h, [&](Handle& hh) {
        if (rowptr) std::memcpy(rowptr, hh.h_M_ptr.data(), sizeof(int32_t) * hh.h_M_ptr.size());
        if (colidx) std::memcpy(colidx, hh.h_M_idx.data(), sizeof(int32_t) * hh.h_M_idx.size());
        if (values) std::memcpy(values, hh.h_M_val.data(), sizeof(double) * hh.h_M_val.size());
        return BAEMM_OK;
    });
}

int baemm_mass_dot(baemm_handle h, double alpha, const double* X, int64_t ldX, double beta, double* Y, int64_t ldY,
                   int64_t nrhs)
{
    return guard(h, [&](Handle& hh) {
        if (!X || !Y || nrhs < 0 || ldX < nrhs || ldY < nrhs) throw ArgError{"invalid mass_dot arguments"};
        std::vector<double> row(nrhs);
        for (int64_t v = 0; v < hh.nv; ++v) {
            std::fill(row.begin(), row.end(), 0.0);
            for (int32_t e = hh.h_M_ptr[v]; e < hh.h_M_ptr[v + 1]; ++e) {
                const double m = hh.h_M_val[e];
                const double* x = X + (int64_t)hh.h_M_idx[e] * ldX;
                for (int64_t k = 0; k < nrhs; ++k) row[k] += m * x[k];
            }
            double* y = Y + v * ldY;
            for (int64_t k = 0; k < nrhs; ++k) y[k] = alpha * row[k] + (beta != 0.0 ? beta * y[k] : 0.0);
        }
        return BAEMM_OK;
    });
}

int baemm_load_parameters(baemm_handle h, const double* kappa, const double* coeff, int64_t W, int64_t wcs,
                          double factor)
{
    return guard(h, [&](Handle& hh) {
        load_params(hh, kappa, coeff, W, wcs, factor);
        return BAEMM_OK;
    });
}

int baemm_get_parameters(baemm_handle h, float* kappa, float* coeff, int64_t* W, int64_t* wcs, int64_t* K)
{
    return guard(h, [&](Handle& hh) {
        const Params& p = hh.prm;
        if (kappa) std::memcpy(kappa, p.kappa.data(), sizeof(float) * p.K);
        if (coeff) std::memcpy(coeff, p.c.data(), sizeof(cfloat) * 4 * p.K);
        if (W) *W = p.W;
        if (wcs) *wcs = p.wcs;
        if (K) *K = p.K;
        return BAEMM_OK;
    });
}

int baemm_load_boundary_operators(baemm_handle h, const double* kappa, const double* coeff, int64_t W, int64_t wcs)
{
    return baemm_load_parameters(h, kappa, coeff, W, wcs, 1.0);
}

int baemm_unload_boundary_operators(baemm_handle h)
{
    return guard(h, [&](Handle& hh) {
        BAEMM_HIP(hipStreamSynchronize(hh.stream));
        return BAEMM_OK;
    });
}

int baemm_apply_boundary(baemm_handle h, const double* alpha, const double* B, int64_t ldB, const double* beta,
                         double* C, int64_t ldC)
{
    return guard(h, [&](Handle& hh) {
        const int64_t W = hh.prm.W;
        host_apply(hh, alpha, B, ldB, hh.nv, beta, C, ldC, hh.nv, W, W, [&](const ApplyArgs& a, hipStream_t s) {
            apply_boundary_dev(hh, a, s);
            return hh.prm.kernel_on() || hh.prm.mass_on();
        });
        return BAEMM_OK;
    });
}

int baemm_apply_far_field(baemm_handle h, const double* alpha, const double* B, int64_t ldB, const double* beta,
                          double* F, int64_t ldF)
{
    return guard(h, [&](Handle& hh) {
        const int64_t W = hh.prm.W;
        host_apply(hh, alpha, B, ldB, hh.nv, beta, F, ldF, hh.m, W, W, [&](const ApplyArgs& a, hipStream_t s) {
            apply_far_dev(hh, a, s);
            return hh.prm.any(1) || hh.prm.any(2);
        });
        return BAEMM_OK;
    });
}

int baemm_apply_near_field(baemm_handle h, const double* alpha, const double* B, int64_t ldB, const double* beta,
                           double* U, int64_t ldU, const double* pts, int64_t P)
{
    return guard(h, [&](Handle& hh) {
        if (!pts || P < 1) throw ArgError{"evaluation points must be given"};
        const int64_t W = hh.prm.W;
        const int64_t P_pad = ((P + kTile - 1) / kTile) * kTile;
        std::vector<float> p4(4 * P_pad, 0.0f);
        for (int64_t i = 0; i < P_pad; ++i) {
            const int64_t src = std::min(i, P - 1);
            for (int d = 0; d < 3; ++d) p4[4 * i + d] = (float)pts[3 * src + d];
        }
        hh.ws_pts.reserve(sizeof(float) * p4.size());
        BAEMM_HIP(hipMemcpyAsync(hh.ws_pts.ptr, p4.data(), sizeof(float) * p4.size(), hipMemcpyHostToDevice,
                                 hh.stream));
        host_apply(hh, alpha, B, ldB, hh.nv, beta, U, ldU, P, W, W, [&](const ApplyArgs& a, hipStream_t s) {
            apply_near_dev(hh, a, hh.ws_pts.as<float4>(), P, s);
            return hh.prm.any(1) || hh.prm.any(2);
        });
        return BAEMM_OK;
    });
}

int baemm_create_herglotz_wave(baemm_handle h, const double* alpha, const double* g, int64_t ldg, const double* beta,
                               double* C, int64_t ldC)
{
    return guard(h, [&](Handle& hh) {
        const int64_t W = hh.prm.W;
        host_apply(hh, alpha, g, ldg, hh.m, beta, C, ldC, hh.nv, W, W, [&](const ApplyArgs& a, hipStream_t s) {
            herglotz_dev(hh, a, s);
            return hh.prm.any(1) || hh.prm.any(2);
        });
        return BAEMM_OK;
    });
}

int baemm_create_incident_wave(baemm_handle h, const double* alpha, const double* dirs, int64_t inc_count,
                               const double* beta, double* C, int64_t ldC, int wave_type)
{
    return guard(h, [&](Handle& hh) {
        const Params& p = hh.prm;
        if (!p.loaded) throw StateError{"parameters are not loaded"};
        if (!C) throw ArgError{"C must not be NULL"};
        check_ld(ldC, p.W, "C");
        hipStream_t s = hh.stream;
        const int64_t W = p.W;
        hh.ws_out.reserve(sizeof(cdouble) * hh.nv * W);
        const cdouble be = cplx(beta, 0.0);
        if (be != cdouble(0.0, 0.0))
            BAEMM_HIP(copy_rows_async(hh.ws_out.ptr, sizeof(cdouble) * W, C, sizeof(cdouble) * ldC,
                                       sizeof(cdouble) * W, hh.nv, hipMemcpyHostToDevice, s));
        ApplyArgs a{cplx(alpha, 1.0), be, nullptr, 0, kC128, hh.ws_out.ptr, W, kC128};
        incident_dev(hh, a, dirs, inc_count, wave_type, s);
        if (p.any(1) || p.any(2))
            BAEMM_HIP(copy_rows_async(C, sizeof(cdouble) * ldC, hh.ws_out.ptr, sizeof(cdouble) * W,
                                       sizeof(cdouble) * W, hh.nv, hipMemcpyDeviceToHost, s));
        BAEMM_HIP(hipStreamSynchronize(s));
        return BAEMM_OK;
    });
}

int baemm_autotune_segments(baemm_handle h, int repeats, int* best_segments, double* best_ms)
{
    return guard(h, [&](Handle& hh) {
        if (!hh.prm.loaded) throw StateError{"operators are not loaded"};
        const int64_t W = hh.prm.W;
        const size_t bytes = sizeof(cdouble) * hh.nv * W;
        hh.io[4].reserve(bytes);
        hh.io[5].reserve(bytes);
        std::vector<cdouble> Bh((size_t)hh.nv * W);
        for (size_t k = 0; k < Bh.size(); ++k) Bh[k] = cdouble(std::sin(0.37 * k), std::cos(0.11 * k));
        hipStream_t s = hh.stream;
        BAEMM_HIP(hipMemcpyAsync(hh.io[4].ptr, Bh.data(), bytes, hipMemcpyHostToDevice, s));
        BAEMM_HIP(hipStreamSynchronize(s));
        ApplyArgs a{cdouble(1.0, 0.0), cdouble(0.0, 0.0), hh.io[4].ptr, W, kC128, hh.io[5].ptr, W, kC128};
        const int saved = hh.segments;
        const int reps = std::max(1, repeats);
        const bool mixed = hh.precision != BAEMM_PRECISION_FP32;   // the MFMA kernels
        const int64_t tiles = hh.nt_pad / (mixed ? kMixTile : kTile);
        const int64_t max_seg = mixed ? hh.nt_pad / 64 : hh.nt_pad / kSrcBlock;
        hipEvent_t e0, e1;
        BAEMM_HIP(hipEventCreate(&e0));
        BAEMM_HIP(hipEventCreate(&e1));
        int best = 0;
        double bt = 1e300;
        std::vector<int> tried;
        try {
            // candidate source-segment counts: launches of about 512 .. 8192 blocks
            for (int64_t blocks = 512; blocks <= 8192; blocks *= 2) {
                const int S = (int)std::max<int64_t>(1, std::min<int64_t>(max_seg, (blocks + tiles - 1) / tiles));
                if (std::find(tried.begin(), tried.end(), S) != tried.end()) continue;
                tried.push_back(S);
                hh.segments = S;
                apply_boundary_dev(hh, a, s);                  // sizes the partial buffer
                BAEMM_HIP(hipEventRecord(e0, s));
                for (int r = 0; r < reps; ++r) apply_boundary_dev(hh, a, s);
                BAEMM_HIP(hipEventRecord(e1, s));
                BAEMM_HIP(hipEventSynchronize(e1));
                float ms = 0.0f;
                BAEMM_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (ms / reps < bt) {
                    bt = ms / reps;
                    best = S;
                }
            }
        } catch (...) {
            hh.segments = saved;
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        hh.segments = best;
        if (best_segments) *best_segments = best;
        if (best_ms) *best_ms = bt;
        return BAEMM_OK;
    });
}

int baemm_apply_boundary_dev(baemm_handle h, const double* alpha, const void* B, int64_t ldB, int B_type,
                             const double* beta, void* C, int64_t ldC, int C_type, void* stream)
{
    return guard(h, [&](Handle& hh) {
        if (!hh.prm.loaded) throw StateError{"operators are not loaded"};
        if (!B || !C) throw ArgError{"B and C must not be NULL"};
        if ((B_type != kC128 && B_type != kC64) || (C_type != kC128 && C_type != kC64))
            throw ArgError{"unknown element type"};
        check_ld(ldB, hh.prm.W, "B");
        check_ld(ldC, hh.prm.W, "C");
        ApplyArgs a{cplx(alpha, 1.0), cplx(beta, 0.0), B, ldB, B_type, C, ldC, C_type};
        apply_boundary_dev(hh, a, stream_of(hh, stream));
        return BAEMM_OK;
    });
}

int baemm_apply_far_field_dev(baemm_handle h, const double* alpha, const void* B, int64_t ldB, int B_type,
                              const double* beta, void* F, int64_t ldF, int F_type, void* stream)
{
    return guard(h, [&](Handle& hh) {
        if (!hh.prm.loaded) throw StateError{"parameters are not loaded"};
        if (!B || !F) throw ArgError{"B and F must not be NULL"};
        if ((B_type != kC128 && B_type != kC64) || (F_type != kC128 && F_type != kC64))
            throw ArgError{"unknown element type"};
        check_ld(ldB, hh.prm.W, "B");
        check_ld(ldF, hh.prm.W, "F");
        ApplyArgs a{cplx(alpha, 1.0), cplx(beta, 0.0), B, ldB, B_type, F, ldF, F_type};
        apply_far_dev(hh, a, stream_of(hh, stream));
        return BAEMM_OK;
    });
}

// ---------------------------------------------------------------- P0 buffers

static int p0_read(baemm_handle h, int which, const float* input, int64_t ld_input, int64_t wave_count)
{
    return guard(h, [&](Handle& hh) {
        if (!input || wave_count < 1 || ld_input < wave_count) throw ArgError{"invalid input, ld or wave_count"};
        P0State& st = hh.p0;
        std::vector<cfloat>& buf = st.host[which];
        buf.resize((size_t)(hh.nt * wave_count));
        const cfloat* in = reinterpret_cast<const cfloat*>(input);
        for (int64_t i = 0; i < hh.nt; ++i)
            std::copy(in + i * ld_input, in + i * ld_input + wave_count, buf.begin() + i * wave_count);
        st.W[which] = wave_count;
        st.host_current[which] = true;
        return BAEMM_OK;
    });
}

static int p0_write(baemm_handle h, int which, float* output, int64_t ld_output)
{
    return guard(h, [&](Handle& hh) {
        if (!p0_sync(hh, which, hh.stream))
            throw StateError{which == 0 ? "no B buffer: no operator apply yet"
                                        : "no C buffer: no boundary-operator apply since the last other operator"};
        const P0State& st = hh.p0;
        const int64_t W = st.W[which];
        if (!output || ld_output < W) throw ArgError{"invalid output or ld smaller than the wave count"};
        cfloat* out = reinterpret_cast<cfloat*>(output);
        for (int64_t i = 0; i < hh.nt; ++i)
            std::copy(st.host[which].begin() + i * W, st.host[which].begin() + (i + 1) * W, out + i * ld_output);
        return BAEMM_OK;
    });
}

int baemm_read_b(baemm_handle h, const float* input, int64_t ld_input, int64_t wave_count)
{
    return p0_read(h, 0, input, ld_input, wave_count);
}

int baemm_write_b(baemm_handle h, float* output, int64_t ld_output) { return p0_write(h, 0, output, ld_output); }

int baemm_read_c(baemm_handle h, const float* input, int64_t ld_input, int64_t wave_count)
{
    return p0_read(h, 1, input, ld_input, wave_count);
}

int baemm_write_c(baemm_handle h, float* output, int64_t ld_output) { return p0_write(h, 1, output, ld_output); }

int baemm_p0_buffer(baemm_handle h, int which, float** data, int64_t* ld)
{
    return guard(h, [&](Handle& hh) {
        if ((which != 0 && which != 1) || !data) throw ArgError{"which must be 0 (B) or 1 (C), data not NULL"};
        if (!p0_sync(hh, which, hh.stream))
            throw StateError{which == 0 ? "no B buffer: no operator apply yet"
                                        : "no C buffer: no boundary-operator apply since the last other operator"};
        *data = reinterpret_cast<float*>(hh.p0.host[which].data());
        if (ld) *ld = hh.p0.W[which];
        return BAEMM_OK;
    });
}

int baemm_get_flags(baemm_handle h, int* flags)
{
    return guard(h, [&](Handle& hh) {
        if (!flags) throw ArgError{"flags is NULL"};
        for (int j = 0; j < 4; ++j) {
            flags[2 * j] = hh.prm.re[j] ? 1 : 0;
            flags[2 * j + 1] = hh.prm.im[j] ? 1 : 0;
        }
        return BAEMM_OK;
    });
}

int baemm_device_info(baemm_handle h, char* buf, int64_t size)
{
    return guard(h, [&](Handle& hh) {
        if (!buf || size < 1) throw ArgError{"buffer is NULL or empty"};
        hipDeviceProp_t prop;
        BAEMM_HIP(hipGetDeviceProperties(&prop, hh.device));
        int drv = 0, rt = 0;
        BAEMM_HIP(hipDriverGetVersion(&drv));
        BAEMM_HIP(hipRuntimeGetVersion(&rt));
        char text[1024];
        std::snprintf(text, sizeof(text),
                      "Helmholtz_HIP::DeviceInfo\n"
                      "DEVICE_VENDOR                  = AMD\n"
                      "DEVICE_NAME                    = %s\n"
                      "DEVICE_VERSION                 = %s (%d compute units)\n"
                      "DRIVER_VERSION                 = %d (HIP runtime %d)\n"
                      "DEVICE_MAX_MEM_ALLOC_SIZE      = %zu\n"
                      "DEVICE_GLOBAL_MEM_SIZE         = %zu\n"
                      "DEVICE_LOCAL_MEM_SIZE          = %zu\n"
                      "DEVICE_MAX_CONSTANT_BUFFER_SIZE = %zu\n",
                      prop.name[0] ? prop.name : "AMD Instinct GPU", prop.gcnArchName, prop.multiProcessorCount, drv, rt,
                      (size_t)prop.totalGlobalMem,
                      (size_t)prop.totalGlobalMem, (size_t)prop.sharedMemPerBlock, (size_t)prop.totalConstMem);
        std::snprintf(buf, (size_t)size, "%s", text);
        return BAEMM_OK;
    });
}

int baemm_get_coordinates(baemm_handle h, float* vertex_coords, float* triangle_coords)
{
    return guard(h, [&](Handle& hh) {
        if (vertex_coords)
            for (int64_t k = 0; k < 3 * hh.nv; ++k) vertex_coords[k] = (float)hh.coords[k];
        if (triangle_coords)
            for (int64_t t = 0; t < hh.nt; ++t)
                for (int c = 0; c < 3; ++c) {
                    const int64_t v = hh.tris[3 * t + c];
                    for (int d = 0; d < 3; ++d) triangle_coords[12 * t + 4 * c + d] = (float)hh.coords[3 * v + d];
                    triangle_coords[12 * t + 4 * c + 3] = 0.0f;
                }
        return BAEMM_OK;
    });
}

int baemm_set_timing(baemm_handle h, int on)
{
    return guard(h, [&](Handle& hh) {
        hh.timing = on != 0;
        hh.timed_ms = 0.0;
        hh.timed_launches = 0;
        return BAEMM_OK;
    });
}

int baemm_get_timing(baemm_handle h, double* total_ms, int64_t* launches)
{
    return guard(h, [&](Handle& hh) {
        if (total_ms) *total_ms = hh.timed_ms;
        if (launches) *launches = hh.timed_launches;
        return BAEMM_OK;
    });
}

int baemm_hw_sincos(int device, const float* rev, int64_t n, float* sin_out, float* cos_out)
{
    if (n < 0 || (n > 0 && (!rev || !sin_out || !cos_out))) return BAEMM_ERR_ARG;
    if (n == 0) return BAEMM_OK;
    void* d = nullptr;
    int st = BAEMM_OK;
    try {
        int ndev = 0;
        BAEMM_HIP(hipGetDeviceCount(&ndev));
        BAEMM_HIP(hipSetDevice((device >= 0 && device < ndev) ? device : 0));
        BAEMM_HIP(hipMalloc(&d, sizeof(float) * 3 * n));
        float* dr = static_cast<float*>(d);
        BAEMM_HIP(hipMemcpy(dr, rev, sizeof(float) * n, hipMemcpyHostToDevice));
        hw_sincos_dev(dr, n, dr + n, dr + 2 * n, nullptr);
        BAEMM_HIP(hipMemcpy(sin_out, dr + n, sizeof(float) * n, hipMemcpyDeviceToHost));
        BAEMM_HIP(hipMemcpy(cos_out, dr + 2 * n, sizeof(float) * n, hipMemcpyDeviceToHost));
    } catch (const HipError& e) {
        set_thread_error(std::string("HIP error '") + hipGetErrorString(e.err) + "' in " + e.what);
        st = BAEMM_ERR_HIP;
    }
    if (d) (void)hipFree(d);
    return st;
}

int baemm_dev_alloc(void** ptr, size_t bytes)
{
    if (!ptr) return BAEMM_ERR_ARG;
    return hipMalloc(ptr, bytes) == hipSuccess ? BAEMM_OK : BAEMM_ERR_HIP;
}

int baemm_dev_free(void* ptr) { return hipFree(ptr) == hipSuccess ? BAEMM_OK : BAEMM_ERR_HIP; }

int baemm_dev_upload(void* dst, const void* src, size_t bytes)
{
    return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? BAEMM_OK : BAEMM_ERR_HIP;
}

int baemm_dev_download(void* dst, const void* src, size_t bytes)
{
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? BAEMM_OK : BAEMM_ERR_HIP;
}

int baemm_dev_synchronize(void) { return hipDeviceSynchronize() == hipSuccess ? BAEMM_OK : BAEMM_ERR_HIP; }

}  // extern "C"
