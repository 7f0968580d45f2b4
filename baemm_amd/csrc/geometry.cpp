// geometry.cpp -- per-mesh setup: triangle geometry, sparse patterns, upload.
//
// Geometry follows Initialize.hpp:57-305 (centroid, area-weighted normal,
// area, unit normal, single-layer self term by the edge formula) but is
// evaluated in FP64 and rounded once to FP32 for the device (the reference
// evaluates it in FP32, the difference is ~1e-7 relative).
#include <algorithm>
#include <cmath>
#include <numeric>

#include "internal.hpp"

namespace baemm {

static double self_term(const double x[3][3], const double c[3], double area)
{
    double u[3][3];
    for (int k = 0; k < 3; ++k)
        for (int d = 0; d < 3; ++d) u[k][d] = x[k][d] - c[d];
    double sum = 0.0;
    for (int e = 0; e < 3; ++e) {
        const int a = (e + 1) % 3, b = (e + 2) % 3;
        const double uaua = u[a][0] * u[a][0] + u[a][1] * u[a][1] + u[a][2] * u[a][2];
        const double ubub = u[b][0] * u[b][0] + u[b][1] * u[b][1] + u[b][2] * u[b][2];
        const double uaub = u[a][0] * u[b][0] + u[a][1] * u[b][1] + u[a][2] * u[b][2];
        const double La = std::sqrt(uaua), Lb = std::sqrt(ubub);
        const double p = uaua - uaub, q = ubub - uaub;
        const double vinv = 1.0 / std::sqrt(std::fabs(p + q));
        const double sa = p * vinv / La, sb = q * vinv / Lb;
        // h * (atanh(sin a) + atanh(sin b)), h = distance from c to the edge line
        sum += La * std::sqrt(1.0 - sa * sa) * (std::atanh(sa) + std::atanh(sb));
    }
    return sum / area;
}

void build_host_geometry(Handle& h)
{
    const int64_t nt = h.nt, nv = h.nv;
    h.nt_pad = ((nt + kTile - 1) / kTile) * kTile;
    h.m_pad = std::max<int64_t>(kTile, ((h.m + kTile - 1) / kTile) * kTile);
    h.h_mid4.assign(4 * h.nt_pad, 0.0f);
    h.h_nrm4.assign(4 * h.nt_pad, 0.0f);
    h.h_area.assign(nt, 0.0f);
    h.h_sdiag.assign(nt, 0.0f);
    {
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300}, rmax2 = 0.0;
        for (int64_t v = 0; v < nv; ++v) {
            double r2 = 0.0;
            for (int d = 0; d < 3; ++d) {
                const double x = h.coords[3 * v + d];
                lo[d] = std::min(lo[d], x);
                hi[d] = std::max(hi[d], x);
                r2 += x * x;
            }
            rmax2 = std::max(rmax2, r2);
        }
        double diag2 = 0.0;
        for (int d = 0; d < 3; ++d) diag2 += (hi[d] - lo[d]) * (hi[d] - lo[d]);
        h.extent = std::sqrt(std::max(diag2, rmax2));
    }
    std::vector<double> area_d(nt);
    for (int64_t t = 0; t < nt; ++t) {
        double x[3][3], c[3];
        for (int k = 0; k < 3; ++k) {
            const int64_t v = h.tris[3 * t + k];
            for (int d = 0; d < 3; ++d) x[k][d] = h.coords[3 * v + d];
        }
        for (int d = 0; d < 3; ++d) c[d] = (x[0][d] + x[1][d] + x[2][d]) / 3.0;
        double e0[3], e1[3], an[3];
        for (int d = 0; d < 3; ++d) {
            e0[d] = x[1][d] - x[0][d];
            e1[d] = x[2][d] - x[0][d];
        }
        an[0] = 0.5 * (e0[1] * e1[2] - e0[2] * e1[1]);
        an[1] = 0.5 * (e0[2] * e1[0] - e0[0] * e1[2]);
        an[2] = 0.5 * (e0[0] * e1[1] - e0[1] * e1[0]);
        const double a = std::sqrt(an[0] * an[0] + an[1] * an[1] + an[2] * an[2]);
        area_d[t] = a;
        for (int d = 0; d < 3; ++d) {
            h.h_mid4[4 * t + d] = (float)c[d];
            h.h_nrm4[4 * t + d] = (float)(an[d] / a);
        }
        h.h_area[t] = (float)a;
        h.h_sdiag[t] = (float)self_term(x, c, a);
    }
    // Padded rows repeat the last triangle; the pair kernels mask j >= nt.
    for (int64_t t = nt; t < h.nt_pad; ++t)
        for (int d = 0; d < 4; ++d) {
            h.h_mid4[4 * t + d] = h.h_mid4[4 * (nt - 1) + d];
            h.h_nrm4[4 * t + d] = h.h_nrm4[4 * (nt - 1) + d];
        }

    // Vertex -> incident triangles (pattern of AvT, Initialize.hpp:307-315).
    h.h_vt_ptr.assign(nv + 1, 0);
    for (int64_t t = 0; t < nt; ++t)
        for (int k = 0; k < 3; ++k) h.h_vt_ptr[h.tris[3 * t + k] + 1]++;
    for (int64_t v = 0; v < nv; ++v) h.h_vt_ptr[v + 1] += h.h_vt_ptr[v];
    h.h_vt_idx.assign(3 * nt, 0);
    {
        std::vector<int32_t> fill(h.h_vt_ptr.begin(), h.h_vt_ptr.end() - 1);
        for (int64_t t = 0; t < nt; ++t)
            for (int k = 0; k < 3; ++k) h.h_vt_idx[fill[h.tris[3 * t + k]]++] = (int32_t)t;
    }

    // P1 mass matrix: a/6 on the diagonal, a/12 off it, duplicates summed
    // (Initialize.hpp:213-247, 329-334).  Values from the FP32 areas the
    // device uses, accumulated in FP64.
    std::vector<std::vector<std::pair<int32_t, double>>> rows(nv);
    for (int64_t t = 0; t < nt; ++t) {
        const double a = (double)h.h_area[t];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                rows[h.tris[3 * t + i]].push_back({h.tris[3 * t + j], (i == j) ? a / 6.0 : a / 12.0});
    }
    h.h_M_ptr.assign(nv + 1, 0);
    h.h_M_idx.clear();
    h.h_M_val.clear();
    for (int64_t v = 0; v < nv; ++v) {
        auto& r = rows[v];
        std::sort(r.begin(), r.end(), [](auto& a, auto& b) { return a.first < b.first; });
        for (size_t k = 0; k < r.size(); ++k) {
            if (!h.h_M_idx.empty() && (int64_t)h.h_M_idx.size() > h.h_M_ptr[v] && h.h_M_idx.back() == r[k].first) {
                h.h_M_val.back() += r[k].second;
            } else {
                h.h_M_idx.push_back(r[k].first);
                h.h_M_val.push_back(r[k].second);
            }
        }
        h.h_M_ptr[v + 1] = (int32_t)h.h_M_idx.size();
    }
    // areas_lumped_inv = 1 / (AvT (1/3)) = 9 / sum_{t ni v} a_t  (Initialize.hpp:337-353);
    // only used as a PCG preconditioner, where the constant factor is immaterial.
    h.h_lumped_inv.assign(nv, 0.0);
    for (int64_t v = 0; v < nv; ++v) {
        double s = 0.0;
        for (int32_t k = h.h_vt_ptr[v]; k < h.h_vt_ptr[v + 1]; ++k) s += (double)h.h_area[h.h_vt_idx[k]] / 9.0;
        h.h_lumped_inv[v] = (s > 0.0) ? 1.0 / s : 0.0;
    }
}

// Copies on the handle's own (non-blocking) stream, so that creating a handle
// does not wait for work other handles have queued on the device.
template <class T>
static T* upload(const void* src, size_t bytes, hipStream_t s)
{
    void* p = nullptr;
    BAEMM_HIP(hipMalloc(&p, bytes > 0 ? bytes : 16));
    if (bytes) {
        BAEMM_HIP(hipMemcpyAsync(p, src, bytes, hipMemcpyHostToDevice, s));
        BAEMM_HIP(hipStreamSynchronize(s));
    }
    return static_cast<T*>(p);
}

void upload_geometry(Handle& h)
{
    BAEMM_HIP(hipSetDevice(h.device));
    const hipStream_t s = h.stream;
    h.d_mid = upload<float4>(h.h_mid4.data(), h.h_mid4.size() * sizeof(float), s);
    h.d_nrm = upload<float4>(h.h_nrm4.data(), h.h_nrm4.size() * sizeof(float), s);
    h.d_area = upload<float>(h.h_area.data(), h.h_area.size() * sizeof(float), s);
    h.d_sdiag = upload<float>(h.h_sdiag.data(), h.h_sdiag.size() * sizeof(float), s);
    std::vector<int32_t> t4(4 * h.nt, 0);
    for (int64_t t = 0; t < h.nt; ++t)
        for (int k = 0; k < 3; ++k) t4[4 * t + k] = h.tris[3 * t + k];
    h.d_tris = upload<int4>(t4.data(), t4.size() * sizeof(int32_t), s);
    h.d_vt_ptr = upload<int32_t>(h.h_vt_ptr.data(), h.h_vt_ptr.size() * sizeof(int32_t), s);
    h.d_vt_idx = upload<int32_t>(h.h_vt_idx.data(), h.h_vt_idx.size() * sizeof(int32_t), s);
    h.d_M_ptr = upload<int32_t>(h.h_M_ptr.data(), h.h_M_ptr.size() * sizeof(int32_t), s);
    h.d_M_idx = upload<int32_t>(h.h_M_idx.data(), h.h_M_idx.size() * sizeof(int32_t), s);
    h.d_M_val = upload<double>(h.h_M_val.data(), h.h_M_val.size() * sizeof(double), s);
    h.d_lumped_inv = upload<double>(h.h_lumped_inv.data(), h.h_lumped_inv.size() * sizeof(double), s);
    std::vector<float> m4(4 * h.m_pad, 0.0f);
    for (int64_t i = 0; i < h.m; ++i)
        for (int d = 0; d < 3; ++d) m4[4 * i + d] = (float)h.meas[3 * i + d];
    h.d_meas = upload<float4>(m4.data(), m4.size() * sizeof(float), s);
}

void free_device(Handle& h)
{
    void* ptrs[] = {h.d_mid, h.d_nrm, h.d_area, h.d_sdiag, h.d_tris, h.d_vt_ptr, h.d_vt_idx,
                    h.d_M_ptr, h.d_M_idx, h.d_M_val, h.d_lumped_inv, h.d_meas};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    DevBuf* bufs[] = {&h.ws_B, &h.ws_out, &h.ws_b32, &h.ws_bop, &h.ws_part, &h.ws_tsum, &h.ws_scale, &h.ws_misc,
                      &h.ws_pts, &h.ws_g, &h.ws_nrm, &h.sws.red, &h.sws.scal, &h.sws.cg, &h.sws.V, &h.sws.w,
                      &h.sws.t, &h.sws.x, &h.sws.r, &h.sws.hc, &h.sws.a, &h.sws.b, &h.sws.c,
                      &h.sws.d, &h.sws.e, &h.sws.cgctr, &h.sws.sc, &h.sws.gm, &h.ws_p0, &h.ws_tp_grad,
                      &h.ws_tp_face, &h.ws_tp_part, &h.ws_tp_out, &h.ws_tp_diag};
    for (DevBuf* b : bufs) b->release();
    for (DevBuf& b : h.io) b.release();
    for (GraphSlot& g : h.graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    h.graphs.clear();
}

// Device bytes a handle holds: geometry plus its grow-only workspaces.
size_t handle_device_bytes(const Handle& h)
{
    size_t g = (size_t)h.nt_pad * (2 * sizeof(float4)) + (size_t)h.nt * (2 * sizeof(float) + sizeof(int4)) +
               h.h_vt_ptr.size() * 4 + h.h_vt_idx.size() * 4 + h.h_M_ptr.size() * 4 + h.h_M_idx.size() * 4 +
               h.h_M_val.size() * 8 + h.h_lumped_inv.size() * 8 + (size_t)h.m_pad * sizeof(float4);
    const DevBuf* bufs[] = {&h.ws_B, &h.ws_out, &h.ws_b32, &h.ws_bop, &h.ws_part, &h.ws_tsum, &h.ws_scale, &h.ws_misc,
                            &h.ws_pts, &h.ws_g, &h.ws_nrm, &h.sws.red, &h.sws.scal, &h.sws.cg, &h.sws.V, &h.sws.w,
                            &h.sws.t, &h.sws.x, &h.sws.r, &h.sws.hc, &h.sws.a, &h.sws.b, &h.sws.c,
                            &h.sws.d, &h.sws.e, &h.sws.cgctr, &h.sws.sc, &h.sws.gm, &h.ws_p0, &h.ws_tp_grad,
                            &h.ws_tp_face, &h.ws_tp_part, &h.ws_tp_out, &h.ws_tp_diag};
    for (const DevBuf* b : bufs) g += b->bytes;
    for (const DevBuf& b : h.io) g += b.bytes;
    return g;
}

// LoadParameters.hpp:198-259; coefficients scaled in FP64 then rounded.
void load_params(Handle& h, const double* kappa, const double* coeff, int64_t W, int64_t wcs, double factor)
{
    if (W < 1 || wcs < 1) throw ArgError{"wave_count and wave_chunk_size must be >= 1"};
    if (W % wcs != 0) throw ArgError{"wave_count must be a multiple of wave_chunk_size"};
    if (!kappa || !coeff) throw ArgError{"kappa and coeff must not be NULL"};
    for (int64_t q = 0; q < W / wcs; ++q)
        if (!(std::fabs(kappa[q]) * h.extent / (2.0 * M_PI) <= kMaxPhaseRev))
            throw ArgError{"kappa times the mesh extent exceeds the FP32 phase range of the pair kernels "
                           "(65,536 revolutions)"};
    Params& p = h.prm;
    p.W = W;
    p.wcs = wcs;
    p.K = W / wcs;
    p.kappa.assign(p.K, 0.0f);
    p.c.assign(4 * p.K, cfloat(0.0f, 0.0f));
    for (int j = 0; j < 4; ++j) p.re[j] = p.im[j] = false;
    const double s = factor / (4.0 * M_PI);
    for (int64_t q = 0; q < p.K; ++q) {
        p.kappa[q] = (float)kappa[q];
        for (int j = 0; j < 4; ++j) {
            double re = coeff[8 * q + 2 * j], im = coeff[8 * q + 2 * j + 1];
            if (j > 0) {
                re *= s;
                im *= s;
            }
            const cfloat z((float)re, (float)im);
            p.c[4 * q + j] = z;
            p.re[j] = p.re[j] || (z.real() != 0.0f);
            p.im[j] = p.im[j] || (z.imag() != 0.0f);
        }
    }
    h.wave_chunk_size = (int)wcs;
    p.loaded = true;
}

std::vector<ChunkCoef> chunk_coefs(const Params& p, int q0, int kg, float scale, const cfloat* rot)
{
    std::vector<ChunkCoef> out(kg);
    for (int g = 0; g < kg; ++g) {
        const int q = q0 + g;
        // coefficients times scale / rot[q] (a factor the caller multiplies back)
        const cdouble f = (double)scale / (rot ? cdouble(rot[q]) : cdouble(1.0, 0.0));
        const cdouble c1 = cdouble(p.c[4 * q + 1]) * f;
        const cdouble c2 = cdouble(p.c[4 * q + 2]) * f;
        const cdouble c3 = cdouble(p.c[4 * q + 3]) * f;
        ChunkCoef cc{};
        cc.kap = p.kappa[q];
        cc.kap_rev = (float)((double)p.kappa[q] / (2.0 * M_PI));
        cc.c1r = (float)c1.real();
        cc.c1i = (float)c1.imag();
        cc.c2r = (float)c2.real();
        cc.c2i = (float)c2.imag();
        cc.c3r = (float)c3.real();
        cc.c3i = (float)c3.imag();
        const double k = (double)p.kappa[q];
        cc.kc2r = (float)(k * c2.real());
        cc.kc2i = (float)(k * c2.imag());
        cc.kc3r = (float)(k * c3.real());
        cc.kc3i = (float)(k * c3.imag());
        out[g] = cc;
    }
    return out;
}

}  // namespace baemm
