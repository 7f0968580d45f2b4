// capi_batch.cpp -- C ABI of the batched persistent apply and its autotuner.
#include <algorithm>
#include <chrono>
#include <map>
#include <mutex>
#include <vector>

#include "batch.hpp"
#include "capi_util.hpp"

namespace baemm {
namespace {

// One work-list workspace per device; batched calls on a device serialise.
constexpr int kMaxDevices = 64;
std::mutex g_batch_mtx;
BatchWorkspace g_bw[kMaxDevices];
BatchStats g_last_stats;

struct BatchLock {
    std::vector<Handle*> hs;
    explicit BatchLock(std::vector<Handle*> v) : hs(std::move(v))
    {
        std::sort(hs.begin(), hs.end());
        for (Handle* h : hs) h->mtx.lock();
    }
    ~BatchLock()
    {
        for (auto it = hs.rbegin(); it != hs.rend(); ++it) (*it)->mtx.unlock();
    }
};

template <class F>
int batch_guard(Handle* first, F&& f)
{
    set_thread_error("");
    try {
        return f();
    } catch (const ArgError& e) {
        set_thread_error(e.msg);
        if (first) first->last_error = e.msg;
        return BAEMM_ERR_ARG;
    } catch (const StateError& e) {
        set_thread_error(e.msg);
        if (first) first->last_error = e.msg;
        return BAEMM_ERR_STATE;
    } catch (const HipError& e) {
        const std::string m = std::string("HIP error '") + hipGetErrorString(e.err) + "' in " + e.what;
        set_thread_error(m);
        if (first) first->last_error = m;
        return BAEMM_ERR_HIP;
    } catch (const std::exception& e) {
        set_thread_error(e.what());
        if (first) first->last_error = e.what();
        return BAEMM_ERR_INTERNAL;
    }
}

// Validates the arguments and builds the entries (handles locked by the caller).
std::vector<BatchEntry> make_entries(int n, const std::vector<Handle*>& hs, const double* alpha, const void* const* B,
                                     const int64_t* ldB, int B_type, const double* beta, void* const* C,
                                     const int64_t* ldC, int C_type)
{
    if ((B_type != kC128 && B_type != kC64) || (C_type != kC128 && C_type != kC64))
        throw ArgError{"unknown element type"};
    std::vector<BatchEntry> es(n);
    const int dev = hs[0]->device;
    for (int k = 0; k < n; ++k) {
        Handle& h = *hs[k];
        if (h.device != dev) throw ArgError{"all handles of a batch must be on one device"};
        if (!h.prm.loaded) throw StateError{"operators are not loaded on handle " + std::to_string(k)};
        if (!B[k] || !C[k]) throw ArgError{"B and C must not be NULL"};
        if (ldB[k] < h.prm.W || ldC[k] < h.prm.W) throw ArgError{"leading dimension smaller than the wave count"};
        const cdouble a = alpha ? cdouble(alpha[2 * k], alpha[2 * k + 1]) : cdouble(1.0, 0.0);
        const cdouble b = beta ? cdouble(beta[2 * k], beta[2 * k + 1]) : cdouble(0.0, 0.0);
        es[k] = BatchEntry{&h, ApplyArgs{a, b, B[k], ldB[k], B_type, C[k], ldC[k], C_type}};
    }
    return es;
}

std::vector<Handle*> handles_of(int n, const baemm_handle* hs)
{
    if (n <= 0 || !hs) throw ArgError{"empty batch"};
    std::vector<Handle*> v(n);
    for (int k = 0; k < n; ++k) {
        if (!hs[k]) throw ArgError{"NULL handle in batch"};
        v[k] = H(hs[k]);
    }
    std::vector<Handle*> u(v);
    std::sort(u.begin(), u.end());
    if (std::adjacent_find(u.begin(), u.end()) != u.end())
        throw ArgError{"a handle appears twice in one batch (its workspaces are per handle)"};
    return v;
}

}  // namespace
}  // namespace baemm

using namespace baemm;

extern "C" {

int baemm_apply_boundary_batch(int n, const baemm_handle* hs, const double* alpha, const void* const* B,
                               const int64_t* ldB, int B_type, const double* beta, void* const* C,
                               const int64_t* ldC, int C_type, void* stream)
{
    Handle* first = (n > 0 && hs && hs[0]) ? H(hs[0]) : nullptr;
    return batch_guard(first, [&]() {
        std::vector<Handle*> v = handles_of(n, hs);
        if (!B || !C || !ldB || !ldC) throw ArgError{"NULL array argument"};
        std::lock_guard<std::mutex> g(g_batch_mtx);
        BatchLock lk(v);
        BAEMM_HIP(hipSetDevice(v[0]->device));
        if (v[0]->device >= kMaxDevices) throw ArgError{"device index too large"};
        std::vector<BatchEntry> es = make_entries(n, v, alpha, B, ldB, B_type, beta, C, ldC, C_type);
        hipStream_t s = stream ? static_cast<hipStream_t>(stream) : v[0]->stream;
        g_last_stats = BatchStats{};
        apply_boundary_batch(es, g_bw[v[0]->device], s, &g_last_stats);
        return BAEMM_OK;
    });
}

int baemm_batch_set_tuning(int mixed_steps, int plain_blocks, int mixed_grid_mult, int plain_grid_mult)
{
    if (mixed_steps < 8 || plain_blocks < 1 || mixed_grid_mult < 0 || plain_grid_mult < 0) return BAEMM_ERR_ARG;
    std::lock_guard<std::mutex> g(g_batch_mtx);
    BatchTuning& t = batch_tuning();
    t.mixed_steps = mixed_steps;
    t.plain_blocks = plain_blocks;
    t.mixed_grid_mult = mixed_grid_mult;
    t.plain_grid_mult = plain_grid_mult;
    return BAEMM_OK;
}

int baemm_batch_get_tuning(int* mixed_steps, int* plain_blocks, int* mixed_grid_mult, int* plain_grid_mult)
{
    std::lock_guard<std::mutex> g(g_batch_mtx);
    const BatchTuning& t = batch_tuning();
    if (mixed_steps) *mixed_steps = t.mixed_steps;
    if (plain_blocks) *plain_blocks = t.plain_blocks;
    if (mixed_grid_mult) *mixed_grid_mult = t.mixed_grid_mult;
    if (plain_grid_mult) *plain_grid_mult = t.plain_grid_mult;
    return BAEMM_OK;
}

int baemm_batch_set_own_launch(double min_rounds)
{
    if (!(min_rounds >= 0.0) || min_rounds > 1e9) return BAEMM_ERR_ARG;
    std::lock_guard<std::mutex> g(g_batch_mtx);
    batch_tuning().own_rounds = min_rounds;
    return BAEMM_OK;
}

int baemm_batch_get_own_launch(double* min_rounds)
{
    if (!min_rounds) return BAEMM_ERR_ARG;
    std::lock_guard<std::mutex> g(g_batch_mtx);
    *min_rounds = batch_tuning().own_rounds;
    return BAEMM_OK;
}

int baemm_batch_get_stats(int64_t* items_mixed, int64_t* items_fp32, int64_t* grid_mixed, int64_t* grid_fp32)
{
    std::lock_guard<std::mutex> g(g_batch_mtx);
    if (items_mixed) *items_mixed = g_last_stats.items[0];
    if (items_fp32) *items_fp32 = g_last_stats.items[1];
    if (grid_mixed) *grid_mixed = g_last_stats.grid[0];
    if (grid_fp32) *grid_fp32 = g_last_stats.grid[1];
    return BAEMM_OK;
}

int baemm_batch_autotune(int n, const baemm_handle* hs, const void* const* B, const int64_t* ldB, int B_type,
                         void* const* C, const int64_t* ldC, int C_type, int repeats, double* best_ms)
{
    Handle* first = (n > 0 && hs && hs[0]) ? H(hs[0]) : nullptr;
    return batch_guard(first, [&]() {
        std::vector<Handle*> v = handles_of(n, hs);
        if (!B || !C || !ldB || !ldC) throw ArgError{"NULL array argument"};
        std::lock_guard<std::mutex> g(g_batch_mtx);
        BatchLock lk(v);
        BAEMM_HIP(hipSetDevice(v[0]->device));
        if (v[0]->device >= kMaxDevices) throw ArgError{"device index too large"};
        std::vector<BatchEntry> es = make_entries(n, v, nullptr, B, ldB, B_type, nullptr, C, ldC, C_type);
        hipStream_t s = v[0]->stream;
        const int reps = std::max(1, repeats);
        BatchTuning& t = batch_tuning();
        const BatchTuning saved = t;
        // Candidates: source-range length per item x persistent blocks per CU.
        const int msteps[] = {64, 128, 256, 512};
        const int pblocks[] = {8, 16, 32};
        const int mults[] = {0, 1, 2};
        hipEvent_t e0, e1;
        BAEMM_HIP(hipEventCreate(&e0));
        BAEMM_HIP(hipEventCreate(&e1));
        BatchTuning best = saved;
        double best_t = 1e300;
        auto time_candidate = [&]() {
            apply_boundary_batch(es, g_bw[v[0]->device], s, nullptr);
            BAEMM_HIP(hipEventRecord(e0, s));
            for (int r = 0; r < reps; ++r) apply_boundary_batch(es, g_bw[v[0]->device], s, nullptr);
            BAEMM_HIP(hipEventRecord(e1, s));
            BAEMM_HIP(hipEventSynchronize(e1));
            float ms_t = 0.0f;
            BAEMM_HIP(hipEventElapsedTime(&ms_t, e0, e1));
            if (ms_t / reps < best_t) {
                best_t = ms_t / reps;
                best = t;
            }
        };
        try {
            apply_boundary_batch(es, g_bw[v[0]->device], s, nullptr);   // warm-up (allocations)
            for (int ms : msteps)
                for (int pb : pblocks)
                    for (int gm : mults) {
                        t = saved;
                        t.mixed_steps = ms;
                        t.plain_blocks = pb;
                        t.mixed_grid_mult = gm;
                        t.plain_grid_mult = gm;
                        t.own_rounds = 1.0;
                        time_candidate();
                    }
            // every instance on its own launch (no persistent launch): the
            // schedule of separate applies, kept where it beats the work queue
            t = saved;
            t.own_rounds = 0.0;
            time_candidate();
        } catch (...) {
            t = saved;
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        t = best;
        if (best_ms) *best_ms = best_t;
        return BAEMM_OK;
    });
}

int baemm_gmres_batch(int n, const baemm_handle* hs, const void* const* B, void* const* X, double cg_tol,
                      double gmres_tol, int64_t* iterations, int* converged, void* stream)
{
    Handle* first = (n > 0 && hs && hs[0]) ? H(hs[0]) : nullptr;
    return batch_guard(first, [&]() {
        std::vector<Handle*> v = handles_of(n, hs);
        if (!B || !X) throw ArgError{"NULL array argument"};
        std::lock_guard<std::mutex> g(g_batch_mtx);
        BatchLock lk(v);
        BAEMM_HIP(hipSetDevice(v[0]->device));
        if (v[0]->device >= kMaxDevices) throw ArgError{"device index too large"};
        std::vector<GmresBatchItem> items;
        for (int k = 0; k < n; ++k) {
            Handle& h = *v[k];
            if (h.device != v[0]->device) throw ArgError{"all handles of a batch must be on one device"};
            if (!h.prm.loaded) throw StateError{"operators are not loaded on handle " + std::to_string(k)};
            if (h.precision == BAEMM_PRECISION_FP32_MFMA)
                throw ArgError{"batched GMRES: instances must use BAEMM_PRECISION_FP32 or BAEMM_PRECISION_MIXED"};
            if (!B[k] || !X[k]) throw ArgError{"B and X must not be NULL"};
            GmresBatchItem it;
            it.h = &h;
            it.B = static_cast<const double2*>(B[k]);
            it.X = static_cast<double2*>(X[k]);
            items.push_back(it);
        }
        hipStream_t s = stream ? static_cast<hipStream_t>(stream) : v[0]->stream;
        gmres_batch(items, g_bw[v[0]->device], gmres_tol, cg_tol, s);
        bool all = true;
        for (int k = 0; k < n; ++k) {
            if (iterations) iterations[k] = items[k].iterations;
            if (converged) converged[k] = items[k].converged ? 1 : 0;
            all = all && items[k].converged;
        }
        return all ? BAEMM_OK : BAEMM_ERR_NOCONV;
    });
}

}  // extern "C"
