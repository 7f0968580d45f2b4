// serve.cpp -- continuous-batching request loop for Gauss-Newton-step requests
// (forward solve + far field, Neumann data, adjoint derivative of the far-field
// misfit), with admission by device-memory need.  No reference counterpart:
// the reference runs one solve at a time from the caller's thread.
//
// Design:
//  * `slots` worker threads.  Each takes the next queued request whose
//    device-memory need fits into the budget next to what is resident (the
//    oldest, or with the shortest-first policy the one of least estimated
//    work, with aging), so a request that arrives while others run starts as
//    soon as a slot and memory are free.  Each request runs on its own handle
//    and HIP stream, so the kernels of the requests in flight overlap on the
//    device.
//  * Mesh geometry and the grow-only solver workspaces (Krylov basis, CG
//    vectors, pair-kernel partials) stay resident in HBM on cached handles,
//    keyed by the caller's mesh id, and are reused by later requests on the
//    same mesh.  Idle cached handles are evicted least-recently-used when a
//    new request needs their memory.
//  * A request's result does not depend on what else is in flight: it runs
//    the same deterministic kernels on its own handle.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "batch.hpp"
#include "capi_util.hpp"

namespace baemm {

size_t handle_device_bytes(const Handle& h);

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    int64_t id = 0;
    int64_t mesh_id = -1;
    std::vector<double> coords, meas, kappa, dirs, fobs;
    std::vector<int64_t> tris;
    int64_t nv = 0, nt = 0, m = 0, K = 0, wcs = 0;
    int wave_type = BAEMM_WAVE_PLANE;
    int precision = BAEMM_PRECISION_FP32;
    double cg_tol = 1e-10, gmres_tol = 1e-6;
    double* F_out = nullptr;
    double* grad_out = nullptr;
    size_t need = 0;
    double work = 0.0;          // cost estimate of the shortest-first policy: N_t^2 * W
    int skipped = 0;            // admissions of younger requests while this one waited
    Clock::time_point t_submit, t_start, t_done;
    int status = -1;            // -1 queued/running, else a BAEMM status
    std::string error;
    int64_t gmres_iterations = 0;
    bool done = false;
};

struct Cached {
    baemm_handle h = nullptr;
    int64_t mesh_id = -1;
    int64_t nv = 0, nt = 0, m = 0;
    size_t bytes = 0;           // resident device bytes (geometry + workspaces)
    bool busy = false;
    Clock::time_point last_used;
};

// Device bytes a request needs on a fresh handle: geometry, the pair-kernel
// staging and partials, the solver workspaces and the entry points' arrays.
size_t estimate_bytes(int64_t nv, int64_t nt, int64_t m, int64_t K, int64_t wcs, int precision, int restart)
{
    const int64_t W = K * wcs;
    const int64_t nt_pad = (nt + 255) / 256 * 256;
    int64_t wcsp = 4;
    while (wcsp < wcs) wcsp *= 2;
    const bool mfma = precision != BAEMM_PRECISION_FP32;
    if (mfma) wcsp = std::max<int64_t>(wcsp, 16);
    const int64_t ldb = K * wcsp;
    size_t geo = (size_t)nt_pad * 56 + (size_t)nt * 12 + (size_t)nv * 8 * 12 + (size_t)(m + 256) * 16;
    size_t b = (size_t)nt_pad * ldb * 8;
    size_t bop = mfma ? 2 * b : 0;
    const int64_t tiles = std::max<int64_t>(1, nt_pad / (mfma ? 128 : 256));
    // about 2,048 work-groups, grown up to twice that count to fill the last
    // round of resident work-groups
    const int64_t S = 2 * std::max<int64_t>(1, std::min<int64_t>((2048 + tiles - 1) / tiles, nt_pad / 64));
    size_t part = (size_t)S * nt_pad * 2 * ldb * 4;
    const size_t vec = (size_t)nv * W * 16;
    size_t gm = (size_t)(restart + 1) * vec + 9 * vec + (size_t)nv * 2 * W * 8 * 4;
    size_t io = 4 * vec + (size_t)m * W * 16 * 2 + (size_t)nv * 24 * 2;
    return geo + b + bop + part + gm + io + (8u << 20);
}

}  // namespace

struct Server {
    int device = 0;
    int slots = 1;
    size_t budget = 0;
    int policy = BAEMM_SERVE_FIFO;
    int max_skips = 8;
    std::mutex mtx;
    std::condition_variable cv;       // queue changes, completions
    std::deque<std::shared_ptr<Request>> queue;
    std::map<int64_t, std::shared_ptr<Request>> all;
    std::list<Cached> cache;
    size_t resident = 0;              // bytes of cached handles (idle and busy)
    size_t reserved = 0;              // need of requests running on handles whose size is not known yet
    size_t peak = 0;
    int64_t next_id = 1;
    int running = 0, max_running = 0;
    bool stop = false;
    std::vector<std::thread> workers;
    // statistics
    std::vector<double> latencies_ms;
    int64_t completed = 0, failed = 0, total_gmres = 0;
    Clock::time_point first_submit, last_done;
    bool any_submit = false;

    // fused mode: worker 0 gathers up to `slots` admitted requests and runs
    // them as one batch whose solves are batched GMRES; the other workers idle
    bool fused = false;
    BatchWorkspace bw;

    void worker(int index);
    void fused_worker();
    bool admit(std::shared_ptr<Request>& rq, Cached*& slot);
    void run(Request& rq, Cached& c);
    void finish(std::shared_ptr<Request>& rq, Cached* c);   // with mtx held
};

// Picks the next queued request that fits; returns false when none does.
// Called with mtx held.
//  * FIFO: the oldest request; one that does not fit holds back the later ones.
//  * Shortest first: requests that have been passed over max_skips times come
//    first, oldest first, and hold back the others while they do not fit (so
//    no request waits forever); then the rest by estimated work N_t^2 * W
//    (the oldest on ties), each tried in turn.  Every queued request older
//    than the admitted one counts one more skip.
bool Server::admit(std::shared_ptr<Request>& out, Cached*& slot)
{
    std::vector<std::deque<std::shared_ptr<Request>>::iterator> cand;
    for (auto it = queue.begin(); it != queue.end(); ++it) cand.push_back(it);
    size_t n_blocking = cand.size();     // FIFO: every candidate holds back the later ones
    if (policy == BAEMM_SERVE_SHORTEST_FIRST) {
        std::stable_partition(cand.begin(), cand.end(), [&](auto it) { return (*it)->skipped >= max_skips; });
        n_blocking = (size_t)std::count_if(cand.begin(), cand.end(), [&](auto it) { return (*it)->skipped >= max_skips; });
        std::stable_sort(cand.begin() + n_blocking, cand.end(),
                         [](auto x, auto y) { return (*x)->work < (*y)->work; });
    }
    for (size_t ci = 0; ci < cand.size(); ++ci) {
        auto it = cand[ci];
        std::shared_ptr<Request> rq = *it;
        // an idle cached handle of the same mesh: only growth beyond its size counts
        Cached* reuse = nullptr;
        if (rq->mesh_id >= 0)
            for (Cached& c : cache)
                if (!c.busy && c.mesh_id == rq->mesh_id && c.nv == rq->nv && c.nt == rq->nt && c.m == rq->m) {
                    reuse = &c;
                    break;
                }
        size_t extra = reuse ? (rq->need > reuse->bytes ? rq->need - reuse->bytes : 0) : rq->need;
        // evict idle handles of other meshes, least recently used first
        while (resident + reserved + extra > budget) {
            Cached* victim = nullptr;
            for (Cached& c : cache)
                if (!c.busy && &c != reuse && (!victim || c.last_used < victim->last_used)) victim = &c;
            if (!victim) break;
            baemm_destroy(victim->h);
            resident -= victim->bytes;
            for (auto ci2 = cache.begin(); ci2 != cache.end(); ++ci2)
                if (&*ci2 == victim) {
                    cache.erase(ci2);
                    break;
                }
        }
        if (resident + reserved + extra > budget) {
            if (running == 0) {
                // nothing runs and every other idle handle is gone: it can never
                // fit, so fail it instead of waiting forever, then look again
                // (the erase invalidates the candidate list)
                rq->status = BAEMM_ERR_ARG;
                rq->error = "request needs more device memory than the server budget";
                rq->done = true;
                rq->t_done = Clock::now();
                ++failed;
                queue.erase(it);
                cv.notify_all();
                return admit(out, slot);
            }
            // keep the order: a blocking request that does not fit holds back the later ones
            if (ci < n_blocking) return false;
            continue;
        }
        if (!reuse) {
            cache.push_back(Cached{});
            reuse = &cache.back();
            reuse->mesh_id = rq->mesh_id;
            reuse->nv = rq->nv;
            reuse->nt = rq->nt;
            reuse->m = rq->m;
        }
        reuse->busy = true;
        reserved += extra;
        rq->need = extra;   // what this request holds in `reserved` while it runs
        for (auto older = queue.begin(); older != it; ++older) ++(*older)->skipped;
        queue.erase(it);
        out = rq;
        slot = reuse;
        return true;
    }
    return false;
}

void Server::run(Request& rq, Cached& c)
{
    int st = BAEMM_OK;
    if (!c.h) {
        st = baemm_create(rq.coords.data(), rq.nv, rq.tris.data(), rq.nt, rq.meas.data(), rq.m, device, &c.h);
        if (st != BAEMM_OK) {
            rq.error = baemm_last_error(nullptr);
            c.h = nullptr;
        }
    }
    if (st == BAEMM_OK) {
        baemm_handle h = c.h;
        baemm_set_precision(h, rq.precision);
        baemm_reset_solver_stats(h);
        const int64_t W = rq.K * rq.wcs;
        // forward solve + far field (FarField.hpp:28-70)
        int s1 = baemm_far_field(h, rq.kappa.data(), rq.K, rq.dirs.data(), rq.wcs, rq.F_out, rq.wave_type, rq.cg_tol,
                                 rq.gmres_tol);
        // Neumann data of the total field (FarField.hpp:148-207)
        std::vector<double> du((size_t)2 * rq.nv * W);
        int s2 = (s1 == BAEMM_OK || s1 == BAEMM_ERR_NOCONV)
                     ? baemm_derivative_ff_helper(h, rq.kappa.data(), rq.K, rq.dirs.data(), rq.wcs, du.data(),
                                                  rq.wave_type, rq.cg_tol, rq.gmres_tol)
                     : s1;
        // misfit g = F - F_obs and its adjoint derivative (FarField.hpp:342-428)
        int s3 = s2;
        if (s2 == BAEMM_OK || s2 == BAEMM_ERR_NOCONV) {
            std::vector<double> g(rq.F_out, rq.F_out + 2 * rq.m * W);
            if (!rq.fobs.empty())
                for (size_t k = 0; k < g.size(); ++k) g[k] -= rq.fobs[k];
            double* dup = du.data();
            s3 = baemm_adjoint_derivative_ff(h, rq.kappa.data(), rq.K, rq.dirs.data(), rq.wcs, g.data(), rq.grad_out,
                                             &dup, rq.wave_type, rq.cg_tol, rq.gmres_tol);
        }
        int64_t its = 0, cyc = 0, cgit = 0, apps = 0;
        double res = 0.0;
        baemm_get_solver_stats(h, &its, &cyc, &res, &cgit, &apps);
        rq.gmres_iterations = its;
        for (int sx : {s1, s2, s3})
            if (sx != BAEMM_OK) {
                st = sx;
                if (sx != BAEMM_ERR_NOCONV) break;
            }
        if (st != BAEMM_OK && st != BAEMM_ERR_NOCONV) rq.error = baemm_last_error(h);
        else if (st == BAEMM_ERR_NOCONV) rq.error = "GMRES did not reach its tolerance";
    }
    rq.status = st;
}

// Bookkeeping of a finished request (mtx held): memory, cache, statistics.
void Server::finish(std::shared_ptr<Request>& rq, Cached* c)
{
    --running;
    reserved -= rq->need;
    const size_t bytes = c->h ? handle_device_bytes(*H(c->h)) : 0;
    resident += bytes;
    resident -= c->bytes;
    c->bytes = bytes;
    peak = std::max(peak, resident + reserved);
    c->busy = false;
    c->last_used = Clock::now();
    if (!c->h || rq->mesh_id < 0) {   // not cacheable: release now
        if (c->h) baemm_destroy(c->h);
        resident -= c->bytes;
        for (auto ci = cache.begin(); ci != cache.end(); ++ci)
            if (&*ci == c) {
                cache.erase(ci);
                break;
            }
    }
    rq->t_done = Clock::now();
    rq->done = true;
    if (rq->status == BAEMM_OK || rq->status == BAEMM_ERR_NOCONV) {
        ++completed;
        latencies_ms.push_back(std::chrono::duration<double, std::milli>(rq->t_done - rq->t_submit).count());
        total_gmres += rq->gmres_iterations;
        last_done = rq->t_done;
    } else {
        ++failed;
    }
}

void Server::worker(int index)
{
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mtx);
    for (;;) {
        if (fused && index == 0) {
            lk.unlock();
            fused_worker();
            lk.lock();
            if (stop && queue.empty()) return;
            continue;
        }
        std::shared_ptr<Request> rq;
        Cached* c = nullptr;
        cv.wait(lk, [&] { return stop || (fused && index == 0) || (!fused && !queue.empty() && admit(rq, c)); });
        if (!rq) {
            if (stop && queue.empty()) return;
            continue;
        }
        ++running;
        max_running = std::max(max_running, running);
        peak = std::max(peak, resident + reserved);
        rq->t_start = Clock::now();
        lk.unlock();
        try {
            run(*rq, *c);
        } catch (const std::exception& e) {   // e.g. host allocation: fail the request, keep the worker
            rq->status = BAEMM_ERR_INTERNAL;
            rq->error = e.what();
        } catch (...) {
            rq->status = BAEMM_ERR_INTERNAL;
            rq->error = "unknown error";
        }
        lk.lock();
        finish(rq, c);
        cv.notify_all();
    }
}

// Fused mode: gathers the queued requests that fit (up to `slots`) and runs
// them as one batch through gn_step_batch, phase by phase, with each of the
// three solves one batched GMRES; requests that arrive meanwhile join the next
// batch.  Returns when fused mode is switched off or the server stops.
void Server::fused_worker()
{
    std::unique_lock<std::mutex> lk(mtx);
    while (fused) {
        cv.wait(lk, [&] { return stop || !fused || !queue.empty(); });
        if (!fused || (stop && queue.empty())) return;
        std::vector<std::pair<std::shared_ptr<Request>, Cached*>> batch;
        for (;;) {
            if ((int)batch.size() >= slots || queue.empty()) break;
            std::shared_ptr<Request> rq;
            Cached* c = nullptr;
            if (!admit(rq, c)) break;
            if (!rq) break;
            batch.emplace_back(rq, c);
        }
        if (batch.empty()) {
            // nothing fits now; admit() fails requests that can never fit
            if (running == 0 && !queue.empty()) cv.wait_for(lk, std::chrono::milliseconds(2));
            continue;
        }
        running += (int)batch.size();
        max_running = std::max(max_running, running);
        peak = std::max(peak, resident + reserved);
        const auto t0 = Clock::now();
        for (auto& e : batch) e.first->t_start = t0;
        lk.unlock();
        std::vector<GnItem> items;
        std::vector<size_t> which;
        for (size_t k = 0; k < batch.size(); ++k) {
            Request& rq = *batch[k].first;
            Cached& c = *batch[k].second;
            int st = BAEMM_OK;
            if (!c.h) {
                st = baemm_create(rq.coords.data(), rq.nv, rq.tris.data(), rq.nt, rq.meas.data(), rq.m, device, &c.h);
                if (st != BAEMM_OK) {
                    rq.error = baemm_last_error(nullptr);
                    rq.status = st;
                    c.h = nullptr;
                    continue;
                }
            }
            baemm_set_precision(c.h, rq.precision);
            baemm_reset_solver_stats(c.h);
            GnItem it;
            it.h = H(c.h);
            it.kappa = rq.kappa.data();
            it.K = rq.K;
            it.wcs = rq.wcs;
            it.dirs = rq.dirs.data();
            it.wave_type = rq.wave_type;
            it.fobs = rq.fobs.empty() ? nullptr : rq.fobs.data();
            it.F_out = rq.F_out;
            it.grad_out = rq.grad_out;
            it.cg_tol = rq.cg_tol;
            it.gmres_tol = rq.gmres_tol;
            items.push_back(it);
            which.push_back(k);
        }
        // requests with other tolerances than the first run in batches of their own
        std::vector<std::vector<size_t>> groups;
        for (size_t k = 0; k < items.size(); ++k) {
            bool placed = false;
            for (auto& g : groups)
                if (items[g[0]].gmres_tol == items[k].gmres_tol && items[g[0]].cg_tol == items[k].cg_tol) {
                    g.push_back(k);
                    placed = true;
                    break;
                }
            if (!placed) groups.push_back({k});
        }
        for (auto& g : groups) {
            std::vector<GnItem> part;
            for (size_t k : g) part.push_back(items[k]);
            int st = BAEMM_OK;
            std::string err;
            try {
                hipStream_t s0 = part[0].h->stream;
                gn_step_batch(part, bw, s0);
            } catch (const ArgError& e) {
                st = BAEMM_ERR_ARG;
                err = e.msg;
            } catch (const HipError& e) {
                st = BAEMM_ERR_HIP;
                err = std::string("HIP error '") + hipGetErrorString(e.err) + "' in " + e.what;
            } catch (const std::exception& e) {
                st = BAEMM_ERR_INTERNAL;
                err = e.what();
            }
            for (size_t i = 0; i < g.size(); ++i) {
                Request& rq = *batch[which[g[i]]].first;
                if (st != BAEMM_OK) {
                    rq.status = st;
                    rq.error = err;
                    continue;
                }
                rq.status = part[i].status;
                if (rq.status == BAEMM_ERR_NOCONV) rq.error = "GMRES did not reach its tolerance";
                int64_t its = 0, cyc = 0, cgit = 0, apps = 0;
                double res = 0.0;
                baemm_get_solver_stats(reinterpret_cast<baemm_handle>(part[i].h), &its, &cyc, &res, &cgit, &apps);
                rq.gmres_iterations = its;
            }
        }
        lk.lock();
        for (auto& e : batch) finish(e.first, e.second);
        cv.notify_all();
    }
}

}  // namespace baemm

using namespace baemm;

struct baemm_server_s {
    Server s;
};

extern "C" {

int baemm_server_create(int device, int slots, int64_t hbm_budget_bytes, baemm_server* out)
{
    if (!out || slots < 1 || slots > 64 || hbm_budget_bytes <= 0) return BAEMM_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        set_thread_error("no HIP device");
        return BAEMM_ERR_HIP;
    }
    auto* srv = new (std::nothrow) baemm_server_s();
    if (!srv) return BAEMM_ERR_INTERNAL;
    Server& s = srv->s;
    s.device = (device >= 0 && device < ndev) ? device : 0;
    s.slots = slots;
    s.budget = (size_t)hbm_budget_bytes;
    try {
        for (int k = 0; k < slots; ++k) s.workers.emplace_back([&s, k] { s.worker(k); });
    } catch (...) {
        {
            std::lock_guard<std::mutex> lk(s.mtx);
            s.stop = true;
        }
        s.cv.notify_all();
        for (auto& t : s.workers)
            if (t.joinable()) t.join();
        delete srv;
        return BAEMM_ERR_INTERNAL;
    }
    *out = srv;
    return BAEMM_OK;
}

int baemm_server_destroy(baemm_server srv)
{
    if (!srv) return BAEMM_ERR_ARG;
    Server& s = srv->s;
    {
        std::unique_lock<std::mutex> lk(s.mtx);
        s.cv.wait(lk, [&] { return s.queue.empty() && s.running == 0; });
        s.stop = true;
    }
    s.cv.notify_all();
    for (auto& t : s.workers) t.join();
    for (Cached& c : s.cache)
        if (c.h) baemm_destroy(c.h);
    (void)hipSetDevice(s.device);
    s.bw.release();
    delete srv;
    return BAEMM_OK;
}

int baemm_server_set_fusion(baemm_server srv, int on)
{
    if (!srv) return BAEMM_ERR_ARG;
    Server& s = srv->s;
    {
        std::lock_guard<std::mutex> lk(s.mtx);
        s.fused = on != 0;
    }
    s.cv.notify_all();
    return BAEMM_OK;
}

int baemm_server_set_policy(baemm_server srv, int policy, int max_skips)
{
    if (!srv || (policy != BAEMM_SERVE_FIFO && policy != BAEMM_SERVE_SHORTEST_FIRST) || max_skips < 0)
        return BAEMM_ERR_ARG;
    Server& s = srv->s;
    {
        std::lock_guard<std::mutex> lk(s.mtx);
        s.policy = policy;
        s.max_skips = max_skips;
    }
    s.cv.notify_all();
    return BAEMM_OK;
}

int baemm_server_submit_gn(baemm_server srv, const baemm_gn_request* rq_in, double* far_field_out,
                           double* gradient_out, int64_t* request_id)
{
    if (!srv || !rq_in || !far_field_out || !gradient_out || !request_id) return BAEMM_ERR_ARG;
    const baemm_gn_request& r = *rq_in;
    if (!r.coords || !r.triangles || r.vertex_count < 3 || r.simplex_count < 1 || !r.meas_directions ||
        r.meas_count < 1 || !r.kappa || r.kappa_count < 1 || !r.inc_directions || r.wave_chunk_size < 1 ||
        r.kappa_count * r.wave_chunk_size > 64) {
        set_thread_error("invalid Gauss-Newton request");
        return BAEMM_ERR_ARG;
    }
    auto rq = std::make_shared<Request>();
    rq->mesh_id = r.mesh_id;
    rq->nv = r.vertex_count;
    rq->nt = r.simplex_count;
    rq->m = r.meas_count;
    rq->K = r.kappa_count;
    rq->wcs = r.wave_chunk_size;
    rq->coords.assign(r.coords, r.coords + 3 * r.vertex_count);
    rq->tris.assign(r.triangles, r.triangles + 3 * r.simplex_count);
    rq->meas.assign(r.meas_directions, r.meas_directions + 3 * r.meas_count);
    rq->kappa.assign(r.kappa, r.kappa + r.kappa_count);
    rq->dirs.assign(r.inc_directions, r.inc_directions + 3 * r.wave_chunk_size);
    const int64_t W = rq->K * rq->wcs;
    if (r.far_field_data) rq->fobs.assign(r.far_field_data, r.far_field_data + 2 * r.meas_count * W);
    rq->wave_type = r.wave_type ? r.wave_type : BAEMM_WAVE_PLANE;
    rq->precision = r.precision;
    rq->cg_tol = r.cg_tol > 0 ? r.cg_tol : 1e-10;
    rq->gmres_tol = r.gmres_tol > 0 ? r.gmres_tol : 1e-6;
    rq->F_out = far_field_out;
    rq->grad_out = gradient_out;
    rq->need = estimate_bytes(rq->nv, rq->nt, rq->m, rq->K, rq->wcs, rq->precision, 30);
    rq->work = (double)rq->nt * (double)rq->nt * (double)(rq->K * rq->wcs);
    Server& s = srv->s;
    {
        std::lock_guard<std::mutex> lk(s.mtx);
        rq->id = s.next_id++;
        rq->t_submit = Clock::now();
        if (!s.any_submit) {
            s.first_submit = rq->t_submit;
            s.any_submit = true;
        }
        s.queue.push_back(rq);
        s.all[rq->id] = rq;
        *request_id = rq->id;
    }
    s.cv.notify_all();
    return BAEMM_OK;
}

int baemm_server_wait(baemm_server srv, int64_t request_id, double* latency_ms, int64_t* gmres_iterations)
{
    if (!srv) return BAEMM_ERR_ARG;
    Server& s = srv->s;
    std::unique_lock<std::mutex> lk(s.mtx);
    auto it = s.all.find(request_id);
    if (it == s.all.end()) return BAEMM_ERR_ARG;
    std::shared_ptr<Request> rq = it->second;
    s.cv.wait(lk, [&] { return rq->done; });
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(rq->t_done - rq->t_submit).count();
    if (gmres_iterations) *gmres_iterations = rq->gmres_iterations;
    if (!rq->error.empty()) set_thread_error(rq->error);
    s.all.erase(it);
    return rq->status;
}

int64_t baemm_server_request_bytes(const baemm_gn_request* r)
{
    if (!r || r->vertex_count < 1 || r->simplex_count < 1 || r->kappa_count < 1 || r->wave_chunk_size < 1) return -1;
    return (int64_t)estimate_bytes(r->vertex_count, r->simplex_count, r->meas_count, r->kappa_count,
                                   r->wave_chunk_size, r->precision, 30);
}

int baemm_server_stats(baemm_server srv, baemm_server_stats_t* out)
{
    if (!srv || !out) return BAEMM_ERR_ARG;
    Server& s = srv->s;
    std::lock_guard<std::mutex> lk(s.mtx);
    std::memset(out, 0, sizeof(*out));
    out->completed = s.completed;
    out->failed = s.failed;
    out->queued = (int64_t)s.queue.size();
    out->running = s.running;
    out->max_in_flight = s.max_running;
    out->resident_bytes = (int64_t)s.resident;
    out->peak_bytes = (int64_t)s.peak;
    out->gmres_iterations = s.total_gmres;
    if (!s.latencies_ms.empty()) {
        std::vector<double> v = s.latencies_ms;
        std::sort(v.begin(), v.end());
        auto pct = [&](double p) {
            // nearest-rank percentile
            size_t k = (size_t)std::ceil(p * v.size());
            k = std::min(std::max<size_t>(k, 1), v.size());
            return v[k - 1];
        };
        out->p50_ms = pct(0.50);
        out->p99_ms = pct(0.99);
        double sum = 0.0;
        for (double x : v) sum += x;
        out->mean_ms = sum / v.size();
        const double span = std::chrono::duration<double>(s.last_done - s.first_submit).count();
        out->solves_per_s = span > 0 ? s.completed / span : 0.0;
    }
    return BAEMM_OK;
}

}  // extern "C"
