// capi_util.hpp -- shared helpers of the extern "C" entry points.
#pragma once

#include <exception>
#include <new>
#include <string>
#include <thread>

#include "internal.hpp"

namespace baemm {

struct StateError {
    std::string msg;
};

inline Handle* H(baemm_handle h) { return reinterpret_cast<Handle*>(h); }

// Message for baemm_last_error(NULL) on the calling thread.
void set_thread_error(const std::string& msg);

// Marks the calling thread as the handle's owner while a guarded call runs.
struct OwnerMark {
    Handle& h;
    explicit OwnerMark(Handle& h_) : h(h_) { h.owner.store(std::this_thread::get_id()); }
    ~OwnerMark() { h.owner.store(std::thread::id()); }
};

template <class F>
inline int guard(baemm_handle h, F&& f)
{
    if (!h) return BAEMM_ERR_ARG;
    Handle& hh = *H(h);
    // A call on h from a callback that a call on h is running (the M and P of
    // baemm_gauss_newton_solve) would wait for itself: refuse it instead.
    if (hh.owner.load() == std::this_thread::get_id()) {
        set_thread_error("call on a handle from inside a callback of a call on the same handle "
                         "(use another handle on the same mesh)");
        return BAEMM_ERR_STATE;
    }
    std::lock_guard<std::mutex> lk(hh.mtx);
    OwnerMark mark(hh);
    hh.last_error.clear();
    try {
        BAEMM_HIP(hipSetDevice(hh.device));
        return f(hh);
    } catch (const ArgError& e) {
        hh.last_error = e.msg;
        return BAEMM_ERR_ARG;
    } catch (const StateError& e) {
        hh.last_error = e.msg;
        return BAEMM_ERR_STATE;
    } catch (const HipError& e) {
        hh.last_error = std::string("HIP error '") + hipGetErrorString(e.err) + "' in " + e.what;
        return BAEMM_ERR_HIP;
    } catch (const std::bad_alloc&) {
        hh.last_error = "host allocation failed";
        return BAEMM_ERR_INTERNAL;
    } catch (const std::exception& e) {
        hh.last_error = e.what();
        return BAEMM_ERR_INTERNAL;
    }
}

inline cdouble cplx(const double* z, double dflt_re)
{
    return z ? cdouble(z[0], z[1]) : cdouble(dflt_re, 0.0);
}

inline hipStream_t stream_of(Handle& h, void* s) { return s ? static_cast<hipStream_t>(s) : h.stream; }

}  // namespace baemm
