// internal.hpp -- shared declarations of libbaemm_hip (not installed).
//
// Device data layout (all resident in HBM for the lifetime of a handle):
//   mid[nt_pad], nrm[nt_pad]   float4 centroids / unit normals (w unused);
//                              padded rows hold the last triangle's data
//   area[nt], sdiag[nt]        float areas and single-layer self terms
//   tris[nt]                   int4 vertex ids (w unused)
//   vt_ptr[nv+1], vt_idx[3nt]  vertex -> incident triangles (AvT pattern)
//   M_ptr[nv+1], M_idx, M_val  P1 mass matrix CSR (double values)
//   meas[m_pad]                float4 measurement directions
// Per-apply workspaces grow on demand and are reused (see Workspace).
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <cstdint>
#include <mutex>
#include <thread>
#include <atomic>
#include <string>
#include <vector>

#include "../../include/baemm.h"

namespace baemm {

using cdouble = std::complex<double>;
using cfloat = std::complex<float>;

constexpr int kTile = 256;          // target rows per plain-kernel block
constexpr int kMixTile = 128;       // target rows per mixed-kernel block (4 waves x 32)
constexpr int kSrcBlock = 64;       // sources staged per LDS block (plain kernels)
constexpr int kMaxCols = 64;        // padded columns handled per kernel launch

struct HipError {
    hipError_t err;
    const char* what;
    int line;
};

#define BAEMM_HIP(x)                                                   \
    do {                                                               \
        hipError_t e_ = (x);                                           \
        if (e_ != hipSuccess) throw ::baemm::HipError{e_, #x, __LINE__}; \
    } while (0)

struct ArgError {
    std::string msg;
};

// Grow-only device buffer.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    void reserve(size_t n) {
        if (n <= bytes) return;
        if (ptr) BAEMM_HIP(hipFree(ptr));
        ptr = nullptr;
        bytes = 0;
        BAEMM_HIP(hipMalloc(&ptr, n));
        bytes = n;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    template <class T>
    T* as() const { return static_cast<T*>(ptr); }
};

// Loaded operator parameters (LoadParameters.hpp:21-81).
struct Params {
    int64_t W = 0, wcs = 0, K = 0;
    std::vector<float> kappa;            // K
    std::vector<cfloat> c;               // K x 4, c_1..3 scaled by factor/(4 pi)
    bool re[4] = {false, false, false, false};
    bool im[4] = {false, false, false, false};
    bool any(int j) const { return re[j] || im[j]; }
    bool kernel_on() const { return any(1) || any(2) || any(3); }
    bool mass_on() const { return any(0); }
    bool loaded = false;
};

struct ChunkCoef {
    float kap;       // kappa
    float kap_rev;   // kappa / (2 pi): the hardware sine takes revolutions
    float c1r, c1i, c2r, c2i, c3r, c3i;
    float kc2r, kc2i, kc3r, kc3i;   // kappa * c2, kappa * c3
};

// Device workspaces of the solvers (grow-only, reused across calls).
struct Workspace {
    DevBuf red, scal, cg;                 // reduction partials, per-column scalars, CG vectors
    DevBuf V, w, t, x, r, hc;             // GMRES basis and vectors, Gram-Schmidt coefficients
    DevBuf a, b, c, d, e;                 // composition temporaries
    DevBuf cgctr;                         // PCG counters: done, iterations, total
    DevBuf sc;                            // per-column scales of the Krylov vectors
    DevBuf gm;                            // GMRES Hessenberg / Givens state on the device
};

struct SolverStats {
    int64_t gmres_iterations = 0;
    int64_t gmres_cycles = 0;
    double gmres_rel_residual = 0.0;
    int64_t cg_iterations = 0;
    int64_t operator_applies = 0;
    bool converged = true;
};

// The reference class's P0 buffers B and C (InputOutput.hpp:5-89): B holds
// the triangle values b = Av B_in of the last operator apply, C the triangle
// values (K + D) b of the last boundary-operator apply before AvT.  Here they
// are filled on demand: the last apply records where its device data lives
// (src), and the host copies are brought up to date when they are read.
struct P0Source {
    bool b = false, c = false;          // device data of B / C available
    const float2* bptr = nullptr;       // staged sources, ldb floats2 per row
    int ldb = 0, WCSP = 0, wcs = 0, K = 0, W = 0;
    const float* part = nullptr;        // segment partials of the pair kernels
    int S = 0, ldp = 0;
    int64_t seg_stride = 0;
    const float* unscale = nullptr;     // per column (mixed precision) or null
    bool diag = false;                  // single-layer diagonal term on
    float rotr[64], roti[64], c1r[64], c1i[64], kap[64];
};

struct P0State {
    std::vector<cfloat> host[2];        // B, C: simplex_count x W[k]
    int64_t W[2] = {0, 0};
    bool host_current[2] = {false, false};   // host[k] holds the buffer's value
    P0Source src;
};

// A captured Arnoldi step (one per Krylov index j) and the key it was captured for.
struct GraphSlot {
    std::vector<char> key;
    hipGraphExec_t exec = nullptr;
};

struct Handle {
    std::mutex mtx;
    std::atomic<std::thread::id> owner{};   // thread in a guarded call on this handle
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t nv = 0, nt = 0, m = 0;
    int64_t nt_pad = 0, m_pad = 0;
    // host copies
    std::vector<double> coords;          // nv x 3
    std::vector<int32_t> tris;           // nt x 3
    std::vector<double> meas;            // m x 3
    std::vector<float> h_mid4, h_nrm4, h_area, h_sdiag;
    std::vector<int32_t> h_vt_ptr, h_vt_idx;
    std::vector<int32_t> h_M_ptr, h_M_idx;
    std::vector<double> h_M_val;
    std::vector<double> h_lumped_inv;    // areas_lumped_inv (Initialize.hpp:337-353)
    // device geometry
    float4* d_mid = nullptr;
    float4* d_nrm = nullptr;
    float* d_area = nullptr;
    float* d_sdiag = nullptr;
    int4* d_tris = nullptr;
    int32_t* d_vt_ptr = nullptr;
    int32_t* d_vt_idx = nullptr;
    int32_t* d_M_ptr = nullptr;
    int32_t* d_M_idx = nullptr;
    double* d_M_val = nullptr;
    double* d_lumped_inv = nullptr;
    float4* d_meas = nullptr;
    float rmin = -1.0f;                  // min centroid distance (mixed scaling), lazily computed
    double extent = 0.0;                 // max(bounding-box diagonal, max |x|) of the vertices
    // settings (GetSetters.hpp)
    bool use_diagonal = true;
    int block_size = 64;
    int wave_chunk_size = 16;
    int precision = BAEMM_PRECISION_FP32;
    int mixed_kernel = 2;                // 0: single-wave MFMA kernel, 1: wave-specialised, 2: automatic
    int segments = 0;                    // 0 = automatic split of the source range
    int gmres_max_iter = 30;
    int gmres_max_restarts = 3;
    Params prm;
    // workspaces
    DevBuf ws_B, ws_out, ws_b32, ws_bop, ws_part, ws_scale, ws_misc, ws_pts, ws_g, ws_nrm;
    DevBuf ws_tsum;                      // per-triangle segment sums of the partials (double2)
    DevBuf io[6];                        // device copies of the host arrays of the solver entry points
    DevBuf ws_p0;                        // P0 buffer read back for ReadB/WriteB/ReadC/WriteC
    // tangent-point energy and metric (tangent_point.hip)
    DevBuf ws_tp_grad;                   // hat-function gradients per triangle (float, 9 each)
    DevBuf ws_tp_face;                   // metric inputs per triangle: du (9), ubar (3) (float)
    DevBuf ws_tp_part;                   // segment partials of the pair sums
    DevBuf ws_tp_out;                    // per-triangle results (double)
    DevBuf ws_tp_diag;                   // diagonal of the tangent-point metric per vertex (double)
    double tp_diag_q = 0.0, tp_diag_p = 0.0;   // (q, p) of ws_tp_diag; p = 0: not computed
    P0State p0;
    Workspace sws;
    SolverStats stats;
    // hipGraph capture of the GMRES Arnoldi steps (on by default: replaying a
    // step saves its launch overheads, 4-15 % of a solve on small meshes)
    bool use_graphs = true;
    bool capturing = false;              // a capture is being recorded on the handle's stream
    // Device stop flag of a GMRES cycle (null outside one): the pair kernels and
    // the PCG of an Arnoldi step return at once when it is set.
    const int* step_skip = nullptr;
    int* pin_ctl = nullptr;              // pinned host copies of the GMRES flags (2 slots x 4 ints, 2 doubles)
    hipEvent_t ev_step[2] = {nullptr, nullptr};
    std::vector<GraphSlot> graphs;
    int64_t graph_captures = 0, graph_launches = 0;
    // profiling of the dominant pair kernel
    bool timing = false;
    double timed_ms = 0.0;
    int64_t timed_launches = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string last_error;
};

// ---- geometry / setup (geometry.cpp) ----
void build_host_geometry(Handle& h);
void upload_geometry(Handle& h);
void free_device(Handle& h);

// Strided host/device copy; a single linear copy when both sides are
// contiguous (pitch = width), which pageable host memory moves far faster than
// row by row.
inline hipError_t copy_rows_async(void* dst, size_t dpitch, const void* src, size_t spitch, size_t width,
                                  size_t height, hipMemcpyKind kind, hipStream_t s)
{
    if (dpitch == width && spitch == width) return hipMemcpyAsync(dst, src, width * height, kind, s);
    return hipMemcpy2DAsync(dst, dpitch, src, spitch, width, height, kind, s);
}

// ---- parameters ----
void load_params(Handle& h, const double* kappa, const double* coeff, int64_t W, int64_t wcs, double factor);
// Coefficients of chunks q0..q0+kg-1 times `scale` and, if rot is given,
// divided by rot[q].
std::vector<ChunkCoef> chunk_coefs(const Params& p, int q0, int kg, float scale, const cfloat* rot = nullptr);

// ---- device ops (ops.hip) ----
// Complex element types of external device arrays.
enum ElemType { kC128 = BAEMM_C128, kC64 = BAEMM_C64 };

struct ApplyArgs {
    cdouble alpha, beta;
    const void* B; int64_t ldB; int Btype;
    void* C; int64_t ldC; int Ctype;
};

// scratch (optional): a fixed device array of scratch_bytes that holds every
// workspace of the apply instead of the handle's grow-only buffers; the call
// then allocates nothing and fails with ArgError before any launch when the
// apply needs more.
void apply_boundary_dev(Handle& h, const ApplyArgs& a, hipStream_t s, char* scratch = nullptr,
                        size_t scratch_bytes = 0);
void apply_far_dev(Handle& h, const ApplyArgs& a, hipStream_t s);
void apply_near_dev(Handle& h, const ApplyArgs& a, const float4* d_pts, int64_t P, hipStream_t s);
void herglotz_dev(Handle& h, const ApplyArgs& a, hipStream_t s);
void incident_dev(Handle& h, const ApplyArgs& a, const double* dirs_host, int64_t inc_count, int wave_type, hipStream_t s);
float compute_rmin(Handle& h, hipStream_t s);
// P0 buffers (InputOutput.hpp): which = 0 (B) or 1 (C).  p0_invalidate: an
// apply rewrote the workspaces the buffers are read from; p0_sync: brings
// host copy `which` up to date (synchronises s); false if the buffer has no
// value (no apply yet, or C after an operator other than the boundary apply).
void p0_invalidate(Handle& h);
bool p0_sync(Handle& h, int which, hipStream_t s);

// Hardware sine and cosine (v_sin_f32 / v_cos_f32, argument in revolutions)
// of rev[0..n) as the pair kernels evaluate them; device pointers.
void hw_sincos_dev(const float* rev, int64_t n, float* sin_out, float* cos_out, hipStream_t s);

// Largest phase kappa r / (2 pi) in revolutions the pair kernels accept: the
// hardware sine and cosine are accurate to about 1e-7 up to there, and the
// FP32 phase itself resolves 2^-8 revolutions at that size.
constexpr double kMaxPhaseRev = 65536.0;

}  // namespace baemm
