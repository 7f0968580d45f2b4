// Helmholtz_HIP.hpp -- header-only C++ drop-in for BAEMM::Helmholtz_OpenCL,
// implemented on the C ABI of libbaemm_hip (baemm.h).
//
// Method names, template parameters (WC, external real/complex/integer types)
// and argument order follow the reference class (Helmholtz_OpenCL.hpp:34-399,
// src/Helmholtz_Common/*.hpp).  `using Helmholtz_OpenCL = Helmholtz_HIP;` lets
// existing callers compile unchanged; link with -lbaemm_hip.
//
// Error behaviour: like the reference (eprint/wprint then return) a failed
// call prints a message to stderr and returns; define BAEMM_HIP_THROW to get a
// std::runtime_error instead.  GaussNewtonSolve returns the outer GMRES flag.
#pragma once

#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "baemm.h"

namespace BAEMM {

// BAEMM_Common.hpp:7-11
enum class WaveType : unsigned char { Plane = 201, Radial = 202 };

namespace detail {

template <class T>
struct is_complex : std::false_type {};
template <class T>
struct is_complex<std::complex<T>> : std::true_type {};

inline void report(baemm_handle h, int st, const char* where)
{
    if (st == BAEMM_OK) return;
    std::string msg = std::string("Helmholtz_HIP::") + where + ": " + baemm_last_error(h) +
                      " (status " + std::to_string(st) + ")";
#ifdef BAEMM_HIP_THROW
    throw std::runtime_error(msg);
#else
    std::fprintf(stderr, "%s\n", msg.c_str());
#endif
}

// Complex external arrays as interleaved doubles (copy only when C_ext is not complex<double>).
template <class C>
struct CIn {
    std::vector<std::complex<double>> buf;
    const double* p = nullptr;
    int64_t ld = 0;
    CIn(const C* a, int64_t rows, int64_t ld_, int64_t cols)
    {
        static_assert(is_complex<C>::value, "complex external type expected");
        if constexpr (std::is_same_v<C, std::complex<double>>) {
            p = reinterpret_cast<const double*>(a);
            ld = ld_;
        } else {
            buf.resize((size_t)rows * cols);
            for (int64_t i = 0; i < rows; ++i)
                for (int64_t k = 0; k < cols; ++k) buf[i * cols + k] = std::complex<double>(a[i * ld_ + k]);
            p = reinterpret_cast<const double*>(buf.data());
            ld = cols;
        }
    }
};

template <class C>
struct COut {
    std::vector<std::complex<double>> buf;
    C* a;
    int64_t rows, ld_, cols;
    double* p = nullptr;
    int64_t ld = 0;
    COut(C* a_, int64_t rows_, int64_t ld__, int64_t cols_, bool read) : a(a_), rows(rows_), ld_(ld__), cols(cols_)
    {
        if constexpr (std::is_same_v<C, std::complex<double>>) {
            p = reinterpret_cast<double*>(a);
            ld = ld_;
        } else {
            buf.assign((size_t)rows * cols, 0.0);
            if (read)
                for (int64_t i = 0; i < rows; ++i)
                    for (int64_t k = 0; k < cols; ++k) buf[i * cols + k] = std::complex<double>(a[i * ld_ + k]);
            p = reinterpret_cast<double*>(buf.data());
            ld = cols;
        }
    }
    ~COut()
    {
        if constexpr (!std::is_same_v<C, std::complex<double>>)
            for (int64_t i = 0; i < rows; ++i)
                for (int64_t k = 0; k < cols; ++k) a[i * ld_ + k] = C(buf[i * cols + k]);
    }
};

template <class R>
std::vector<double> to_double(const R* a, int64_t n)
{
    return std::vector<double>(a, a + n);
}

template <class C>
std::vector<double> coeff_double(const C* c, int64_t n)
{
    std::vector<double> out(2 * n);
    for (int64_t i = 0; i < n; ++i) {
        out[2 * i] = (double)std::real(c[i]);
        out[2 * i + 1] = (double)std::imag(c[i]);
    }
    return out;
}

template <class C>
void scalar(const C& z, double (&out)[2])
{
    out[0] = (double)std::real(z);
    out[1] = (double)std::imag(z);
}

}  // namespace detail

class TangentPoint_HIP;

class Helmholtz_HIP {
public:
    using Int = int32_t;
    using LInt = int64_t;
    using Real = float;
    using Complex = std::complex<float>;

    static constexpr Int VarSize = 0;

    // Public members of the reference (MemberVariables.hpp:64-65).
    Int gmres_max_iter = 30;
    Int gmres_max_restarts = 3;

    // P1 mass matrix view with the Dot<NRHS> interface used by callers
    // (Example_GaussNewton/main.cpp:395-399).
    class MassMatrixView {
    public:
        explicit MassMatrixView(baemm_handle h) : h_(h) {}
        template <Int NRHS = VarSize, class R, class I>
        void Dot(R alpha, const R* X, I ldX, R beta, R* Y, I ldY, I nrhs = NRHS) const
        {
            const int64_t n = baemm_vertex_count(h_);
            const int64_t k = (NRHS > 0) ? NRHS : (int64_t)nrhs;
            std::vector<double> x((size_t)n * k), y((size_t)n * k);
            for (int64_t i = 0; i < n; ++i)
                for (int64_t j = 0; j < k; ++j) {
                    x[i * k + j] = (double)X[i * ldX + j];
                    y[i * k + j] = (double)Y[i * ldY + j];
                }
            detail::report(h_, baemm_mass_dot(h_, (double)alpha, x.data(), k, (double)beta, y.data(), k, k),
                           "MassMatrix().Dot");
            for (int64_t i = 0; i < n; ++i)
                for (int64_t j = 0; j < k; ++j) Y[i * ldY + j] = (R)y[i * k + j];
        }

    private:
        baemm_handle h_;
    };

    // Helmholtz_OpenCL.hpp:109-159 (device 0) and :175-249 (device_num).
    template <typename ExtReal, typename ExtInt>
    Helmholtz_HIP(const ExtReal* vertex_coords, ExtInt vertex_count, const ExtInt* triangles, ExtInt simplex_count,
                  const ExtReal* meas_directions, ExtInt meas_count, ExtInt CPU_thread_count)
        : Helmholtz_HIP(vertex_coords, vertex_count, triangles, simplex_count, meas_directions, meas_count, ExtInt(0),
                        CPU_thread_count)
    {
    }

    template <typename ExtReal, typename ExtInt>
    Helmholtz_HIP(const ExtReal* vertex_coords, ExtInt vertex_count, const ExtInt* triangles, ExtInt simplex_count,
                  const ExtReal* meas_directions, ExtInt meas_count, ExtInt device_num, ExtInt CPU_thread_count)
        : threads_((Int)CPU_thread_count)
    {
        static_assert(std::is_integral_v<ExtInt>, "");
        static_assert(std::is_floating_point_v<ExtReal>, "");
        const std::vector<double> V = detail::to_double(vertex_coords, 3 * (int64_t)vertex_count);
        const std::vector<double> M =
            meas_directions ? detail::to_double(meas_directions, 3 * (int64_t)meas_count) : std::vector<double>();
        std::vector<int64_t> T(triangles, triangles + 3 * (int64_t)simplex_count);
        const int st = baemm_create(V.data(), vertex_count, T.data(), simplex_count, M.empty() ? nullptr : M.data(),
                                    M.empty() ? 0 : meas_count, (int)device_num, &h_);
        detail::report(h_, st, "Helmholtz_HIP");
        if (st != BAEMM_OK) throw std::runtime_error("Helmholtz_HIP: construction failed");
        vertex_coords_ = V;
        triangles_ = T;
        device_ = (int)device_num;
    }

    ~Helmholtz_HIP()
    {
        if (h_) baemm_destroy(h_);
    }

    Helmholtz_HIP(const Helmholtz_HIP&) = delete;
    Helmholtz_HIP& operator=(const Helmholtz_HIP&) = delete;

    baemm_handle Handle() const { return h_; }

    // ---------------------------------------------------------- GetSetters.hpp
    Int ThreadCountCPU() const { return threads_; }
    Int VertexCount() const { return (Int)baemm_vertex_count(h_); }
    Int SimplexCount() const { return (Int)baemm_simplex_count(h_); }
    Int GetMeasCount() const { return (Int)baemm_meas_count(h_); }
    MassMatrixView MassMatrix() const { return MassMatrixView(h_); }

    Int GetWaveCount() const
    {
        if (wave_count_set_ >= 0) return wave_count_set_;
        int64_t W = 0;
        baemm_get_parameters(h_, nullptr, nullptr, &W, nullptr, nullptr);
        return (Int)W;
    }
    // GetSetters.hpp:40-45.  The reference's LoadParameters overwrites the value;
    // until then it is what GetWaveCount reports and the width the preloaded
    // applies expect (an apply whose width differs from the loaded parameters'
    // fails with an argument error).
    void SetWaveCount(Int wave_count_) { wave_count_set_ = wave_count_; }
    Int GetWaveChunkSize() const { return (Int)baemm_get_wave_chunk_size(h_); }
    void SetWaveChunkSize(Int wcs) { detail::report(h_, baemm_set_wave_chunk_size(h_, wcs), "SetWaveChunkSize"); }
    Int GetWaveChunkCount(Int wave_count) const
    {
        const Int wcs = GetWaveChunkSize();
        return (wave_count + wcs - 1) / wcs;
    }
    Int GetBlockSize() const { return (Int)baemm_get_block_size(h_); }
    void SetBlockSize(Int b) { detail::report(h_, baemm_set_block_size(h_, b), "SetBlockSize"); }
    void UseDiagonal(bool use_it) { detail::report(h_, baemm_use_diagonal(h_, use_it ? 1 : 0), "UseDiagonal"); }

    std::vector<Real> GetWaveNumbers() const
    {
        int64_t K = 0;
        baemm_get_parameters(h_, nullptr, nullptr, nullptr, nullptr, &K);
        std::vector<Real> k(K);
        if (K) baemm_get_parameters(h_, k.data(), nullptr, nullptr, nullptr, nullptr);
        return k;
    }
    std::vector<Complex> GetCoefficients() const
    {
        int64_t K = 0;
        baemm_get_parameters(h_, nullptr, nullptr, nullptr, nullptr, &K);
        std::vector<Complex> c(4 * K);
        if (K) baemm_get_parameters(h_, nullptr, reinterpret_cast<float*>(c.data()), nullptr, nullptr, nullptr);
        return c;
    }

    const Real* Areas() const { return tri_data(2); }
    const Real* TriangleAreas() const { return tri_data(2); }
    const Real* TriangleNormals() const { return tri_data(1); }
    const Real* TriangleMidpoints() const { return tri_data(0); }
    const Real* SingleLayerDiagonal() const { return tri_data(3); }

    // Not in the reference: contraction arithmetic of the pair kernels.
    void SetPrecision(int mode) { detail::report(h_, baemm_set_precision(h_, mode), "SetPrecision"); }

    std::string ClassName() const { return "Helmholtz_HIP"; }

    // ------------------------------------------------ Helmholtz_OpenCL.hpp:333-391
    // FP32 copies of the input: V x 3, and T x 12 (three vertices, each x, y, z, 0).
    const float* VertexCoordinates() const
    {
        if (vcoords_.empty()) {
            vcoords_.resize(3 * (size_t)VertexCount());
            baemm_get_coordinates(h_, vcoords_.data(), nullptr);
        }
        return vcoords_.data();
    }

    const float* TriangleCoordinates() const
    {
        if (tcoords_.empty()) {
            tcoords_.resize(12 * (size_t)SimplexCount());
            baemm_get_coordinates(h_, nullptr, tcoords_.data());
        }
        return tcoords_.data();
    }

    std::string DeviceInfo()
    {
        char buf[1024] = {0};
        detail::report(h_, baemm_device_info(h_, buf, sizeof(buf)), "DeviceInfo");
        return std::string(buf);
    }

    // ------------------------------------------------ LoadParameters.hpp:85-95
    void PrintBooleans() const
    {
        static const char* names[8] = {"Re_mass_matrix",  "Im_mass_matrix",  "Re_single_layer", "Im_single_layer",
                                       "Re_double_layer", "Im_double_layer", "Re_adjdbl_layer", "Im_adjdbl_layer"};
        int f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        detail::report(h_, baemm_get_flags(h_, f), "PrintBooleans");
        for (int j = 0; j < 8; ++j) std::printf("%s = %s\n", names[j], f[j] ? "true" : "false");
    }

    // ------------------------------------------------ InputOutput.hpp:5-89
    // The P0 buffers: B = Av B_in of the last operator apply, C = (K + D) b of
    // the last boundary-operator apply, simplex_count x wave_count (Complex).
    Complex& B(const Int i, const Int k) { return p0_ref(0, i, k); }
    const Complex& B(const Int i, const Int k) const { return p0_ref(0, i, k); }
    Complex& C(const Int i, const Int k) { return p0_ref(1, i, k); }
    const Complex& C(const Int i, const Int k) const { return p0_ref(1, i, k); }

    void ReadB(const Complex* input, const Int ld_input, const Int wave_count_)
    {
        detail::report(h_, baemm_read_b(h_, reinterpret_cast<const float*>(input), ld_input, wave_count_), "ReadB");
    }
    void ReadB(const Complex* input, const Int wave_count_) { ReadB(input, wave_count_, wave_count_); }
    void WriteB(Complex* output, const Int ld_output) const
    {
        detail::report(h_, baemm_write_b(h_, reinterpret_cast<float*>(output), ld_output), "WriteB");
    }
    void ReadC(const Complex* input, const Int ld_input, const Int wave_count_)
    {
        detail::report(h_, baemm_read_c(h_, reinterpret_cast<const float*>(input), ld_input, wave_count_), "ReadC");
    }
    void ReadC(const Complex* input, const Int wave_count_) { ReadC(input, wave_count_, wave_count_); }
    void WriteC(Complex* output, const Int ld_output) const
    {
        detail::report(h_, baemm_write_c(h_, reinterpret_cast<float*>(output), ld_output), "WriteC");
    }

    // ----------------------------------------------------- LoadParameters.hpp
    template <typename R_ext, typename C_ext, typename I_ext>
    void LoadParameters(const R_ext* kappa_list, const C_ext* coeff_list, I_ext wave_count_, I_ext wave_chunk_size_,
                        R_ext factor = R_ext(1))
    {
        const int64_t K = (wave_count_ + wave_chunk_size_ - 1) / wave_chunk_size_;
        const std::vector<double> k = detail::to_double(kappa_list, K);
        const std::vector<double> c = detail::coeff_double(coeff_list, 4 * K);
        wave_count_set_ = -1;
        detail::report(h_, baemm_load_parameters(h_, k.data(), c.data(), wave_count_, wave_chunk_size_, (double)factor),
                       "LoadParameters");
    }

    // ----------------------------------------------- ApplyBoundaryOperators.hpp
    template <typename R_ext, typename C_ext, typename I_ext>
    void LoadBoundaryOperators_PL(const R_ext* kappa_list, const C_ext* coeff_list, I_ext wave_count_,
                                  I_ext wave_chunk_size_)
    {
        LoadParameters(kappa_list, coeff_list, wave_count_, wave_chunk_size_, R_ext(1));
    }

    void UnloadBoundaryOperators_PL() { detail::report(h_, baemm_unload_boundary_operators(h_), "Unload"); }

    template <Int WC = VarSize, typename R_ext, typename C_ext, typename I_ext>
    void ApplyBoundaryOperators_PL(const C_ext alpha, const C_ext* B_in, const I_ext ldB_in, const C_ext beta,
                                   C_ext* C_out, const I_ext ldC_out, const R_ext* kappa_list,
                                   const C_ext* coeff_list, const I_ext wave_count_, const I_ext wave_chunk_size_)
    {
        LoadBoundaryOperators_PL(kappa_list, coeff_list, wave_count_, wave_chunk_size_);
        ApplyBoundaryOperators_PL<WC>(alpha, B_in, ldB_in, beta, C_out, ldC_out);
        UnloadBoundaryOperators_PL();
    }

    template <Int WC = VarSize, typename C_ext, typename I_ext>
    void ApplyBoundaryOperators_PL(const C_ext alpha, const C_ext* B_in, const I_ext ldB_in, const C_ext beta,
                                   C_ext* C_out, const I_ext ldC_out)
    {
        const Int W = GetWaveCount();
        if (!check_wc(WC, W, "ApplyBoundaryOperators_PL")) return;
        const int64_t n = VertexCount();
        detail::CIn<C_ext> B(B_in, n, ldB_in, W);
        detail::COut<C_ext> C(C_out, n, ldC_out, W, true);
        double a[2], b[2];
        detail::scalar(alpha, a);
        detail::scalar(beta, b);
        detail::report(h_, baemm_apply_boundary(h_, a, B.p, B.ld, b, C.p, C.ld), "ApplyBoundaryOperators_PL");
    }

    // ----------------------------------------------- ApplyFarFieldOperators.hpp
    template <Int WC = VarSize, typename R_ext, typename C_ext, typename I_ext>
    void ApplyFarFieldOperators_PL(const C_ext alpha, const C_ext* B_in, const I_ext ldB_in, const C_ext beta,
                                   C_ext* C_out, const I_ext ldC_out, const R_ext* kappa_list,
                                   const C_ext* coeff_list, const I_ext wave_count_, const I_ext wave_chunk_size_)
    {
        LoadParameters(kappa_list, coeff_list, wave_count_, wave_chunk_size_);
        ApplyFarFieldOperators_PL<WC>(alpha, B_in, ldB_in, beta, C_out, ldC_out);
    }

    template <Int WC = VarSize, typename C_ext, typename I_ext>
    void ApplyFarFieldOperators_PL(const C_ext alpha, const C_ext* B_in, const I_ext ldB_in, const C_ext beta,
                                   C_ext* C_out, const I_ext ldC_out)
    {
        const Int W = GetWaveCount();
        if (!check_wc(WC, W, "ApplyFarFieldOperators_PL")) return;
        detail::CIn<C_ext> B(B_in, VertexCount(), ldB_in, W);
        detail::COut<C_ext> C(C_out, GetMeasCount(), ldC_out, W, true);
        double a[2], b[2];
        detail::scalar(alpha, a);
        detail::scalar(beta, b);
        detail::report(h_, baemm_apply_far_field(h_, a, B.p, B.ld, b, C.p, C.ld), "ApplyFarFieldOperators_PL");
    }

    // ---------------------------------------------- ApplyNearFieldOperators.hpp
    template <typename R_ext, typename C_ext, typename I_ext>
    void ApplyNearFieldOperators_PL(const C_ext alpha, const C_ext* B_in, const I_ext ldB_in, const C_ext beta,
                                    C_ext* C_out, const I_ext ldC_out, const R_ext* kappa_list,
                                    const C_ext* coeff_list, const I_ext wave_count_, const I_ext wave_chunk_size_,
                                    const R_ext* evaluation_points_, const I_ext evaluation_count_)
    {
        LoadParameters(kappa_list, coeff_list, wave_count_, wave_chunk_size_);
        ApplyNearFieldOperators_PL(alpha, B_in, ldB_in, beta, C_out, ldC_out, evaluation_points_, evaluation_count_);
    }

    template <typename R_ext, typename C_ext, typename I_ext>
    void ApplyNearFieldOperators_PL(const C_ext alpha, const C_ext* B_in, const I_ext ldB_in, const C_ext beta,
                                    C_ext* C_out, const I_ext ldC_out, const R_ext* evaluation_points_,
                                    const I_ext evaluation_count_)
    {
        const Int W = GetWaveCount();
        detail::CIn<C_ext> B(B_in, VertexCount(), ldB_in, W);
        detail::COut<C_ext> C(C_out, evaluation_count_, ldC_out, W, true);
        const std::vector<double> P = detail::to_double(evaluation_points_, 3 * (int64_t)evaluation_count_);
        double a[2], b[2];
        detail::scalar(alpha, a);
        detail::scalar(beta, b);
        detail::report(h_, baemm_apply_near_field(h_, a, B.p, B.ld, b, C.p, C.ld, P.data(), evaluation_count_),
                       "ApplyNearFieldOperators_PL");
    }

    // ------------------------------------------ CreateIncidentWave / Herglotz
    template <typename R_ext, typename C_ext, typename I_ext>
    void CreateIncidentWave_PL(const C_ext alpha, const R_ext* incident_directions, const I_ext inc_count,
                               const C_ext beta, C_ext* C_out, const I_ext ldC_out, const R_ext* kappa_list,
                               const C_ext* coeff_list, const I_ext wave_count_, const I_ext wave_chunk_size_,
                               WaveType type = WaveType::Plane)
    {
        if (inc_count > wave_chunk_size_)
            std::fprintf(stderr, "Helmholtz_HIP::CreateIncidentWave_PL: excess incident directions will be ignored\n");
        LoadParameters(kappa_list, coeff_list, wave_count_, wave_chunk_size_, R_ext(4.0 * M_PI));
        const Int W = GetWaveCount();
        const std::vector<double> d = detail::to_double(incident_directions, 3 * (int64_t)inc_count);
        detail::COut<C_ext> C(C_out, VertexCount(), ldC_out, W, true);
        double a[2], b[2];
        detail::scalar(alpha, a);
        detail::scalar(beta, b);
        detail::report(h_, baemm_create_incident_wave(h_, a, d.data(), inc_count, b, C.p, C.ld, (int)type),
                       "CreateIncidentWave_PL");
    }

    template <typename R_ext, typename C_ext, typename I_ext>
    void CreateHerglotzWave_PL(const C_ext alpha, const C_ext* B_in, const I_ext ldB_in, const C_ext beta,
                               C_ext* C_out, const I_ext ldC_out, const R_ext* kappa_list, const C_ext* coeff_list,
                               const I_ext wave_count_, const I_ext wave_chunk_size_)
    {
        LoadParameters(kappa_list, coeff_list, wave_count_, wave_chunk_size_, R_ext(4.0 * M_PI / GetMeasCount()));
        const Int W = GetWaveCount();
        detail::CIn<C_ext> B(B_in, GetMeasCount(), ldB_in, W);
        detail::COut<C_ext> C(C_out, VertexCount(), ldC_out, W, true);
        double a[2], b[2];
        detail::scalar(alpha, a);
        detail::scalar(beta, b);
        detail::report(h_, baemm_create_herglotz_wave(h_, a, B.p, B.ld, b, C.p, C.ld), "CreateHerglotzWave_PL");
    }

    // ------------------------------------------------------ ApplyMassInverse.hpp
    template <Int NRHS = VarSize, typename X_T, typename Y_T, typename R_ext>
    void ApplyMassInverse(const X_T* X, const Int ldX, Y_T* Y, const Int ldY, const R_ext cg_tol,
                          const Int nrhs = NRHS)
    {
        const int64_t n = VertexCount();
        if constexpr (detail::is_complex<Y_T>::value) {
            detail::CIn<X_T> x(X, n, ldX, nrhs);
            detail::COut<Y_T> y(Y, n, ldY, nrhs, false);
            detail::report(h_, baemm_apply_mass_inverse_complex(h_, x.p, x.ld, y.p, y.ld, nrhs, (double)cg_tol),
                           "ApplyMassInverse");
        } else {
            std::vector<double> x((size_t)n * nrhs), y((size_t)n * nrhs);
            for (int64_t i = 0; i < n; ++i)
                for (int64_t k = 0; k < nrhs; ++k) x[i * nrhs + k] = (double)X[i * ldX + k];
            detail::report(h_, baemm_apply_mass_inverse(h_, x.data(), nrhs, y.data(), nrhs, nrhs, (double)cg_tol),
                           "ApplyMassInverse");
            for (int64_t i = 0; i < n; ++i)
                for (int64_t k = 0; k < nrhs; ++k) Y[i * ldY + k] = (Y_T)y[i * nrhs + k];
        }
    }

    // ApplyMassInverse.hpp:161-197: Y = diag(areas_lumped_inv) X, row by row;
    // real or complex columns.
    template <Int NRHS = VarSize, typename X_T, typename Y_T>
    void ApplyLumpedMassInverse(const X_T* X, const Int ldX, Y_T* Y, const Int ldY, const Int nrhs = NRHS)
    {
        constexpr int c = detail::is_complex<X_T>::value ? 2 : 1;
        static_assert(detail::is_complex<Y_T>::value == (c == 2), "X and Y must both be real or both complex");
        const int64_t n = VertexCount(), k = (int64_t)c * nrhs;
        std::vector<double> x((size_t)n * k), y((size_t)n * k);
        for (int64_t i = 0; i < n; ++i)
            for (int64_t j = 0; j < nrhs; ++j) {
                if constexpr (c == 2) {
                    x[i * k + 2 * j] = (double)X[i * ldX + j].real();
                    x[i * k + 2 * j + 1] = (double)X[i * ldX + j].imag();
                } else {
                    x[i * k + j] = (double)X[i * ldX + j];
                }
            }
        detail::report(h_, baemm_apply_lumped_mass_inverse(h_, x.data(), k, y.data(), k, k), "ApplyLumpedMassInverse");
        for (int64_t i = 0; i < n; ++i)
            for (int64_t j = 0; j < nrhs; ++j) {
                if constexpr (c == 2) {
                    using R = typename Y_T::value_type;
                    Y[i * ldY + j] = Y_T((R)y[i * k + 2 * j], (R)y[i * k + 2 * j + 1]);
                } else {
                    Y[i * ldY + j] = (Y_T)y[i * k + j];
                }
            }
    }

    // ------------------------------------------------------------- FarField.hpp
    template <Int WC, typename I_ext, typename R_ext, typename C_ext>
    void FarField(const R_ext* kappa_, const I_ext wave_chunk_count_, R_ext* inc_directions,
                  const I_ext wave_chunk_size_, C_ext* Y_out, const WaveType type, const R_ext cg_tol,
                  const R_ext gmres_tol)
    {
        FarField_parameters<WC>(kappa_, wave_chunk_count_, inc_directions, wave_chunk_size_, Y_out, type, kappa_,
                                cg_tol, gmres_tol);
    }

    template <Int WC = VarSize, typename I_ext, typename R_ext, typename C_ext>
    void FarField_parameters(const R_ext* kappa_, const I_ext wave_chunk_count_, R_ext* inc_directions,
                             const I_ext wave_chunk_size_, C_ext* Y_out, const WaveType type, const R_ext* eta,
                             const R_ext cg_tol, const R_ext gmres_tol)
    {
        const int64_t K = wave_chunk_count_, wcs = wave_chunk_size_, W = K * wcs;
        if (!check_wc(WC, (Int)W, "FarField")) return;
        sync_limits();
        const std::vector<double> k = detail::to_double(kappa_, K), e = detail::to_double(eta, K);
        const std::vector<double> d = detail::to_double(inc_directions, 3 * wcs);
        detail::COut<C_ext> Y(Y_out, GetMeasCount(), W, W, false);
        solve_report(baemm_far_field_parameters(h_, k.data(), K, d.data(), wcs, Y.p, (int)type, e.data(),
                                                (double)cg_tol, (double)gmres_tol),
                     "FarField");
    }

    template <Int WC = VarSize, typename I_ext, typename R_ext, typename C_ext>
    void Derivative_FF_Helper(const R_ext* kappa_, const I_ext wave_chunk_count_, const R_ext* inc_directions,
                              const I_ext wave_chunk_size_, C_ext* du_dn, const WaveType type, const R_ext cg_tol,
                              const R_ext gmres_tol)
    {
        const int64_t K = wave_chunk_count_, wcs = wave_chunk_size_, W = K * wcs;
        sync_limits();
        const std::vector<double> k = detail::to_double(kappa_, K), d = detail::to_double(inc_directions, 3 * wcs);
        detail::COut<C_ext> du(du_dn, VertexCount(), W, W, false);
        solve_report(baemm_derivative_ff_helper(h_, k.data(), K, d.data(), wcs, du.p, (int)type, (double)cg_tol,
                                                (double)gmres_tol),
                     "Derivative_FF_Helper");
    }

    template <Int WC = VarSize, typename I_ext, typename R_ext, typename C_ext>
    void Derivative_FF(const R_ext* kappa_, const I_ext wave_chunk_count_, const R_ext* inc_directions,
                       const I_ext wave_chunk_size_, const R_ext* X_in, C_ext* Y_out, C_ext*& du_dn,
                       const WaveType type, const R_ext cg_tol, const R_ext gmres_tol)
    {
        static_assert(std::is_same_v<C_ext, std::complex<double>>, "du_dn is exchanged as complex<double>");
        const int64_t K = wave_chunk_count_, wcs = wave_chunk_size_, W = K * wcs;
        if (!check_wc(WC, (Int)W, "Derivative_FF")) return;
        sync_limits();
        const std::vector<double> k = detail::to_double(kappa_, K), d = detail::to_double(inc_directions, 3 * wcs);
        const std::vector<double> X = detail::to_double(X_in, 3 * (int64_t)VertexCount());
        double* du = reinterpret_cast<double*>(du_dn);
        solve_report(baemm_derivative_ff(h_, k.data(), K, d.data(), wcs, X.data(), reinterpret_cast<double*>(Y_out),
                                         &du, (int)type, (double)cg_tol, (double)gmres_tol),
                     "Derivative_FF");
        du_dn = reinterpret_cast<C_ext*>(du);
    }

    template <Int WC, typename I_ext, typename R_ext, typename C_ext>
    void AdjointDerivative_FF(const R_ext* kappa_, const I_ext wave_chunk_count_, const R_ext* inc_directions,
                              const I_ext wave_chunk_size_, const C_ext* g, R_ext* Y_out, C_ext*& du_dn,
                              const WaveType type, const R_ext cg_tol, const R_ext gmres_tol)
    {
        static_assert(std::is_same_v<C_ext, std::complex<double>>, "du_dn is exchanged as complex<double>");
        const int64_t K = wave_chunk_count_, wcs = wave_chunk_size_, W = K * wcs;
        if (!check_wc(WC, (Int)W, "AdjointDerivative_FF")) return;
        sync_limits();
        const std::vector<double> k = detail::to_double(kappa_, K), d = detail::to_double(inc_directions, 3 * wcs);
        std::vector<double> Y(3 * (size_t)VertexCount());
        double* du = reinterpret_cast<double*>(du_dn);
        solve_report(baemm_adjoint_derivative_ff(h_, k.data(), K, d.data(), wcs, reinterpret_cast<const double*>(g),
                                                 Y.data(), &du, (int)type, (double)cg_tol, (double)gmres_tol),
                     "AdjointDerivative_FF");
        du_dn = reinterpret_cast<C_ext*>(du);
        for (size_t i = 0; i < Y.size(); ++i) Y_out[i] = (R_ext)Y[i];
    }

    template <Int WC, typename I_ext, typename R_ext, typename C_ext, typename M_T, typename P_T>
    bool GaussNewtonSolve(const R_ext* kappa_, const I_ext wave_chunk_count_, const R_ext* inc_directions,
                          const I_ext wave_chunk_size_, M_T& M, P_T& P, const R_ext alpha, const R_ext* B_in,
                          const I_ext ldB, const R_ext beta, R_ext* X_out, const I_ext ldX, C_ext*& du_dn,
                          const WaveType type, const R_ext cg_tol, const R_ext gmres_tol_inner,
                          const R_ext gmres_tol_outer)
    {
        static_assert(std::is_same_v<R_ext, double>, "GaussNewtonSolve callbacks take double arrays");
        static_assert(std::is_same_v<C_ext, std::complex<double>>, "du_dn is exchanged as complex<double>");
        const int64_t K = wave_chunk_count_, wcs = wave_chunk_size_, W = K * wcs;
        if (!check_wc(WC, (Int)W, "GaussNewtonSolve")) return false;
        sync_limits();
        const std::vector<double> k = detail::to_double(kappa_, K), d = detail::to_double(inc_directions, 3 * wcs);
        auto tramp_M = [](void* u, const double* x, double* y) { (*static_cast<M_T*>(u))(x, y); };
        auto tramp_P = [](void* u, const double* x, double* y) { (*static_cast<P_T*>(u))(x, y); };
        double* du = reinterpret_cast<double*>(du_dn);
        int ok = 0;
        detail::report(h_,
                       baemm_gauss_newton_solve(h_, k.data(), K, d.data(), wcs, tramp_M, &M, tramp_P, &P, alpha, B_in,
                                                ldB, beta, X_out, ldX, &du, (int)type, cg_tol, gmres_tol_inner,
                                                gmres_tol_outer, &ok),
                       "GaussNewtonSolve");
        du_dn = reinterpret_cast<C_ext*>(du);
        return ok != 0;
    }

    template <Int WC = VarSize, typename I_ext, typename R_ext, typename C_ext>
    void BoundaryPotential(const R_ext* kappa_, C_ext* coeff, C_ext* wave, C_ext* phi, const I_ext wave_chunk_count_,
                           const I_ext wave_chunk_size_, const R_ext cg_tol, const R_ext gmres_tol)
    {
        static_assert(std::is_same_v<C_ext, std::complex<double>>, "complex<double> expected");
        const int64_t K = wave_chunk_count_, wcs = wave_chunk_size_;
        sync_limits();
        const std::vector<double> k = detail::to_double(kappa_, K);
        solve_report(baemm_boundary_potential(h_, k.data(), K, nullptr, reinterpret_cast<double*>(coeff),
                                              reinterpret_cast<const double*>(wave), reinterpret_cast<double*>(phi),
                                              wcs, (double)cg_tol, (double)gmres_tol),
                     "BoundaryPotential");
    }

    template <Int WC, typename I_ext, typename R_ext, typename C_ext>
    void DirichletToNeumann(const R_ext* kappa_, C_ext* wave, C_ext* neumann_trace, const I_ext wave_chunk_count_,
                            const I_ext wave_chunk_size_, const R_ext cg_tol, const R_ext gmres_tol)
    {
        static_assert(std::is_same_v<C_ext, std::complex<double>>, "complex<double> expected");
        const int64_t K = wave_chunk_count_, wcs = wave_chunk_size_;
        sync_limits();
        const std::vector<double> k = detail::to_double(kappa_, K);
        solve_report(baemm_dirichlet_to_neumann(h_, k.data(), K, reinterpret_cast<const double*>(wave),
                                                reinterpret_cast<double*>(neumann_trace), wcs, (double)cg_tol,
                                                (double)gmres_tol),
                     "DirichletToNeumann");
    }

private:
    baemm_handle h_ = nullptr;
    Int threads_ = 1;
    Int wave_count_set_ = -1;   // SetWaveCount's value until the next LoadParameters (-1: none)
    std::vector<double> vertex_coords_;
    std::vector<int64_t> triangles_;
    int device_ = 0;
    friend class TangentPoint_HIP;
    mutable std::vector<Real> tri_[4];
    mutable std::vector<float> vcoords_, tcoords_;
    mutable Complex p0_dummy_;

    const Real* tri_data(int which) const
    {
        const int64_t nt = SimplexCount();
        if (tri_[which].empty()) {
            tri_[which].resize((which <= 1 ? 4 : 1) * nt);
            float* p[4] = {nullptr, nullptr, nullptr, nullptr};
            p[which] = tri_[which].data();
            baemm_get_triangle_data(h_, p[0], p[1], p[2], p[3]);
        }
        return tri_[which].data();
    }

    // Element (i, k) of the handle's host copy of P0 buffer `which` (kept by
    // the library; brought up to date with the device on each access).
    Complex& p0_ref(int which, Int i, Int k) const
    {
        float* data = nullptr;
        int64_t ld = 0;
        const int st = baemm_p0_buffer(h_, which, &data, &ld);
        if (st != BAEMM_OK || !data) {
            detail::report(h_, st == BAEMM_OK ? BAEMM_ERR_STATE : st, which == 0 ? "B" : "C");
            p0_dummy_ = Complex(0.0f, 0.0f);
            return p0_dummy_;
        }
        return reinterpret_cast<Complex*>(data)[ld * (int64_t)i + k];
    }

    bool check_wc(Int WC, Int W, const char* where) const
    {
        // ApplyBoundaryOperators.hpp:100-105: a compile-time wave count that
        // differs from the loaded one is an error and the call does nothing.
        if (WC > VarSize && WC != W) {
            std::fprintf(stderr, "Helmholtz_HIP::%s: WC != wave_count. Doing nothing.\n", where);
            return false;
        }
        return true;
    }

    void sync_limits() { baemm_set_gmres_limits(h_, gmres_max_iter, gmres_max_restarts); }

    void solve_report(int st, const char* where)
    {
        if (st == BAEMM_ERR_NOCONV) {
            std::fprintf(stderr, "Helmholtz_HIP::%s: GMRES did not reach its tolerance.\n", where);
            return;
        }
        detail::report(h_, st, where);
    }
};

// Drop-in name of the reference class.
using Helmholtz_OpenCL = Helmholtz_HIP;

// The tangent-point energy and metric that the Gauss-Newton example takes from
// Repulsor (Example_GaussNewton/main.cpp:301-346: TangentPointEnergy0,
// TangentPointMetric0, not part of BAEMM), on the mesh of a Helmholtz_HIP,
// with the call shapes the example uses, the mesh argument dropped:
//   tpe.Differential(*M, alpha, beta, B, ldB)  ->  tp.Differential(alpha, beta, B, ldB)
//   tpm.MultiplyMetric(*M, alpha, X, ldX, beta, Y, ldY, DIM)
//                                             ->  tp.MultiplyMetric(alpha, X, ldX, beta, Y, ldY, DIM)
//   tpm.MultiplyPreconditioner(...)           ->  tp.MultiplyPreconditioner(...)
// Every pair of distinct triangles is summed exactly (Repulsor approximates far
// pairs), and the preconditioner is the metric's diagonal.  DIM must be 3.
// It keeps a handle of its own on the same mesh and device, so that its
// methods may be called from the M and P callbacks of GaussNewtonSolve, which
// run while the Helmholtz_HIP handle is busy with the solve.
class TangentPoint_HIP {
public:
    using Int = Helmholtz_HIP::Int;

    TangentPoint_HIP(const Helmholtz_HIP& H, double q, double p) : q_(q), p_(p)
    {
        const int st = baemm_create(H.vertex_coords_.data(), (int64_t)H.vertex_coords_.size() / 3,
                                    H.triangles_.data(), (int64_t)H.triangles_.size() / 3, nullptr, 0, H.device_, &h_);
        if (st != BAEMM_OK) {
            detail::report(nullptr, st, "TangentPoint_HIP");
            throw std::runtime_error("TangentPoint_HIP: construction failed");
        }
        detail::report(h_, baemm_tp_metric_prepare(h_, q_, p_), "TangentPoint_HIP");
    }

    ~TangentPoint_HIP()
    {
        if (h_) baemm_destroy(h_);
    }

    TangentPoint_HIP(const TangentPoint_HIP&) = delete;
    TangentPoint_HIP& operator=(const TangentPoint_HIP&) = delete;

    double Value() const
    {
        double E = 0.0;
        detail::report(h_, baemm_tp_energy(h_, q_, p_, &E, nullptr), "Value");
        return E;
    }

    // B = alpha dE/dx + beta B (vertex_count x 3, leading dimension ldB)
    template <typename R, typename I>
    void Differential(const R alpha, const R beta, R* B, const I ldB) const
    {
        const int64_t n = baemm_vertex_count(h_);
        std::vector<double> g((size_t)(3 * n));
        const int st = baemm_tp_energy(h_, q_, p_, nullptr, g.data());
        if (st != BAEMM_OK) {
            detail::report(h_, st, "Differential");
            return;
        }
        for (int64_t v = 0; v < n; ++v)
            for (int c = 0; c < 3; ++c) {
                R& b = B[v * (int64_t)ldB + c];
                b = (R)((double)alpha * g[3 * v + c] + (beta == R(0) ? 0.0 : (double)beta * (double)b));
            }
    }

    // Y = alpha M X + beta Y
    template <typename R, typename I>
    void MultiplyMetric(const R alpha, const R* X, const I ldX, const R beta, R* Y, const I ldY, const I nrhs) const
    {
        apply(false, (double)alpha, X, ldX, beta, Y, ldY, nrhs, "MultiplyMetric");
    }

    // Y = alpha P X + beta Y, P X = X / d row by row (d: the diagonal of M)
    template <typename R, typename I>
    void MultiplyPreconditioner(const R alpha, const R* X, const I ldX, const R beta, R* Y, const I ldY,
                                const I nrhs) const
    {
        apply(true, (double)alpha, X, ldX, beta, Y, ldY, nrhs, "MultiplyPreconditioner");
    }

private:
    template <typename R, typename I>
    void apply(bool jacobi, double alpha, const R* X, I ldX, R beta, R* Y, I ldY, I nrhs, const char* where) const
    {
        if (nrhs != 3) {
            std::fprintf(stderr, "TangentPoint_HIP::%s: nrhs must be 3.\n", where);
            return;
        }
        const int64_t n = baemm_vertex_count(h_);
        std::vector<double> x((size_t)(3 * n)), y((size_t)(3 * n), 0.0);
        for (int64_t v = 0; v < n; ++v)
            for (int c = 0; c < 3; ++c) x[3 * v + c] = (double)X[v * (int64_t)ldX + c];
        if (alpha != 0.0) {
            // the metric scaled by alpha; the preconditioner's x / (scale d) with scale = 1 / alpha
            const int st = jacobi ? baemm_tp_metric_jacobi_host(h_, q_, p_, 1.0 / alpha, x.data(), y.data())
                                  : baemm_tp_metric_apply_host(h_, q_, p_, alpha, x.data(), y.data());
            if (st != BAEMM_OK) {
                detail::report(h_, st, where);
                return;
            }
        }
        for (int64_t v = 0; v < n; ++v)
            for (int c = 0; c < 3; ++c) {
                R& o = Y[v * (int64_t)ldY + c];
                o = (R)(y[3 * v + c] + (beta == R(0) ? 0.0 : (double)beta * (double)o));
            }
    }

    baemm_handle h_ = nullptr;
    double q_, p_;
};

}  // namespace BAEMM
