// Exercises the whole C++ drop-in surface (include/Helmholtz_HIP.hpp, the
// methods of Helmholtz_OpenCL.hpp:34-399) on a small sphere and checks
// relations that must hold between the calls: overloads that must agree,
// M (M^{-1} X) = X, the equations the solvers solve, the Gauss-Newton
// residual.  Prints one line per check and exits 1 if any fails.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "Helmholtz_HIP.hpp"

using C64 = std::complex<double>;
using C32 = std::complex<float>;

namespace {

int failures = 0;

void check(const char* name, double value, double limit)
{
    const bool ok = std::isfinite(value) && value <= limit;
    std::printf("CHECK %-44s %.3e <= %.1e %s\n", name, value, limit, ok ? "PASS" : "FAIL");
    if (!ok) ++failures;
}

template <class T>
double rel(const std::vector<T>& a, const std::vector<T>& b)
{
    double num = 0, den = 0;
    for (size_t k = 0; k < a.size(); ++k) {
        num += std::norm(a[k] - b[k]);
        den += std::norm(b[k]);
    }
    return std::sqrt(num / std::max(den, 1e-300));
}

void icosphere(int level, std::vector<double>& V, std::vector<int64_t>& T)
{
    const double p = (1.0 + std::sqrt(5.0)) / 2.0;
    V = {-1, p, 0, 1, p, 0, -1, -p, 0, 1, -p, 0, 0, -1, p, 0, 1, p,
         0, -1, -p, 0, 1, -p, p, 0, -1, p, 0, 1, -p, 0, -1, -p, 0, 1};
    T = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
         3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    auto normalize = [&](int64_t i) {
        double* x = &V[3 * i];
        const double r = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
        x[0] /= r, x[1] /= r, x[2] /= r;
    };
    for (int64_t i = 0; i < 12; ++i) normalize(i);
    for (int l = 0; l < level; ++l) {
        std::map<std::pair<int64_t, int64_t>, int64_t> mid;
        auto midpoint = [&](int64_t a, int64_t b) {
            const auto key = std::make_pair(std::min(a, b), std::max(a, b));
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            const int64_t m = (int64_t)V.size() / 3;
            for (int d = 0; d < 3; ++d) V.push_back(0.5 * (V[3 * a + d] + V[3 * b + d]));
            normalize(m);
            mid.emplace(key, m);
            return m;
        };
        std::vector<int64_t> S;
        for (size_t t = 0; t < T.size(); t += 3) {
            const int64_t a = T[t], b = T[t + 1], c = T[t + 2];
            const int64_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            const int64_t n[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
            S.insert(S.end(), n, n + 12);
        }
        T.swap(S);
    }
}

// Identity-scaled metric of GaussNewtonSolve: y = c x (M) and y = x / c (P).
struct Scale {
    double c;
    void operator()(const double* x, double* y) const
    {
        for (int64_t k = 0; k < n; ++k) y[k] = c * x[k];
    }
    int64_t n;
};

}  // namespace

int main()
{
    try {
        std::vector<double> V, M;
        std::vector<int64_t> T, TM;
        icosphere(3, V, T);
        icosphere(2, M, TM);
        const int64_t nv = (int64_t)V.size() / 3, nt = (int64_t)T.size() / 3, m = (int64_t)M.size() / 3;
        BAEMM::Helmholtz_OpenCL H(V.data(), nv, T.data(), nt, M.data(), m, int64_t(0), int64_t(4));
        check("VertexCount", std::abs((double)H.VertexCount() - nv), 0);
        check("SimplexCount", std::abs((double)H.SimplexCount() - nt), 0);
        check("GetMeasCount", std::abs((double)H.GetMeasCount() - m), 0);
        double area = 0;
        for (int64_t t = 0; t < nt; ++t) area += H.TriangleAreas()[t];
        check("sum of TriangleAreas vs 4 pi (coarse sphere)", std::abs(area - 4 * M_PI) / (4 * M_PI), 2e-2);

        // ---- boundary operators
        const int64_t W = 8, wcs = 4, K = 2;
        const double kappa[K] = {2.0, 3.5};
        const C64 coeff[4 * K] = {{0.5, 0}, {0, -2.0}, {1, 0}, {0, 0}, {0.25, 0.1}, {1, -0.5}, {-0.75, 1}, {0.3, -0.9}};
        std::vector<C64> B((size_t)nv * W), C1((size_t)nv * W), C2((size_t)nv * W, C64(1, 1));
        for (size_t k = 0; k < B.size(); ++k) B[k] = C64(std::sin(0.3 * k), std::cos(0.7 * k));
        H.ApplyBoundaryOperators_PL<0>(C64(1, 0), B.data(), W, C64(0, 0), C1.data(), W, kappa, coeff, W, wcs);
        H.LoadBoundaryOperators_PL(kappa, coeff, W, wcs);
        check("GetWaveCount after load", std::abs((double)H.GetWaveCount() - W), 0);
        H.SetWaveCount(5);
        check("SetWaveCount then GetWaveCount", std::abs((double)H.GetWaveCount() - 5), 0);
        H.LoadBoundaryOperators_PL(kappa, coeff, W, wcs);
        check("LoadParameters overrides SetWaveCount", std::abs((double)H.GetWaveCount() - W), 0);
        check("GetWaveNumbers[1]", std::abs(H.GetWaveNumbers()[1] - 3.5), 1e-6);
        H.ApplyBoundaryOperators_PL<8>(C64(1, 0), B.data(), W, C64(0, 0), C2.data(), W);
        check("loaded vs one-shot apply", rel(C2, C1), 0);
        std::vector<C64> C3 = C2;
        H.ApplyBoundaryOperators_PL<16>(C64(1, 0), B.data(), W, C64(0, 0), C3.data(), W);   // WC != W: no-op
        check("WC != wave count leaves C untouched", rel(C3, C2), 0);
        H.UnloadBoundaryOperators_PL();
        std::vector<C32> Bf(B.begin(), B.end()), Cf((size_t)nv * W);
        const C32 coeff_f[4 * K] = {C32(coeff[0]), C32(coeff[1]), C32(coeff[2]), C32(coeff[3]),
                                    C32(coeff[4]), C32(coeff[5]), C32(coeff[6]), C32(coeff[7])};
        const float kappa_f[K] = {2.0f, 3.5f};
        H.ApplyBoundaryOperators_PL<0>(C32(1, 0), Bf.data(), W, C32(0, 0), Cf.data(), W, kappa_f, coeff_f, W, wcs);
        std::vector<C64> Cfd(Cf.begin(), Cf.end());
        check("complex<float> vs complex<double> apply", rel(Cfd, C1), 2e-6);

        // ---- P0 buffers of the last boundary apply (InputOutput.hpp)
        {
            std::vector<C32> Bp((size_t)nt * W), Cp((size_t)nt * W);
            H.WriteB(Bp.data(), W);
            H.WriteC(Cp.data(), W);
            std::vector<C32> b_ref((size_t)nt * W);
            const float* a = H.TriangleAreas();
            for (int64_t t = 0; t < nt; ++t)
                for (int64_t k = 0; k < W; ++k)
                    b_ref[t * W + k] = C32(a[t] / 3.0f) * (Bf[T[3 * t] * W + k] + Bf[T[3 * t + 1] * W + k] +
                                                           Bf[T[3 * t + 2] * W + k]);
            check("WriteB = Av B of the last apply", rel(Bp, b_ref), 1e-6);
            // AvT C + c0 M B reproduces the apply's result
            std::vector<C64> Cv((size_t)nv * W, C64(0, 0));
            for (int64_t t = 0; t < nt; ++t)
                for (int c = 0; c < 3; ++c)
                    for (int64_t k = 0; k < W; ++k) Cv[T[3 * t + c] * W + k] += C64(a[t] / 3.0) * C64(Cp[t * W + k]);
            std::vector<double> Br((size_t)nv * 2 * W), Mr((size_t)nv * 2 * W, 0.0);
            for (size_t k = 0; k < B.size(); ++k) {
                Br[2 * k] = Bf[k].real();
                Br[2 * k + 1] = Bf[k].imag();
            }
            H.MassMatrix().Dot<16>(1.0, Br.data(), int64_t(16), 0.0, Mr.data(), int64_t(16));
            for (int64_t v = 0; v < nv; ++v)
                for (int64_t k = 0; k < W; ++k)
                    Cv[v * W + k] += C64(coeff_f[4 * (k / wcs)]) * C64(Mr[v * 16 + 2 * k], Mr[v * 16 + 2 * k + 1]);
            check("AvT WriteC + c0 M B = the apply's result", rel(Cv, Cfd), 1e-6);
            check("B(i,k) = WriteB", std::abs(H.B(5, 3) - Bp[5 * W + 3]) + std::abs(H.B(nt - 1, W - 1) - Bp.back()), 0);
            check("C(i,k) = WriteC", std::abs(H.C(7, 2) - Cp[7 * W + 2]) + std::abs(H.C(0, 0) - Cp[0]), 0);
            // ReadB / ReadC replace the buffers (ld > wave_count on input)
            std::vector<C32> in((size_t)nt * 5);
            for (size_t k = 0; k < in.size(); ++k) in[k] = C32(0.5f * k, -1.0f * k);
            H.ReadB(in.data(), 5, 3);
            H.ReadC(in.data(), 5, 2);
            std::vector<C32> ob((size_t)nt * 4), oc((size_t)nt * 2);
            H.WriteB(ob.data(), 4);
            H.WriteC(oc.data(), 2);
            double d = 0;
            for (int64_t t = 0; t < nt; ++t) {
                for (int k = 0; k < 3; ++k) d += std::abs(ob[t * 4 + k] - in[t * 5 + k]);
                for (int k = 0; k < 2; ++k) d += std::abs(oc[t * 2 + k] - in[t * 5 + k]);
            }
            check("ReadB/ReadC then WriteB/WriteC round trip", d, 0);
            check("B(i,k) after ReadB", std::abs(H.B(9, 2) - in[9 * 5 + 2]), 0);
            H.B(9, 2) = C32(4.0f, 2.0f);
            check("B(i,k) is a writable reference", std::abs(H.B(9, 2) - C32(4.0f, 2.0f)), 0);
            // vertex and triangle coordinates, device information, flags
            double dv = 0;
            for (int64_t k = 0; k < 3 * nv; ++k) dv += std::abs(H.VertexCoordinates()[k] - (float)V[k]);
            for (int64_t t = 0; t < nt; ++t)
                for (int c = 0; c < 3; ++c) {
                    for (int e = 0; e < 3; ++e)
                        dv += std::abs(H.TriangleCoordinates()[12 * t + 4 * c + e] - (float)V[3 * T[3 * t + c] + e]);
                    dv += std::abs(H.TriangleCoordinates()[12 * t + 4 * c + 3]);
                }
            check("VertexCoordinates / TriangleCoordinates", dv, 0);
            const std::string info = H.DeviceInfo();
            std::printf("%s", info.c_str());
            check("DeviceInfo names the device", info.find("DEVICE_NAME") != std::string::npos ? 0.0 : 1.0, 0);
            H.PrintBooleans();
        }

        // ---- far field, potential, waves
        std::vector<C64> F((size_t)m * W), U(7 * W);
        H.ApplyFarFieldOperators_PL<0>(C64(1, 0), B.data(), W, C64(0, 0), F.data(), W, kappa, coeff, W, wcs);
        double fn = 0;
        for (auto& z : F) fn += std::norm(z);
        check("far field finite and nonzero", fn > 0 ? 0.0 : 1.0, 0);
        std::vector<double> P = {2, 0, 0, 0, 2.5, 0, 0, 0, -3, 1.5, 1.5, 1.5, -2, 1, 0, 0, -2, 1, 3, 3, 0};
        H.ApplyNearFieldOperators_PL(C64(1, 0), B.data(), W, C64(0, 0), U.data(), W, kappa, coeff, W, wcs, P.data(),
                                     int64_t(7));
        double un = 0;
        for (auto& z : U) un += std::norm(z);
        check("potential finite and nonzero", un > 0 ? 0.0 : 1.0, 0);
        const double dirs[3 * 4] = {0, 0, 1, 0.6, 0.8, 0, 0, -0.6, 0.8, -1, 0, 0};
        std::vector<C64> Winc((size_t)nv * W);
        H.CreateIncidentWave_PL(C64(1, 0), dirs, int64_t(4), C64(0, 0), Winc.data(), W, kappa, coeff, W, wcs,
                                BAEMM::WaveType::Plane);
        std::vector<C64> G((size_t)m * W), Hw((size_t)nv * W);
        for (size_t k = 0; k < G.size(); ++k) G[k] = C64(std::cos(0.2 * k), 0.1);
        H.CreateHerglotzWave_PL(C64(1, 0), G.data(), W, C64(0, 0), Hw.data(), W, kappa, coeff, W, wcs);
        double hn = 0;
        for (auto& z : Hw) hn += std::norm(z);
        check("Herglotz wave finite and nonzero", hn > 0 ? 0.0 : 1.0, 0);
        {
            std::vector<float> tmp((size_t)2 * nt * W);
            check("WriteC after a Herglotz wave: no value",
                  baemm_write_c(H.Handle(), tmp.data(), W) == BAEMM_ERR_STATE ? 0.0 : 1.0, 0);
        }

        // ---- mass matrix and its inverse
        std::vector<double> X((size_t)nv * 3), Y((size_t)nv * 3), Z((size_t)nv * 3, 0.0);
        for (size_t k = 0; k < X.size(); ++k) X[k] = std::sin(1.7 * k);
        H.ApplyMassInverse<3>(X.data(), 3, Y.data(), 3, 1e-12);
        H.MassMatrix().Dot<3>(1.0, Y.data(), int64_t(3), 0.0, Z.data(), int64_t(3));
        check("M (M^-1 X) = X, real", rel(Z, X), 1e-8);
        std::vector<C64> Xc((size_t)nv * 2), Yc((size_t)nv * 2);
        for (size_t k = 0; k < Xc.size(); ++k) Xc[k] = C64(std::sin(0.9 * k), std::cos(0.4 * k));
        H.ApplyMassInverse<2>(Xc.data(), 2, Yc.data(), 2, 1e-12);
        std::vector<double> Yr((size_t)nv * 4), Zr((size_t)nv * 4, 0.0), Xr((size_t)nv * 4);
        for (int64_t v = 0; v < nv; ++v)
            for (int c = 0; c < 2; ++c) {
                Yr[v * 4 + 2 * c] = Yc[v * 2 + c].real();
                Yr[v * 4 + 2 * c + 1] = Yc[v * 2 + c].imag();
                Xr[v * 4 + 2 * c] = Xc[v * 2 + c].real();
                Xr[v * 4 + 2 * c + 1] = Xc[v * 2 + c].imag();
            }
        H.MassMatrix().Dot<4>(1.0, Yr.data(), int64_t(4), 0.0, Zr.data(), int64_t(4));
        check("M (M^-1 X) = X, complex", rel(Zr, Xr), 1e-8);
        // lumped inverse: 1 / (AvT (1/3)), i.e. three over the row sums of the mass matrix
        std::vector<double> ones((size_t)nv, 1.0), rows((size_t)nv, 0.0);
        H.MassMatrix().Dot<1>(1.0, ones.data(), int64_t(1), 0.0, rows.data(), int64_t(1));
        std::vector<double> YL((size_t)nv * 3);
        H.ApplyLumpedMassInverse<3>(X.data(), 3, YL.data(), 3);
        std::vector<C64> YLc((size_t)nv * 2);
        H.ApplyLumpedMassInverse<2>(Xc.data(), 2, YLc.data(), 2);
        double dl = 0.0;
        for (int64_t v = 0; v < nv; ++v) {
            for (int c = 0; c < 3; ++c) dl = std::max(dl, std::abs(YL[v * 3 + c] * rows[v] / 3.0 - X[v * 3 + c]));
            for (int c = 0; c < 2; ++c) dl = std::max(dl, std::abs(YLc[v * 2 + c] * (rows[v] / 3.0) - Xc[v * 2 + c]));
        }
        check("ApplyLumpedMassInverse = 3 / mass row sums, real and complex", dl, 1e-12);

        // ---- solvers
        const double cg_tol = 1e-12, gmres_tol = 1e-9;
        H.gmres_max_iter = 60;
        H.gmres_max_restarts = 4;
        const int64_t Ks = 1, wcs_s = 4, Ws = 4;
        const double kap_s[1] = {2.5};
        double dirs_s[3 * 4] = {0, 0, 1, 0.6, 0.8, 0, 0, -0.6, 0.8, -1, 0, 0};
        std::vector<C64> FF((size_t)m * Ws), FF2((size_t)m * Ws);
        H.FarField<4>(kap_s, Ks, dirs_s, wcs_s, FF.data(), BAEMM::WaveType::Plane, cg_tol, gmres_tol);
        H.FarField_parameters<4>(kap_s, Ks, dirs_s, wcs_s, FF2.data(), BAEMM::WaveType::Plane, kap_s, cg_tol,
                                 gmres_tol);
        check("FarField vs FarField_parameters(eta = kappa)", rel(FF2, FF), 0);

        // BoundaryPotential: (1/2 M - i eta S + K) phi = wave
        std::vector<C64> coef_bp(4), wave((size_t)nv * Ws), phi((size_t)nv * Ws), back((size_t)nv * Ws);
        for (size_t k = 0; k < wave.size(); ++k) wave[k] = C64(std::cos(0.5 * k), std::sin(0.25 * k));
        H.BoundaryPotential<4>(kap_s, coef_bp.data(), wave.data(), phi.data(), Ks, wcs_s, cg_tol, gmres_tol);
        H.ApplyBoundaryOperators_PL<4>(C64(1, 0), phi.data(), Ws, C64(0, 0), back.data(), Ws, kap_s, coef_bp.data(), Ws,
                                       wcs_s);
        check("BoundaryPotential solves its equation", rel(back, wave), 1e-5);
        // DirichletToNeumann: (1/2 M - i kappa S + K') psi = wave
        std::vector<C64> psi((size_t)nv * Ws);
        H.DirichletToNeumann<4>(kap_s, wave.data(), psi.data(), Ks, wcs_s, cg_tol, gmres_tol);
        const C64 coef_dtn[4] = {{0.5, 0}, {0, -kap_s[0]}, {0, 0}, {1, 0}};
        H.ApplyBoundaryOperators_PL<4>(C64(1, 0), psi.data(), Ws, C64(0, 0), back.data(), Ws, kap_s, coef_dtn, Ws,
                                       wcs_s);
        check("DirichletToNeumann solves its equation", rel(back, wave), 1e-5);

        // Derivative_FF with du_dn computed inside (calloc) and given
        std::vector<double> h(V);
        std::vector<C64> DFh((size_t)m * Ws), DFh2((size_t)m * Ws), du((size_t)nv * Ws);
        C64* du_null = nullptr;
        H.Derivative_FF<4>(kap_s, Ks, dirs_s, wcs_s, h.data(), DFh.data(), du_null, BAEMM::WaveType::Plane, cg_tol,
                           gmres_tol);
        check("Derivative_FF allocates du_dn when NULL", du_null ? 0.0 : 1.0, 0);
        H.Derivative_FF_Helper<4>(kap_s, Ks, dirs_s, wcs_s, du.data(), BAEMM::WaveType::Plane, cg_tol, gmres_tol);
        std::vector<C64> du_alloc(du_null, du_null + nv * Ws);
        check("calloc'd du_dn equals Derivative_FF_Helper", rel(du_alloc, du), 0);
        C64* du_ptr = du.data();
        H.Derivative_FF<4>(kap_s, Ks, dirs_s, wcs_s, h.data(), DFh2.data(), du_ptr, BAEMM::WaveType::Plane, cg_tol,
                           gmres_tol);
        check("Derivative_FF with given du_dn", rel(DFh2, DFh), 1e-12);
        std::free(du_null);

        // GaussNewtonSolve: c X + Mass DF* DF X = B with M = c I, P = I / c
        const double c = 0.05;
        Scale Mop{c, 3 * nv}, Pop{1.0 / c, 3 * nv};
        std::vector<double> Bg((size_t)nv * 3), Xg((size_t)nv * 3, 0.0);
        for (size_t k = 0; k < Bg.size(); ++k) Bg[k] = std::cos(0.37 * k);
        C64* du_gn = du.data();
        const bool ok = H.GaussNewtonSolve<4>(kap_s, Ks, dirs_s, wcs_s, Mop, Pop, 1.0, Bg.data(), int64_t(3), 0.0,
                                              Xg.data(), int64_t(3), du_gn, BAEMM::WaveType::Plane, cg_tol, gmres_tol,
                                              1e-8);
        check("GaussNewtonSolve reports success", ok ? 0.0 : 1.0, 0);
        std::vector<C64> DFX((size_t)m * Ws);
        std::vector<double> Zg((size_t)nv * 3), MZ((size_t)nv * 3, 0.0);
        H.Derivative_FF<4>(kap_s, Ks, dirs_s, wcs_s, Xg.data(), DFX.data(), du_gn, BAEMM::WaveType::Plane, cg_tol,
                           gmres_tol);
        H.AdjointDerivative_FF<4>(kap_s, Ks, dirs_s, wcs_s, DFX.data(), Zg.data(), du_gn, BAEMM::WaveType::Plane,
                                  cg_tol, gmres_tol);
        H.MassMatrix().Dot<3>(1.0, Zg.data(), int64_t(3), 0.0, MZ.data(), int64_t(3));
        std::vector<double> res((size_t)nv * 3);
        for (size_t k = 0; k < res.size(); ++k) res[k] = c * Xg[k] + MZ[k];
        check("Gauss-Newton residual", rel(res, Bg), 1e-5);

        // The Gauss-Newton example's caller side (Example_GaussNewton/main.cpp:301-416)
        // with the tangent-point energy and metric in place of Repulsor's
        BAEMM::TangentPoint_HIP tp(H, 6.0, 12.0);
        const double E = tp.Value();
        check("tangent-point energy positive", (E > 0.0 && std::isfinite(E)) ? 0.0 : 1.0, 0);
        std::vector<double> D1((size_t)nv * 3), D2((size_t)nv * 3, 1.0), D3((size_t)nv * 3);
        tp.Differential(1.0, 0.0, D1.data(), int64_t(3));
        tp.Differential(2.0, 0.5, D2.data(), int64_t(3));
        for (size_t k = 0; k < D3.size(); ++k) D3[k] = 2.0 * D1[k] + 0.5;
        check("Differential alpha dE + beta B", rel(D2, D3), 1e-14);
        std::vector<double> u((size_t)nv * 3), w((size_t)nv * 3), Mu((size_t)nv * 3), Mw((size_t)nv * 3);
        for (size_t k = 0; k < u.size(); ++k) {
            u[k] = std::sin(0.91 * k + 0.3);
            w[k] = std::cos(1.37 * k);
        }
        tp.MultiplyMetric(1.0, u.data(), int64_t(3), 0.0, Mu.data(), int64_t(3), int64_t(3));
        tp.MultiplyMetric(1.0, w.data(), int64_t(3), 0.0, Mw.data(), int64_t(3), int64_t(3));
        double uMw = 0, wMu = 0, uMu = 0, wMw = 0;
        for (size_t k = 0; k < u.size(); ++k) {
            uMw += u[k] * Mw[k];
            wMu += w[k] * Mu[k];
            uMu += u[k] * Mu[k];
            wMw += w[k] * Mw[k];
        }
        check("tangent-point metric symmetric", std::abs(uMw - wMu) / std::sqrt(uMu * wMw), 1e-8);
        std::vector<double> P1((size_t)nv * 3), P2((size_t)nv * 3, 3.0), P3((size_t)nv * 3);
        tp.MultiplyPreconditioner(1.0, u.data(), int64_t(3), 0.0, P1.data(), int64_t(3), int64_t(3));
        tp.MultiplyPreconditioner(2.0, u.data(), int64_t(3), -1.0, P2.data(), int64_t(3), int64_t(3));
        for (size_t k = 0; k < P3.size(); ++k) P3[k] = 2.0 * P1[k] - 3.0;
        check("MultiplyPreconditioner alpha P X + beta Y", rel(P2, P3), 1e-14);
        // regpar so that both terms of the operator count on smooth fields: a third of
        // the data term's gain on a smooth deformation s of the surface
        std::vector<double> sm((size_t)nv * 3), Ms((size_t)nv * 3);
        for (int64_t v = 0; v < nv; ++v) {
            sm[3 * v] = V[3 * v] * V[3 * v + 1];
            sm[3 * v + 1] = V[3 * v + 2] * V[3 * v + 2];
            sm[3 * v + 2] = std::sin(V[3 * v]);
        }
        tp.MultiplyMetric(1.0, sm.data(), int64_t(3), 0.0, Ms.data(), int64_t(3), int64_t(3));
        std::vector<C64> DFu((size_t)m * Ws);
        std::vector<double> Zu((size_t)nv * 3), MZu((size_t)nv * 3, 0.0);
        H.Derivative_FF<4>(kap_s, Ks, dirs_s, wcs_s, sm.data(), DFu.data(), du_gn, BAEMM::WaveType::Plane, cg_tol,
                           gmres_tol);
        H.AdjointDerivative_FF<4>(kap_s, Ks, dirs_s, wcs_s, DFu.data(), Zu.data(), du_gn, BAEMM::WaveType::Plane,
                                  cg_tol, gmres_tol);
        H.MassMatrix().Dot<3>(1.0, Zu.data(), int64_t(3), 0.0, MZu.data(), int64_t(3));
        double nd = 0, nm = 0;
        for (size_t k = 0; k < sm.size(); ++k) {
            nd += MZu[k] * MZu[k];
            nm += Ms[k] * Ms[k];
        }
        const double regpar = 0.3 * std::sqrt(nd / nm);
        auto Atp = [&](const double* X, double* Y) {
            tp.MultiplyMetric(regpar, X, int64_t(3), 0.0, Y, int64_t(3), int64_t(3));
        };
        auto Ptp = [&](const double* X, double* Y) {
            tp.MultiplyPreconditioner(1.0 / regpar, X, int64_t(3), 0.0, Y, int64_t(3), int64_t(3));
        };
        H.gmres_max_iter = 100;
        H.gmres_max_restarts = 8;
        std::vector<double> Bt((size_t)nv * 3, 0.0), Xt((size_t)nv * 3, 0.0);
        tp.Differential(regpar, 0.0, Bt.data(), int64_t(3));
        for (size_t k = 0; k < Bt.size(); ++k) Bt[k] += 0.1 * u[k];
        const bool ok_tp = H.GaussNewtonSolve<4>(kap_s, Ks, dirs_s, wcs_s, Atp, Ptp, 1.0, Bt.data(), int64_t(3), 0.0,
                                                 Xt.data(), int64_t(3), du_gn, BAEMM::WaveType::Plane, cg_tol,
                                                 gmres_tol, 1e-8);
        check("GaussNewtonSolve with the tangent-point metric reports success", ok_tp ? 0.0 : 1.0, 0);
        std::vector<double> Zt((size_t)nv * 3), MZt((size_t)nv * 3, 0.0), AX((size_t)nv * 3);
        std::vector<C64> DFXt((size_t)m * Ws);
        H.Derivative_FF<4>(kap_s, Ks, dirs_s, wcs_s, Xt.data(), DFXt.data(), du_gn, BAEMM::WaveType::Plane, cg_tol,
                           gmres_tol);
        H.AdjointDerivative_FF<4>(kap_s, Ks, dirs_s, wcs_s, DFXt.data(), Zt.data(), du_gn, BAEMM::WaveType::Plane,
                                  cg_tol, gmres_tol);
        H.MassMatrix().Dot<3>(1.0, Zt.data(), int64_t(3), 0.0, MZt.data(), int64_t(3));
        Atp(Xt.data(), AX.data());
        for (size_t k = 0; k < AX.size(); ++k) AX[k] += MZt[k];
        check("Gauss-Newton residual with the tangent-point metric", rel(AX, Bt), 1e-5);
    } catch (const std::exception& e) {
        std::printf("exception: %s\n", e.what());
        return 2;
    }
    std::printf("%s: %d failed\n", failures ? "FAIL" : "OK", failures);
    return failures ? 1 : 0;
}
