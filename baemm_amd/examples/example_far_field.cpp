// Far field, its shape derivative and the adjoint derivative of a sound-soft
// obstacle through the C++ drop-in class -- the workflow of the reference's
// Example_OpenCL/main.cpp:25-249 (FarField, Derivative_FF with h = vertex
// coordinates, AdjointDerivative_FF of DF h), on the HIP library.
//
// Usage: example_far_field [mesh.txt [meas.txt]] [--level L] [--waves W] [--out DIR]
//   mesh.txt / meas.txt are in the reference's text mesh format
//   (ReadMeshFromFile.hpp:30-97).  Without files an icosphere of the given
//   subdivision level (default 5, 20480 triangles) is the obstacle and the
//   vertices of a level-4 icosphere (2562 points) are the measurement
//   directions.  Prints the L2 norm of the far field, the norms of DF h and
//   DF* DF h, and the wall time of each call.  With --out, writes the three
//   results as the reference example does (Example_OpenCL/main.cpp:191,220,233):
//   DIR/F_New.txt (far field, meas_count x W), DIR/DF_New.txt (DF h) and
//   DIR/DFT_New.txt (DF* DF h, vertex_count x 3).  One row per line; a complex
//   entry is written as its real and imaginary parts, so a row of W complex
//   values holds 2 W numbers (the reference prints with its Tools library's
//   ToString, which is not part of its repository).
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "Helmholtz_HIP.hpp"

using Complex = std::complex<double>;

namespace {

bool read_mesh(const std::string& path, std::vector<double>& V, std::vector<int64_t>& T)
{
    std::ifstream f(path);
    if (!f.good()) return false;
    std::string key;
    int64_t dom = 0, amb = 0, nv = 0, nt = 0;
    f >> key >> dom >> key >> amb >> key >> nv >> key >> nt;
    if (!f.good() || dom != 2 || amb != 3) return false;
    V.resize(3 * nv);
    T.resize(3 * nt);
    for (auto& v : V) f >> v;
    for (auto& t : T) f >> t;
    return !f.fail();
}

// Unit icosphere by repeated 4-way subdivision with projection to the sphere.
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
        S.reserve(4 * T.size());
        for (size_t t = 0; t < T.size(); t += 3) {
            const int64_t a = T[t], b = T[t + 1], c = T[t + 2];
            const int64_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            const int64_t n[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
            S.insert(S.end(), n, n + 12);
        }
        T.swap(S);
    }
}

double fro(const std::vector<Complex>& a)
{
    double s = 0;
    for (const auto& z : a) s += std::norm(z);
    return std::sqrt(s);
}

void write_complex(const std::string& path, const Complex* a, int64_t rows, int64_t cols)
{
    std::FILE* f = std::fopen(path.c_str(), "w");
    if (!f) throw std::runtime_error("cannot write " + path);
    for (int64_t i = 0; i < rows; ++i) {
        for (int64_t k = 0; k < cols; ++k)
            std::fprintf(f, "%s%.17g %.17g", k ? " " : "", a[i * cols + k].real(), a[i * cols + k].imag());
        std::fputc('\n', f);
    }
    std::fclose(f);
}

void write_real(const std::string& path, const double* a, int64_t rows, int64_t cols)
{
    std::FILE* f = std::fopen(path.c_str(), "w");
    if (!f) throw std::runtime_error("cannot write " + path);
    for (int64_t i = 0; i < rows; ++i) {
        for (int64_t k = 0; k < cols; ++k) std::fprintf(f, "%s%.17g", k ? " " : "", a[i * cols + k]);
        std::fputc('\n', f);
    }
    std::fclose(f);
}

double seconds_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

int run(int argc, char** argv)
{
    std::vector<std::string> files;
    std::string out_dir;
    int level = 5;
    int wave_count = 16;
    for (int i = 1; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--level") && i + 1 < argc)
            level = std::atoi(argv[++i]);
        else if (!std::strcmp(argv[i], "--waves") && i + 1 < argc)
            wave_count = std::atoi(argv[++i]);
        else if (!std::strcmp(argv[i], "--out") && i + 1 < argc)
            out_dir = argv[++i];
        else
            files.push_back(argv[i]);
    }
    constexpr int wave_chunk_size = 16;
    if (wave_count % wave_chunk_size != 0 || wave_count <= 0) {
        std::fprintf(stderr, "--waves must be a positive multiple of %d\n", wave_chunk_size);
        return 2;
    }
    const int wave_chunk_count = wave_count / wave_chunk_size;

    std::vector<double> V, M;
    std::vector<int64_t> T, TM;
    if (!files.empty()) {
        if (!read_mesh(files[0], V, T)) {
            std::fprintf(stderr, "File %s not found or invalid. Exiting.\n", files[0].c_str());
            return 1;
        }
    } else {
        icosphere(level, V, T);
    }
    if (files.size() > 1) {
        if (!read_mesh(files[1], M, TM)) {
            std::fprintf(stderr, "File %s not found or invalid. Exiting.\n", files[1].c_str());
            return 1;
        }
    } else {
        icosphere(4, M, TM);
    }
    const int64_t nv = (int64_t)V.size() / 3, nt = (int64_t)T.size() / 3, m = (int64_t)M.size() / 3;
    std::printf("mesh: %lld vertices, %lld triangles; %lld measurement directions; %d waves\n", (long long)nv,
                (long long)nt, (long long)m, wave_count);

    BAEMM::Helmholtz_OpenCL H(V.data(), nv, T.data(), nt, M.data(), m, int64_t(0), int64_t(8));

    // Wave numbers and incident directions of Example_OpenCL/main.cpp:97-168.
    std::vector<double> kappa(wave_chunk_count);
    for (int i = 0; i < wave_chunk_count; ++i) kappa[i] = (1 + 2 * (i + 2)) * M_PI;
    const double inc[16][3] = {{0.22663516023574246, 0.4654289185373517, 0.8555772472045235},
                               {-0.3474510994575331, -0.7864451523456505, -0.5106679506663584},
                               {-0.9541836471144971, -0.27451273343735705, -0.11906438073591856},
                               {0.43775029208027166, -0.8876694056416434, 0.1428905457734996},
                               {-0.44907802104219113, -0.8139693295211834, 0.36848726113078095},
                               {0.47666884531973414, 0.3192753380458682, -0.8190543757390276},
                               {-0.02194389227182042, 0.8818743114784301, -0.47097363444932655},
                               {0.786615210967289, -0.1684579992437276, 0.5940188653281036},
                               {0.42244725439759406, -0.5580931680624308, -0.7141920841160128},
                               {0.6792359217126083, 0.7269051927489542, 0.10122945919953094},
                               {-0.3904321861768485, 0.03454064212408624, -0.919983506394991},
                               {-0.3009524704632311, 0.8792870632871999, 0.36916374531886265},
                               {-0.006080463353741372, -0.38604914872328944, 0.9224581739761577},
                               {0.9742753958166932, -0.0853848807263007, -0.20855904499585115},
                               {-0.8068074072061289, 0.5272347230598002, -0.2666183686046023},
                               {-0.7266826985737691, 0.12599246204091846, 0.675320779409617}};
    std::vector<double> dirs(&inc[0][0], &inc[0][0] + 3 * wave_chunk_size);

    H.UseDiagonal(true);
    const double cg_tol = 1e-5, gmres_tol = 1e-4;

    std::vector<Complex> F((size_t)m * wave_count);
    auto t0 = std::chrono::steady_clock::now();
    H.FarField<0>(kappa.data(), wave_chunk_count, dirs.data(), wave_chunk_size, F.data(), BAEMM::WaveType::Plane,
                  cg_tol, gmres_tol);
    std::printf("FarField: %.3f s\n", seconds_since(t0));
    const double factor = 2.0 * M_PI / double(m * wave_count);
    std::printf("L^2-norm of C = %.10e\n", std::sqrt(factor) * fro(F));
    if (!out_dir.empty()) write_complex(out_dir + "/F_New.txt", F.data(), m, wave_count);

    std::vector<double> h(V), w((size_t)3 * nv);
    std::vector<Complex> DFh((size_t)m * wave_count);
    Complex* du_dn = nullptr;
    t0 = std::chrono::steady_clock::now();
    H.Derivative_FF<0>(kappa.data(), wave_chunk_count, dirs.data(), wave_chunk_size, h.data(), DFh.data(), du_dn,
                       BAEMM::WaveType::Plane, cg_tol, gmres_tol);
    std::printf("Derivative_FF: %.3f s\n", seconds_since(t0));
    std::printf("|DF h| = %.10e\n", fro(DFh));
    if (!out_dir.empty()) write_complex(out_dir + "/DF_New.txt", DFh.data(), m, wave_count);

    t0 = std::chrono::steady_clock::now();
    H.AdjointDerivative_FF<0>(kappa.data(), wave_chunk_count, dirs.data(), wave_chunk_size, DFh.data(), w.data(),
                              du_dn, BAEMM::WaveType::Plane, cg_tol, gmres_tol);
    std::printf("AdjointDerivative_FF: %.3f s\n", seconds_since(t0));
    double s = 0;
    for (double x : w) s += x * x;
    std::printf("|DF* DF h| = %.10e\n", std::sqrt(s));
    if (!out_dir.empty()) write_real(out_dir + "/DFT_New.txt", w.data(), nv, 3);

    if (du_dn != nullptr) std::free(du_dn);
    return 0;
}

int main(int argc, char** argv)
{
    try {
        return run(argc, argv);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
}
