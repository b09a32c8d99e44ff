// Output pipeline. See outputs.hpp.
#include "outputs.hpp"

#include <cmath>
#include <cstdio>
#include <map>
#include <sstream>

namespace dg {

std::string fmt6(double v) {
    std::ostringstream os;
    os << v;
    return os.str();
}

bool writeVtu(const std::string& path, const MshMesh& mesh, const double* state, double c0) {
    std::ofstream out(path);
    if (!out) return false;
    const int64_t K = mesh.numElements();
    const int Np = mesh.nodesPerElement, dim = mesh.dim, nv = dim + 1;
    const int64_t N = K * Np;
    std::map<int64_t, int64_t> index;
    int64_t i = 0;
    for (const auto& kv : mesh.nodes) index[kv.first] = i++;
    out << "<?xml version=\"1.0\"?>\n";
    out << "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
    out << "  <UnstructuredGrid>\n";
    out << "    <Piece NumberOfPoints=\"" << mesh.nodes.size() << "\" NumberOfCells=\"" << K << "\">\n";
    out << "      <Points>\n        <DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n";
    out.precision(17);
    for (const auto& kv : mesh.nodes) out << kv.second[0] << " " << kv.second[1] << " " << kv.second[2] << "\n";
    out << "        </DataArray>\n      </Points>\n";
    out << "      <Cells>\n        <DataArray type=\"Int64\" Name=\"connectivity\" format=\"ascii\">\n";
    for (int64_t e = 0; e < K; ++e) {
        for (int v = 0; v < nv; ++v) out << index[mesh.elementNodes[e * Np + v]] << (v + 1 < nv ? " " : "\n");
    }
    out << "        </DataArray>\n        <DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n";
    for (int64_t e = 0; e < K; ++e) out << (e + 1) * nv << "\n";
    out << "        </DataArray>\n        <DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n";
    for (int64_t e = 0; e < K; ++e) out << (dim == 3 ? 10 : 5) << "\n";
    out << "        </DataArray>\n      </Cells>\n      <CellData>\n";
    std::vector<double> p(K, 0.0), rho(K, 0.0), v(3 * K, 0.0);
    for (int64_t e = 0; e < K; ++e) {
        double sp = 0, sr = 0, sx = 0, sy = 0, sz = 0;
        for (int a = 0; a < Np; ++a) {
            const int64_t n = e * Np + a;
            sp += state[n];
            sr += state[n] / (c0 * c0);
            sx += state[N + n];
            sy += state[2 * N + n];
            sz += state[3 * N + n];
        }
        p[e] = sp / Np;
        rho[e] = sr / Np;
        v[3 * e] = sx / Np;
        v[3 * e + 1] = sy / Np;
        v[3 * e + 2] = sz / Np;
    }
    auto arr = [&](const char* name, int comps, const std::vector<double>& d) {
        out << "        <DataArray type=\"Float64\" Name=\"" << name << "\" NumberOfComponents=\"" << comps
            << "\" format=\"ascii\">\n";
        for (size_t k = 0; k < d.size(); ++k) out << d[k] << ((k + 1) % comps == 0 ? "\n" : " ");
        out << "        </DataArray>\n";
    };
    arr("Pressure [Pa]", 1, p);
    arr("Density [kg/m³]", 1, rho);
    arr("Velocity [m/s]", 3, v);
    out << "      </CellData>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n";
    return static_cast<bool>(out);
}

bool writePvd(const std::string& path, double tStart, double tEnd, double dt, double tRate) {
    std::ofstream file(path);
    if (!file) return false;
    file << "<VTKFile type=\"Collection\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n";
    file << "  <Collection>\n";
    for (double t = tStart, step = 0, tDisplay = 0; t <= tEnd; t += dt, tDisplay += dt, ++step) {
        if (tDisplay >= tRate || step == 0) {
            tDisplay = 0;
            file << "    <DataSet timestep=\"" << fmt6(t) << "\" part=\"0\" file=\"results/result"
                 << static_cast<int>(step) << ".vtu\"/>\n";
        }
    }
    file << "  </Collection>\n</VTKFile>\n";
    return static_cast<bool>(file);
}

void fftRadix2(std::vector<double>& re, std::vector<double>& im) {
    const size_t n = re.size();
    for (size_t i = 1, j = 0; i < n; ++i) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            std::swap(re[i], re[j]);
            std::swap(im[i], im[j]);
        }
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        const double ang = -2.0 * M_PI / static_cast<double>(len);
        for (size_t i = 0; i < n; i += len)
            for (size_t k = 0; k < len / 2; ++k) {
                const double wr = std::cos(ang * k), wi = std::sin(ang * k);
                const size_t a = i + k, b = i + k + len / 2;
                const double xr = re[b] * wr - im[b] * wi, xi = re[b] * wi + im[b] * wr;
                re[b] = re[a] - xr;
                im[b] = im[a] - xi;
                re[a] += xr;
                im[a] += xi;
            }
    }
}

bool writeSpectra(const std::vector<float>& series, double dt, const std::string& basename) {
    if (series.size() < 2) return false;
    const size_t nb = static_cast<size_t>(std::pow(2.0, std::floor(std::log(static_cast<double>(series.size())) /
                                                                   std::log(2.0))));
    std::vector<double> t(nb), x(nb);
    double mean = 0.0;
    for (size_t i = 0; i < nb; ++i) {
        t[i] = i * dt;
        x[i] = series[i];
        mean += x[i];
    }
    mean /= static_cast<double>(nb);
    for (auto& v : x) v -= mean;
    std::vector<double> re = x, im(nb, 0.0);
    fftRadix2(re, im);
    {
        std::ofstream f(basename + "_fft.txt");
        if (!f) return false;
        const double periods = t[nb - 1] - t[0];
        f << "Frequency Norm Real Imag SPL\n";
        for (size_t i = 1; i < nb / 2; ++i) {
            const double freq = i / periods;
            const double norm = 2.0 * std::sqrt(re[i] * re[i] + im[i] * im[i]) / nb;
            const double spl = 20.0 * std::log10(norm / 2.0e-5);
            f << fmt6(freq) << " " << fmt6(norm) << " " << fmt6(re[i]) << " " << fmt6(im[i]) << " " << fmt6(spl)
              << " \n";
        }
    }
    {
        std::ofstream f(basename + "_psd.txt");
        if (!f) return false;
        const double sampleRate = 1.0 / (t[nb - 1] - t[0]);
        const double norm = 1.0 / (static_cast<double>(nb) * nb) / sampleRate;
        std::vector<double> psd(nb / 2 + 1);
        for (size_t i = 1; i < nb / 2 + 1; ++i) psd[i] = 2.0 * (re[i] * re[i] + im[i] * im[i]) * norm;
        if (nb % 2 == 0) psd[nb / 2] *= 0.5;
        f << "Frequency Norm PSD\n";
        for (size_t i = 1; i < nb / 2 + 1; ++i)
            f << fmt6(i * sampleRate) << " " << fmt6(psd[i]) << " " << fmt6(10.0 * std::log10(psd[i] / (2.0e-5 * 2.0e-5)))
              << " \n";
    }
    return true;
}

}  // namespace dg
