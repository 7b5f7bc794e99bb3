// The reference's ASCII VTU writer (save, post.h:58-180) as a parallel
// formatter.  The reference appends every number to one std::string on one
// thread; here each section is cut into cell (or node) ranges that threads
// format into their own buffers, which are then written in order, so the file
// is the same byte for byte.  Number formats are the reference's:
//   std::to_string(double)                       -> "%f"   (node coordinates)
//   std::fixed << std::setprecision(16) << x     -> "%.16f" (cell data)
//   std::to_string(unsigned)                     -> "%u"   (indices)
#include "afx_post.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <functional>
#include <thread>
#include <vector>

namespace afx {
namespace {

inline void put_f(std::string& s, const char* fmt, double x) {
    char buf[512];   // "%.16f" of |x| up to 1e308 needs 326 characters
    const int n = std::snprintf(buf, sizeof buf, fmt, x);
    s.append(buf, (size_t)std::max(n, 0));
}
inline void put_u(std::string& s, unsigned long long v) {
    char buf[32];
    const int n = std::snprintf(buf, sizeof buf, "%llu", v);
    s.append(buf, (size_t)n);
}

// solution accessors of core.h:132-163 on q(4*i+k)
struct Sol {
    const double* q;
    double gamma, R;
    double rho(size_t i) const { return q[4 * i]; }
    double rhou(size_t i) const { return q[4 * i + 1]; }
    double rhov(size_t i) const { return q[4 * i + 2]; }
    double rhoe(size_t i) const { return q[4 * i + 3]; }
    double p(size_t i) const { return (gamma - 1) * (rhoe(i) - 0.5 / rho(i) * (rhou(i) * rhou(i) + rhov(i) * rhov(i))); }
    double u(size_t i) const { return rhou(i) / rho(i); }
    double v(size_t i) const { return rhov(i) / rho(i); }
    double T(size_t i) const { return p(i) / (R * rho(i)); }
    double c(size_t i) const { return std::sqrt(gamma * p(i) / rho(i)); }
    double mach(size_t i) const { return std::sqrt(rhou(i) * rhou(i) + rhov(i) * rhov(i)) / (rho(i) * c(i)); }
};

// formats items [0, n) with fn(i, out) on up to `threads` threads, in order
void format_ranges(size_t n, unsigned threads, const std::function<void(size_t, std::string&)>& fn,
                   std::vector<std::string>& parts) {
    const size_t nt = std::max<size_t>(1, std::min<size_t>(threads, (n + 4095) / 4096));
    parts.assign(nt, std::string());
    std::vector<std::thread> pool;
    for (size_t t = 0; t < nt; ++t)
        pool.emplace_back([&, t] {
            const size_t a = n * t / nt, b = n * (t + 1) / nt;
            std::string& s = parts[t];
            s.reserve((b - a) * 24);
            for (size_t i = a; i < b; ++i) fn(i, s);
        });
    for (auto& th : pool) th.join();
}

}  // namespace

std::string write_vtu(const std::string& path, const Mesh& m, const double* q, const double* wall_dist, double gamma,
                      double R, unsigned threads) {
    if (!threads) threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::ofstream out(path, std::ios::binary);
    if (!out) return "cannot open " + path + " for writing";
    const size_t nn = m.nodesX.size(), nc = m.nRealCells;
    const Sol S{q, gamma, R};
    std::vector<std::string> parts;
    auto emit = [&](size_t n, const std::function<void(size_t, std::string&)>& fn) {
        format_ranges(n, threads, fn, parts);
        for (const std::string& s : parts) out.write(s.data(), (std::streamsize)s.size());
    };
    auto text = [&](const std::string& s) { out.write(s.data(), (std::streamsize)s.size()); };

    text("<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"BigEndian\">\n");
    text("  <UnstructuredGrid>\n");
    text("    <Piece NumberOfPoints=\"" + std::to_string(nn) + "\" NumberOfCells=\"" + std::to_string(nc) + "\">\n");
    text("      <Points>\n");
    text("        <DataArray type=\"Float32\" NumberOfComponents=\"3\" Format=\"ascii\">\n");
    emit(nn, [&](size_t i, std::string& s) {
        s += "          ";
        put_f(s, "%f", m.nodesX[i]);
        s += ' ';
        put_f(s, "%f", m.nodesY[i]);
        s += " 0.0\n";
    });
    text("        </DataArray>\n");
    text("      </Points>\n");
    text("      <Cells>\n");
    text("        <DataArray type=\"Int32\" Name=\"connectivity\" Format=\"ascii\">\n");
    text("          ");
    emit(nc, [&](size_t i, std::string& s) {
        const int nv = m.cellsIsTriangle[i] ? 3 : 4;
        for (int k = 0; k < nv; ++k) {
            put_u(s, m.cellsNodes[4 * i + k]);
            s += "  ";
        }
        s += "\n          ";
    });
    text("        </DataArray>\n");
    text("        <DataArray type=\"Int32\" Name=\"offsets\" Format=\"ascii\">\n");
    text("          ");
    {
        // running offsets: each range starts from the prefix count of its cells
        std::vector<unsigned long long> off(nc);
        unsigned long long cur = 0;
        for (size_t i = 0; i < nc; ++i) off[i] = (cur += m.cellsIsTriangle[i] ? 3 : 4);
        emit(nc, [&](size_t i, std::string& s) {
            put_u(s, off[i]);
            s += "  ";
        });
    }
    text("\n");
    text("        </DataArray>\n");
    text("        <DataArray type=\"Int32\" Name=\"types\" Format=\"ascii\">\n");
    text("          ");
    emit(nc, [&](size_t i, std::string& s) { s += m.cellsIsTriangle[i] ? "5  " : "9  "; });
    text("\n");
    text("        </DataArray>\n");
    text("      </Cells>\n");
    text("      <CellData Scalars=\"scalars\">\n");
    auto scalar = [&](const char* name, const std::function<double(size_t)>& f) {
        text(std::string("        <DataArray type=\"Float32\" Name=\"") + name + "\" Format=\"ascii\">\n");
        emit(nc, [&](size_t i, std::string& s) {
            s += "          ";
            put_f(s, "%.16f", f(i));
            s += '\n';
        });
        text("        </DataArray>\n");
    };
    scalar("Wall Distance", [&](size_t i) { return wall_dist[i]; });
    scalar("Mach", [&](size_t i) { return S.mach(i); });
    scalar("Density", [&](size_t i) { return S.rho(i); });
    scalar("Pressure", [&](size_t i) { return S.p(i); });
    scalar("Temperature", [&](size_t i) { return S.T(i); });
    text("        <DataArray type=\"Float32\" Name=\"Velocity\" NumberOfComponents=\"3\" Format=\"ascii\">\n");
    emit(nc, [&](size_t i, std::string& s) {
        s += "          ";
        put_f(s, "%.16f", S.u(i));
        s += ' ';
        put_f(s, "%.16f", S.v(i));
        s += " 0.0\n";
    });
    text("        </DataArray>\n");
    text("      </CellData>\n");
    text("    </Piece>\n");
    text("  </UnstructuredGrid>\n");
    text("</VTKFile>");
    out.close();
    if (!out) return "write to " + path + " failed";
    return std::string();
}

}  // namespace afx
