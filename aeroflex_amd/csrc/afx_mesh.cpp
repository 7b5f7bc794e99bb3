// Host mesh: MSH 4.1 reader/writer, reference-order face construction,
// benchmark mesh generators and RCM ordering.
#include "afx_mesh.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <unordered_map>

namespace afx {

int Mesh::name_index(uint32_t key) const {
    for (size_t i = 0; i < name_keys.size(); ++i)
        if (name_keys[i] == key) return (int)i;
    return -1;
}

double SplitMix64::normal() {
    // Box-Muller; u1 in (0,1]
    double u1 = 1.0 - uniform();
    double u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

// ----------------------------------------------------------------- reader --
namespace {

struct LineReader {
    std::ifstream in;
    std::string line;
    long lineno = 0;
    explicit LineReader(const std::string& p) : in(p) {
        if (!in) throw std::runtime_error("cannot open mesh file: " + p);
    }
    bool next() {
        if (!std::getline(in, line)) return false;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        ++lineno;
        return true;
    }
    void need() {
        if (!next()) throw std::runtime_error("unexpected end of mesh file");
    }
};

// whitespace tokenizer into doubles / longs
static void tok_longs(const std::string& s, std::vector<long>& out) {
    out.clear();
    const char* p = s.c_str();
    char* end;
    while (*p) {
        while (*p == ' ' || *p == '\t') ++p;
        if (!*p) break;
        long v = std::strtol(p, &end, 10);
        if (end == p) throw std::runtime_error("bad integer in mesh file: " + s);
        out.push_back(v);
        // skip rest of token (e.g. fractional part of a float read as int)
        p = end;
        while (*p && *p != ' ' && *p != '\t') ++p;
    }
}
static void tok_doubles(const std::string& s, std::vector<double>& out) {
    out.clear();
    const char* p = s.c_str();
    char* end;
    while (*p) {
        while (*p == ' ' || *p == '\t') ++p;
        if (!*p) break;
        double v = std::strtod(p, &end);
        if (end == p) throw std::runtime_error("bad number in mesh file: " + s);
        out.push_back(v);
        p = end;
    }
}

}  // namespace

Elements read_msh41(const std::string& path) {
    Elements el;
    LineReader r(path);
    std::unordered_map<long, uint32_t> curve_phys;      // curve entity tag -> key
    std::unordered_map<long, uint32_t> node_index;      // node tag -> index
    std::vector<long> li;
    std::vector<double> ld;
    bool have_nodes = false;
    while (r.next()) {
        if (r.line.empty() || r.line[0] != '$') continue;
        std::string sec = r.line.substr(1);
        if (sec == "MeshFormat") {
            r.need();
            if (r.line.compare(0, 3, "4.1") != 0) throw std::runtime_error("only MSH 4.1 ASCII is supported");
        } else if (sec == "PhysicalNames") {
            r.need();
            long n = std::strtol(r.line.c_str(), nullptr, 10);
            for (long k = 0; k < n; ++k) {
                r.need();
                std::istringstream ss(r.line);
                long dim, tag;
                std::string nm;
                ss >> dim >> tag;
                std::getline(ss, nm);
                nm.erase(std::remove(nm.begin(), nm.end(), '"'), nm.end());
                nm.erase(0, nm.find_first_not_of(' '));
                while (!nm.empty() && nm.back() == ' ') nm.pop_back();
                el.name_keys.push_back((uint32_t)(tag + 1000 * dim));
                el.names.push_back(nm);
            }
        } else if (sec == "Entities") {
            r.need();
            tok_longs(r.line, li);
            long np = li[0], nc = li[1], ns = li[2], nv = li[3];
            for (long k = 0; k < np; ++k) r.need();
            for (long k = 0; k < nc; ++k) {
                r.need();
                tok_doubles(r.line, ld);   // tag, bbox(6), nphys, phys..., nbnd, ...
                long tag = (long)ld[0];
                long nphys = (long)ld[7];
                if (nphys >= 1) curve_phys[tag] = (uint32_t)((long)ld[8] + 1000);
            }
            for (long k = 0; k < ns + nv; ++k) r.need();
        } else if (sec == "Nodes") {
            r.need();
            tok_longs(r.line, li);
            long nblocks = li[0];
            el.x.reserve(li[1]); el.y.reserve(li[1]);
            std::vector<long> tags;
            for (long b = 0; b < nblocks; ++b) {
                r.need();
                tok_longs(r.line, li);
                long n = li[3];
                tags.resize(n);
                for (long k = 0; k < n; ++k) { r.need(); tags[k] = std::strtol(r.line.c_str(), nullptr, 10); }
                for (long k = 0; k < n; ++k) {
                    r.need();
                    tok_doubles(r.line, ld);
                    node_index[tags[k]] = (uint32_t)el.x.size();
                    el.x.push_back(ld[0]); el.y.push_back(ld[1]);
                }
            }
            have_nodes = true;
        } else if (sec == "Elements") {
            if (!have_nodes) throw std::runtime_error("$Elements before $Nodes");
            r.need();
            tok_longs(r.line, li);
            long nblocks = li[0];
            for (long b = 0; b < nblocks; ++b) {
                r.need();
                tok_longs(r.line, li);
                long dim = li[0], etag = li[1], n = li[3];
                uint32_t key = 0;
                if (dim == 1) {
                    auto it = curve_phys.find(etag);
                    if (it == curve_phys.end()) throw std::runtime_error("boundary curve without physical tag");
                    key = it->second;
                }
                for (long k = 0; k < n; ++k) {
                    r.need();
                    tok_longs(r.line, li);
                    size_t nn = li.size() - 1;
                    auto map = [&](long t) {
                        auto it = node_index.find(t);
                        if (it == node_index.end()) throw std::runtime_error("element references unknown node");
                        return it->second;
                    };
                    if (dim == 1 && nn == 2) {
                        el.bseg_nodes.push_back(map(li[1]));
                        el.bseg_nodes.push_back(map(li[2]));
                        el.bseg_key.push_back(key);
                    } else if (dim == 2 && nn != 2) {
                        if (nn != 3 && nn != 4) throw std::runtime_error("only linear triangles and quads are supported");
                        for (size_t v = 0; v < 4; ++v) el.cell_nodes.push_back(v < nn ? map(li[1 + v]) : 0u);
                        el.cell_nv.push_back((uint8_t)nn);
                    }
                }
            }
        }
    }
    return el;
}

void write_msh41(const Elements& el, const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) throw std::runtime_error("cannot write " + path);
    // one curve entity per boundary physical name (dim 1), one surface (dim 2)
    std::vector<size_t> curve_names, surf_names;
    for (size_t i = 0; i < el.name_keys.size(); ++i)
        (el.name_keys[i] / 1000 == 1 ? curve_names : surf_names).push_back(i);
    std::fprintf(f, "$MeshFormat\n4.1 0 8\n$EndMeshFormat\n$PhysicalNames\n%zu\n", el.name_keys.size());
    for (size_t i = 0; i < el.name_keys.size(); ++i)
        std::fprintf(f, "%u %u \"%s\"\n", el.name_keys[i] / 1000, el.name_keys[i] % 1000, el.names[i].c_str());
    std::fprintf(f, "$EndPhysicalNames\n$Entities\n0 %zu 1 0\n", curve_names.size());
    for (size_t k = 0; k < curve_names.size(); ++k)
        std::fprintf(f, "%zu 0 0 0 0 0 0 1 %u 0\n", k + 1, el.name_keys[curve_names[k]] % 1000);
    unsigned stag = surf_names.empty() ? 0u : el.name_keys[surf_names[0]] % 1000;
    std::fprintf(f, "1 0 0 0 0 0 0 %d %u 0\n$EndEntities\n", surf_names.empty() ? 0 : 1, stag);
    size_t nn = el.x.size();
    std::fprintf(f, "$Nodes\n1 %zu 1 %zu\n2 1 0 %zu\n", nn, nn, nn);
    for (size_t i = 0; i < nn; ++i) std::fprintf(f, "%zu\n", i + 1);
    for (size_t i = 0; i < nn; ++i) std::fprintf(f, "%.17g %.17g 0\n", el.x[i], el.y[i]);
    std::fprintf(f, "$EndNodes\n");
    size_t nb = el.bseg_key.size(), nc = el.cell_nv.size();
    // element blocks follow the input order: one block per run of boundary
    // segments with the same physical key, then one per run of cells of the
    // same type, so that reading the file back reproduces the face and ghost
    // order exactly
    std::vector<std::pair<size_t, size_t>> bruns, cruns;
    for (size_t b = 0; b < nb;) {
        size_t e = b;
        while (e < nb && el.bseg_key[e] == el.bseg_key[b]) ++e;
        bruns.push_back({b, e});
        b = e;
    }
    for (size_t c = 0; c < nc;) {
        size_t e = c;
        while (e < nc && el.cell_nv[e] == el.cell_nv[c]) ++e;
        cruns.push_back({c, e});
        c = e;
    }
    std::fprintf(f, "$Elements\n%zu %zu 1 %zu\n", bruns.size() + cruns.size(), nb + nc, nb + nc);
    size_t tag = 1;
    for (auto& run : bruns) {
        uint32_t key = el.bseg_key[run.first];
        size_t ent = 0;
        for (size_t k = 0; k < curve_names.size(); ++k)
            if (el.name_keys[curve_names[k]] == key) ent = k + 1;
        if (ent == 0) { std::fclose(f); throw std::runtime_error("boundary segment key without a physical name"); }
        std::fprintf(f, "1 %zu 1 %zu\n", ent, run.second - run.first);
        for (size_t b = run.first; b < run.second; ++b)
            std::fprintf(f, "%zu %u %u\n", tag++, el.bseg_nodes[2 * b] + 1, el.bseg_nodes[2 * b + 1] + 1);
    }
    for (auto& run : cruns) {
        int nv = el.cell_nv[run.first];
        std::fprintf(f, "2 1 %d %zu\n", nv == 3 ? 2 : 3, run.second - run.first);
        for (size_t c = run.first; c < run.second; ++c) {
            std::fprintf(f, "%zu", tag++);
            for (int v = 0; v < nv; ++v) std::fprintf(f, " %u", el.cell_nodes[4 * c + v] + 1);
            std::fprintf(f, "\n");
        }
    }
    std::fprintf(f, "$EndElements\n");
    std::fclose(f);
}

// ----------------------------------------------------------- construction --
namespace {
struct EdgeKeyHash {
    size_t operator()(uint64_t k) const { return (size_t)(k * 0x9E3779B97F4A7C15ull >> 17) ^ (size_t)k; }
};
inline uint64_t ekey(uint32_t a, uint32_t b) {
    uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
    return ((uint64_t)lo << 32) | hi;
}
}  // namespace

Mesh build_mesh(const Elements& el) {
    Mesh m;
    m.name_keys = el.name_keys;
    m.names = el.names;
    m.nodesX = el.x;
    m.nodesY = el.y;
    const uint32_t nC = (uint32_t)el.cell_nv.size();
    const uint32_t nB = (uint32_t)el.bseg_key.size();
    m.nRealCells = nC;
    m.cellsNodes.resize(4 * (size_t)nC);
    m.cellsIsTriangle.resize(nC);
    for (uint32_t c = 0; c < nC; ++c) {
        bool tri = el.cell_nv[c] == 3;
        m.cellsIsTriangle[c] = tri ? 1 : 0;
        for (int k = 0; k < 4; ++k) m.cellsNodes[4 * (size_t)c + k] = (tri && k == 3) ? 0u : el.cell_nodes[4 * (size_t)c + k];
    }

    // faces on first sight, owner = creating cell (mesh.h:317-341)
    std::unordered_map<uint64_t, uint32_t, EdgeKeyHash> eref;
    eref.reserve((size_t)nC * 2 + 16);
    size_t fguess = (size_t)nC * 2 + nB;
    m.edgesNodes.reserve(2 * fguess);
    m.edgesCells.reserve(2 * fguess);
    for (uint32_t c = 0; c < nC; ++c) {
        uint32_t n = m.cellsIsTriangle[c] ? 3 : 4;
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t j = (i < n - 1) ? i + 1 : 0;
            uint32_t a = m.cellsNodes[4 * (size_t)c + i], b = m.cellsNodes[4 * (size_t)c + j];
            auto ins = eref.emplace(ekey(a, b), (uint32_t)(m.edgesNodes.size() / 2));
            if (ins.second) {
                m.edgesNodes.push_back(a); m.edgesNodes.push_back(b);
                m.edgesCells.push_back(c); m.edgesCells.push_back(0);
            }
        }
    }
    const uint32_t nF = (uint32_t)(m.edgesNodes.size() / 2);
    auto find = [&](uint32_t a, uint32_t b) {
        auto it = eref.find(ekey(a, b));
        return it == eref.end() ? kNoEdge : it->second;
    };
    // cellsEdges and neighbours (mesh.h:344-375)
    m.cellsEdges.resize(4 * (size_t)nC);
    for (uint32_t c = 0; c < nC; ++c) {
        const uint32_t* cn = &m.cellsNodes[4 * (size_t)c];
        uint32_t ce[4];
        ce[0] = find(cn[0], cn[1]);
        ce[1] = find(cn[1], cn[2]);
        if (m.cellsIsTriangle[c]) { ce[2] = find(cn[2], cn[0]); ce[3] = kNoEdge; }
        else { ce[2] = find(cn[2], cn[3]); ce[3] = find(cn[3], cn[0]); }
        for (int k = 0; k < 4; ++k) {
            m.cellsEdges[4 * (size_t)c + k] = ce[k];
            if (ce[k] != kNoEdge && c != m.edgesCells[2 * (size_t)ce[k]]) m.edgesCells[2 * (size_t)ce[k] + 1] = c;
        }
    }
    // metrics (mesh.h:378-453)
    m.cellsAreas.assign(nC, 0.0);
    m.cellsCentersX.assign(nC, 0.0);
    m.cellsCentersY.assign(nC, 0.0);
    for (uint32_t c = 0; c < nC; ++c) {
        uint32_t n = m.cellsIsTriangle[c] ? 3 : 4;
        double cx = 0., cy = 0.;
        for (uint32_t j = 0; j < n; ++j) {
            cx += m.nodesX[m.cellsNodes[4 * (size_t)c + j]] / ((double)n);
            cy += m.nodesY[m.cellsNodes[4 * (size_t)c + j]] / ((double)n);
        }
        m.cellsCentersX[c] = cx;
        m.cellsCentersY[c] = cy;
    }
    m.edgesLengths.resize(nF); m.edgesNormalsX.resize(nF); m.edgesNormalsY.resize(nF);
    m.edgesCentersX.resize(nF); m.edgesCentersY.resize(nF);
    for (uint32_t f = 0; f < nF; ++f) {
        uint32_t a = m.edgesNodes[2 * (size_t)f], b = m.edgesNodes[2 * (size_t)f + 1];
        double xf = (m.nodesX[b] + m.nodesX[a]) * 0.5;
        double yf = (m.nodesY[b] + m.nodesY[a]) * 0.5;
        double dex = m.nodesX[b] - m.nodesX[a];
        double dey = m.nodesY[b] - m.nodesY[a];
        double l = std::sqrt(dex * dex + dey * dey);
        double nx = -dey / l, ny = dex / l;
        uint32_t o = m.edgesCells[2 * (size_t)f];
        double dot = nx * (xf - m.cellsCentersX[o]) + ny * (yf - m.cellsCentersY[o]);
        if (dot < 0) { nx *= -1.; ny *= -1.; }
        m.edgesCentersX[f] = xf; m.edgesCentersY[f] = yf;
        m.edgesLengths[f] = l; m.edgesNormalsX[f] = nx; m.edgesNormalsY[f] = ny;
    }
    for (uint32_t c = 0; c < nC; ++c) {
        const uint32_t* cn = &m.cellsNodes[4 * (size_t)c];
        double x1 = m.nodesX[cn[0]], x2 = m.nodesX[cn[1]], x3 = m.nodesX[cn[2]];
        double y1 = m.nodesY[cn[0]], y2 = m.nodesY[cn[1]], y3 = m.nodesY[cn[2]];
        double a = 0.5 * std::fabs(x1 * (y2 - y3) + x2 * (y3 - y1) + x3 * (y1 - y2));
        if (!m.cellsIsTriangle[c]) {
            double x4 = m.nodesX[cn[3]], y4 = m.nodesY[cn[3]];
            a = a + 0.5 * std::fabs(x1 * (y3 - y4) + x3 * (y4 - y1) + x4 * (y1 - y3));
        }
        m.cellsAreas[c] = a;
    }
    // ghosts (mesh.h:744-787)
    m.boundaryEdges.reserve(nB);
    for (uint32_t b = 0; b < nB; ++b) {
        uint32_t a = el.bseg_nodes[2 * (size_t)b], bb = el.bseg_nodes[2 * (size_t)b + 1];
        uint32_t nmin = std::min(a, bb), nmax = std::max(a, bb);
        uint32_t e = find(nmin, nmax);
        if (e == kNoEdge)
            throw std::runtime_error("invalid edge ref (" + std::to_string(nmin) + ", " + std::to_string(nmax) + ")");
        uint32_t cell = m.edgesCells[2 * (size_t)e];
        double dx = m.edgesCentersX[e] - m.cellsCentersX[cell];
        double dy = m.edgesCentersY[e] - m.cellsCentersY[cell];
        double dist = std::sqrt(dx * dx + dy * dy);
        m.cellsAreas.push_back(m.cellsAreas[cell]);
        m.cellsCentersX.push_back(m.edgesCentersX[e] + dist * m.edgesNormalsX[e]);
        m.cellsCentersY.push_back(m.edgesCentersY[e] + dist * m.edgesNormalsY[e]);
        m.cellsIsTriangle.push_back(1);
        m.cellsNodes.push_back(nmin); m.cellsNodes.push_back(nmax);
        m.cellsNodes.push_back(0); m.cellsNodes.push_back(0);
        m.edgesCells[2 * (size_t)e + 1] = m.n_cells() - 1;
        m.boundaryEdges.push_back(e);
        m.boundaryEdgesIntTag.push_back(el.bseg_key[b]);
        int ni = m.name_index(el.bseg_key[b]);
        if (ni < 0) throw std::runtime_error("boundary segment with unnamed physical tag " + std::to_string(el.bseg_key[b]));
        m.boundaryNameIdx.push_back(ni);
    }
    return m;
}

// ------------------------------------------------------------- generators --
namespace {
double naca_half_thickness(double x, double t) {
    // 4-digit thickness distribution with the closed-trailing-edge coefficient
    if (x < 0) x = 0;
    if (x > 1) x = 1;
    return 5.0 * t * (0.2969 * std::sqrt(x) - 0.1260 * x - 0.3516 * x * x + 0.2843 * x * x * x - 0.1036 * x * x * x * x);
}

// O-grid node coordinates: airfoil point at parameter theta (0 = TE, pi = LE,
// upper surface for theta in (0, pi)), farfield circle about mid-chord.
void ogrid_nodes(int ni, int nj, double r_far, double t, double first_h, std::vector<double>& X, std::vector<double>& Y,
                 std::vector<double>& sr) {
    X.assign((size_t)ni * nj, 0.0);
    Y.assign((size_t)ni * nj, 0.0);
    // geometric radial distribution: s_0 = 0, s_1 = h/L, sum to 1
    // find ratio r with first_h * (r^(nj-1) - 1)/(r - 1) = L where L ~ r_far
    double L = r_far;
    double lo = 1.0 + 1e-12, hi = 2.0;
    for (int it = 0; it < 200; ++it) {
        double r = 0.5 * (lo + hi);
        double s = first_h * (std::pow(r, nj - 1) - 1.0) / (r - 1.0);
        if (s > L) hi = r; else lo = r;
    }
    double r = 0.5 * (lo + hi);
    sr.assign(nj, 0.0);
    for (int j = 1; j < nj; ++j) sr[j] = first_h * (std::pow(r, j) - 1.0) / (r - 1.0) / L;
    sr[nj - 1] = 1.0;
    const double PI = 3.141592653589793;
    for (int i = 0; i < ni; ++i) {
        double th = 2.0 * PI * i / ni;
        double xa = 0.5 * (1.0 + std::cos(th));
        double ya = naca_half_thickness(xa, t) * (th <= PI ? 1.0 : -1.0);
        if (i == 0) ya = 0.0;
        double xo = 0.5 + r_far * std::cos(th);
        double yo = r_far * std::sin(th);
        for (int j = 0; j < nj; ++j) {
            double s = sr[j];
            X[(size_t)j * ni + i] = xa + s * (xo - xa);
            Y[(size_t)j * ni + i] = ya + s * (yo - ya);
        }
    }
}

void add_names(Elements& el) {
    el.name_keys = {1001, 1002, 2003};
    el.names = {"farfield", "wall", "internal"};
}

// in-circle test: > 0 if d lies inside the circumcircle of ccw triangle abc
double incircle(double ax, double ay, double bx, double by, double cx, double cy, double dx, double dy) {
    double adx = ax - dx, ady = ay - dy, bdx = bx - dx, bdy = by - dy, cdx = cx - dx, cdy = cy - dy;
    double ad = adx * adx + ady * ady, bd = bdx * bdx + bdy * bdy, cd = cdx * cdx + cdy * cdy;
    return adx * (bdy * cd - bd * cdy) - ady * (bdx * cd - bd * cdx) + ad * (bdx * cdy - bdy * cdx);
}
double orient(double ax, double ay, double bx, double by, double cx, double cy) {
    return (bx - ax) * (cy - ay) - (by - ay) * (cx - ax);
}
}  // namespace

Elements gen_naca_ogrid(int ni, int nj, double r_far, double thickness, double first_h) {
    Elements el;
    std::vector<double> sr;
    ogrid_nodes(ni, nj, r_far, thickness, first_h, el.x, el.y, sr);
    add_names(el);
    for (int j = 0; j < nj - 1; ++j) {
        for (int i = 0; i < ni; ++i) {
            uint32_t a = (uint32_t)(j * ni + i), b = (uint32_t)(j * ni + (i + 1) % ni);
            uint32_t c = (uint32_t)((j + 1) * ni + (i + 1) % ni), d = (uint32_t)((j + 1) * ni + i);
            uint32_t t1[4] = {a, b, c, 0}, t2[4] = {a, c, d, 0};
            el.cell_nodes.insert(el.cell_nodes.end(), t1, t1 + 4); el.cell_nv.push_back(3);
            el.cell_nodes.insert(el.cell_nodes.end(), t2, t2 + 4); el.cell_nv.push_back(3);
        }
    }
    for (int i = 0; i < ni; ++i) {   // wall segments (j = 0)
        el.bseg_nodes.push_back((uint32_t)i); el.bseg_nodes.push_back((uint32_t)((i + 1) % ni));
        el.bseg_key.push_back(1002);
    }
    for (int i = 0; i < ni; ++i) {   // farfield segments (j = nj-1)
        el.bseg_nodes.push_back((uint32_t)((nj - 1) * ni + i)); el.bseg_nodes.push_back((uint32_t)((nj - 1) * ni + (i + 1) % ni));
        el.bseg_key.push_back(1001);
    }
    return el;
}

Elements gen_jittered_airfoil(int ni, int nj, double r_far, double thickness, double first_h,
                              double jitter, double quad_frac, uint64_t seed) {
    Elements el;
    std::vector<double> sr;
    ogrid_nodes(ni, nj, r_far, thickness, first_h, el.x, el.y, sr);
    add_names(el);
    SplitMix64 rng(seed);
    const double PI = 3.141592653589793;
    // jitter interior nodes: sigma = jitter * h, h = the smallest altitude
    // (area / longest edge) of the node's four grid quads before jittering,
    // Gaussian clipped at 2 sigma (the grid spacings alone overestimate the
    // room in the skewed cells at a sharp trailing edge); the cells built
    // below are checked for a single orientation and convexity
    auto quad_alt = [&](int i, int j) {   // grid quad with lower-left node (i, j)
        const size_t a = (size_t)j * ni + i, b = (size_t)j * ni + (i + 1) % ni;
        const size_t c = (size_t)(j + 1) * ni + (i + 1) % ni, d = (size_t)(j + 1) * ni + i;
        const double area = 0.5 * std::fabs((el.x[c] - el.x[a]) * (el.y[d] - el.y[b]) -
                                            (el.y[c] - el.y[a]) * (el.x[d] - el.x[b]));
        const size_t v[4] = {a, b, c, d};
        double lmax = 0.0;
        for (int e = 0; e < 4; ++e) {
            const double dx = el.x[v[(e + 1) % 4]] - el.x[v[e]], dy = el.y[v[(e + 1) % 4]] - el.y[v[e]];
            lmax = std::max(lmax, std::sqrt(dx * dx + dy * dy));
        }
        return area / lmax;
    };
    std::vector<double> halt((size_t)ni * nj, 0.0);
    for (int j = 1; j < nj - 1; ++j)
        for (int i = 0; i < ni; ++i) {
            const int im = (i + ni - 1) % ni;
            halt[(size_t)j * ni + i] = std::min(std::min(quad_alt(im, j - 1), quad_alt(i, j - 1)),
                                                std::min(quad_alt(im, j), quad_alt(i, j)));
        }
    for (int j = 1; j < nj - 1; ++j) {
        for (int i = 0; i < ni; ++i) {
            size_t k = (size_t)j * ni + i;
            double h = halt[k];
            double gx = rng.normal(), gy = rng.normal();
            gx = std::max(-2.0, std::min(2.0, gx));
            gy = std::max(-2.0, std::min(2.0, gy));
            el.x[k] += jitter * h * gx;
            el.y[k] += jitter * h * gy;
        }
    }
    (void)PI;
    for (int j = 0; j < nj - 1; ++j) {
        for (int i = 0; i < ni; ++i) {
            uint32_t a = (uint32_t)(j * ni + i), b = (uint32_t)(j * ni + (i + 1) % ni);
            uint32_t c = (uint32_t)((j + 1) * ni + (i + 1) % ni), d = (uint32_t)((j + 1) * ni + i);
            double u = rng.uniform();
            if (u < quad_frac) {
                uint32_t q[4] = {a, b, c, d};
                el.cell_nodes.insert(el.cell_nodes.end(), q, q + 4); el.cell_nv.push_back(4);
                continue;
            }
            // diagonal a-c unless d lies inside circumcircle(a,b,c) -> use b-d
            double sgn = orient(el.x[a], el.y[a], el.x[b], el.y[b], el.x[c], el.y[c]) > 0 ? 1.0 : -1.0;
            double ic = sgn * incircle(el.x[a], el.y[a], el.x[b], el.y[b], el.x[c], el.y[c], el.x[d], el.y[d]);
            if (ic > 0) {
                uint32_t t1[4] = {a, b, d, 0}, t2[4] = {b, c, d, 0};
                el.cell_nodes.insert(el.cell_nodes.end(), t1, t1 + 4); el.cell_nv.push_back(3);
                el.cell_nodes.insert(el.cell_nodes.end(), t2, t2 + 4); el.cell_nv.push_back(3);
            } else {
                uint32_t t1[4] = {a, b, c, 0}, t2[4] = {a, c, d, 0};
                el.cell_nodes.insert(el.cell_nodes.end(), t1, t1 + 4); el.cell_nv.push_back(3);
                el.cell_nodes.insert(el.cell_nodes.end(), t2, t2 + 4); el.cell_nv.push_back(3);
            }
        }
    }
    // every cell must keep the orientation of the unjittered grid and quads
    // must stay convex: all corner cross products of one sign
    int sign0 = 0;
    for (size_t c = 0; c < el.cell_nv.size(); ++c) {
        const int nv = el.cell_nv[c];
        const uint32_t* v = &el.cell_nodes[4 * c];
        for (int k = 0; k < nv; ++k) {
            const uint32_t a = v[k], b = v[(k + 1) % nv], d = v[(k + 2) % nv];
            const double o = orient(el.x[a], el.y[a], el.x[b], el.y[b], el.x[d], el.y[d]);
            const int sg = o > 0 ? 1 : (o < 0 ? -1 : 0);
            if (sign0 == 0) sign0 = sg;
            if (sg == 0 || sg != sign0)
                throw std::runtime_error("jittered mesh has a folded or non-convex cell; lower the jitter");
        }
    }
    for (int i = 0; i < ni; ++i) {
        el.bseg_nodes.push_back((uint32_t)i); el.bseg_nodes.push_back((uint32_t)((i + 1) % ni));
        el.bseg_key.push_back(1002);
    }
    for (int i = 0; i < ni; ++i) {
        el.bseg_nodes.push_back((uint32_t)((nj - 1) * ni + i)); el.bseg_nodes.push_back((uint32_t)((nj - 1) * ni + (i + 1) % ni));
        el.bseg_key.push_back(1001);
    }
    return el;
}

// -------------------------------------------------------------------- RCM --
std::vector<uint32_t> rcm_order(const Mesh& m) {
    const uint32_t n = m.nRealCells;
    // adjacency through interior faces
    std::vector<uint32_t> deg(n, 0);
    const uint32_t nF = m.n_faces();
    for (uint32_t f = 0; f < nF; ++f) {
        uint32_t a = m.edgesCells[2 * (size_t)f], b = m.edgesCells[2 * (size_t)f + 1];
        if (a < n && b < n && a != b) { deg[a]++; deg[b]++; }
    }
    std::vector<uint64_t> ptr(n + 1, 0);
    for (uint32_t i = 0; i < n; ++i) ptr[i + 1] = ptr[i] + deg[i];
    std::vector<uint32_t> adj(ptr[n]);
    std::vector<uint64_t> fill(ptr.begin(), ptr.end() - 1);
    for (uint32_t f = 0; f < nF; ++f) {
        uint32_t a = m.edgesCells[2 * (size_t)f], b = m.edgesCells[2 * (size_t)f + 1];
        if (a < n && b < n && a != b) { adj[fill[a]++] = b; adj[fill[b]++] = a; }
    }
    // Cuthill-McKee: each component starts at a minimum-degree unvisited node
    // (ties by index); neighbours are appended in increasing (degree, index).
    std::vector<uint32_t> by_deg(n);
    for (uint32_t i = 0; i < n; ++i) by_deg[i] = i;
    std::stable_sort(by_deg.begin(), by_deg.end(), [&](uint32_t x, uint32_t y) { return deg[x] < deg[y]; });
    std::vector<char> seen(n, 0);
    std::vector<uint32_t> order;
    order.reserve(n);
    std::vector<uint32_t> nb;
    for (uint32_t s : by_deg) {
        if (seen[s]) continue;
        seen[s] = 1;
        size_t head = order.size();
        order.push_back(s);
        while (head < order.size()) {
            uint32_t v = order[head++];
            nb.clear();
            for (uint64_t k = ptr[v]; k < ptr[v + 1]; ++k)
                if (!seen[adj[k]]) { seen[adj[k]] = 1; nb.push_back(adj[k]); }
            std::stable_sort(nb.begin(), nb.end(), [&](uint32_t x, uint32_t y) {
                return deg[x] != deg[y] ? deg[x] < deg[y] : x < y;
            });
            order.insert(order.end(), nb.begin(), nb.end());
        }
    }
    std::reverse(order.begin(), order.end());
    return order;
}

}  // namespace afx
