// Tile construction for the fused residual kernel (see afx_tiles.hpp).
#include "afx_tiles.hpp"

#include <algorithm>
#include <stdexcept>

namespace afx {

TileLayout build_tiles(const Mesh& m, const std::vector<uint32_t>& rcm, int32_t target) {
    if (target < 16 || target > 4096) throw std::invalid_argument("tile size must be in [16, 4096]");
    const int32_t nC = (int32_t)m.nRealCells;
    const int32_t nT = (int32_t)m.n_cells();
    TileLayout L;
    L.target = target;

    // face-graph adjacency of real cells (interior faces only)
    std::vector<int32_t> deg(nC, 0);
    const int32_t nF = (int32_t)m.n_faces();
    for (int32_t f = 0; f < nF; ++f) {
        uint32_t a = m.edgesCells[2 * (size_t)f], b = m.edgesCells[2 * (size_t)f + 1];
        if ((int32_t)a < nC && (int32_t)b < nC && a != b) { deg[a]++; deg[b]++; }
    }
    std::vector<int64_t> ptr(nC + 1, 0);
    for (int32_t i = 0; i < nC; ++i) ptr[i + 1] = ptr[i] + deg[i];
    std::vector<int32_t> adj(ptr[nC]);
    {
        std::vector<int64_t> fill(ptr.begin(), ptr.end() - 1);
        for (int32_t f = 0; f < nF; ++f) {
            uint32_t a = m.edgesCells[2 * (size_t)f], b = m.edgesCells[2 * (size_t)f + 1];
            if ((int32_t)a < nC && (int32_t)b < nC && a != b) { adj[fill[a]++] = (int32_t)b; adj[fill[b]++] = (int32_t)a; }
        }
    }

    // breadth-first growth from seeds in RCM order
    std::vector<int32_t> tile_of(nC, -1);
    std::vector<std::vector<int32_t>> cells_of;
    for (uint32_t s : rcm) {
        if (tile_of[s] >= 0) continue;
        const int32_t t = (int32_t)cells_of.size();
        cells_of.emplace_back();
        std::vector<int32_t>& tc = cells_of.back();
        size_t head = 0;
        tc.push_back((int32_t)s);
        tile_of[s] = t;
        while (head < tc.size() && (int32_t)tc.size() < target) {
            int32_t v = tc[head++];
            for (int64_t k = ptr[v]; k < ptr[v + 1] && (int32_t)tc.size() < target; ++k) {
                int32_t w = adj[k];
                if (tile_of[w] < 0) { tile_of[w] = t; tc.push_back(w); }
            }
        }
    }
    // a tile whose frontier ran out early is a pocket enclosed by earlier
    // tiles: merge pockets of at most target/8 cells into their smallest
    // neighbouring tile (a block per 1-2 cells would idle most of its lanes)
    const int32_t pocket = std::max(1, target / 8);
    for (size_t t = 0; t < cells_of.size(); ++t) {
        if (cells_of[t].empty() || (int32_t)cells_of[t].size() > pocket) continue;
        int32_t best = -1;
        for (int32_t c : cells_of[t])
            for (int64_t k = ptr[c]; k < ptr[c + 1]; ++k) {
                const int32_t u = tile_of[adj[k]];
                if (u == (int32_t)t) continue;
                if (best < 0 || cells_of[u].size() < cells_of[best].size() ||
                    (cells_of[u].size() == cells_of[best].size() && u < best))
                    best = u;
            }
        if (best < 0) continue;
        for (int32_t c : cells_of[t]) { tile_of[c] = best; cells_of[best].push_back(c); }
        cells_of[t].clear();
    }
    std::vector<int32_t> order;
    order.reserve(nC);
    L.cell0.push_back(0);
    for (auto& tc : cells_of) {
        if (tc.empty()) continue;
        // quadrilaterals first inside the tile, so that most wavefronts see
        // cells with the same number of faces
        std::stable_partition(tc.begin(), tc.end(), [&](int32_t c) { return m.cellsIsTriangle[c] == 0; });
        order.insert(order.end(), tc.begin(), tc.end());
        L.cell0.push_back((int32_t)order.size());
        L.ntiles++;
    }
    L.cell_new2old = order;

    // per-tile local tables
    std::vector<int32_t> stamp(nT, -1), loc(nT, 0);
    std::vector<int32_t> fstamp(nF, -1), floc(nF, 0);
    L.loff.push_back(0);
    L.foff.push_back(0);
    L.goff.push_back(0);
    std::vector<int32_t> Lc, Fc;
    for (int32_t t = 0; t < L.ntiles; ++t) {
        Lc.clear();
        Fc.clear();
        const int32_t b0 = L.cell0[t], b1 = L.cell0[t + 1];
        auto add_cell = [&](int32_t c) {
            if (stamp[c] != t) { stamp[c] = t; loc[c] = (int32_t)Lc.size(); Lc.push_back(c); }
        };
        auto other_of = [&](uint32_t f, int32_t c) {
            uint32_t a = m.edgesCells[2 * (size_t)f], b = m.edgesCells[2 * (size_t)f + 1];
            return (int32_t)((int32_t)a == c ? b : a);
        };
        for (int32_t i = b0; i < b1; ++i) add_cell(order[i]);
        // H1: real neighbours of T
        for (int32_t i = b0; i < b1; ++i) {
            int32_t c = order[i];
            for (int s = 0; s < 4; ++s) {
                uint32_t f = m.cellsEdges[4 * (size_t)c + s];
                if (f == kNoEdge) continue;
                int32_t o = other_of(f, c);
                if (o < nC) add_cell(o);
            }
        }
        const int32_t nG = (int32_t)Lc.size();
        // second ring and ghosts of G (state only)
        for (int32_t i = 0; i < nG; ++i) {
            int32_t c = Lc[i];
            for (int s = 0; s < 4; ++s) {
                uint32_t f = m.cellsEdges[4 * (size_t)c + s];
                if (f == kNoEdge) continue;
                add_cell(other_of(f, c));
            }
        }
        // faces: touching T first, then touching H1 only
        int32_t nft = 0;
        for (int32_t i = 0; i < nG; ++i) {
            int32_t c = Lc[i];
            for (int s = 0; s < 4; ++s) {
                uint32_t f = m.cellsEdges[4 * (size_t)c + s];
                if (f == kNoEdge) continue;
                if (fstamp[f] != t) { fstamp[f] = t; floc[f] = (int32_t)Fc.size(); Fc.push_back((int32_t)f); }
            }
            if (i == b1 - b0 - 1) nft = (int32_t)Fc.size();
        }
        if ((int32_t)Lc.size() > 4095 || (int32_t)Fc.size() > 32767) throw std::runtime_error("tile too large");
        // slots of G cells in reference face order
        for (int32_t i = 0; i < nG; ++i) {
            int32_t c = Lc[i];
            uint32_t fs[4];
            int n = 0;
            for (int s = 0; s < 4; ++s) {
                uint32_t f = m.cellsEdges[4 * (size_t)c + s];
                if (f != kNoEdge) fs[n++] = f;
            }
            std::sort(fs, fs + n);
            uint64_t word = 0;
            for (int s = 0; s < 4; ++s) {
                uint64_t v = 0xFFFFu;
                if (s < n) {
                    uint32_t side = ((int32_t)m.edgesCells[2 * (size_t)fs[s]] == c) ? 0u : 1u;
                    v = (uint64_t)(floc[fs[s]] | (side << 15));
                }
                word |= v << (16 * s);
            }
            L.slots.push_back(word);
        }
        int32_t nbf = 0;
        for (int32_t f : Fc) {
            uint32_t a = m.edgesCells[2 * (size_t)f], b = m.edgesCells[2 * (size_t)f + 1];
            L.fcell.push_back((uint32_t)loc[a] | ((uint32_t)loc[b] << 16));
            L.fref.push_back(f);
            nbf += (int32_t)b >= nC || (int32_t)a >= nC;
        }
        L.nbf.push_back(nbf);
        for (int32_t c : Lc) L.L_ref.push_back(c);
        L.ng.push_back(nG);
        L.nft.push_back(nft);
        L.loff.push_back((int32_t)L.L_ref.size());
        L.foff.push_back((int32_t)L.fref.size());
        L.goff.push_back((int32_t)L.slots.size());
        L.lmax = std::max(L.lmax, (int32_t)Lc.size());
        L.gmax = std::max(L.gmax, nG);
        L.fmax = std::max(L.fmax, (int32_t)Fc.size());
        L.ftmax = std::max(L.ftmax, nft);
    }
    return L;
}

}  // namespace afx
