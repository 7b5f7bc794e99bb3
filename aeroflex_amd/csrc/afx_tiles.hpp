// Cell tiles for the fused residual kernel.
//
// A tile is a compact set of real cells (grown breadth-first on the face
// graph from seeds taken in RCM order).  The device numbers cells tile by
// tile, so tile t owns the contiguous range [cell0[t], cell0[t+1]).  For each
// tile the builder lists
//   L  = local cells: the tile's cells T, then the real cells of its first
//        halo ring H1, then the second ring H2 and the ghost cells touching
//        T u H1 (states only);
//   G  = T u H1, the cells whose gradients and limiters the tile computes
//        (local indices 0..nG-1);
//   faces touching G, those touching T first (local indices 0..nFT-1);
//   per G cell up to four local face slots in reference face order.
// One workgroup then evaluates the residual of its tile entirely in LDS.
#pragma once
#include <cstdint>
#include <vector>

#include "afx_mesh.hpp"

namespace afx {

struct TileLayout {
    int32_t ntiles = 0;
    int32_t target = 0;
    // device real-cell order (new -> reference index), tile-major
    std::vector<int32_t> cell_new2old;
    std::vector<int32_t> cell0;          // ntiles + 1
    std::vector<int32_t> loff;           // ntiles + 1 (into L_ref)
    std::vector<int32_t> ng;             // ntiles
    std::vector<int32_t> foff;           // ntiles + 1 (into per-face arrays)
    std::vector<int32_t> nft;            // ntiles
    std::vector<int32_t> nbf;            // ntiles: faces of the tile's list with a ghost side
    std::vector<int32_t> goff;           // ntiles + 1 (into slots)
    std::vector<int32_t> L_ref;          // reference cell index of each local cell
    std::vector<uint64_t> slots;         // per G cell: 4 x u16 (bit 15 = neighbour side, 0xFFFF pad)
    std::vector<uint32_t> fcell;         // per local face: c0 | c1 << 16 (local, < 4096)
    std::vector<int32_t> fref;           // per local face: reference face index
    int32_t lmax = 0, gmax = 0, fmax = 0, ftmax = 0;
};

// Builds the tiles (target cells per tile) from a reference-order mesh.
// rcm gives the seed order (perm[new] = old).
TileLayout build_tiles(const Mesh& m, const std::vector<uint32_t>& rcm, int32_t target);

}  // namespace afx
