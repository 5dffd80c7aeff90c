// Host-side preparation of the tile kernels: Hilbert cell order, the geometry check, and
// the per-tile packed mesh (local cells, halo, nodes, faces) streamed by one work-group.
#pragma once
#include <cstdint>
#include <vector>

#include "../../include/fvmr.h"

namespace fvmr {

// Local ids inside a tile are 10-bit fields; kLocalNone marks "no cell / not local".
constexpr int kLocalBits = 10;
constexpr int kLocalMask = (1 << kLocalBits) - 1;
constexpr int kLocalNone = kLocalMask;

// Per-tile header, kTileInfo ints:
//   c0, nT (tile cells c0 .. c0+nT-1), nG (cells with a gradient: tile + first halo ring),
//   nL (cells with a state: + second ring), nN (nodes), nF (faces), nFF (faces of tile cells,
//   listed first), off_halo, off_node, off_face, off_lcell, off_gcell, case of a batch
enum TileInfoField { kTiC0, kTiNT, kTiNG, kTiNL, kTiNN, kTiNF, kTiNFF, kTiOffH, kTiOffN, kTiOffF, kTiOffL, kTiOffG,
                     kTiCase, kTileInfo };

// Gradient-cell slot word: local face (10 bits) | local cell across (10) | cell is the
// face's R (bit 20) | periodic face (bit 21).
constexpr int kGsOutShift = 10;
constexpr int kGsIsR = 1 << 20;
constexpr int kGsPeriodic = 1 << 21;

// Face record (int4): x = node a | node b << 10 | boundary code << 20 (0 interior, else
// type - 5999); y = L | R << 10 | pair face << 20; z = float bits of the length; w = edge.
constexpr int kBcShift = 20;
constexpr int kBcBase = 5999;

struct TileCaps {
    int cells;      // tile cells (nT)
    int gcells;     // nG
    int lcells;     // nL
    int nodes;      // nN
    int faces;      // nF
    int ffaces;     // nFF
};

struct TileMesh {
    int ntiles = 0;
    std::vector<int> info;        // kTileInfo per tile
    std::vector<int> halo;        // device cell id of local cells nT .. nL-1
    std::vector<double> nxy;      // node coordinates, 2 per local node
    std::vector<int> lnodes;      // per local cell: 3 local node ids (10 bits each)
    std::vector<int> gslot;       // per gradient cell: 3 slot words + dtcode (tile cells)
    std::vector<double> gcoef;    // per gradient cell: Green-Gauss weight of slots 0-2, volume
    std::vector<double> lcxy;     // per local cell: centroid x, y (the caller's, as lnodes)
    std::vector<double> fnrm;     // per face: normal x, y (the caller's, as face)
    std::vector<int> face;        // 4 ints per face
    std::vector<int> dtface;      // per device cell: its 3 slot faces, local ids in its tile
    std::vector<int> case_tile_begin;  // tiles of case c: [c] .. [c+1]-1
    int max_nL = 0, max_nG = 0, max_nF = 0, max_nFF = 0, max_nN = 0;
    int64_t sum_nL = 0, sum_nG = 0, sum_nF = 0, sum_nN = 0;
};

// Device-order inputs of the tile builder (all in the device numbering).
struct TileInput {
    int nc, nf;
    const int *cslot_x;   // [3][nc] edge | flags (bit 30 cell is R, bit 28 periodic)
    const int *cslot_y;   // [3][nc] cell across the face, -1 none
    const int *cnode;     // [3][nc] global node id of vertex j
    const int *eL, *eR, *etype, *epair;  // per edge
    const int *en0, *en1; // edge nodes (reference edge order)
    const double *elen;   // edge length
    const double *node_x, *node_y;
    const double *vol;    // cell volume
    const int *dtcode;    // per cell
    const double *cx, *cy;  // cell centroids
    const double *mx, *my;  // edge midpoints
    const double *nx, *ny;  // edge normals
    int ncases;             // cases of a batch: device cells case_begin[c] .. [c+1]-1 (tiles never
    const int *case_begin;  // cross a case boundary)
};

// Cells sorted along a Hilbert curve over their centroids (ties by index).
std::vector<int> hilbert_order(int nc, const double *cx, const double *cy);

// True when the caller's centroids, midpoints, normals and lengths are exactly what the
// tile kernels recompute from the node coordinates (centroid ((x0+x1)+x2)/3, midpoint
// (xa+xb)/2, normal (dy,-dx)/L) and every length is a float value.
bool standard_geometry(const fvmr_mesh_soa *m, const char **why);

// Cuts the device order into tiles no larger than caps; returns 0 or an FVMR_ERR_* code.
int build_tiles(const TileInput &in, const TileCaps &caps, TileMesh &out);

}  // namespace fvmr
