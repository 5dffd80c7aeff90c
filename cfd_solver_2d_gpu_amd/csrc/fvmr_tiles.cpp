// Host-side preparation of the tile kernels (see fvmr_tiles.h).
//
// A tile is a run of consecutive device cells.  Its work-group needs, besides the tile
// cells T, the first halo ring H1 (cells across a face of T, periodic partners included:
// their limited gradients enter the face fluxes of T) and the second ring H2 (cells across
// a face of H1: their states enter the gradients of H1).  Faces are listed in the order
// faces of T (the fluxes the tile evaluates), faces of H1, then the periodic partners of
// listed faces (geometry only).  Everything a tile reads is packed contiguously per tile, so
// a work-group streams its records; only the halo states are gathered.
#include "fvmr_tiles.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "fvmr_internal.h"

namespace fvmr {

namespace {

// Hilbert index of (x, y) on an n x n grid, n = 2^order (the textbook xy -> d walk).
uint64_t hilbert_d(uint32_t x, uint32_t y, int order) {
    const uint32_t n = 1u << order;
    uint64_t d = 0;
    for (uint32_t s = n >> 1; s > 0; s >>= 1) {
        const uint32_t rx = (x & s) ? 1u : 0u;
        const uint32_t ry = (y & s) ? 1u : 0u;
        d += (uint64_t)s * s * ((3u * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                x = n - 1 - x;
                y = n - 1 - y;
            }
            std::swap(x, y);
        }
    }
    return d;
}

bool same_bits(double a, double b) { return std::memcmp(&a, &b, sizeof(double)) == 0; }

// Green-Gauss weight g_C = |r_F - r_f| / max(|r_F - r_C|, EPSILON) of cell c across edge e to
// cell o, the neighbour's centroid shifted by the periodic translation on periodic faces
// (GradientGreenGaussKernel, FVMR/src/space/gradient/GradientGPU.cu:223-425); 0.5 without a
// cell across.
double gg_weight(const TileInput &in, int c, int e, int o, bool periodic) {
    if (o < 0) return 0.5;
    const double fx = in.mx[e], fy = in.my[e];
    double Fx = in.cx[o], Fy = in.cy[o];
    const int pe = in.epair[e];
    if (periodic && pe >= 0) {
        Fx += fx - in.mx[pe];
        Fy += fy - in.my[pe];
    }
    const double ax = fx - Fx, ay = fy - Fy, bx = in.cx[c] - Fx, by = in.cy[c] - Fy;
    const double dFf = std::sqrt(ax * ax + ay * ay);
    double dFC = std::sqrt(bx * bx + by * by);
    if (dFC < kEPS) dFC = kEPS;
    return dFf / dFC;
}

}  // namespace

std::vector<int> hilbert_order(int nc, const double *cx, const double *cy) {
    double x0 = INFINITY, x1 = -INFINITY, y0 = INFINITY, y1 = -INFINITY;
    for (int c = 0; c < nc; c++) {
        x0 = std::min(x0, cx[c]); x1 = std::max(x1, cx[c]);
        y0 = std::min(y0, cy[c]); y1 = std::max(y1, cy[c]);
    }
    const int order = 16;
    const double span = std::max(std::max(x1 - x0, y1 - y0), 1e-300);
    const double scale = ((1u << order) - 1) / span;
    std::vector<std::pair<uint64_t, int>> key(nc);
    for (int c = 0; c < nc; c++) {
        const uint32_t ix = (uint32_t)std::min<double>((cx[c] - x0) * scale, (1u << order) - 1);
        const uint32_t iy = (uint32_t)std::min<double>((cy[c] - y0) * scale, (1u << order) - 1);
        key[c] = {hilbert_d(ix, iy, order), c};
    }
    std::sort(key.begin(), key.end());
    std::vector<int> order_out(nc);
    for (int i = 0; i < nc; i++) order_out[i] = key[i].second;
    return order_out;
}

bool standard_geometry(const fvmr_mesh_soa *m, const char **why) {
    const double *X = m->node_xy[0], *Y = m->node_xy[1];
    for (int c = 0; c < m->num_element; c++) {
        double sx = 0.0, sy = 0.0;
        for (int j = 0; j < 3; j++) {
            sx += X[m->element_nodes[j][c]];
            sy += Y[m->element_nodes[j][c]];
        }
        if (!same_bits(sx / 3.0, m->element_xy[0][c]) || !same_bits(sy / 3.0, m->element_xy[1][c])) {
            *why = "element centroids are not (x0+x1+x2)/3 of the node coordinates";
            return false;
        }
    }
    for (int e = 0; e < m->num_edge; e++) {
        const int a = m->edge_nodes[0][e], b = m->edge_nodes[1][e];
        const double len = m->edge_length[e];
        if (!same_bits((double)(float)len, len) || !(len > 0.0)) {
            *why = "edge lengths are not positive float values";
            return false;
        }
        if (!same_bits((X[a] + X[b]) / 2.0, m->edge_xy[0][e]) || !same_bits((Y[a] + Y[b]) / 2.0, m->edge_xy[1][e])) {
            *why = "edge midpoints are not (xa+xb)/2";
            return false;
        }
        const double dx = X[b] - X[a], dy = Y[b] - Y[a];
        if (!same_bits(dy / len, m->edge_normal[0][e]) || !same_bits(-dx / len, m->edge_normal[1][e])) {
            *why = "edge normals are not (dy,-dx)/length";
            return false;
        }
    }
    return true;
}

int build_tiles(const TileInput &in, const TileCaps &caps, TileMesh &out) {
    const int nc = in.nc, nf = in.nf;
    if (caps.cells < 1 || caps.gcells < 8 || caps.lcells < 16 || caps.nodes < 16 || caps.faces < 16 || caps.ffaces < 4 ||
        caps.lcells > kLocalNone || caps.nodes > kLocalNone || caps.faces > kLocalNone)
        return set_error(FVMR_ERR_INVALID_ARG, "bad tile capacities");
    for (int e = 0; e < nf; e++) {
        const int t = in.etype[e];
        if (t != -1 && (t <= kBcBase || t - kBcBase > 127))
            return set_error(FVMR_ERR_UNSUPPORTED, "boundary type outside the tile kernels' code range");
    }
    out = TileMesh();
    out.dtface.assign(nc, 0);
    // stamps: lid valid when lst == current tile
    std::vector<int> lst(nc, -1), lid(nc, 0), fst(nf, -1), fid(nf, 0);
    int nv = 0;
    for (int c = 0; c < 3 * nc; c++) nv = std::max(nv, in.cnode[c] + 1);
    for (int e = 0; e < nf; e++) nv = std::max(nv, std::max(in.en0[e], in.en1[e]) + 1);
    std::vector<int> nst(nv, -1), nid(nv, 0);
    std::vector<int> cells, faces, nodes;
    int stamp = 0;

    // Builds the sets of the candidate tile [c0, c0+s); false when a capacity is exceeded.
    auto try_tile = [&](int c0, int s, int &nG, int &nFF) -> bool {
        stamp++;
        cells.clear(); faces.clear(); nodes.clear();
        auto add_cell = [&](int c) {
            if (lst[c] == stamp) return;
            lst[c] = stamp; lid[c] = (int)cells.size(); cells.push_back(c);
        };
        auto add_face = [&](int e) {
            if (fst[e] == stamp) return;
            fst[e] = stamp; fid[e] = (int)faces.size(); faces.push_back(e);
        };
        auto add_node = [&](int v) {
            if (nst[v] == stamp) return;
            nst[v] = stamp; nid[v] = (int)nodes.size(); nodes.push_back(v);
        };
        for (int i = 0; i < s; i++) add_cell(c0 + i);
        for (int i = 0; i < s; i++)
            for (int j = 0; j < 3; j++) {
                const int o = in.cslot_y[(size_t)j * nc + c0 + i];
                if (o >= 0) add_cell(o);
            }
        nG = (int)cells.size();
        if (nG > caps.gcells) return false;
        for (int i = s; i < nG; i++)
            for (int j = 0; j < 3; j++) {
                const int o = in.cslot_y[(size_t)j * nc + cells[i]];
                if (o >= 0) add_cell(o);
            }
        if ((int)cells.size() > caps.lcells) return false;
        for (int i = 0; i < s; i++)
            for (int j = 0; j < 3; j++) add_face(in.cslot_x[(size_t)j * nc + c0 + i] & ((1 << 28) - 1));
        nFF = (int)faces.size();
        if (nFF > caps.ffaces) return false;
        for (int i = s; i < nG; i++)
            for (int j = 0; j < 3; j++) add_face(in.cslot_x[(size_t)j * nc + cells[i]] & ((1 << 28) - 1));
        for (size_t k = 0; k < faces.size(); k++) {
            const int pe = in.epair[faces[k]];
            if (pe >= 0) add_face(pe);
            if ((int)faces.size() > caps.faces) return false;
        }
        if ((int)faces.size() > caps.faces) return false;
        for (int c : cells)
            for (int j = 0; j < 3; j++) add_node(in.cnode[(size_t)j * nc + c]);
        for (int e : faces) { add_node(in.en0[e]); add_node(in.en1[e]); }
        return (int)nodes.size() <= caps.nodes;
    };

    out.case_tile_begin.assign(1, 0);
    int cs = 0;
    for (int c0 = 0; c0 < nc;) {
        while (cs + 1 < in.ncases && c0 >= in.case_begin[cs + 1]) {
            cs++;
            out.case_tile_begin.push_back(out.ntiles);
        }
        const int case_end = (cs + 1 < in.ncases) ? in.case_begin[cs + 1] : nc;
        int s = std::min(caps.cells, case_end - c0), nG = 0, nFF = 0;
        while (!try_tile(c0, s, nG, nFF)) {
            if (s == 1) return set_error(FVMR_ERR_UNSUPPORTED, "one cell and its halo exceed the tile capacity");
            s = std::max(1, s - std::max(1, s / 8));
        }
        const int nL = (int)cells.size(), nF = (int)faces.size(), nN = (int)nodes.size();
        const int info[kTileInfo] = {c0, s, nG, nL, nN, nF, nFF,
                                     (int)out.halo.size(), (int)(out.nxy.size() / 2), (int)(out.face.size() / 4),
                                     (int)out.lnodes.size(), (int)(out.gcoef.size() / 4), cs};
        out.info.insert(out.info.end(), info, info + kTileInfo);
        for (int i = s; i < nL; i++) out.halo.push_back(cells[i]);
        for (int v : nodes) { out.nxy.push_back(in.node_x[v]); out.nxy.push_back(in.node_y[v]); }
        for (int i = 0; i < nL; i++) {
            const int c = cells[i];
            out.lcxy.push_back(in.cx[c]);
            out.lcxy.push_back(in.cy[c]);
            int w = 0;
            for (int j = 0; j < 3; j++) w |= nid[in.cnode[(size_t)j * nc + c]] << (kLocalBits * j);
            out.lnodes.push_back(w);
        }
        for (int i = 0; i < nG; i++) {
            const int c = cells[i];
            double gw[3];
            for (int j = 0; j < 3; j++) {
                const int x = in.cslot_x[(size_t)j * nc + c], o = in.cslot_y[(size_t)j * nc + c];
                const int e = x & ((1 << 28) - 1);
                int w = fid[e] | ((o >= 0 ? lid[o] : kLocalNone) << kGsOutShift);
                if (x & (1 << 30)) w |= kGsIsR;
                if (x & (1 << 28)) w |= kGsPeriodic;
                out.gslot.push_back(w);
                gw[j] = gg_weight(in, c, e, o, (x & (1 << 28)) != 0);
            }
            out.gslot.push_back(i < s ? in.dtcode[c] : 0);
            if (i < s) {
                int w = 0;
                for (int j = 0; j < 3; j++) w |= fid[in.cslot_x[(size_t)j * nc + c] & ((1 << 28) - 1)] << (kLocalBits * j);
                out.dtface[c] = w;
            }
            out.gcoef.insert(out.gcoef.end(), {gw[0], gw[1], gw[2], in.vol[c]});
        }
        for (int k = 0; k < nF; k++) {
            const int e = faces[k];
            const int t = in.etype[e];
            const int bc = (t == -1) ? 0 : t - kBcBase;
            auto loc = [&](int c) { return (c >= 0 && lst[c] == stamp) ? lid[c] : kLocalNone; };
            const int pe = in.epair[e];
            const int pl = (pe >= 0 && fst[pe] == stamp) ? fid[pe] : kLocalNone;
            const float lf = (float)in.elen[e];
            int lbits;
            std::memcpy(&lbits, &lf, sizeof(int));
            out.face.push_back(nid[in.en0[e]] | (nid[in.en1[e]] << kLocalBits) | (bc << kBcShift));
            out.face.push_back(loc(in.eL[e]) | (loc(in.eR[e]) << kLocalBits) | (pl << (2 * kLocalBits)));
            out.face.push_back(lbits);
            out.face.push_back(e);
            out.fnrm.push_back(in.nx[e]);
            out.fnrm.push_back(in.ny[e]);
        }
        out.max_nL = std::max(out.max_nL, nL); out.max_nG = std::max(out.max_nG, nG);
        out.max_nF = std::max(out.max_nF, nF); out.max_nFF = std::max(out.max_nFF, nFF);
        out.max_nN = std::max(out.max_nN, nN);
        out.sum_nL += nL; out.sum_nG += nG; out.sum_nF += nF; out.sum_nN += nN;
        out.ntiles++;
        c0 += s;
    }
    while ((int)out.case_tile_begin.size() <= in.ncases) out.case_tile_begin.push_back(out.ntiles);
    return FVMR_OK;
}

}  // namespace fvmr
