// Host mesh layer of the FVMR residual path: SU2 ingest, connectivity, boundary
// tagging, periodic pairing and the reference SoA export, plus deterministic
// synthetic meshes and the field initialisers.
//
// The connectivity is the reference's (FVMR/src/legacy/FVM_2D.cpp:36-198) with the
// O(n_cells * n_edges) linear search getEdgeByNodeIDs replaced by a hash table
// keyed on the unordered node pair.  Edges are still numbered in first-appearance
// order over (element i, local edges (n0,n1),(n1,n2),(n2,n0)), the first owner is
// L and the second R, so every index matches the reference bit for bit.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/fvmr.h"
#include "fvmr_internal.h"

namespace fvmr {

thread_local std::string g_last_error;

int set_error(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

namespace {

// Open-addressing map (unordered node-id pair) -> edge index.
class EdgeTable {
public:
    explicit EdgeTable(size_t expected) {
        size_t cap = 16;
        while (cap < expected * 2 + 16) cap <<= 1;
        keys_.assign(cap, 0);
        vals_.assign(cap, -1);
        mask_ = cap - 1;
    }
    static uint64_t key(int a, int b) {
        uint32_t lo = (uint32_t)std::min(a, b), hi = (uint32_t)std::max(a, b);
        return ((uint64_t)hi << 32) | lo;
    }
    // index of the slot holding k, or of the empty slot where it would go
    size_t slot(uint64_t k) const {
        uint64_t h = k * 0x9E3779B97F4A7C15ull;
        size_t i = (size_t)(h >> 20) & mask_;
        while (vals_[i] != -1 && keys_[i] != k) i = (i + 1) & mask_;
        return i;
    }
    int find(int a, int b) const {
        size_t i = slot(key(a, b));
        return vals_[i];
    }
    void insert(int a, int b, int v) {
        uint64_t k = key(a, b);
        size_t i = slot(k);
        keys_[i] = k;
        vals_[i] = v;
    }

private:
    std::vector<uint64_t> keys_;
    std::vector<int> vals_;
    size_t mask_;
};

struct Marker {
    std::string name;
    std::vector<std::pair<int, int>> e;  // 1-based node ids
};

struct RawLists {
    std::vector<double> x, y;
    std::vector<int> id1;                 // file id + 1
    int max_id = 1;
    std::vector<std::array<int, 3>> tri;  // 1-based ids
    std::vector<Marker> markers;
    // kept for fvmr_mesh_write_su2
    std::vector<int> etype;               // 5 or 9
    std::vector<std::array<int, 4>> econn; // 0-based file ids
};

bool is_periodic(int t) { return FVMR_BC_PERIODIC_0 <= t && t <= FVMR_BC_PERIODIC_0 + 9; }

// BoundaryManager::boundaryNameToType (FVMR/src/boundary_condition/BoundaryManager.cpp:15-45).
// "wall_isothermal" is a name of this library (the reference declares the type only).
int name_to_type(const std::string &n, int equation, bool &ok) {
    ok = true;
    if (n == "wall" || n == "obstacle" || n == "airfoil" || n == "foil")
        return equation == FVMR_EQ_EULER ? FVMR_BC_WALL_NONVISCOUS : FVMR_BC_WALL_ADIABAT;
    if (n == "wall_nonViscous" || n == "wall_slippery" || n == "slippery" || n == "slippery_wall")
        return FVMR_BC_WALL_NONVISCOUS;
    if (n == "inlet") return FVMR_BC_INLET;
    if (n == "outlet") return FVMR_BC_OUTLET;
    if (n == "inf" || n == "infinity" || n == "far" || n == "farfield") return FVMR_BC_INF;
    if (n == "dsr") return FVMR_BC_DSR;
    if (n == "symmetry") return FVMR_BC_SYMMETRY;
    if (n.compare(0, 8, "periodic") == 0 && n.size() >= 8) {
        char c = n.back();
        if (c < '0' || c > '9') { ok = false; return -1; }
        return FVMR_BC_PERIODIC_0 + (c - '0');
    }
    if (n == "wall_isothermal") return FVMR_BC_WALL_ISOTHERMAL;
    ok = false;
    return -1;
}

}  // namespace
}  // namespace fvmr

struct BoundarySet {
    std::string name;
    int type = -1;
    int ID = 0;
    std::vector<int> edges;
};

struct fvmr_mesh {
    fvmr::RawLists raw;
    std::vector<int> elem_ids3;  // post-swap 1-based node ids
    std::vector<BoundarySet> sets;
    // SoA storage
    std::vector<double> node_x, node_y;
    std::vector<int> en[4], ee[4], enb[4];
    std::vector<double> ex, ey, evol;
    std::vector<int> gn0, gn1, gset, gpair, gL, gR;
    std::vector<double> glen, gx, gy, gnx, gny;
    std::vector<int> btype;
    fvmr_mesh_soa soa{};
    std::vector<double> cont_U[4];  // element states of a continue file (fvmr_mesh_read_continue)
};

namespace fvmr {
namespace {

int build(fvmr_mesh &m, int equation) {
    RawLists &r = m.raw;
    const int nv = (int)r.x.size();
    if (nv == 0) return set_error(FVMR_ERR_MESH, "null nodes exception");
    if (r.tri.empty()) return set_error(FVMR_ERR_MESH, "null elements exception");
    // FVM_2D::iniPNodeTable (FVM_2D.cpp:25-34)
    std::vector<int> id2node(r.max_id + 1, -1);
    for (int i = 0; i < nv; i++) {
        if (r.id1[i] < 0 || r.id1[i] > r.max_id) return set_error(FVMR_ERR_MESH, "bad node id");
        id2node[r.id1[i]] = i;
    }
    const int ne = (int)r.tri.size();
    std::vector<std::array<int, 3>> el = r.tri;  // node ids, 1-based
    std::vector<double> area(ne);
    for (int ie = 0; ie < ne; ie++)
        for (int k = 0; k < 3; k++) {
            int id = el[ie][k];
            if (id < 0 || id > r.max_id || id2node[id] < 0)
                return set_error(FVMR_ERR_MESH, "element references an unknown node");
        }
    // CCW fix in double (FVM_2D.cpp:51-70)
    for (int ie = 0; ie < ne; ie++) {
        double xn[3], yn[3];
        for (int i = 0; i < 3; i++) { int v = id2node[el[ie][i]]; xn[i] = r.x[v]; yn[i] = r.y[v]; }
        double a = 0.5 * (xn[0] * (yn[1] - yn[2]) + xn[1] * (yn[2] - yn[0]) + xn[2] * (yn[0] - yn[1]));
        if (a < 0) { area[ie] = -a; std::swap(el[ie][1], el[ie][2]); }
        else area[ie] = a;
    }
    // edge registration in first-appearance order (FVM_2D.cpp:72-81,116-134)
    EdgeTable table((size_t)ne * 2);
    std::vector<int> e0, e1, eL, eR;
    e0.reserve((size_t)ne * 3 / 2 + 16); e1.reserve(e0.capacity());
    eL.reserve(e0.capacity()); eR.reserve(e0.capacity());
    auto reg = [&](int a, int b, int ie) {
        int e = table.find(a, b);
        if (e < 0) {
            table.insert(a, b, (int)e0.size());
            e0.push_back(a); e1.push_back(b); eL.push_back(ie); eR.push_back(-1);
        } else {
            eR[e] = ie;
        }
    };
    for (int ie = 0; ie < ne; ie++) {
        reg(el[ie][0], el[ie][1], ie);
        reg(el[ie][1], el[ie][2], ie);
        reg(el[ie][2], el[ie][0], ie);
    }
    const int nf = (int)e0.size();
    // SoA export (OldDataConverter.cpp:35-136)
    m.node_x = r.x; m.node_y = r.y;
    for (int k = 0; k < 4; k++) { m.en[k].assign(ne, -1); m.ee[k].assign(ne, -1); m.enb[k].assign(ne, -1); }
    m.ex.resize(ne); m.ey.resize(ne); m.evol = area;
    m.elem_ids3.resize(3 * (size_t)ne);
    for (int ie = 0; ie < ne; ie++) {
        // centroid in double (FVM_2D.cpp:166-174)
        double sx = 0, sy = 0;
        for (int i = 0; i < 3; i++) { int v = id2node[el[ie][i]]; sx += r.x[v]; sy += r.y[v]; }
        m.ex[ie] = sx / 3.0;
        m.ey[ie] = sy / 3.0;
        int pe[3] = {table.find(el[ie][0], el[ie][1]), table.find(el[ie][1], el[ie][2]), table.find(el[ie][2], el[ie][0])};
        for (int j = 0; j < 3; j++) {
            m.en[j][ie] = id2node[el[ie][j]];
            m.ee[j][ie] = pe[j];
            m.elem_ids3[3 * (size_t)ie + j] = el[ie][j];
            // Element_2D::findNeighbor (Element_2D.cpp:19-48)
            if (eL[pe[j]] == ie) m.enb[j][ie] = eR[pe[j]];
            else if (eR[pe[j]] == ie) m.enb[j][ie] = eL[pe[j]];
        }
    }
    m.gn0.resize(nf); m.gn1.resize(nf); m.gset.assign(nf, -1); m.gpair.assign(nf, -1);
    m.gL = eL; m.gR = eR;
    m.glen.resize(nf); m.gx.resize(nf); m.gy.resize(nf); m.gnx.resize(nf); m.gny.resize(nf);
    for (int i = 0; i < nf; i++) {
        int a = id2node[e0[i]], b = id2node[e1[i]];
        m.gn0[i] = a; m.gn1[i] = b;
        double x0 = r.x[a], y0 = r.y[a], x1 = r.x[b], y1 = r.y[b];
        // FVM_2D::iniEdges_lengths (FVM_2D.cpp:186-196): rounded through float
        m.glen[i] = (double)(float)std::sqrt((x0 - x1) * (x0 - x1) + (y0 - y1) * (y0 - y1));
        m.gx[i] = (x0 + x1) / 2.0;
        m.gy[i] = (y0 + y1) / 2.0;
        double dx = x1 - x0, dy = y1 - y0;
        m.gnx[i] = dy / m.glen[i];
        m.gny[i] = -dx / m.glen[i];
    }
    // boundary sets (SU2MeshReader.cpp:201-247)
    struct PP { int bType, s0, s1; };
    std::vector<PP> pairs;
    m.sets.resize(r.markers.size());
    for (size_t i = 0; i < r.markers.size(); i++) {
        bool ok;
        BoundarySet &bs = m.sets[i];
        bs.name = r.markers[i].name;
        bs.type = name_to_type(bs.name, equation, ok);
        if (!ok) return set_error(FVMR_ERR_MESH, "unknown boundary type: " + bs.name);
        bs.ID = (int)i + 1;
        bs.edges.reserve(r.markers[i].e.size());
        for (auto &pr : r.markers[i].e) {
            int e = table.find(pr.first, pr.second);
            if (e < 0) return set_error(FVMR_ERR_MESH, "boundary edge not found: pEdge == nullptr");
            m.gset[e] = bs.ID;
            bs.edges.push_back(e);
        }
        if (is_periodic(bs.type)) {
            int idx = -1;
            for (size_t k = 0; k < pairs.size(); k++) if (pairs[k].bType == bs.type) idx = (int)k;
            if (idx != -1) pairs[idx].s1 = bs.ID;
            else pairs.push_back({bs.type, bs.ID, -1});
        }
    }
    // BoundaryManager_2D::checkPeriodPairs (BoundaryManager_2D.cpp:183-255)
    for (auto &pp : pairs) {
        if (pp.s0 == -1 || pp.s1 == -1)
            return set_error(FVMR_ERR_MESH, "b_error_uninitialized: periodic boundaries do not appear in pairs");
        const BoundarySet &b0 = m.sets[pp.s0 - 1], &b1 = m.sets[pp.s1 - 1];
        if (b0.edges.size() != b1.edges.size())
            return set_error(FVMR_ERR_MESH, "b_error_edge_num_not_match: periodic edge counts differ");
        const int n = (int)b0.edges.size();
        if (n == 0) continue;
        double tx = m.gx[b1.edges[n - 1]] - m.gx[b0.edges[0]];
        double ty = m.gy[b1.edges[n - 1]] - m.gy[b0.edges[0]];
        for (int ie = 0; ie < n; ie++) {
            double rx = m.gx[b1.edges[n - 1 - ie]] - (m.gx[b0.edges[ie]] + tx);
            double ry = m.gy[b1.edges[n - 1 - ie]] - (m.gy[b0.edges[ie]] + ty);
            if ((rx * rx + ry * ry) / (tx * tx + ty * ty) >= 0.0001)
                return set_error(FVMR_ERR_MESH, "b_error_not_translatable: periodic boundary not translatable");
        }
    }
    // periodic elementR and pair (OldDataConverter.cpp:155-172, BoundaryManager_2D.cpp:304-331)
    for (auto &bs : m.sets) {
        if (!is_periodic(bs.type)) continue;
        int other = -1;
        for (auto &pp : pairs) if (pp.bType == bs.type) { other = (pp.s0 == bs.ID) ? pp.s1 : pp.s0; break; }
        const BoundarySet &b1 = m.sets[other - 1];
        const int n = (int)bs.edges.size();
        for (int e : bs.edges) {
            // getEdgeIndex: first position of e in the set its setID names
            const BoundarySet &own = m.sets[m.gset[e] - 1];
            int idx0 = -1;
            for (int k = 0; k < (int)own.edges.size(); k++) if (own.edges[k] == e) { idx0 = k; break; }
            int pe = b1.edges[n - 1 - idx0];
            m.gR[e] = m.gL[pe];
            m.gpair[e] = pe;
        }
    }
    m.btype.resize(m.sets.size());
    for (size_t i = 0; i < m.sets.size(); i++) m.btype[i] = m.sets[i].type;

    fvmr_mesh_soa &s = m.soa;
    s.num_node = nv; s.num_element = ne; s.num_edge = nf; s.num_boundary_set = (int)m.sets.size();
    s.node_xy[0] = m.node_x.data(); s.node_xy[1] = m.node_y.data();
    for (int k = 0; k < 4; k++) {
        s.element_nodes[k] = m.en[k].data(); s.element_edges[k] = m.ee[k].data(); s.element_neighbors[k] = m.enb[k].data();
    }
    s.element_xy[0] = m.ex.data(); s.element_xy[1] = m.ey.data(); s.element_volume = m.evol.data();
    s.edge_nodes[0] = m.gn0.data(); s.edge_nodes[1] = m.gn1.data();
    s.edge_setID = m.gset.data(); s.edge_periodicPair = m.gpair.data();
    s.edge_elementL = m.gL.data(); s.edge_elementR = m.gR.data(); s.edge_length = m.glen.data();
    s.edge_xy[0] = m.gx.data(); s.edge_xy[1] = m.gy.data();
    s.edge_normal[0] = m.gnx.data(); s.edge_normal[1] = m.gny.data();
    s.boundary_type = m.btype.data();
    return FVMR_OK;
}

void add_element(RawLists &r, int type, int a, int b, int c, int d) {
    // 0-based file ids; SU2MeshReader.cpp:91-131 with convertRectToTriangle
    r.etype.push_back(type);
    r.econn.push_back({a, b, c, d});
    if (type == 5) {
        r.tri.push_back({a + 1, b + 1, c + 1});
    } else {
        r.tri.push_back({a + 1, b + 1, d + 1});
        r.tri.push_back({d + 1, b + 1, c + 1});
    }
}

// SU2 tokenizer of SU2MeshReader::readMesh: '=' is its own word, whitespace splits.
int split_words(const char *line, const char **w, int *wl, int maxw) {
    int n = 0;
    const char *p = line;
    while (*p && n < maxw) {
        while (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n') p++;
        if (!*p) break;
        if (*p == '=') { w[n] = p; wl[n] = 1; n++; p++; continue; }
        const char *s = p;
        while (*p && *p != ' ' && *p != '\t' && *p != '\r' && *p != '\n' && *p != '=') p++;
        w[n] = s; wl[n] = (int)(p - s); n++;
    }
    return n;
}

bool word_is(const char *w, int l, const char *lit) { return (int)strlen(lit) == l && strncmp(w, lit, l) == 0; }

double word_d(const char *w, int l) {
    char buf[64];
    int n = std::min(l, 63);
    memcpy(buf, w, n); buf[n] = 0;
    return strtod(buf, nullptr);
}

int read_su2(const char *path, double scale, RawLists &r) {
    FILE *f = fopen(path, "rb");
    if (!f) return set_error(FVMR_ERR_IO, std::string("cannot open file ") + path);
    enum { S0, SNDIME, SNPOIN, SNELEM, SNMARK } st = S0;
    std::vector<char> line(1 << 16);
    const char *w[16]; int wl[16];
    int rc = FVMR_OK;
    while (fgets(line.data(), (int)line.size(), f)) {
        int n = split_words(line.data(), w, wl, 16);
        if (n == 0 || word_is(w[0], wl[0], "%")) continue;
        if (word_is(w[0], wl[0], "NDIME")) st = SNDIME;
        else if (word_is(w[0], wl[0], "NPOIN")) st = SNPOIN;
        else if (word_is(w[0], wl[0], "NELEM")) st = SNELEM;
        else if (word_is(w[0], wl[0], "NMARK")) st = SNMARK;
        if (st == SNDIME) {
            if (!(n == 3 && word_is(w[2], wl[2], "2"))) { rc = set_error(FVMR_ERR_MESH, "SU2 mesh dimension is not 2D"); break; }
        } else if (st == SNPOIN && w[0][0] != 'N') {
            if (n < 3) { rc = set_error(FVMR_ERR_MESH, "SU2 node line without id"); break; }
            r.x.push_back(word_d(w[0], wl[0]));
            r.y.push_back(word_d(w[1], wl[1]));
            int id = (n == 3) ? (int)word_d(w[2], wl[2]) + 1 : (int)word_d(w[3], wl[3]) + 1;
            r.id1.push_back(id);
            r.max_id = std::max(r.max_id, id);
        } else if (st == SNELEM && w[0][0] != 'N') {
            if (word_is(w[0], wl[0], "5") && n >= 4) {
                add_element(r, 5, (int)word_d(w[1], wl[1]), (int)word_d(w[2], wl[2]), (int)word_d(w[3], wl[3]), -1);
            } else if (word_is(w[0], wl[0], "9") && n >= 5) {
                add_element(r, 9, (int)word_d(w[1], wl[1]), (int)word_d(w[2], wl[2]), (int)word_d(w[3], wl[3]), (int)word_d(w[4], wl[4]));
            }
        } else if (st == SNMARK) {
            if (word_is(w[0], wl[0], "NMARK") || word_is(w[0], wl[0], "MARKER_ELEMS")) {
            } else if (word_is(w[0], wl[0], "MARKER_TAG")) {
                r.markers.push_back(Marker{n > 2 ? std::string(w[2], wl[2]) : std::string(), {}});
            } else {
                if (r.markers.empty() || n < 3) { rc = set_error(FVMR_ERR_MESH, "SU2 marker edge before MARKER_TAG"); break; }
                r.markers.back().e.push_back({(int)word_d(w[1], wl[1]) + 1, (int)word_d(w[2], wl[2]) + 1});
            }
        }
    }
    fclose(f);
    if (rc) return rc;
    // SU2MeshReader::scaleMesh (SU2MeshReader.cpp:157-173)
    double res = scale - 1.0;
    if (!(res > -0.01 && res < 0.01)) {
        if (scale < 0.0) return set_error(FVMR_ERR_INVALID_ARG, "mesh scale factor should be positive");
        for (auto &v : r.x) v *= scale;
        for (auto &v : r.y) v *= scale;
    }
    return FVMR_OK;
}

// splitmix64 (Vigna), used for the reproducible node jitter.
uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
double unit_rand(uint64_t seed, uint64_t k) {  // uniform in [-1, 1)
    return (double)(splitmix64(seed * 0x100000001B3ull + k) >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

void push_node(RawLists &r, double x, double y) {
    int id = (int)r.x.size();
    r.x.push_back(x); r.y.push_back(y); r.id1.push_back(id + 1);
    r.max_id = std::max(r.max_id, id + 1);
}

int gen_periodic_square(const fvmr_gen_params &g, RawLists &r) {
    const int ni = g.ni, nj = g.nj;
    const double x0 = g.a, L = g.b > 0 ? g.b : 10.0;
    auto id = [&](int i, int j) { return j * (ni + 1) + i; };
    for (int j = 0; j <= nj; j++)
        for (int i = 0; i <= ni; i++) push_node(r, x0 + L * i / ni, x0 + L * j / nj);
    for (int j = 0; j < nj; j++)
        for (int i = 0; i < ni; i++) add_element(r, 9, id(i, j), id(i + 1, j), id(i + 1, j + 1), id(i, j + 1));
    Marker bot{"periodic_0", {}}, top{"periodic_0", {}}, rgt{"periodic_1", {}}, lft{"periodic_1", {}};
    for (int i = 0; i < ni; i++) bot.e.push_back({id(i, 0) + 1, id(i + 1, 0) + 1});
    for (int i = ni - 1; i >= 0; i--) top.e.push_back({id(i + 1, nj) + 1, id(i, nj) + 1});
    for (int j = 0; j < nj; j++) rgt.e.push_back({id(ni, j) + 1, id(ni, j + 1) + 1});
    for (int j = nj - 1; j >= 0; j--) lft.e.push_back({id(0, j + 1) + 1, id(0, j) + 1});
    r.markers = {bot, top, rgt, lft};
    return FVMR_OK;
}

// [0,Lx]x[0,Ly] channel; bottom wall y_b = max(0, (x - x0) tan(angle)); columns mapped
// y = y_b + eta (Ly - y_b); interior nodes jittered by +-jitter of the local spacing.
int gen_wedge(const fvmr_gen_params &g, RawLists &r, bool ramp) {
    const int ni = g.ni, nj = g.nj;
    const double Lx = g.a > 0 ? g.a : 3.0, Ly = g.b > 0 ? g.b : 1.0;
    const double xr = g.c, ang = ramp ? g.d * M_PI / 180.0 : 0.0;
    const double ta = std::tan(ang);
    auto id = [&](int i, int j) { return j * (ni + 1) + i; };
    for (int j = 0; j <= nj; j++)
        for (int i = 0; i <= ni; i++) {
            double x = Lx * i / ni;
            double yb = ramp ? std::max(0.0, (x - xr) * ta) : 0.0;
            double eta = (double)j / nj;
            double y = yb + eta * (Ly - yb);
            if (i > 0 && i < ni && j > 0 && j < nj && g.jitter > 0) {
                uint64_t k = (uint64_t)id(i, j);
                x += unit_rand(g.seed, 2 * k) * g.jitter * (Lx / ni);
                y += unit_rand(g.seed, 2 * k + 1) * g.jitter * ((Ly - yb) / nj);
            }
            push_node(r, x, y);
        }
    for (int j = 0; j < nj; j++)
        for (int i = 0; i < ni; i++) add_element(r, 9, id(i, j), id(i + 1, j), id(i + 1, j + 1), id(i, j + 1));
    Marker inl{"inlet", {}}, out{"outlet", {}}, bot{"slippery", {}}, top{"slippery", {}};
    for (int j = nj - 1; j >= 0; j--) inl.e.push_back({id(0, j + 1) + 1, id(0, j) + 1});
    for (int j = 0; j < nj; j++) out.e.push_back({id(ni, j) + 1, id(ni, j + 1) + 1});
    for (int i = 0; i < ni; i++) bot.e.push_back({id(i, 0) + 1, id(i + 1, 0) + 1});
    for (int i = ni - 1; i >= 0; i--) top.e.push_back({id(i + 1, nj) + 1, id(i, nj) + 1});
    r.markers = {inl, out, bot, top};
    return FVMR_OK;
}

// O-grid: ni cells around (seam shares nodes), nj cells radially from r0 to r1,
// geometric growth ratio q (q <= 1 gives uniform spacing).
int gen_ogrid(const fvmr_gen_params &g, RawLists &r, const std::vector<double> *surf_x,
              const std::vector<double> *surf_y, const char *wall_name) {
    const int ni = g.ni, nj = g.nj;
    const double r0 = g.a > 0 ? g.a : 0.5, r1 = g.b > 0 ? g.b : 30.0, q = g.c;
    std::vector<double> s(nj + 1);
    if (q > 1.0) {
        double h0 = (r1 - r0) * (q - 1.0) / (std::pow(q, nj) - 1.0);
        for (int j = 0; j <= nj; j++) s[j] = h0 * (std::pow(q, j) - 1.0) / (q - 1.0) / (r1 - r0);
    } else {
        for (int j = 0; j <= nj; j++) s[j] = (double)j / nj;
    }
    s[nj] = 1.0;
    auto id = [&](int i, int j) { return j * ni + (i % ni); };
    for (int j = 0; j <= nj; j++)
        for (int i = 0; i < ni; i++) {
            double th = 2.0 * M_PI * i / ni;
            double xo = r1 * std::cos(th), yo = r1 * std::sin(th);
            double xi, yi;
            if (surf_x) { xi = (*surf_x)[i]; yi = (*surf_y)[i]; }
            else { xi = r0 * std::cos(th); yi = r0 * std::sin(th); }
            push_node(r, xi + s[j] * (xo - xi), yi + s[j] * (yo - yi));
        }
    for (int j = 0; j < nj; j++)
        for (int i = 0; i < ni; i++) add_element(r, 9, id(i, j), id(i + 1, j), id(i + 1, j + 1), id(i, j + 1));
    Marker wall{wall_name, {}}, far{"farfield", {}};
    for (int i = 0; i < ni; i++) wall.e.push_back({id(i + 1, 0) + 1, id(i, 0) + 1});
    for (int i = 0; i < ni; i++) far.e.push_back({id(i, nj) + 1, id(i + 1, nj) + 1});
    r.markers = {wall, far};
    return FVMR_OK;
}

// NACA 4-digit symmetric/cambered surface, closed trailing edge, cosine spacing;
// ni points ordered counter-clockwise starting at the trailing edge (upper side first), so
// that point i lies on the ray of the farfield point at angle 2 pi i / ni.
void naca_surface(int digits, int ni, std::vector<double> &xs, std::vector<double> &ys) {
    double m = (digits / 1000) / 100.0, pp = ((digits / 100) % 10) / 10.0, t = (digits % 100) / 100.0;
    xs.resize(ni); ys.resize(ni);
    for (int i = 0; i < ni; i++) {
        double th = 2.0 * M_PI * i / ni;             // 0 at TE, pi at LE
        double x = 0.5 * (1.0 + std::cos(th));       // 1 -> 0 -> 1
        double yt = 5.0 * t * (0.2969 * std::sqrt(x) - 0.1260 * x - 0.3516 * x * x + 0.2843 * x * x * x - 0.1036 * x * x * x * x);
        double yc = 0, dyc = 0;
        if (m > 0 && pp > 0) {
            if (x < pp) { yc = m / (pp * pp) * (2 * pp * x - x * x); dyc = 2 * m / (pp * pp) * (pp - x); }
            else { yc = m / ((1 - pp) * (1 - pp)) * ((1 - 2 * pp) + 2 * pp * x - x * x); dyc = 2 * m / ((1 - pp) * (1 - pp)) * (pp - x); }
        }
        double thc = std::atan(dyc);
        double sgn = (i <= ni / 2) ? 1.0 : -1.0;     // upper surface first (counter-clockwise)
        xs[i] = x - sgn * yt * std::sin(thc) - 0.5;  // centred on mid-chord
        ys[i] = yc + sgn * yt * std::cos(thc);
    }
}

}  // namespace
}  // namespace fvmr

using namespace fvmr;

extern "C" {

const char *fvmr_last_error(void) { return fvmr::g_last_error.c_str(); }

int fvmr_params_default(fvmr_params *p) {
    if (!p) return set_error(FVMR_ERR_INVALID_ARG, "null params");
    memset(p, 0, sizeof(*p));
    // FVMR/src/global/GlobalPara.cpp:4-90 (limiter/reconstruction/gradient chosen as the GPU path supports)
    p->gamma = 1.4;
    p->R = 287.06;
    p->Pr = 0.73;
    p->Pr_viscous = 0.72;
    p->T0 = 288.16;
    p->mu0 = 17.9e-6;
    p->CFL = 0.6;
    p->equation = FVMR_EQ_EULER;
    p->flux_scheme = FVMR_SOL_ROE;
    p->reconstruct = FVMR_REC_LINEAR;
    p->gradient = FVMR_GRA_GREENGAUSS;
    p->limiter = FVMR_LIM_BARTH;
    p->time_advance = FVMR_EVO_EXPLICIT;
    p->viscosity_model = FVMR_MU_SUTHERLAND;
    p->precision = FVMR_PREC_FP64;
    p->renumber = FVMR_RENUMBER_HILBERT;
    p->compat_inv2x2 = 0;
    p->kernels = FVMR_KERNELS_AUTO;
    p->tile_cells = 0;
    p->tile_threads = 0;
    p->reserved = 0;
    p->wall_temperature = 288.16;
    p->venkat_K = 5.0;
    p->dsr_shock_speed = 20.0;
    return FVMR_OK;
}

int fvmr_ruvp_from_mach(double Ma, double AoA, double T0, double p0, double gamma, double R, double ruvp[4]) {
    // AirParameterConverter::get_ruvp_by_Ma_AoA_T0_p0_gamma_R (AirParameterConverter.cpp:3-24)
    if (!ruvp) return set_error(FVMR_ERR_INVALID_ARG, "null ruvp");
    const double PI = 3.1415926535897;
    double T = T0 / (1.0 + (gamma - 1.0) / 2.0 * Ma * Ma);
    double p = p0 * std::pow(T / T0, gamma / (gamma - 1.0));
    double c = std::sqrt(gamma * R * T);
    ruvp[0] = p / (R * T);
    ruvp[1] = c * Ma * std::cos(AoA / 180 * PI);
    ruvp[2] = c * Ma * std::sin(AoA / 180 * PI);
    ruvp[3] = p;
    return FVMR_OK;
}

// ContinueFileReader::readContinueFile_1 (FVMR/src/input/ContinueFileReader.cpp:19-158) and
// process (:457-494): nodes (ID, x, y), elements (ID, node IDs, U), boundary sets (ID, name,
// then edge IDs into the first-appearance edge list).  The edges are those of FVM_2D::iniEdges
// over the element lists, so the mesh is built once without sets to number the edges, and the
// set edge IDs become node pairs for the final build.
int fvmr_mesh_read_continue(const char *path, int equation, fvmr_mesh **out, double *t, int *istep, double *CFL) {
    if (!path || !out) return set_error(FVMR_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    FILE *f = fopen(path, "rb");
    if (!f) return set_error(FVMR_ERR_IO, std::string("cannot open file ") + path);
    fvmr_mesh *m = new fvmr_mesh();
    fvmr::RawLists &r = m->raw;
    struct SetIn { int id; std::string name; std::vector<int> edges; };
    std::vector<SetIn> sets;
    double t0 = 0.0, cfl0 = 0.0;
    int step0 = 0;
    int state = -1, rc = FVMR_OK;
    std::vector<char> line(1 << 16);
    const char *w[16]; int wl[16];
    while (fgets(line.data(), (int)line.size(), f)) {
        const int n = fvmr::split_words(line.data(), w, wl, 16);
        if (n == 0) continue;
        if (fvmr::word_is(w[0], wl[0], "t,")) { state = 0; continue; }
        if (fvmr::word_is(w[0], wl[0], "nodes:")) { state = 1; continue; }
        if (fvmr::word_is(w[0], wl[0], "edges:")) { state = 2; continue; }
        if (fvmr::word_is(w[0], wl[0], "elements:")) { state = 3; continue; }
        if (fvmr::word_is(w[0], wl[0], "boundaries:")) { state = 4; continue; }
        if (state == 0) {
            if (n < 2) { rc = set_error(FVMR_ERR_MESH, "continue file: bad header line"); break; }
            t0 = fvmr::word_d(w[0], wl[0]);
            step0 = (int)fvmr::word_d(w[1], wl[1]);
            if (n >= 3) cfl0 = fvmr::word_d(w[2], wl[2]);
        } else if (state == 1) {
            if (n < 3) { rc = set_error(FVMR_ERR_MESH, "continue file: bad node line"); break; }
            const int id = (int)fvmr::word_d(w[0], wl[0]);
            r.id1.push_back(id);
            r.max_id = std::max(r.max_id, id);
            r.x.push_back(fvmr::word_d(w[1], wl[1]));
            r.y.push_back(fvmr::word_d(w[2], wl[2]));
        } else if (state == 3) {
            if (n < 8) { rc = set_error(FVMR_ERR_MESH, "continue file: bad element line"); break; }
            const int a = (int)fvmr::word_d(w[1], wl[1]), b = (int)fvmr::word_d(w[2], wl[2]), c = (int)fvmr::word_d(w[3], wl[3]);
            r.tri.push_back({a, b, c});
            r.etype.push_back(5);
            r.econn.push_back({a - 1, b - 1, c - 1, -1});
            for (int k = 0; k < 4; k++) m->cont_U[k].push_back(fvmr::word_d(w[4 + k], wl[4 + k]));
        } else if (state == 4) {
            // a set header "ID name" (second word not a number) or a line of edge IDs
            if (n >= 2 && !isdigit((unsigned char)w[1][0])) {
                sets.push_back({(int)fvmr::word_d(w[0], wl[0]), std::string(w[1], wl[1]), {}});
            } else {
                if (sets.empty()) { rc = set_error(FVMR_ERR_MESH, "continue file: edge IDs before a set"); break; }
                for (int k = 0; k < n; k++) sets.back().edges.push_back((int)fvmr::word_d(w[k], wl[k]));
            }
        }
    }
    fclose(f);
    if (!rc && (r.x.empty() || r.tri.empty())) rc = set_error(FVMR_ERR_MESH, "continue file without nodes or elements");
    if (!rc) {
        // number the edges (first appearance) and turn the set edge IDs into node-ID pairs
        fvmr_mesh tmp;
        tmp.raw = r;
        tmp.raw.markers.clear();
        rc = fvmr::build(tmp, equation);
        for (size_t s = 0; !rc && s < sets.size(); s++) {
            fvmr::Marker mk;
            mk.name = sets[s].name;
            for (int id : sets[s].edges) {
                const int e = id - 1;
                if (e < 0 || e >= (int)tmp.gn0.size()) { rc = set_error(FVMR_ERR_MESH, "continue file: edge ID out of range"); break; }
                mk.e.push_back({r.id1[tmp.gn0[e]], r.id1[tmp.gn1[e]]});
            }
            r.markers.push_back(mk);
        }
    }
    if (!rc) rc = fvmr::build(*m, equation);
    if (rc) { delete m; return rc; }
    if (t) *t = t0;
    if (istep) *istep = step0;
    if (CFL) *CFL = cfl0;
    *out = m;
    return FVMR_OK;
}

int fvmr_mesh_continue_state(const fvmr_mesh *m, double *const U[4]) {
    if (!m || !U) return set_error(FVMR_ERR_INVALID_ARG, "null argument");
    if (m->cont_U[0].size() != m->evol.size()) return set_error(FVMR_ERR_INVALID_ARG, "the mesh was not read from a continue file");
    for (int k = 0; k < 4; k++) {
        if (!U[k]) return set_error(FVMR_ERR_INVALID_ARG, "null state component");
        std::copy(m->cont_U[k].begin(), m->cont_U[k].end(), U[k]);
    }
    return FVMR_OK;
}

// FieldWriter::writeContinueFile_1 (FVMR/src/output/FieldWriter.cpp:264-326) in the same layout,
// numbers with 17 significant digits so that a restart reproduces the state bit for bit.
int fvmr_write_continue(const fvmr_mesh *m, const char *path, double t, int istep, double CFL, const double *const U[4]) {
    if (!m || !path || !U) return set_error(FVMR_ERR_INVALID_ARG, "null argument");
    for (int k = 0; k < 4; k++)
        if (!U[k]) return set_error(FVMR_ERR_INVALID_ARG, "null state component");
    FILE *f = fopen(path, "w");
    if (!f) return set_error(FVMR_ERR_IO, std::string("cannot open file ") + path);
    const fvmr::RawLists &r = m->raw;
    fprintf(f, "t, istep, CFL\n%.17g %d %.17g\nnodes: ID, x, y\n", t, istep, CFL);
    for (size_t i = 0; i < r.x.size(); i++) fprintf(f, "%d %.17g %.17g\n", r.id1[i], r.x[i], r.y[i]);
    fprintf(f, "elements: ID, nodes, U\n");
    const size_t ne = m->evol.size();
    for (size_t ie = 0; ie < ne; ie++)
        fprintf(f, "%zu %d %d %d %.17g %.17g %.17g %.17g\n", ie + 1, m->elem_ids3[3 * ie], m->elem_ids3[3 * ie + 1],
                m->elem_ids3[3 * ie + 2], U[0][ie], U[1][ie], U[2][ie], U[3][ie]);
    fprintf(f, "boundaries: ID, name; edgeIDs\n");
    for (const BoundarySet &bs : m->sets) {
        fprintf(f, "%d %s\n", bs.ID, bs.name.c_str());
        bool at_line_end = false;
        for (size_t k = 0; k < bs.edges.size(); k++) {
            fprintf(f, "%d", bs.edges[k] + 1);
            at_line_end = (k % 10 == 9);
            fputc(at_line_end ? '\n' : ' ', f);
        }
        if (!at_line_end) fputc('\n', f);
    }
    const bool ok = !ferror(f);
    fclose(f);
    return ok ? FVMR_OK : set_error(FVMR_ERR_IO, std::string("write failed: ") + path);
}

int fvmr_mesh_read_su2(const char *path, int equation, double scale, fvmr_mesh **out) {
    if (!path || !out) return set_error(FVMR_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    fvmr_mesh *m = new fvmr_mesh();
    int rc = read_su2(path, scale, m->raw);
    if (!rc) rc = build(*m, equation);
    if (rc) { delete m; return rc; }
    *out = m;
    return FVMR_OK;
}

int fvmr_mesh_from_lists(int num_node, const double *x, const double *y, const int *node_ids,
                         int num_elem, const int *elem_types, const int *elem_conn,
                         int num_marker, const char *const *marker_names,
                         const int *marker_offsets, const int *marker_edge_nodes,
                         int equation, fvmr_mesh **out) {
    if (!out || num_node <= 0 || num_elem <= 0 || !x || !y || !node_ids || !elem_types || !elem_conn)
        return set_error(FVMR_ERR_INVALID_ARG, "invalid mesh lists");
    if (num_marker > 0 && (!marker_names || !marker_offsets || !marker_edge_nodes))
        return set_error(FVMR_ERR_INVALID_ARG, "invalid marker lists");
    *out = nullptr;
    fvmr_mesh *m = new fvmr_mesh();
    RawLists &r = m->raw;
    r.x.assign(x, x + num_node); r.y.assign(y, y + num_node);
    r.id1.resize(num_node);
    for (int i = 0; i < num_node; i++) { r.id1[i] = node_ids[i] + 1; r.max_id = std::max(r.max_id, r.id1[i]); }
    for (int i = 0; i < num_elem; i++) {
        const int *c = elem_conn + 4 * (size_t)i;
        if (elem_types[i] != 5 && elem_types[i] != 9) { delete m; return set_error(FVMR_ERR_MESH, "unsupported element type"); }
        add_element(r, elem_types[i], c[0], c[1], c[2], c[3]);
    }
    for (int k = 0; k < num_marker; k++) {
        Marker mk{marker_names[k], {}};
        for (int j = marker_offsets[k]; j < marker_offsets[k + 1]; j++)
            mk.e.push_back({marker_edge_nodes[2 * (size_t)j] + 1, marker_edge_nodes[2 * (size_t)j + 1] + 1});
        r.markers.push_back(mk);
    }
    int rc = build(*m, equation);
    if (rc) { delete m; return rc; }
    *out = m;
    return FVMR_OK;
}

// Families (lengths in g->a..d):
//   PERIODIC_SQUARE: a = x0 (lower-left corner x = y), b = side length
//   WEDGE_CHANNEL:   a = Lx, b = Ly, c = ramp start x0, d = ramp angle (deg)
//   RECT_CHANNEL:    a = Lx, b = Ly
//   CYLINDER_OGRID:  a = cylinder radius, b = farfield radius, c = radial growth ratio
//   NACA_OGRID:      b = farfield radius (chords), c = radial growth ratio, naca_digits
int fvmr_mesh_generate(int family, const fvmr_gen_params *g, int equation, fvmr_mesh **out) {
    if (!g || !out || g->ni <= 0 || g->nj <= 0) return set_error(FVMR_ERR_INVALID_ARG, "invalid generator parameters");
    *out = nullptr;
    fvmr_mesh *m = new fvmr_mesh();
    int rc;
    switch (family) {
    case FVMR_GEN_PERIODIC_SQUARE: rc = gen_periodic_square(*g, m->raw); break;
    case FVMR_GEN_WEDGE_CHANNEL: rc = gen_wedge(*g, m->raw, true); break;
    case FVMR_GEN_RECT_CHANNEL: rc = gen_wedge(*g, m->raw, false); break;
    case FVMR_GEN_CYLINDER_OGRID:
        rc = gen_ogrid(*g, m->raw, nullptr, nullptr, equation == FVMR_EQ_NS && g->reserved == 1 ? "wall_isothermal" : "wall");
        break;
    case FVMR_GEN_NACA_OGRID: {
        std::vector<double> sx, sy;
        naca_surface(g->naca_digits > 0 ? g->naca_digits : 12, g->ni, sx, sy);
        fvmr_gen_params g2 = *g;
        rc = gen_ogrid(g2, m->raw, &sx, &sy, "airfoil");
        break;
    }
    default: rc = set_error(FVMR_ERR_INVALID_ARG, "unknown mesh family");
    }
    if (!rc) rc = build(*m, equation);
    if (rc) { delete m; return rc; }
    *out = m;
    return FVMR_OK;
}

int fvmr_mesh_write_su2(const fvmr_mesh *m, const char *path) {
    if (!m || !path) return set_error(FVMR_ERR_INVALID_ARG, "null argument");
    FILE *f = fopen(path, "w");
    if (!f) return set_error(FVMR_ERR_IO, std::string("cannot open file ") + path);
    const RawLists &r = m->raw;
    fprintf(f, "%% written by fvmr\nNDIME= 2\nNELEM= %zu\n", r.etype.size());
    for (size_t i = 0; i < r.etype.size(); i++) {
        if (r.etype[i] == 5) fprintf(f, "5 %d %d %d %zu\n", r.econn[i][0], r.econn[i][1], r.econn[i][2], i);
        else fprintf(f, "9 %d %d %d %d %zu\n", r.econn[i][0], r.econn[i][1], r.econn[i][2], r.econn[i][3], i);
    }
    fprintf(f, "NPOIN= %zu\n", r.x.size());
    for (size_t i = 0; i < r.x.size(); i++) fprintf(f, "%.17g %.17g %d\n", r.x[i], r.y[i], r.id1[i] - 1);
    fprintf(f, "NMARK= %zu\n", r.markers.size());
    for (auto &mk : r.markers) {
        fprintf(f, "MARKER_TAG= %s\nMARKER_ELEMS= %zu\n", mk.name.c_str(), mk.e.size());
        for (auto &e : mk.e) fprintf(f, "3 %d %d\n", e.first - 1, e.second - 1);
    }
    fclose(f);
    return FVMR_OK;
}

int fvmr_mesh_soa_view(const fvmr_mesh *m, fvmr_mesh_soa *out) {
    if (!m || !out) return set_error(FVMR_ERR_INVALID_ARG, "null argument");
    *out = m->soa;
    return FVMR_OK;
}

int fvmr_mesh_boundary_set(const fvmr_mesh *m, int s, const char **name, int *type, const int **edges, int *num_edges) {
    if (!m || s < 0 || s >= (int)m->sets.size()) return set_error(FVMR_ERR_INVALID_ARG, "boundary set out of range");
    if (name) *name = m->sets[s].name.c_str();
    if (type) *type = m->sets[s].type;
    if (edges) *edges = m->sets[s].edges.data();
    if (num_edges) *num_edges = (int)m->sets[s].edges.size();
    return FVMR_OK;
}

int fvmr_mesh_element_node_ids(const fvmr_mesh *m, const int **ids3) {
    if (!m || !ids3) return set_error(FVMR_ERR_INVALID_ARG, "null argument");
    *ids3 = m->elem_ids3.data();
    return FVMR_OK;
}

void fvmr_mesh_free(fvmr_mesh *m) { delete m; }

// FieldInitializer (FVMR/src/input/FieldInitializer.cpp:129-229)
int fvmr_field_init(const fvmr_mesh_soa *m, const fvmr_params *p, int type, const double *args, double *const U[4]) {
    if (!m || !p || !U) return set_error(FVMR_ERR_INVALID_ARG, "null argument");
    if (type != FVMR_INIT_UNIFORM && !args) return set_error(FVMR_ERR_INVALID_ARG, "initial condition needs args");
    const double gamma = p->gamma;
    auto ruvp2U = [gamma](const double r[4], double Ue[4]) {
        // PhysicalKernel.h:20-27
        Ue[0] = r[0];
        Ue[1] = r[0] * r[1];
        Ue[2] = r[0] * r[2];
        Ue[3] = r[3] / (gamma - 1) + 0.5 * r[0] * (r[1] * r[1] + r[2] * r[2]);
    };
    for (int ie = 0; ie < m->num_element; ie++) {
        double Ue[4];
        const double x = m->element_xy[0][ie], y = m->element_xy[1][ie];
        switch (type) {
        case FVMR_INIT_UNIFORM: ruvp2U(p->inf_ruvp, Ue); break;
        case FVMR_INIT_ISENTROPIC_VORTEX: {
            // setInitialIsentropicVortex (FieldInitializer.cpp:139-177); T_inf is taken as p_inf
            const double xc = args[0], yc = args[1], chi = args[2];
            const double chi2 = chi * chi, ga1 = gamma - 1.0;
            const double two_pi = 2.0 * kPI_REF, pi2 = kPI_REF * kPI_REF;
            const double c_du = chi / two_pi;
            const double c_dT = ga1 * chi2 / (8. * gamma * pi2);
            double dx = x - xc, dy = y - yc;
            double r2 = dx * dx + dy * dy;
            double omr2 = 1.0 - r2;
            double cd = std::exp(0.5 * omr2);
            double du = c_du * cd * (-dy), dv = c_du * cd * dx;
            double u = p->inf_ruvp[1] + du, v = p->inf_ruvp[2] + dv;
            double dT = -c_dT * std::exp(omr2);
            double rho = std::pow(p->inf_ruvp[3] + dT, 1. / ga1);
            double pr = rho * (p->inf_ruvp[3] + dT);
            double r[4] = {rho, u, v, pr};
            ruvp2U(r, Ue);
            break;
        }
        case FVMR_INIT_SHOCK_TUBE: {
            // setInitialShockTube (FieldInitializer.cpp:179-218)
            double sx = args[0], sy = args[1], nx = args[2], ny = args[3];
            if (nx == 0.0 && ny == 0.0) nx = 1.0;
            double li = 1.0 / std::sqrt(nx * nx + ny * ny);
            nx *= li; ny *= li;
            bool inlet = (nx == 1.0) ? (x < sx) : !((x - sx) * nx + (y - sy) * ny > 0);
            ruvp2U(inlet ? p->inlet_ruvp : p->outlet_ruvp, Ue);
            break;
        }
        case FVMR_INIT_DSR: {
            // setInitialDoubleShockReflection + isUpStreamOfShock_forElement (CBoundaryDoubleShockReflect.cpp:17-32)
            double ang = args[2] * kPI_REF / 180.0;
            double nx = -std::sin(ang), ny = std::cos(ang);
            bool up = ((x - args[0]) * nx + (y - args[1]) * ny > 0);
            ruvp2U(up ? p->inlet_ruvp : p->outlet_ruvp, Ue);
            break;
        }
        default: return set_error(FVMR_ERR_INVALID_ARG, "Invalid initialCondition type");
        }
        for (int k = 0; k < 4; k++) U[k][ie] = Ue[k];
    }
    return FVMR_OK;
}

}  // extern "C"
