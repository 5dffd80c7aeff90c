/*
 * fvmr_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
 *
 * A CPU restatement, in FP64, of the reference FVMR explicit residual-and-update
 * path (OliveTIen/CFD_Solver_2D_GPU).  It is the checker for the HIP library:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it, and never as the thing measured or shipped.
 *
 * Every function cites the reference file:line it follows (paths relative to the
 * reference repository).  Where the reference GPU and CPU paths differ, the GPU
 * form is followed (division by -V, farfield states never exit, restriction
 * repairs NaN cells).  Deliberate deviations, all listed with their reasons in
 * the design notes of the repository:
 *   - dt accumulation is the sequential CPU form (the GPU kernel races);
 *   - restriction reads neighbour values of the pre-restriction field (Jacobi),
 *     where the CPU loop is order dependent and the GPU kernel races;
 *   - SSP-RK3 (Shu-Osher) replaces the inconsistent CPU RK3;
 *   - the LSQ 2x2 inverse is the correct one unless compat_inv2x2 is set;
 *   - the Venkatakrishnan limiter, constant viscosity and the isothermal wall
 *     are new (absent from the reference).
 *
 * Parity status of each function is stated in the repository's design notes;
 * functions whose only check is a closed form say so in the tests.
 *
 * Build: see oracle/Makefile (g++ -O2 -ffp-contract=off, no fast-math).  The `fast` target
 * (-O3 -march=native -fopenmp, still without contraction or fast-math) is the CPU baseline of the
 * benchmark: its loops run over cells or edges with every write owned by one iteration, and the
 * time-step sums become per-cell sums over the cell's edges in ascending edge order (the order in
 * which the sequential edge loop adds them), so every thread count gives the same bits.
 */
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>
#include <algorithm>
#include <array>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/fvmr.h"

namespace {

thread_local std::string g_err;
int fail(int code, const std::string &msg) { g_err = msg; return code; }

/* FVMR/src/math/CommonValue.h:16-31 */
constexpr double kPI = 3.1415926535897;
constexpr double kEPS = 1e-10;
constexpr double kRHO_MAX = 300;
constexpr double kRHO_MIN = kEPS;
constexpr double kP_MAX = 400 * 101325.0;
constexpr double kP_MIN = kEPS;
/* FVMR/src/space/viscous_flux/ViscousFluxGPU.h:11 */
constexpr double kS_SUTHERLAND = 110.4;

/* dst = src over n doubles (a loop the OpenMP build shares out; a copy in the plain build) */
inline void copy_n(double *dst, const double *src, int n) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) dst[i] = src[i];
}
inline void copy_field(std::vector<double> *dst, const std::vector<double> *src, int n) {
    for (int k = 0; k < 4; k++) {
        dst[k].resize(n);
        copy_n(dst[k].data(), src[k].data(), n);
    }
}

inline double dmin(double a, double b) { return (a < b) ? a : b; }   /* BasicAlgorithm.h min */
inline double dmax(double a, double b) { return (a > b) ? a : b; }   /* BasicAlgorithm.h max */
inline double dabs(double a) { return (a > 0) ? a : (-a); }          /* BasicAlgorithm.h abs */
inline double distance2D(double x1, double y1, double x2, double y2) { /* BasicAlgorithmGPU.h:32-36 */
    double dx = x2 - x1, dy = y2 - y1;
    return std::sqrt(dx * dx + dy * dy);
}

/* FVMR/src/math/PhysicalKernel.h:11-18 */
inline void U2ruvp(const double U[4], double ruvp[4], double gamma) {
    ruvp[0] = U[0];
    ruvp[1] = U[1] / U[0];
    ruvp[2] = U[2] / U[0];
    ruvp[3] = ruvp[0] * (gamma - 1) * (U[3] / U[0] - (ruvp[1] * ruvp[1] + ruvp[2] * ruvp[2]) * 0.5);
}
/* FVMR/src/math/PhysicalKernel.h:20-27 */
inline void ruvp2U(const double ruvp[4], double U[4], double gamma) {
    U[0] = ruvp[0];
    U[1] = ruvp[0] * ruvp[1];
    U[2] = ruvp[0] * ruvp[2];
    U[3] = ruvp[3] / (gamma - 1) + 0.5 * ruvp[0] * (ruvp[1] * ruvp[1] + ruvp[2] * ruvp[2]);
}
/* FVMR/src/space/restrict/RestrictGPU.h:11-19 */
inline bool outOfRange(const double ruvp[4]) {
    return ruvp[0] < kRHO_MIN || ruvp[0] > kRHO_MAX || ruvp[3] < kP_MIN || ruvp[3] > kP_MAX;
}

/* ====================================================================== */
/* Mesh preprocessing, restating SU2MeshReader + FVM_2D + BoundaryManager_2D */
/* + OldDataConverter with the reference's linear-search edge lookup.       */
/* ====================================================================== */
struct OEdge { int n0, n1, L, R, setID; };
struct OBoundary { std::string name; int type, ID; std::vector<int> edges; };

struct OMesh {
    /* AoS as FVM_2D keeps it */
    std::vector<double> nx_, ny_;
    std::vector<int> node_id;            /* file id + 1 */
    std::vector<int> id2node;            /* pNodeTable */
    std::vector<std::array<int, 3>> elem; /* node ids (1-based), post CCW swap */
    std::vector<double> area, cx, cy;
    std::vector<std::array<int, 3>> pEdges;
    std::vector<OEdge> edges;
    std::vector<OBoundary> bsets;
    struct PP { int bType, s0, s1; };
    std::vector<PP> pairs;
    /* SoA as OldDataConverter produces it */
    std::vector<double> node_x, node_y;
    std::vector<int> en[4], ee[4], enb[4];
    std::vector<double> ex, ey, evol;
    std::vector<int> gn0, gn1, gset, gpair, gL, gR;
    std::vector<double> glen, gx, gy, gnx, gny;
    std::vector<int> btype;
    std::vector<int> elem_ids3;           /* post-swap node ids, 3 per element */
    fvmr_mesh_soa soa{};
};

/* BoundaryManager::boundaryNameToType (FVMR/src/boundary_condition/BoundaryManager.cpp:15-45) */
int boundaryNameToType(const std::string &n, int equation, bool &ok) {
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
    if (n.substr(0, 8) == "periodic") {
        char c = n[n.size() - 1];
        if (c < '0' || c > '9') { ok = false; return -1; }
        return FVMR_BC_PERIODIC_0 + (c - '0');
    }
    if (n == "wall_isothermal") return FVMR_BC_WALL_ISOTHERMAL;  /* new name, no reference counterpart */
    ok = false;
    return -1;
}

/* FVM_2D::getEdgeByNodeIDs (FVMR/src/legacy/FVM_2D.cpp:105-114): linear search */
int getEdgeByNodeIDs(const OMesh &m, int n0, int n1) {
    for (size_t i = 0; i < m.edges.size(); i++) {
        if (m.edges[i].n0 == n0 && m.edges[i].n1 == n1) return (int)i;
        if (m.edges[i].n0 == n1 && m.edges[i].n1 == n0) return (int)i;
    }
    return -1;
}

struct RawMarker { std::string name; std::vector<std::pair<int, int>> e; };

int buildMesh(OMesh &m, const std::vector<double> &xs, const std::vector<double> &ys,
              const std::vector<int> &ids1, int maxNodeID,
              const std::vector<std::array<int, 3>> &elems, const std::vector<RawMarker> &markers,
              int equation) {
    const int nv = (int)xs.size();
    if (nv == 0) return fail(FVMR_ERR_MESH, "null nodes");
    if (elems.empty()) return fail(FVMR_ERR_MESH, "null elements");
    m.nx_ = xs; m.ny_ = ys; m.node_id = ids1;
    /* FVM_2D::iniPNodeTable (FVM_2D.cpp:25-34) */
    m.id2node.assign(maxNodeID + 1, -1);
    for (int i = 0; i < nv; i++) {
        if (ids1[i] < 0 || ids1[i] > maxNodeID) return fail(FVMR_ERR_MESH, "bad node id");
        m.id2node[ids1[i]] = i;
    }
    auto node = [&](int id) -> int {
        if (id < 0 || id >= (int)m.id2node.size() || m.id2node[id] < 0) return -1;
        return m.id2node[id];
    };
    m.elem = elems;
    const int ne = (int)elems.size();
    m.area.resize(ne);
    for (int ie = 0; ie < ne; ie++)
        for (int k = 0; k < 3; k++)
            if (node(m.elem[ie][k]) < 0) return fail(FVMR_ERR_MESH, "element references unknown node");
    /* FVM_2D::iniEdges (FVM_2D.cpp:36-82): CCW fix in double, then edge registration */
    for (int ie = 0; ie < ne; ie++) {
        double xn[3], yn[3];
        for (int i = 0; i < 3; i++) { xn[i] = m.nx_[node(m.elem[ie][i])]; yn[i] = m.ny_[node(m.elem[ie][i])]; }
        double a = 0.5 * (xn[0] * (yn[1] - yn[2]) + xn[1] * (yn[2] - yn[0]) + xn[2] * (yn[0] - yn[1]));
        if (a < 0) { m.area[ie] = -a; std::swap(m.elem[ie][1], m.elem[ie][2]); }
        else m.area[ie] = a;
    }
    /* FVM_2D::iniEdges_registerSingle (FVM_2D.cpp:116-134) */
    auto reg = [&](int n0, int n1, int ie) {
        int e = getEdgeByNodeIDs(m, n0, n1);
        if (e < 0) m.edges.push_back(OEdge{n0, n1, ie, -1, -1});
        else m.edges[e].R = ie;
    };
    for (int ie = 0; ie < ne; ie++) {
        int n0 = m.elem[ie][0], n1 = m.elem[ie][1], n2 = m.elem[ie][2];
        reg(n0, n1, ie); reg(n1, n2, ie); reg(n2, n0, ie);
    }
    /* FVM_2D::iniElement_xy_pEdges (FVM_2D.cpp:147-183) */
    m.cx.resize(ne); m.cy.resize(ne); m.pEdges.resize(ne);
    for (int ie = 0; ie < ne; ie++) {
        double sx = 0, sy = 0;
        for (int i = 0; i < 3; i++) { sx += m.nx_[node(m.elem[ie][i])]; sy += m.ny_[node(m.elem[ie][i])]; }
        m.cx[ie] = sx / 3.0; m.cy[ie] = sy / 3.0;
        m.pEdges[ie][0] = getEdgeByNodeIDs(m, m.elem[ie][0], m.elem[ie][1]);
        m.pEdges[ie][1] = getEdgeByNodeIDs(m, m.elem[ie][1], m.elem[ie][2]);
        m.pEdges[ie][2] = getEdgeByNodeIDs(m, m.elem[ie][2], m.elem[ie][0]);
    }
    const int nf = (int)m.edges.size();
    /* FVM_2D::iniEdges_lengths (FVM_2D.cpp:186-196) with Edge_2D::getLength (Edge_2D.cpp:19-30):
       the length is rounded through float even when myfloat is double. */
    m.glen.resize(nf);
    for (int i = 0; i < nf; i++) {
        double x0 = m.nx_[node(m.edges[i].n0)], y0 = m.ny_[node(m.edges[i].n0)];
        double x1 = m.nx_[node(m.edges[i].n1)], y1 = m.ny_[node(m.edges[i].n1)];
        m.glen[i] = (double)(float)std::sqrt((x0 - x1) * (x0 - x1) + (y0 - y1) * (y0 - y1));
    }
    /* SU2MeshReader::process boundary part (SU2MeshReader.cpp:201-247) */
    m.bsets.resize(markers.size());
    for (size_t i = 0; i < markers.size(); i++) {
        bool ok;
        m.bsets[i].name = markers[i].name;
        m.bsets[i].type = boundaryNameToType(markers[i].name, equation, ok);
        if (!ok) return fail(FVMR_ERR_MESH, "unknown boundary type: " + markers[i].name);
        m.bsets[i].ID = (int)i + 1;
        for (auto &pr : markers[i].e) {
            int e = getEdgeByNodeIDs(m, pr.first, pr.second);
            if (e < 0) return fail(FVMR_ERR_MESH, "boundary edge not found");
            m.edges[e].setID = m.bsets[i].ID;
            m.bsets[i].edges.push_back(e);
        }
        int bType = m.bsets[i].type, bID = m.bsets[i].ID;
        if (FVMR_BC_PERIODIC_0 <= bType && bType <= FVMR_BC_PERIODIC_0 + 9) {
            int idx = -1;
            for (size_t k = 0; k < m.pairs.size(); k++) if (m.pairs[k].bType == bType) idx = (int)k;
            if (idx != -1) m.pairs[idx].s1 = bID;
            else m.pairs.push_back({bType, bID, -1});
        }
    }
    /* BoundaryManager_2D::checkPeriodPairs (BoundaryManager_2D.cpp:183-255) */
    auto mid_x = [&](int e) { return (m.nx_[node(m.edges[e].n0)] + m.nx_[node(m.edges[e].n1)]) / 2.0; };
    auto mid_y = [&](int e) { return (m.ny_[node(m.edges[e].n0)] + m.ny_[node(m.edges[e].n1)]) / 2.0; };
    for (auto &pp : m.pairs) {
        if (pp.s0 == -1 || pp.s1 == -1) return fail(FVMR_ERR_MESH, "periodic boundary not paired");
        const auto &b0 = m.bsets[pp.s0 - 1], &b1 = m.bsets[pp.s1 - 1];
        if (b0.edges.size() != b1.edges.size()) return fail(FVMR_ERR_MESH, "periodic edge counts differ");
        const int n = (int)b0.edges.size();
        if (n == 0) continue;
        double tx = mid_x(b1.edges[n - 1]) - mid_x(b0.edges[0]);
        double ty = mid_y(b1.edges[n - 1]) - mid_y(b0.edges[0]);
        for (int ie = 0; ie < n; ie++) {
            double x1r = mid_x(b0.edges[ie]) + tx, y1r = mid_y(b0.edges[ie]) + ty;
            double x1 = mid_x(b1.edges[n - 1 - ie]), y1 = mid_y(b1.edges[n - 1 - ie]);
            double rx = x1 - x1r, ry = y1 - y1r;
            double rs = rx * rx + ry * ry, ts = tx * tx + ty * ty;
            if (rs / ts >= 0.0001) return fail(FVMR_ERR_MESH, "periodic boundary not translatable");
        }
    }
    /* OldDataConverter::ConvertNode (OldDataConverter.cpp:35-43) */
    m.node_x = m.nx_; m.node_y = m.ny_;
    /* OldDataConverter::ConvertElement (OldDataConverter.cpp:45-97) */
    for (int k = 0; k < 4; k++) { m.en[k].assign(ne, -1); m.ee[k].assign(ne, -1); m.enb[k].assign(ne, -1); }
    m.ex = m.cx; m.ey = m.cy; m.evol = m.area;
    m.elem_ids3.resize(3 * (size_t)ne);
    for (int i = 0; i < ne; i++) {
        for (int j = 0; j < 3; j++) {
            m.en[j][i] = node(m.elem[i][j]);
            m.ee[j][i] = m.pEdges[i][j];
            m.elem_ids3[3 * (size_t)i + j] = m.elem[i][j];
        }
        /* Element_2D::findNeighbor (Element_2D.cpp:19-48): AoS pElement_R is null for periodic edges */
        for (int j = 0; j < 3; j++) {
            int e = m.pEdges[i][j];
            if (e < 0) continue;
            if (m.edges[e].L == i) m.enb[j][i] = m.edges[e].R;
            else if (m.edges[e].R == i) m.enb[j][i] = m.edges[e].L;
        }
    }
    /* OldDataConverter::ConvertEdge (OldDataConverter.cpp:99-136) */
    m.gn0.resize(nf); m.gn1.resize(nf); m.gset.resize(nf); m.gpair.assign(nf, -1);
    m.gL.resize(nf); m.gR.resize(nf); m.gx.resize(nf); m.gy.resize(nf); m.gnx.resize(nf); m.gny.resize(nf);
    for (int i = 0; i < nf; i++) {
        m.gn0[i] = node(m.edges[i].n0); m.gn1[i] = node(m.edges[i].n1);
        m.gset[i] = m.edges[i].setID;
        m.gL[i] = m.edges[i].L; m.gR[i] = m.edges[i].R;
        int a = m.gn0[i], b = m.gn1[i];
        m.gx[i] = (m.node_x[a] + m.node_x[b]) / 2.0;
        m.gy[i] = (m.node_y[a] + m.node_y[b]) / 2.0;
        double dx = m.node_x[b] - m.node_x[a], dy = m.node_y[b] - m.node_y[a];
        m.gnx[i] = dy / m.glen[i];
        m.gny[i] = -dx / m.glen[i];
    }
    /* OldDataConverter::ConvertBoundary (OldDataConverter.cpp:138-177) with
       BoundaryManager_2D::get_pairEdge_periodic (BoundaryManager_2D.cpp:304-323) */
    for (auto &bs : m.bsets) {
        if (!(FVMR_BC_PERIODIC_0 <= bs.type && bs.type <= FVMR_BC_PERIODIC_0 + 9)) continue;
        int other = -1;
        for (auto &pp : m.pairs) if (pp.bType == bs.type) { other = (pp.s0 == bs.ID) ? pp.s1 : pp.s0; break; }
        if (other < 0) return fail(FVMR_ERR_MESH, "pair not found");
        const auto &b1 = m.bsets[other - 1];
        const int n = (int)bs.edges.size();
        for (int e : bs.edges) {
            int idx0 = -1;
            const auto &own = m.bsets[m.edges[e].setID - 1];
            for (int k = 0; k < (int)own.edges.size(); k++) if (own.edges[k] == e) { idx0 = k; break; }
            if (idx0 < 0) return fail(FVMR_ERR_MESH, "edge not found in its set");
            int pe = b1.edges[n - 1 - idx0];
            m.gR[e] = m.gL[pe];
            m.gpair[e] = pe;
        }
    }
    m.btype.resize(m.bsets.size());
    for (size_t i = 0; i < m.bsets.size(); i++) m.btype[i] = m.bsets[i].type;

    fvmr_mesh_soa &s = m.soa;
    s.num_node = nv; s.num_element = ne; s.num_edge = nf; s.num_boundary_set = (int)m.bsets.size();
    s.node_xy[0] = m.node_x.data(); s.node_xy[1] = m.node_y.data();
    for (int k = 0; k < 4; k++) { s.element_nodes[k] = m.en[k].data(); s.element_edges[k] = m.ee[k].data(); s.element_neighbors[k] = m.enb[k].data(); }
    s.element_xy[0] = m.ex.data(); s.element_xy[1] = m.ey.data(); s.element_volume = m.evol.data();
    s.edge_nodes[0] = m.gn0.data(); s.edge_nodes[1] = m.gn1.data();
    s.edge_setID = m.gset.data(); s.edge_periodicPair = m.gpair.data();
    s.edge_elementL = m.gL.data(); s.edge_elementR = m.gR.data(); s.edge_length = m.glen.data();
    s.edge_xy[0] = m.gx.data(); s.edge_xy[1] = m.gy.data();
    s.edge_normal[0] = m.gnx.data(); s.edge_normal[1] = m.gny.data();
    s.boundary_type = m.btype.data();
    return FVMR_OK;
}

/* StringProcessor::replaceCharInString + splitString as used by SU2MeshReader::readMesh */
std::vector<std::string> su2_words(const std::string &line) {
    std::string t;
    for (char c : line) { if (c == '=') t += " = "; else t += c; }
    std::vector<std::string> w;
    std::istringstream is(t);
    std::string x;
    while (is >> x) w.push_back(x);
    return w;
}

/* SU2MeshReader::readMesh (FVMR/src/input/SU2MeshReader.cpp:21-155) + scaleMesh (:157-173) */
int readSU2(const char *path, int equation, double scale, OMesh &m) {
    std::ifstream in(path);
    if (!in) return fail(FVMR_ERR_IO, std::string("cannot open file ") + path);
    enum { S0, SNDIME, SNPOIN, SNELEM, SNMARK } st = S0;
    std::vector<double> xs, ys; std::vector<int> ids; int maxNode = 1;
    std::vector<std::array<int, 3>> el;
    std::vector<RawMarker> mk;
    std::string line;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        auto w = su2_words(line);
        if (w.empty() || w[0] == "%") continue;
        if (w[0] == "NDIME") st = SNDIME;
        else if (w[0] == "NPOIN") st = SNPOIN;
        else if (w[0] == "NELEM") st = SNELEM;
        else if (w[0] == "NMARK") st = SNMARK;
        if (st == SNDIME) {
            if (!(w.size() == 3 && w[2] == "2")) return fail(FVMR_ERR_MESH, "SU2 mesh dimension is not 2D");
        } else if (st == SNPOIN && w[0][0] != 'N') {
            if (w.size() < 3) return fail(FVMR_ERR_MESH, "node line without id");
            xs.push_back(std::stod(w[0])); ys.push_back(std::stod(w[1]));
            int id = (w.size() == 3) ? (int)std::stod(w[2]) + 1 : (int)std::stod(w[3]) + 1;
            ids.push_back(id); maxNode = std::max(maxNode, id);
        } else if (st == SNELEM && w[0][0] != 'N') {
            if (w[0] == "5") {
                el.push_back({(int)std::stod(w[1]) + 1, (int)std::stod(w[2]) + 1, (int)std::stod(w[3]) + 1});
            } else if (w[0] == "9") {
                int a = (int)std::stod(w[1]) + 1, b = (int)std::stod(w[2]) + 1;
                int c = (int)std::stod(w[3]) + 1, d = (int)std::stod(w[4]) + 1;
                el.push_back({a, b, d});
                el.push_back({d, b, c});
            }
        } else if (st == SNMARK) {
            if (w[0] == "NMARK" || w[0] == "MARKER_ELEMS") {
            } else if (w[0] == "MARKER_TAG") {
                mk.push_back(RawMarker{w.size() > 2 ? w[2] : std::string(), {}});
            } else {
                if (mk.empty() || w.size() < 3) return fail(FVMR_ERR_MESH, "marker edge before MARKER_TAG");
                mk.back().e.push_back({std::stoi(w[1]) + 1, std::stoi(w[2]) + 1});
            }
        }
    }
    double res = scale - 1.0;
    if (!(res > -0.01 && res < 0.01)) {
        if (scale < 0) return fail(FVMR_ERR_INVALID_ARG, "mesh scale factor should be positive");
        for (auto &x : xs) x *= scale;
        for (auto &y : ys) y *= scale;
    }
    return buildMesh(m, xs, ys, ids, maxNode, el, mk, equation);
}

/* ====================================================================== */
/* Solver                                                                   */
/* ====================================================================== */
struct OSolver {
    int nv, ne, nf, nb;
    std::vector<double> node_x, node_y;
    std::vector<int> en[4], ee[4], enb[4];
    std::vector<double> ex, ey, evol;
    std::vector<int> gset, gpair, gL, gR;
    std::vector<double> glen, gx, gy, gnx, gny;
    std::vector<int> btype;
    fvmr_params p;
    std::vector<double> U[4], Ux[4], Uy[4], Flux[4], ruvp[4], Uold[4], alphaC, eFlux[4];
    double last_dt = 0;
    int n_restricted = 0;
    std::vector<int> dt_off, dt_edge;     /* per cell, the edges adding to its alphaC, ascending */
    std::vector<double> e_alpha, e_aV;    /* per edge: alpha and the viscous term */
};

/* FVMR/src/space/FluxGPU.cu:34-61: linear reconstruction with out-of-range fallback */
void reconstruct(const OSolver &s, const double *const *U, double x, double y, int ie, double Ud[4]) {
    for (int i = 0; i < 4; i++) {
        if (s.p.reconstruct == FVMR_REC_LINEAR)
            Ud[i] = U[i][ie] + s.Ux[i][ie] * (x - s.ex[ie]) + s.Uy[i][ie] * (y - s.ey[ie]);
        else
            Ud[i] = U[i][ie];   /* constant, and the reference fallback for MUSCL (FluxGPU.cu:41-44) */
    }
    double r[4];
    U2ruvp(Ud, r, s.p.gamma);
    if (outOfRange(r)) for (int i = 0; i < 4; i++) Ud[i] = U[i][ie];
}

/* FVMR/src/space/FluxGPU.cu:63-77 */
void ghost_slip(const double UL[4], double UR[4], double nx, double ny) {
    double uxL = UL[1] / UL[0], uyL = UL[2] / UL[0];
    double unL = uxL * nx + uyL * ny;
    double utL = uxL * ny - uyL * nx;
    double unR = -unL, utR = utL;
    double uxR = unR * nx + utR * ny;
    double uyR = unR * ny - utR * nx;
    UR[0] = UL[0];
    UR[1] = UR[0] * uxR;
    UR[2] = UR[0] * uyR;
    UR[3] = UL[3];
}
/* FVMR/src/space/FluxGPU.cu:79-85: U_R starts zeroed, so the momenta stay 0 */
void ghost_adiabatic(const double UL[4], double UR[4]) {
    UR[0] = UL[0];
    UR[1] = -0.0;   /* -U_R[1] of a zeroed U_R */
    UR[2] = -0.0;
    UR[3] = UL[3];
}
/* New: no-slip isothermal wall ghost: u_R = -u_L, p_R = p_L, T_R = 2 T_w - T_L (face T = T_w). */
void ghost_isothermal(const double UL[4], double UR[4], double Tw, double R, double gamma) {
    double r[4];
    U2ruvp(UL, r, gamma);
    double TL = r[3] / (r[0] * R);
    double TR = 2.0 * Tw - TL;
    if (TR < 0.5 * Tw) TR = 0.5 * Tw;
    double g[4] = {r[3] / (R * TR), -r[1], -r[2], r[3]};
    ruvp2U(g, UR, gamma);
}
/* FVMR/src/space/FluxGPU.cu:87-167 (GPU form: out-of-range only reported) */
void ghost_farfield(const double UL[4], const double ruvp_inf[4], double UR[4], double nx, double ny, double gamma) {
    double ga1 = gamma - 1;
    double rho_f = ruvp_inf[0], u_f = ruvp_inf[1], v_f = ruvp_inf[2], p_f = ruvp_inf[3];
    double e[4];
    U2ruvp(UL, e, gamma);
    double rho_e = e[0], u_e = e[1], v_e = e[2], p_e = e[3];
    double a2_f = gamma * p_f / rho_f;
    double af = std::sqrt(a2_f);
    double ae = std::sqrt(gamma * p_e / rho_e);
    double qnf = u_f * nx + v_f * ny;
    double qne = u_e * nx + v_e * ny;
    double rf = qnf - 2.0 * af / ga1;
    double re = qne + 2.0 * ae / ga1;
    double qn = 0.5 * (re + rf);
    double as = ga1 * (re - rf) / 4.0;
    double rho0 = 1, u0 = 0, v0 = 0, p0 = 1;
    if (qn <= -as) {
        rho0 = rho_f; u0 = u_f; v0 = v_f; p0 = p_f;
    } else if (qn < 0) {
        double son = p_f / std::pow(rho_f, gamma);
        double qtx = u_f - qnf * nx, qty = v_f - qnf * ny;
        rho0 = std::pow((as * as / son / gamma), 1.0 / ga1);
        u0 = qtx + qn * nx; v0 = qty + qn * ny;
        p0 = as * as * rho0 / gamma;
    } else if (qn < as) {
        double son = p_e / std::pow(rho_e, gamma);
        double qtx = u_e - qne * nx, qty = v_e - qne * ny;
        rho0 = std::pow((as * as / son / gamma), 1.0 / ga1);
        u0 = qtx + qn * nx; v0 = qty + qn * ny;
        p0 = as * as * rho0 / gamma;
    } else {
        rho0 = rho_e; u0 = u_e; v0 = v_e; p0 = p_e;
    }
    double r0[4] = {rho0, u0, v0, p0};
    ruvp2U(r0, UR, gamma);
}

/* FVMR/src/space/convection/LocalLaxFriedrichsGPU.cu:4-49 */
void llf(const double UL[4], const double UR[4], double nx, double ny, double length, double flux[4], double gamma) {
    double L[4], R[4];
    U2ruvp(UL, L, gamma); U2ruvp(UR, R, gamma);
    double rL = L[0], uL = L[1], vL = L[2], pL = L[3];
    double rR = R[0], uR = R[1], vR = R[2], pR = R[3];
    double unL = nx * uL + ny * vL, unR = nx * uR + ny * vR;
    double FnL[4], FnR[4];
    FnL[0] = rL * unL; FnL[1] = rL * unL * uL + nx * pL; FnL[2] = rL * unL * vL + ny * pL;
    double hl = pL * gamma / (rL * (gamma - 1.)) + 0.5 * (uL * uL + vL * vL);
    FnL[3] = rL * unL * hl;
    FnR[0] = rR * unR; FnR[1] = rR * unR * uR + nx * pR; FnR[2] = rR * unR * vR + ny * pR;
    double hr = pR * gamma / (rR * (gamma - 1.)) + 0.5 * (uR * uR + vR * vR);
    FnR[3] = rR * unR * hr;
    double lam = dmax(std::fabs(unL) + std::sqrt(gamma * pL / rL), std::fabs(unR) + std::sqrt(gamma * pR / rR));
    for (int i = 0; i < 4; i++) {
        flux[i] = 0.5 * (FnL[i] + FnR[i] - lam * (UR[i] - UL[i]));
        flux[i] *= length;
    }
}

/* FVMR/src/space/convection/RoeGPU.cu:4-159 (no entropy fix, Etm as written) */
void roe(const double UL[4], const double UR[4], double nx, double ny, double area, double F[4], double gamma) {
    double L[4], R[4];
    U2ruvp(UL, L, gamma); U2ruvp(UR, R, gamma);
    double rL = L[0], uL = L[1], vL = L[2], pL = L[3];
    double rR = R[0], uR = R[1], vR = R[2], pR = R[3];
    double ulnormaln = (nx * uL + ny * vL + 0.0);
    double urnormaln = (nx * uR + ny * vR + 0.0);
    double rulnormaln = rL * ulnormaln, rurnormaln = rR * urnormaln;
    double V2L = uL * uL + vL * vL, V2R = uR * uR + vR * vR;
    double ga1 = gamma - 1.0;
    double HL = pL * gamma / (rL * ga1) + 0.5 * V2L;
    double HR = pR * gamma / (rR * ga1) + 0.5 * V2R;
    F[0] = 0.5 * (rulnormaln + rurnormaln) * area;
    F[1] = 0.5 * (rulnormaln * uL + rurnormaln * uR + nx * (pL + pR)) * area;
    F[2] = 0.5 * (rulnormaln * vL + rurnormaln * vR + ny * (pL + pR)) * area;
    F[3] = 0.5 * (rulnormaln * HL + rurnormaln * HR - 0.0 * (pL + pR)) * area;
    /* RoeDissapationTerm2d (RoeGPU.cu:66-159) */
    double rrorl = std::sqrt(rR / rL);
    double rrorlp1 = rrorl + 1.0;
    double rm = std::sqrt(rR * rL);
    double um = (uL + uR * rrorl) / rrorlp1;
    double vm = (vL + vR * rrorl) / rrorlp1;
    double Hm = (HL + HR * rrorl) / rrorlp1;
    double Vm2 = um * um + vm * vm;
    double am2 = ga1 * (Hm - 0.5 * Vm2);
    double am = std::sqrt(am2);
    double unormaln = um * nx + vm * ny;
    double eig1 = std::fabs(unormaln);
    double eig2 = std::fabs(unormaln + am);
    double eig3 = std::fabs(unormaln - am);
    double ER = HR - pR / rR, EL = HL - pL / rL;
    double Etm = Hm / gamma + ga1 / gamma * 0.5 * Vm2;
    double dW[4];
    dW[0] = rR - rL;
    dW[1] = rm * (uR - uL) + dW[0] * um;
    dW[2] = rm * (vR - vL) + dW[0] * vm;
    dW[3] = rm * (ER - EL) + dW[0] * Etm;
    double astar = (eig2 + eig3) / 2.0;
    double Mstar = (eig2 - eig3) / 2.0 / am;
    double dunormaln = (nx * (uR - uL) + ny * (vR - vL));
    double dUroe = Mstar * dunormaln + (astar - eig1) * (pR - pL) / rm / am2;
    double dProe = Mstar * (pR - pL) + (astar - eig1) * rm * dunormaln;
    double d0 = 0.5 * (eig1 * dW[0] + dUroe * rm) * area;
    double d1 = 0.5 * (eig1 * dW[1] + dUroe * rm * um + dProe * nx) * area;
    double d2 = 0.5 * (eig1 * dW[2] + dUroe * rm * vm + dProe * ny) * area;
    double d3 = 0.5 * (eig1 * dW[3] + dUroe * rm * Hm + dProe * (unormaln - 0.0)) * area;
    F[0] -= 1.0 * d0; F[1] -= 1.0 * d1; F[2] -= 1.0 * d2; F[3] -= 1.0 * d3;
}

/* FVMR/src/space/FluxGPU.cu:190-205 */
void riemann(const OSolver &s, const double UL[4], const double UR[4], double nx, double ny, double len, double F[4]) {
    if (s.p.flux_scheme == FVMR_SOL_LLF) llf(UL, UR, nx, ny, len, F, s.p.gamma);
    else roe(UL, UR, nx, ny, len, F, s.p.gamma);
}

inline bool is_periodic(int t) { return FVMR_BC_PERIODIC_0 <= t && t <= FVMR_BC_PERIODIC_0 + 9; }

/* edge_convection_flux_device_kernel (FVMR/src/space/FluxGPU.cu:207-283) */
void edge_flux(const OSolver &s, const double *const *U, double t_rk, int e, double F[4]) {
    int setID = s.gset[e];
    int bType = (setID != -1) ? s.btype[setID - 1] : -1;
    double nx = s.gnx[e], ny = s.gny[e], xe = s.gx[e], ye = s.gy[e], len = s.glen[e];
    int L = s.gL[e], R = s.gR[e];
    double right = s.p.dsr_shock_x + (ye + s.p.dsr_shock_speed * t_rk) / std::sqrt(3.0);
    bool upstream = xe < right;
    double UL[4] = {0, 0, 0, 0}, UR[4] = {0, 0, 0, 0};
    reconstruct(s, U, xe, ye, L, UL);
    switch (bType) {
    case FVMR_BC_SYMMETRY:
    case FVMR_BC_WALL_NONVISCOUS: ghost_slip(UL, UR, nx, ny); break;
    case FVMR_BC_WALL_ADIABAT: ghost_adiabatic(UL, UR); break;
    case FVMR_BC_WALL_ISOTHERMAL: ghost_isothermal(UL, UR, s.p.wall_temperature, s.p.R, s.p.gamma); break;
    case FVMR_BC_INLET: ghost_farfield(UL, s.p.inlet_ruvp, UR, nx, ny, s.p.gamma); break;
    case FVMR_BC_OUTLET: ghost_farfield(UL, s.p.outlet_ruvp, UR, nx, ny, s.p.gamma); break;
    case FVMR_BC_INF: ghost_farfield(UL, s.p.inf_ruvp, UR, nx, ny, s.p.gamma); break;
    case FVMR_BC_DSR:
        if (upstream) ghost_farfield(UL, s.p.inlet_ruvp, UR, nx, ny, s.p.gamma);
        else ghost_farfield(UL, s.p.outlet_ruvp, UR, nx, ny, s.p.gamma);
        break;
    default:
        if (R != -1) {
            /* get_UR_inner_and_periodic_device (FluxGPU.cu:169-188) */
            if (setID != -1) {
                int pe = s.gpair[e];
                if (pe >= 0 && pe < s.nf) reconstruct(s, U, s.gx[pe], s.gy[pe], R, UR);
            } else {
                reconstruct(s, U, xe, ye, R, UR);
            }
        }
    }
    riemann(s, UL, UR, nx, ny, len, F);
}

/* Sutherland helpers (FVMR/src/space/viscous_flux/ViscousFluxGPU.h:26-32) */
inline double sutherland_C1(double mu0, double T0) { return mu0 / std::pow(T0, 1.5) * (T0 + kS_SUTHERLAND); }
inline double mu_of_T(const OSolver &s, double T) {
    if (s.p.viscosity_model == FVMR_MU_CONSTANT) return s.p.mu0;
    return sutherland_C1(s.p.mu0, s.p.T0) * std::pow(T, 1.5) / (T + 110.4);
}

/* viscous_flux_device_kernel (FVMR/src/space/viscous_flux/ViscousFluxGPU.cu:139-229) */
void edge_viscous(const OSolver &s, const double *const *U, int e, double F[4]) {
    const int L = s.gL[e], R = s.gR[e];
    const double ga1 = s.p.gamma - 1.0;
    const double Cv = s.p.R / ga1;
    const double Cp = Cv * s.p.gamma;
    bool hasR = !(R < 0 || R >= s.ne);
    double gx[4], gy[4], Um[4];
    for (int k = 0; k < 4; k++) {
        if (hasR) {
            gx[k] = 0.5 * (s.Ux[k][L] + s.Ux[k][R]);
            gy[k] = 0.5 * (s.Uy[k][L] + s.Uy[k][R]);
            Um[k] = 0.5 * (U[k][L] + U[k][R]);
        } else {
            gx[k] = s.Ux[k][L]; gy[k] = s.Uy[k][L]; Um[k] = U[k][L];
        }
    }
    /* get_physical_variable_and_gradient_host_device (ViscousFluxGPU.cu:97-137) */
    double one_on_rho = 1.0 / Um[0];
    double u = Um[1] * one_on_rho, v = Um[2] * one_on_rho, E = Um[3] * one_on_rho;
    double dudx = one_on_rho * (gx[1] - u * gx[0]);
    double dvdx = one_on_rho * (gx[2] - v * gx[0]);
    double dudy = one_on_rho * (gy[1] - u * gy[0]);
    double dvdy = one_on_rho * (gy[2] - v * gy[0]);
    double dEdx = one_on_rho * (gx[3] - E * gx[0]);
    double dEdy = one_on_rho * (gy[3] - E * gy[0]);
    double one_on_Cv = 1.0 / Cv;
    double V2 = u * u + v * v;
    double T = one_on_Cv * (E - 0.5 * V2);
    double dTdx = (dEdx - u * dudx - v * dvdx) * one_on_Cv;
    double dTdy = (dEdy - u * dudy - v * dvdy) * one_on_Cv;
    double mu = mu_of_T(s, T);
    /* get_Reynolds_stress_2d_device_inline (ViscousFluxGPU.cu:126-136) */
    const double NMD = -2.0 / 3.0;
    double div = dudx + dvdy;
    double txx = mu * (NMD * div + 2.0 * dudx);
    double tyy = mu * (NMD * div + 2.0 * dvdy);
    double txy = mu * (dudy + dvdx);
    double akmu = Cp * (mu / s.p.Pr_viscous);
    double Phix = akmu * dTdx + u * txx + v * txy;
    double Phiy = akmu * dTdy + u * txy + v * tyy;
    double nx = s.gnx[e], ny = s.gny[e], area = s.glen[e];
    F[0] = 0.0;
    F[1] = (txx * nx + txy * ny) * area;
    F[2] = (txy * nx + tyy * ny) * area;
    F[3] = (Phix * nx + Phiy * ny) * area;
}

/* GradientGreenGaussKernel (FVMR/src/space/gradient/GradientGPU.cu:223-425) */
void gradient_gg(OSolver &s, const double *const *U) {
#pragma omp parallel for schedule(static)
    for (int c = 0; c < s.ne; c++) {
        double volumeC = s.evol[c];
        int eid[4], etype[4], esign[4] = {0, 0, 0, 0}, out[4] = {-1, -1, -1, -1};
        for (int i = 0; i < 4; i++) {
            eid[i] = s.ee[i][c];
            if (eid[i] <= -1 || eid[i] >= s.nf) eid[i] = -1;
            etype[i] = (eid[i] == -1) ? -1 : s.gset[eid[i]];
        }
        for (int i = 0; i < 4; i++) {
            if (eid[i] == -1) continue;
            if (c != s.gR[eid[i]]) { esign[i] = 1; out[i] = s.gR[eid[i]]; }
            else { esign[i] = -1; out[i] = s.gL[eid[i]]; }
        }
        double Sx[4] = {0, 0, 0, 0}, Sy[4] = {0, 0, 0, 0}, gC[4] = {0, 0, 0, 0};
        for (int i = 0; i < 4; i++) {
            int e = eid[i];
            if (e == -1) continue;
            Sx[i] = s.gnx[e] * s.glen[e];
            Sy[i] = s.gny[e] * s.glen[e];
            if (out[i] != -1) {
                double fx = s.gx[e], fy = s.gy[e], Cx = s.ex[c], Cy = s.ey[c];
                double Fx = s.ex[out[i]], Fy = s.ey[out[i]];
                if (etype[i] != -1) {
                    int pe = s.gpair[e];
                    double shx = fx - s.gx[pe], shy = fy - s.gy[pe];
                    Fx += shx; Fy += shy;
                }
                double dFf = distance2D(Fx, Fy, fx, fy);
                double dFC = distance2D(Fx, Fy, Cx, Cy);
                if (dFC < kEPS) dFC = kEPS;
                gC[i] = dFf / dFC;
            } else {
                gC[i] = 0.5;
            }
        }
        for (int j = 0; j < 4; j++) {
            double sum0 = 0, sum1 = 0;
            for (int i = 0; i < 4; i++) {
                if (eid[i] == -1) continue;
                double phiC = U[j][c];
                double phif;
                if (out[i] != -1) {
                    double phiF = U[j][out[i]];
                    double g = gC[i], g1 = 1.0 - g;
                    phif = g * phiC + g1 * phiF;
                } else {
                    phif = phiC;
                }
                sum0 += phif * Sx[i] * esign[i];
                sum1 += phif * Sy[i] * esign[i];
            }
            s.Ux[j][c] = 1 / volumeC * sum0;
            s.Uy[j][c] = 1 / volumeC * sum1;
        }
    }
}

/* inv_2x2 (FVMR/src/math/MatrixKernel.h:45-59); compat reproduces the determinant bug */
void inv2x2(double *m, bool compat) {
    double det = m[0] * m[3] - m[1] * m[2];
    if (compat) {
        if (det >= 0 && det < kEPS) det = kEPS;
        else if (det > -kEPS) det = -kEPS;
    } else {
        if (det >= 0 && det < kEPS) det = kEPS;
        else if (det < 0 && det > -kEPS) det = -kEPS;
    }
    double inv_det = 1.0 / det;
    double a = m[3] * inv_det, b = -m[1] * inv_det, c = -m[2] * inv_det, d = m[0] * inv_det;
    m[0] = a; m[1] = b; m[2] = c; m[3] = d;
}

/* GradientLeastSquare_2 (FVMR/src/space/gradient/Gradient.cpp:160-246) */
int gradient_lsq(OSolver &s, const double *const *U) {
    int no_neighbour = 0;
#pragma omp parallel for schedule(static) reduction(| : no_neighbour)
    for (int c = 0; c < s.ne; c++) {
        int vn[3], n = 0;
        for (int j = 0; j < 3; j++) if (s.enb[j][c] != -1) vn[n++] = s.enb[j][c];
        if (n <= 0) { no_neighbour = 1; continue; }
        double dX[3][2], dU[3][4];
        for (int k = 0; k < n; k++) {
            dX[k][0] = s.ex[vn[k]] - s.ex[c];
            dX[k][1] = s.ey[vn[k]] - s.ey[c];
            for (int v = 0; v < 4; v++) dU[k][v] = U[v][vn[k]] - U[v][c];
        }
        /* mul_ixj_jxk(2, n, 2, dXtrans, dX) */
        double A[4];
        for (int x = 0; x < 2; x++)
            for (int y = 0; y < 2; y++) {
                double acc = 0.0;
                for (int z = 0; z < n; z++) acc += dX[z][x] * dX[z][y];
                A[x * 2 + y] = acc;
            }
        inv2x2(A, s.p.compat_inv2x2 != 0);
        /* invdXdX_dX = A * dXtrans (2 x n) */
        double B[2][3];
        for (int x = 0; x < 2; x++)
            for (int y = 0; y < n; y++) {
                double acc = 0.0;
                for (int z = 0; z < 2; z++) acc += A[x * 2 + z] * dX[y][z];
                B[x][y] = acc;
            }
        /* dUdX = B * dU (2 x 4) */
        for (int v = 0; v < 4; v++) {
            double gx = 0.0, gy = 0.0;
            for (int z = 0; z < n; z++) { gx += B[0][z] * dU[z][v]; }
            for (int z = 0; z < n; z++) { gy += B[1][z] * dU[z][v]; }
            s.Ux[v][c] = gx;
            s.Uy[v][c] = gy;
        }
    }
    if (no_neighbour) return fail(FVMR_ERR_NUMERIC, "invalid numOfValidNeighbor");
    return FVMR_OK;
}

/* LimiterBarth (FVMR/src/space/gradient/Gradient.cpp:58-114; GradientGPU.cu:453-502) */
void limiter_barth(OSolver &s, const double *const *U) {
#pragma omp parallel for schedule(static)
    for (int c = 0; c < s.ne; c++) {
        int vn[3], n = 0;
        for (int j = 0; j < 3; j++) if (s.enb[j][c] != -1) vn[n++] = s.enb[j][c];
        for (int v = 0; v < 4; v++) {
            double Uc = U[v][c], lo = Uc, hi = Uc;
            for (int k = 0; k < n; k++) { lo = dmin(lo, U[v][vn[k]]); hi = dmax(hi, U[v][vn[k]]); }
            double phi = 1.0;
            for (int j = 0; j < 3; j++) {
                int nd = s.en[j][c];
                double Un = Uc + s.Ux[v][c] * (s.node_x[nd] - s.ex[c]) + s.Uy[v][c] * (s.node_y[nd] - s.ey[c]);
                double rel = Un - Uc, hrel = hi - Uc, lrel = lo - Uc;
                double pn = 1.0;
                if (rel > 0) { rel += kEPS; pn = dmin(1, hrel / rel); }
                else if (rel < 0) { rel -= kEPS; pn = dmin(1, lrel / rel); }
                phi = dmin(phi, pn);
            }
            s.Ux[v][c] *= phi;
            s.Uy[v][c] *= phi;
        }
    }
}

/* New: Venkatakrishnan (1993) limiter on the same neighbour/vertex sets as Barth,
   with eps^2 = (K h)^3, h = sqrt(area). */
void limiter_venkat(OSolver &s, const double *const *U) {
#pragma omp parallel for schedule(static)
    for (int c = 0; c < s.ne; c++) {
        int vn[3], n = 0;
        for (int j = 0; j < 3; j++) if (s.enb[j][c] != -1) vn[n++] = s.enb[j][c];
        double h = std::sqrt(s.evol[c]);
        double Kh = s.p.venkat_K * h;
        double eps2 = Kh * Kh * Kh;
        for (int v = 0; v < 4; v++) {
            double Uc = U[v][c], lo = Uc, hi = Uc;
            for (int k = 0; k < n; k++) { lo = dmin(lo, U[v][vn[k]]); hi = dmax(hi, U[v][vn[k]]); }
            double phi = 1.0;
            for (int j = 0; j < 3; j++) {
                int nd = s.en[j][c];
                double d2 = s.Ux[v][c] * (s.node_x[nd] - s.ex[c]) + s.Uy[v][c] * (s.node_y[nd] - s.ey[c]);
                double pn = 1.0;
                if (d2 != 0.0) {
                    double d1 = (d2 > 0) ? (hi - Uc) : (lo - Uc);
                    double num = (d1 * d1 + eps2) * d2 + 2.0 * d2 * d2 * d1;
                    double den = d1 * d1 + 2.0 * d2 * d2 + d1 * d2 + eps2;
                    pn = num / (den * d2);
                }
                phi = dmin(phi, pn);
            }
            s.Ux[v][c] *= phi;
            s.Uy[v][c] *= phi;
        }
    }
}

/* GPU::Space::Gradient::Gradient_2 (GradientGPU.cu:513-551) / U2NITS Gradient (Gradient.cpp:116-158) */
int gradient(OSolver &s, const double *const *U) {
    int rc = FVMR_OK;
    if (s.p.gradient == FVMR_GRA_LEASTSQUARE) rc = gradient_lsq(s, U);
    else if (s.p.gradient == FVMR_GRA_GREENGAUSS) gradient_gg(s, U);
    else return fail(FVMR_ERR_INVALID_ARG, "invalid gradient type");
    if (rc) return rc;
    if (s.p.limiter == FVMR_LIM_BARTH) limiter_barth(s, U);
    else if (s.p.limiter == FVMR_LIM_VENKATAKRISHNAN) limiter_venkat(s, U);
    else if (s.p.limiter != FVMR_LIM_NONE) return fail(FVMR_ERR_INVALID_ARG, "invalid limiter type");
    return FVMR_OK;
}

/* calculateFunctionF_device (EvolveGPU.cu:40-51): gradient, edge flux, viscous, gather, /(-V) */
int rhs(OSolver &s, const double *const *U, double t_rk) {
    int rc = gradient(s, U);
    if (rc) return rc;
    for (int k = 0; k < 4; k++) std::fill(s.Flux[k].begin(), s.Flux[k].end(), 0.0);
#pragma omp parallel for schedule(static)
    for (int e = 0; e < s.nf; e++) {
        double F[4];
        edge_flux(s, U, t_rk, e, F);
        for (int k = 0; k < 4; k++) s.eFlux[k][e] = F[k];
    }
    if (s.p.equation == FVMR_EQ_NS) {
#pragma omp parallel for schedule(static)
        for (int e = 0; e < s.nf; e++) {
            double F[4];
            edge_viscous(s, U, e, F);
            for (int k = 0; k < 4; k++) s.eFlux[k][e] -= F[k];
        }
    }
    /* add_edge_flux_to_element_flux_kernel (FluxGPU.cu:285-323) */
#pragma omp parallel for schedule(static)
    for (int c = 0; c < s.ne; c++) {
        for (int i = 0; i < 4; i++) {
            int e = s.ee[i][c];
            if (e < 0 || e >= s.nf) continue;
            int sign = (c != s.gR[e]) ? 1 : -1;
            for (int k = 0; k < 4; k++) s.Flux[k][c] += sign * s.eFlux[k][e];
        }
    }
    /* vector_weighted_divide_kernel (BasicAlgorithmGPU.cu:4-9) with weight -1 */
    for (int k = 0; k < 4; k++)
#pragma omp parallel for schedule(static)
        for (int c = 0; c < s.ne; c++) s.Flux[k][c] /= -1.0 * s.evol[c];
    return FVMR_OK;
}

/* update_alphaC_host + get_minimus_dt_by_alphaC (FVMR/src/time/CalculateDt.cpp:10-85) */
double global_dt(OSolver &s, const double *const *U) {
    for (int k = 0; k < 4; k++)
#pragma omp parallel for schedule(static)
        for (int c = 0; c < s.ne; c++) {
            /* CIteration::update_host_ruvp_Uold (CIteration.cpp:163-180) */
            double Uc[4] = {U[0][c], U[1][c], U[2][c], U[3][c]}, r[4];
            U2ruvp(Uc, r, s.p.gamma);
            s.ruvp[k][c] = r[k];
        }
    const double gamma = s.p.gamma;
    const double value_1 = dmax(4.0 / 3.0, gamma / s.p.Pr);
    const bool ns = s.p.equation == FVMR_EQ_NS;
    /* per edge: the spectral radius and the viscous term */
    auto edge_terms = [&](int i, double &alpha, double &aV) {
        int L = s.gL[i], R = s.gR[i];
        double rho, u, v, p;
        if (R != -1) {
            rho = 0.5 * (s.ruvp[0][L] + s.ruvp[0][R]);
            u = 0.5 * (s.ruvp[1][L] + s.ruvp[1][R]);
            v = 0.5 * (s.ruvp[2][L] + s.ruvp[2][R]);
            p = 0.5 * (s.ruvp[3][L] + s.ruvp[3][R]);
        } else {
            rho = s.ruvp[0][L]; u = s.ruvp[1][L]; v = s.ruvp[2][L]; p = s.ruvp[3][L];
        }
        double dx = s.gnx[i] * s.glen[i];
        double dy = s.gny[i] * s.glen[i];
        const double length = s.glen[i];
        double unormal = u * dx + v * dy;
        alpha = std::fabs(unormal) + std::sqrt(gamma * p / rho) * std::sqrt(length);
        aV = 0.0;
        if (ns) {
            double T = p / s.p.R / rho;
            double mu = mu_of_T(s, T);
            aV = 2.0 * length * value_1 * mu / rho;
        }
    };
    int nthreads = 1;
#ifdef _OPENMP
    nthreads = omp_get_max_threads();
#endif
    if (nthreads == 1) {
        std::fill(s.alphaC.begin(), s.alphaC.end(), 0.0);
        for (int i = 0; i < s.nf; i++) {
            int L = s.gL[i], R = s.gR[i];
            double alpha, aV;
            edge_terms(i, alpha, aV);
            s.alphaC[L] += alpha;
            if (R != -1) s.alphaC[R] += alpha;
            if (ns) {
                s.alphaC[L] += aV / s.evol[L];
                if (R != -1) s.alphaC[R] += aV / s.evol[R];
            }
        }
    } else {
        /* the same sums, cell-owned: each cell adds its edges' terms in ascending edge order */
        if (s.dt_off.empty()) {
            s.dt_off.assign(s.ne + 1, 0);
            for (int i = 0; i < s.nf; i++) {
                s.dt_off[s.gL[i] + 1]++;
                if (s.gR[i] != -1) s.dt_off[s.gR[i] + 1]++;
            }
            for (int c = 0; c < s.ne; c++) s.dt_off[c + 1] += s.dt_off[c];
            s.dt_edge.assign(s.dt_off[s.ne], 0);
            std::vector<int> fill(s.dt_off.begin(), s.dt_off.end() - 1);
            for (int i = 0; i < s.nf; i++) {
                s.dt_edge[fill[s.gL[i]]++] = i;
                if (s.gR[i] != -1) s.dt_edge[fill[s.gR[i]]++] = i;
            }
            s.e_alpha.assign(s.nf, 0.0);
            s.e_aV.assign(s.nf, 0.0);
        }
#pragma omp parallel for schedule(static)
        for (int i = 0; i < s.nf; i++) edge_terms(i, s.e_alpha[i], s.e_aV[i]);
#pragma omp parallel for schedule(static)
        for (int c = 0; c < s.ne; c++) {
            double a = 0.0;
            for (int k = s.dt_off[c]; k < s.dt_off[c + 1]; k++) {
                const int i = s.dt_edge[k];
                a += s.e_alpha[i];
                if (ns) a += s.e_aV[i] / s.evol[c];
            }
            s.alphaC[c] = a;
        }
    }
    /* the minimum skips NaN cells, as the device reduction's operator_min
       (BasicAlgorithmGPU.h:38-40, "if (a > b) a = b") does for a finite running value */
    double dt = 1e10;
#pragma omp parallel
    {
        double dt_t = 1e10;
#pragma omp for schedule(static) nowait
        for (int c = 0; c < s.ne; c++) {
            double v = s.p.CFL * s.evol[c] / s.alphaC[c];
            if (dt_t > v) dt_t = v;
        }
#pragma omp critical
        if (dt > dt_t) dt = dt_t;
    }
    return dt;
}

/* modifyElementFieldU2d (Restrict.cpp:7-90 / RestrictGPU.cu:6-84), Jacobi reads */
int restrict_field(OSolver &s, double *const *U) {
    std::vector<double> snap[4];
    for (int k = 0; k < 4; k++) {
        snap[k].resize(s.ne);
        copy_n(snap[k].data(), U[k], s.ne);
    }
    int count = 0;
#pragma omp parallel for schedule(static) reduction(+ : count)
    for (int c = 0; c < s.ne; c++) {
        double Uc[4] = {snap[0][c], snap[1][c], snap[2][c], snap[3][c]}, r[4];
        U2ruvp(Uc, r, s.p.gamma);
        bool bad = false;
        for (int i = 0; i < 4; i++) if (std::isnan(r[i])) bad = true;
        if (outOfRange(r)) bad = true;
        if (!bad) continue;
        count++;
        int nvalid = 0;
        double Usum[4] = {0, 0, 0, 0}, vs = 0.0;
        for (int i = 0; i < 4; i++) {
            int nb = s.enb[i][c];
            if (nb == -1) continue;
            nvalid++;
            for (int j = 0; j < 4; j++) {
                double vol = s.evol[nb];
                Usum[j] += snap[j][nb] * vol;
                vs += vol;
            }
        }
        if (nvalid == 0) continue;
        for (int j = 0; j < 4; j++) U[j][c] = Usum[j] / vs;
    }
    return count;
}

}  /* namespace */

/* ====================================================================== */
/* C API for the tests (ctypes)                                            */
/* ====================================================================== */
extern "C" {

const char *orc_last_error(void) { return g_err.c_str(); }

void *orc_mesh_read_su2(const char *path, int equation, double scale) {
    OMesh *m = new OMesh();
    if (readSU2(path, equation, scale, *m) != FVMR_OK) { delete m; return nullptr; }
    return m;
}

void *orc_mesh_from_lists(int num_node, const double *x, const double *y, const int *node_ids,
                          int num_elem, const int *elem_types, const int *elem_conn,
                          int num_marker, const char *const *names, const int *offs,
                          const int *edge_nodes, int equation) {
    std::vector<double> xs(x, x + num_node), ys(y, y + num_node);
    std::vector<int> ids(num_node);
    int maxNode = 1;
    for (int i = 0; i < num_node; i++) { ids[i] = node_ids[i] + 1; maxNode = std::max(maxNode, ids[i]); }
    std::vector<std::array<int, 3>> el;
    for (int i = 0; i < num_elem; i++) {
        const int *c = elem_conn + 4 * (size_t)i;
        if (elem_types[i] == 5) el.push_back({c[0] + 1, c[1] + 1, c[2] + 1});
        else if (elem_types[i] == 9) { el.push_back({c[0] + 1, c[1] + 1, c[3] + 1}); el.push_back({c[3] + 1, c[1] + 1, c[2] + 1}); }
        else { g_err = "unsupported element type"; return nullptr; }
    }
    std::vector<RawMarker> mk(num_marker);
    for (int k = 0; k < num_marker; k++) {
        mk[k].name = names[k];
        for (int j = offs[k]; j < offs[k + 1]; j++) mk[k].e.push_back({edge_nodes[2 * (size_t)j] + 1, edge_nodes[2 * (size_t)j + 1] + 1});
    }
    OMesh *m = new OMesh();
    if (buildMesh(*m, xs, ys, ids, maxNode, el, mk, equation) != FVMR_OK) { delete m; return nullptr; }
    return m;
}

void orc_mesh_free(void *m) { delete (OMesh *)m; }
void orc_mesh_soa(void *m, fvmr_mesh_soa *out) { *out = ((OMesh *)m)->soa; }
const int *orc_mesh_element_node_ids(void *m) { return ((OMesh *)m)->elem_ids3.data(); }
int orc_mesh_num_sets(void *m) { return (int)((OMesh *)m)->bsets.size(); }
const char *orc_mesh_set_name(void *m, int s) { return ((OMesh *)m)->bsets[s].name.c_str(); }
int orc_mesh_set_type(void *m, int s) { return ((OMesh *)m)->bsets[s].type; }
int orc_mesh_set_size(void *m, int s) { return (int)((OMesh *)m)->bsets[s].edges.size(); }
const int *orc_mesh_set_edges(void *m, int s) { return ((OMesh *)m)->bsets[s].edges.data(); }

/* FieldInitializer (FVMR/src/input/FieldInitializer.cpp:129-177,179-229) */
int orc_field_init(const fvmr_mesh_soa *m, const fvmr_params *p, int type, const double *args, double *const *U) {
    const double gamma = p->gamma;
    for (int ie = 0; ie < m->num_element; ie++) {
        double Ue[4];
        double x = m->element_xy[0][ie], y = m->element_xy[1][ie];
        if (type == FVMR_INIT_UNIFORM) {
            ruvp2U(p->inf_ruvp, Ue, gamma);
        } else if (type == FVMR_INIT_ISENTROPIC_VORTEX) {
            double xc = args[0], yc = args[1], chi = args[2];
            double chi2 = chi * chi, ga1 = gamma - 1.0;
            const double two_pi = 2.0 * kPI, pi2 = kPI * kPI;
            double c_du = chi / two_pi;
            double c_dT = ga1 * chi2 / (8. * gamma * pi2);
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
            ruvp2U(r, Ue, gamma);
        } else if (type == FVMR_INIT_SHOCK_TUBE) {
            double sx = args[0], sy = args[1], nx = args[2], ny = args[3];
            if (nx == 0.0 && ny == 0.0) nx = 1.0;
            double li = 1.0 / std::sqrt(nx * nx + ny * ny);
            nx *= li; ny *= li;
            bool inlet;
            if (nx == 1.0) inlet = x < sx;
            else inlet = !((x - sx) * nx + (y - sy) * ny > 0);
            ruvp2U(inlet ? p->inlet_ruvp : p->outlet_ruvp, Ue, gamma);
        } else if (type == FVMR_INIT_DSR) {
            double ang = args[2] * kPI / 180.0;
            double nx = -std::sin(ang), ny = std::cos(ang);
            bool up = ((x - args[0]) * nx + (y - args[1]) * ny > 0);
            ruvp2U(up ? p->inlet_ruvp : p->outlet_ruvp, Ue, gamma);
        } else {
            g_err = "invalid initialCondition type";
            return FVMR_ERR_INVALID_ARG;
        }
        for (int k = 0; k < 4; k++) U[k][ie] = Ue[k];
    }
    return FVMR_OK;
}

void *orc_solver_create(const fvmr_mesh_soa *m, const fvmr_params *p) {
    OSolver *s = new OSolver();
    s->nv = m->num_node; s->ne = m->num_element; s->nf = m->num_edge; s->nb = m->num_boundary_set;
    s->node_x.assign(m->node_xy[0], m->node_xy[0] + s->nv);
    s->node_y.assign(m->node_xy[1], m->node_xy[1] + s->nv);
    for (int k = 0; k < 4; k++) {
        s->en[k].assign(m->element_nodes[k], m->element_nodes[k] + s->ne);
        s->ee[k].assign(m->element_edges[k], m->element_edges[k] + s->ne);
        s->enb[k].assign(m->element_neighbors[k], m->element_neighbors[k] + s->ne);
        s->U[k].assign(s->ne, 0.0); s->Ux[k].assign(s->ne, 0.0); s->Uy[k].assign(s->ne, 0.0);
        s->Flux[k].assign(s->ne, 0.0); s->ruvp[k].assign(s->ne, 0.0); s->Uold[k].assign(s->ne, 0.0);
        s->eFlux[k].assign(s->nf, 0.0);
    }
    s->ex.assign(m->element_xy[0], m->element_xy[0] + s->ne);
    s->ey.assign(m->element_xy[1], m->element_xy[1] + s->ne);
    s->evol.assign(m->element_volume, m->element_volume + s->ne);
    s->gset.assign(m->edge_setID, m->edge_setID + s->nf);
    s->gpair.assign(m->edge_periodicPair, m->edge_periodicPair + s->nf);
    s->gL.assign(m->edge_elementL, m->edge_elementL + s->nf);
    s->gR.assign(m->edge_elementR, m->edge_elementR + s->nf);
    s->glen.assign(m->edge_length, m->edge_length + s->nf);
    s->gx.assign(m->edge_xy[0], m->edge_xy[0] + s->nf);
    s->gy.assign(m->edge_xy[1], m->edge_xy[1] + s->nf);
    s->gnx.assign(m->edge_normal[0], m->edge_normal[0] + s->nf);
    s->gny.assign(m->edge_normal[1], m->edge_normal[1] + s->nf);
    s->btype.assign(m->boundary_type, m->boundary_type + s->nb);
    s->alphaC.assign(s->ne, 0.0);
    s->p = *p;
    return s;
}
void orc_solver_free(void *h) { delete (OSolver *)h; }

/* Threads of the OpenMP build (the plain build has one); returns the count in use. */
int orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

void orc_set_state(void *h, const double *const *U) {
    OSolver *s = (OSolver *)h;
    for (int k = 0; k < 4; k++) std::copy(U[k], U[k] + s->ne, s->U[k].begin());
}
void orc_get_state(void *h, double *const *U) {
    OSolver *s = (OSolver *)h;
    for (int k = 0; k < 4; k++) std::copy(s->U[k].begin(), s->U[k].end(), U[k]);
}

/* limited gradient of the current state */
int orc_gradient(void *h, double *const *Ux, double *const *Uy) {
    OSolver *s = (OSolver *)h;
    const double *Up[4] = {s->U[0].data(), s->U[1].data(), s->U[2].data(), s->U[3].data()};
    int rc = gradient(*s, Up);
    if (rc) return rc;
    for (int k = 0; k < 4; k++) {
        std::copy(s->Ux[k].begin(), s->Ux[k].end(), Ux[k]);
        std::copy(s->Uy[k].begin(), s->Uy[k].end(), Uy[k]);
    }
    return FVMR_OK;
}

/* dU/dt of the current state (t_rk is the double-Mach-reflection boundary time) */
int orc_stage_rhs(void *h, double t_rk, double *const *out) {
    OSolver *s = (OSolver *)h;
    const double *Up[4] = {s->U[0].data(), s->U[1].data(), s->U[2].data(), s->U[3].data()};
    int rc = rhs(*s, Up, t_rk);
    if (rc) return rc;
    for (int k = 0; k < 4; k++) std::copy(s->Flux[k].begin(), s->Flux[k].end(), out[k]);
    return FVMR_OK;
}

/* edge fluxes (inviscid minus viscous) of the current state, after orc_stage_rhs */
void orc_edge_flux(void *h, double *const *out) {
    OSolver *s = (OSolver *)h;
    for (int k = 0; k < 4; k++) std::copy(s->eFlux[k].begin(), s->eFlux[k].end(), out[k]);
}

double orc_global_dt(void *h) {
    OSolver *s = (OSolver *)h;
    const double *Up[4] = {s->U[0].data(), s->U[1].data(), s->U[2].data(), s->U[3].data()};
    return global_dt(*s, Up);
}

/* CIteration::iteration_host (CIteration.cpp:28-80) with the GPU restriction; SSP-RK3 for 3103,
 * restricted once after its last stage as the reference restricts once per iteration */
int orc_step(void *h, double *t, double T, fvmr_step_info *info) {
    OSolver *s = (OSolver *)h;
    const int ne = s->ne;
    copy_field(s->Uold, s->U, ne);
    const double *Up[4] = {s->U[0].data(), s->U[1].data(), s->U[2].data(), s->U[3].data()};
    double dt = global_dt(*s, Up);
    if (*t + dt > T) dt = T - *t;
    double t_rk = *t;
    *t += dt;
    int rc;
    double *Uw[4] = {s->U[0].data(), s->U[1].data(), s->U[2].data(), s->U[3].data()};
    std::vector<double> rhs1[4];
    int nres = 0;
    if (s->p.time_advance == FVMR_EVO_EXPLICIT) {
        if ((rc = rhs(*s, Up, t_rk))) return rc;
        copy_field(rhs1, s->Flux, ne);
        /* vector_weighted_add_kernel (BasicAlgorithmGPU.cu:18-23) */
        for (int k = 0; k < 4; k++)
#pragma omp parallel for schedule(static)
            for (int c = 0; c < ne; c++) s->U[k][c] += dt * s->Flux[k][c];
        nres += restrict_field(*s, Uw);
    } else if (s->p.time_advance == FVMR_EVO_RK3) {
        const double a[3] = {0.0, 0.75, 1.0 / 3.0};
        const double b[3] = {1.0, 0.25, 2.0 / 3.0};
        for (int st = 0; st < 3; st++) {
            if ((rc = rhs(*s, Up, t_rk))) return rc;
            if (st == 0) copy_field(rhs1, s->Flux, ne);
            for (int k = 0; k < 4; k++)
#pragma omp parallel for schedule(static)
                for (int c = 0; c < ne; c++) {
                    double stage = s->U[k][c] + dt * s->Flux[k][c];
                    s->U[k][c] = (st == 0) ? stage : a[st] * s->Uold[k][c] + b[st] * stage;
                }
        }
        /* once per iteration, after the evolve (CIteration.cpp:79) */
        nres += restrict_field(*s, Uw);
    } else {
        return fail(FVMR_ERR_UNSUPPORTED, "invalid temporal_choice");
    }
    copy_field(s->Flux, rhs1, ne);
    s->last_dt = dt;
    s->n_restricted = nres;
    if (info) {
        info->t = *t; info->dt = dt; info->n_restricted = nres; info->steps_done = 1;
        if (info->residual_norm) {
            for (int k = 0; k < 4; k++) {
                double norm = 0;
                for (int c = 0; c < ne; c++) {
                    double f = dabs(s->Flux[k][c]);
                    if (info->residual_norm == FVMR_NORM_1) norm += f;
                    else if (info->residual_norm == FVMR_NORM_2) norm += f * f;
                    else norm = dmax(norm, f);
                }
                if (info->residual_norm == FVMR_NORM_2) norm = std::sqrt(norm);
                info->residual[k] = norm;
            }
        }
    }
    return FVMR_OK;
}

/* get_residual_host (FVMR/src/output/residual/Residual.cpp:4-43) over the last step's rhs */
void orc_residual(void *h, int norm_type, double *out) {
    OSolver *s = (OSolver *)h;
    for (int k = 0; k < 4; k++) {
        double norm = 0;
        for (int c = 0; c < s->ne; c++) {
            double f = dabs(s->Flux[k][c]);
            if (norm_type == FVMR_NORM_1) norm += f;
            else if (norm_type == FVMR_NORM_2) norm += f * f;
            else norm = dmax(norm, f);
        }
        if (norm_type == FVMR_NORM_2) norm = std::sqrt(norm);
        out[k] = norm;
    }
}

int orc_restrict(void *h) {
    OSolver *s = (OSolver *)h;
    double *Uw[4] = {s->U[0].data(), s->U[1].data(), s->U[2].data(), s->U[3].data()};
    return restrict_field(*s, Uw);
}

/* single-face kernels for closed-form tests */
void orc_roe(const double *UL, const double *UR, double nx, double ny, double len, double gamma, double *F) { roe(UL, UR, nx, ny, len, F, gamma); }
void orc_llf(const double *UL, const double *UR, double nx, double ny, double len, double gamma, double *F) { llf(UL, UR, nx, ny, len, F, gamma); }
void orc_farfield(const double *UL, const double *ruvp_inf, double nx, double ny, double gamma, double *UR) { ghost_farfield(UL, ruvp_inf, UR, nx, ny, gamma); }
void orc_slip(const double *UL, double nx, double ny, double *UR) { ghost_slip(UL, UR, nx, ny); }
void orc_isothermal(const double *UL, double Tw, double R, double gamma, double *UR) { ghost_isothermal(UL, UR, Tw, R, gamma); }
void orc_u2ruvp(const double *U, double gamma, double *r) { U2ruvp(U, r, gamma); }

/* FieldWriter::update_nodeField_ruvp (FVMR/src/output/FieldWriter.cpp:371-418): node ruvp =
 * mean of the ruvp of the cells listing the node, accumulated in element order, slot order;
 * then update_nodeField_other_variables (:420-452): Cp = (p - p_inf) / (0.5 rho_inf V_inf^2)
 * (ReferenceData::update_farfield, :544-553) and Ma = sqrt(V^2 / (gamma p / rho)).
 * Cp or Ma may be NULL. */
void orc_node_field(const fvmr_mesh_soa *m, const double *const *U, double gamma, const double *ruvp_inf,
                    double *const *node_ruvp, double *Cp, double *Ma) {
    const int nv = m->num_node, ne = m->num_element;
    std::vector<int> cnt(nv, 0);
    for (int k = 0; k < 4; k++) std::fill(node_ruvp[k], node_ruvp[k] + nv, 0.0);
    for (int ie = 0; ie < ne; ie++) {
        double Ue[4] = {U[0][ie], U[1][ie], U[2][ie], U[3][ie]}, r[4];
        U2ruvp(Ue, r, gamma);
        for (int j = 0; j < 4; j++) {
            const int v = m->element_nodes[j][ie];
            if (v < 0 || v >= nv) continue;
            cnt[v]++;
            for (int k = 0; k < 4; k++) node_ruvp[k][v] += r[k];
        }
    }
    for (int v = 0; v < nv; v++) {
        if (cnt[v] == 0) continue;
        for (int k = 0; k < 4; k++) node_ruvp[k][v] /= cnt[v];
    }
    const double V2f = ruvp_inf[1] * ruvp_inf[1] + ruvp_inf[2] * ruvp_inf[2];
    const double p_inf = ruvp_inf[3], p_dyn = 0.5 * ruvp_inf[0] * V2f;
    for (int v = 0; v < nv; v++) {
        const double rho = node_ruvp[0][v], u = node_ruvp[1][v], w = node_ruvp[2][v], p = node_ruvp[3][v];
        if (Cp) Cp[v] = (p - p_inf) / p_dyn;
        if (Ma) {
            const double V2 = u * u + w * w;
            const double a2 = gamma * p / rho;
            Ma[v] = std::sqrt(V2 / a2);
        }
    }
}

/* getColor_device (FVMR/src/gui/cuda/CudaGrid_kernels.cu:22-65): input clamped to [0, 1]; the
 * first control interval [w_i, w_i+1] holding it; with n its position in the interval and
 * m = 0.5 (c_i + c_i+1), colour = (1-n)^3 c_i + 3 (1-n)^2 n m + 3 (1-n) n^2 m + n^3 c_i+1 in that
 * order; purple (252, 3, 253) / 255 when no interval holds it.  cmap: {r, g, b, position}. */
static void color_of(int npts, const float *cmap, float x, float out[3]) {
    if (x < 0.0f) x = 0.0f;
    if (x > 1.0f) x = 1.0f;
    for (int i = 0; i + 1 < npts; ++i) {
        const float *a = cmap + 4 * i, *b = cmap + 4 * (i + 1);
        if (x >= a[3] && x <= b[3]) {
            const float n = (x - a[3]) / (b[3] - a[3]);
            const float omn = 1 - n, omn2 = omn * omn, omn3 = omn2 * omn, n2 = n * n, n3 = n2 * n;
            for (int k = 0; k < 3; k++) {
                const float mean = 0.5f * (a[k] + b[k]);
                out[k] = omn3 * a[k] + 3 * omn2 * n * mean + 3 * omn * n2 * mean + n3 * b[k];
            }
            return;
        }
    }
    out[0] = 252.0f / 255.0f;
    out[1] = 3.0f / 255.0f;
    out[2] = 253.0f / 255.0f;
}

/* GPU::setVerticesColor (FVMR/src/gui/cuda/CudaGrid_kernels.cu:185-232): node data zeroed; per
 * element and vertex, node = max(node, q_e * n_v) with q_e the float primitive `var` and n_v the
 * node's element count (accumulateElementDataToNodeData_k :77-95, GPUSolver2.cpp:185-196); node *=
 * (float)(1.0 / n_v), 1 when that exceeds 1e5 (divideNodeData_k :97-104); x = (node - lo) /
 * |hi - lo| (0 for an empty range) coloured into vertices[6 v + 3 .. 5] (nodeDataToVertexColor_k
 * :106-123). */
void orc_vertex_colors(const fvmr_mesh_soa *m, const double *const *U, double gamma, int var, float lo, float hi,
                       int npts, const float *cmap, float *vertices) {
    const int nv = m->num_node, ne = m->num_element;
    std::vector<int> cnt(nv, 0);
    for (int ie = 0; ie < ne; ie++)
        for (int j = 0; j < 3; j++) cnt[m->element_nodes[j][ie]]++;
    std::vector<float> node(nv, 0.0f);
    for (int ie = 0; ie < ne; ie++) {
        double Ue[4] = {U[0][ie], U[1][ie], U[2][ie], U[3][ie]}, r[4];
        U2ruvp(Ue, r, gamma);
        for (int j = 0; j < 3; j++) {
            const int v = m->element_nodes[j][ie];
            float in = (float)r[var];
            in *= cnt[v];
            const float value = node[v];
            node[v] = (value > in) ? value : in;
        }
    }
    for (int v = 0; v < nv; v++) {
        float div = 1.0 / cnt[v];
        if (div > 1e5) div = 1.0f;
        node[v] *= div;
        const float len = std::fabs(hi - lo);
        const float x = (len == 0) ? 0.0f : (node[v] - lo) / len;
        color_of(npts, cmap, x, vertices + (size_t)6 * v + 3);
    }
}

/* FieldWriter::ReferenceData::calculate_edgeSet_force (FVMR/src/output/FieldWriter.cpp:174-205):
 * per boundary set, over its edges in set order, Cp_e = 0.5 (Cp_0 + Cp_1),
 * dC = Cp_e (1/S_ref) L; CD += dC nx; CL += dC ny.  Set s lists edges off[s] .. off[s+1]-1. */
void orc_surface_force(const fvmr_mesh_soa *m, const double *Cp, int nsets, const int *off, const int *edges,
                       double S_ref, double *CD, double *CL) {
    const double one_on_Sref = 1 / S_ref;
    for (int s = 0; s < nsets; s++) {
        CD[s] = 0.0;
        CL[s] = 0.0;
        for (int k = off[s]; k < off[s + 1]; k++) {
            const int e = edges[k];
            const int n0 = m->edge_nodes[0][e], n1 = m->edge_nodes[1][e];
            const double nx = m->edge_normal[0][e], ny = m->edge_normal[1][e], dS = m->edge_length[e];
            const double cp = 0.5 * (Cp[n0] + Cp[n1]);
            const double dC = cp * one_on_Sref * dS;
            CD[s] += dC * nx;
            CL[s] += dC * ny;
        }
    }
}
void orc_ruvp2u(const double *r, double gamma, double *U) { ruvp2U(r, U, gamma); }
void orc_inv2x2(double *m, int compat) { inv2x2(m, compat != 0); }

}  /* extern "C" */
