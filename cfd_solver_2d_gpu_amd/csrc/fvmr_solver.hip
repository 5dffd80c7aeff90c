// Device path of the FVMR explicit residual-and-update step for MI355X (gfx950).
//
// One step (CIteration::iteration_device, FVMR/src/time/CIteration.cpp:89-161) runs one of two
// kernel families over the same device numbering (cells in Hilbert, RCM or caller order):
//
// Tile family (default; the mesh is cut into tiles of consecutive cells, fvmr_tiles.cpp):
//   k_tile_dt       one work-group per tile: spectral radius of every face of the tile once,
//                   alphaC per cell in the reference edge order, block minimum of CFL V / alphaC.
//   k_dt_finish     one work-group: global minimum, clamp to T - t, t advanced on the device.
//   k_tile          one work-group per tile and RK stage: states of the tile and two halo
//                   rings, nodes and face records staged in LDS; centroids and normals
//                   recomputed from the nodes; limited gradients of the tile and the first
//                   ring in LDS; the canonical flux of every face of the tile once; per cell
//                   the signed sum in slot order, /(-V), RK update.
//   k_restrict      one work-group: flagged cells repaired from their neighbours (Jacobi).
//
// Cell family (any geometry; packed per-cell / per-edge records, a neighbour costs 1-2 lines):
//   k_gradient<DT>  3 wave64 x 64 cells: wave j gathers face slot j; wave 0 forms the
//                   gradient and limiter (and, in stage 1, the dt block minimum).
//   k_stage         3 wave64 x 64 cells: wave j evaluates the canonical flux of face slot j;
//                   wave 0 sums in slot order, divides by -V, applies the RK update.
//   k_dt_finish, k_restrict as above.
//
// Records: state {rho, rho u, rho v, rho E} (32 B, three rotating buffers); gradient
// {Ux[4], Uy[4]}; cell geo {cx, cy, V, 0}; edge geo {nx, ny, L, 0}, {mx, my}; edge int
// {L, R, BC type, periodic pair}; G = double (plain FP64) or float (mixed precision).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/fvmr.h"
#include "fvmr_internal.h"
#include "fvmr_tiles.h"

using fvmr::set_error;

namespace {

constexpr int kBlock = 256;      // work-group of the 1-thread-per-cell kernels
#ifndef FVMR_FAST_DIV
#define FVMR_FAST_DIV 1          // FP64 fdiv / frcp by refined hardware reciprocal (see below)
#endif
constexpr int kMaxDtSlots = 6;   // own 3 edges + pair edges of periodic faces
constexpr int kNumProf = 5;
constexpr int kCellsPerGroup = 64;              // cells of one cell-kernel work-group
constexpr int kGroup3 = 3 * kCellsPerGroup;     // one wave64 per face slot

#define HIP_OK(expr)                                                                   \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return set_error(FVMR_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Tuning knobs (compile time): minimum waves per SIMD requested for the cell kernels
// (0: compiler default) and the XCD-aware block remap.
#ifndef FVMR_STAGE_WPE
#define FVMR_STAGE_WPE 0
#endif
#ifndef FVMR_GRAD_WPE
#define FVMR_GRAD_WPE 0
#endif
#ifndef FVMR_XCD_REMAP
#define FVMR_XCD_REMAP 1
#endif
// 1: the stage kernel k_tile loads each local cell's centroid and each face's normal (packed
// per tile on the host, equal bit for bit to what it would recompute) instead of recomputing
// them from the node coordinates (about 1% faster at 1M cells).  The time-step kernel always
// recomputes its normals: it is bound by its loads, and loading them cost it 6%.
#ifndef FVMR_TILE_GEOM_LOADED
#define FVMR_TILE_GEOM_LOADED 1
#endif
#if FVMR_STAGE_WPE > 0
#define FVMR_STAGE_ATTR __attribute__((amdgpu_waves_per_eu(FVMR_STAGE_WPE)))
#else
#define FVMR_STAGE_ATTR
#endif
#if FVMR_GRAD_WPE > 0
#define FVMR_GRAD_ATTR __attribute__((amdgpu_waves_per_eu(FVMR_GRAD_WPE)))
#else
#define FVMR_GRAD_ATTR
#endif

template <class T> struct alignas(4 * sizeof(T)) V4 { T x, y, z, w; };
template <class T> struct alignas(2 * sizeof(T)) V2 { T x, y; };
template <class T> struct alignas(4 * sizeof(T)) Grad { T x[4]; T y[4]; };
using S4 = V4<double>;

// flags of a face slot (cslot.x): the cell is the edge's R; the edge belongs to a boundary
// set (its BC type and periodic pair are in eint); the edge is periodic
constexpr int kSlotCellIsR = 1 << 30;
constexpr int kSlotBoundary = 1 << 29;
constexpr int kSlotPeriodic = 1 << 28;
constexpr int kSlotEdgeMask = (1 << 28) - 1;

// ---------------------------------------------------------------------------
// Device-side arguments (plain pointers, passed by value).
// ---------------------------------------------------------------------------
template <class G>
struct Mesh {
    int nc, nf;
    const V4<G> *cgeo;     // cx, cy, volume, 0
    const int2 *cslot;     // [3][nc] face slot j: x = edge | flags, y = cell across the face (-1 none)
    const int *cnode;      // [3][nc] vertex j
    const V2<G> *nxy;      // node coordinates
    const V4<G> *egeo;     // nx, ny, length, 0 (normal L -> R)
    const V2<G> *emid;     // midpoint
    const int4 *eint;      // L, R (periodic: pair's L, boundary: -1), BC type, periodic pair
    const int *dtcode;     // per cell: alphaC contributions in reference edge order
};

// Values a case of a batch may set on its own: boundary states and CFL.
struct CaseBC {
    double inf[4], inlet[4], outlet[4];
    double CFL, pad[3];
};

struct Phys {
    double gamma, R, Pr, Pr_visc, mu0, sutherland_C1, wall_T, venkat_K;
    double dsr_x, dsr_y, dsr_speed;
    double g_ga1, inv_g;           // gamma/(gamma-1), 1/gamma
    CaseBC bc;                     // the handle's (first case's) boundary states and CFL
    int equation, scheme, reconstruct, gradient, limiter, mu_model, compat_inv;
    int has_periodic;              // the mesh has periodic faces
};

// Device scalars of a handle: t, dt, T, t at the start of the step.
struct Scalars {
    double t, dt, T, t_rk;
    unsigned int counter;
    unsigned int spare;
    int bad_count;              // sc[0]: length of the handle's list of flagged cells
    int n_restricted;           // cells of this case restricted in the current step
};

// Blocks are dispatched round-robin over the 8 XCDs; give each XCD a contiguous
// range of logical blocks so that neighbouring cells share an L2.
__device__ __forceinline__ int xcd_block() {
    if (!FVMR_XCD_REMAP) return blockIdx.x;
    const int nb = gridDim.x, b = blockIdx.x;
    const int q = nb >> 3, r = nb & 7, x = b & 7;
    return x * q + min(x, r) + (b >> 3);
}

template <class T> __device__ __forceinline__ T tmin(T a, T b) { return (a < b) ? a : b; }
template <class T> __device__ __forceinline__ T tmax(T a, T b) { return (a > b) ? a : b; }

// Division, reciprocal and square root of the face and gradient arithmetic: IEEE in FP64;
// the hardware's approximate forms (about an ulp) in the FP32 of mixed precision.
template <class T> __device__ __forceinline__ T fdiv(T a, T b) { return a / b; }
template <> __device__ __forceinline__ float fdiv<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
template <class T> __device__ __forceinline__ T frcp(T a) { return T(1) / a; }
template <> __device__ __forceinline__ float frcp<float>(float a) { return __builtin_amdgcn_rcpf(a); }
#if FVMR_FAST_DIV
// FP64 reciprocal and quotient without the IEEE sequence's scaling steps: the hardware estimate
// (v_rcp_f64, about 2^-24 relative) refined by Newton-Raphson steps r += r (1 - b r); the
// quotient a r then corrected once with its residual, q += r (a - b q) (Markstein).  On 2^22
// random operand pairs over 24 binades both equal the IEEE results bit for bit; they are not
// meant for operands at the ends of the exponent range (denormals, overflow), which the solver's
// states, volumes and lengths never reach.
template <> __device__ __forceinline__ double frcp<double>(double b) {
    double r = __builtin_amdgcn_rcp(b);
    r = fma(r, fma(-b, r, 1.0), r);
    return fma(r, fma(-b, r, 1.0), r);
}
template <> __device__ __forceinline__ double fdiv<double>(double a, double b) {
    double r = __builtin_amdgcn_rcp(b);
    r = fma(r, fma(-b, r, 1.0), r);
    const double q = a * r;
    return fma(fma(-b, q, a), r, q);
}
#endif
template <class T> __device__ __forceinline__ T fsqrt(T a) { return sqrt(a); }
template <> __device__ __forceinline__ float fsqrt<float>(float a) { return __builtin_amdgcn_sqrtf(a); }

__device__ __forceinline__ void ld(const S4 *p, int i, double u[4]) {
    const S4 v = p[i];
    u[0] = v.x; u[1] = v.y; u[2] = v.z; u[3] = v.w;
}
__device__ __forceinline__ void st(S4 *p, int i, const double u[4]) {
    S4 v;
    v.x = u[0]; v.y = u[1]; v.z = u[2]; v.w = u[3];
    p[i] = v;
}
// The same store with the nontemporal hint, for the FP64 stage kernels' new state: 61.1 -> 59.5
// us per stage kernel at 1M cells (2 % per step; the kernel boundary has less to write back).
// Not used where it measured slower: the FP32-mixed stage kernel (+1.5 %) and the adjoint's
// products (0.353 -> 0.425 ms per iteration), whose outputs the next kernel gathers.
__device__ __forceinline__ void st_stream(S4 *p, int i, const double u[4]) {
    typedef double d2v __attribute__((ext_vector_type(2)));
    d2v *q = reinterpret_cast<d2v *>(p + i);
    const d2v lo = {u[0], u[1]}, hi = {u[2], u[3]};
    __builtin_nontemporal_store(lo, q);
    __builtin_nontemporal_store(hi, q + 1);
}

// ruvp2U (PhysicalKernelGPU.h:17-22)
template <class T>
__device__ __forceinline__ void ruvp2u(const T r[4], T U[4], T gamma) {
    U[0] = r[0];
    U[1] = r[0] * r[1];
    U[2] = r[0] * r[2];
    U[3] = fdiv(r[3], gamma - 1) + T(0.5) * r[0] * (r[1] * r[1] + r[2] * r[2]);
}
// outOfRange_device (RestrictGPU.h:11-19)
template <class T>
__device__ __forceinline__ bool out_of_range(T rho, T p) {
    return rho < T(fvmr::kRHO_MIN) || rho > T(fvmr::kRHO_MAX) || p < T(fvmr::kP_MIN) || p > T(fvmr::kP_MAX);
}
// U2ruvp (PhysicalKernelGPU.h:11-16) with the divisions by U[0] folded into one
// reciprocal (results differ from the reference by rounding only)
template <class T>
__device__ __forceinline__ void prim(const T U[4], T r[4], T gamma) {
    const T ir = frcp(U[0]);
    r[0] = U[0];
    r[1] = U[1] * ir;
    r[2] = U[2] * ir;
    r[3] = r[0] * (gamma - 1) * (U[3] * ir - (r[1] * r[1] + r[2] * r[2]) * T(0.5));
}
__device__ __forceinline__ bool is_periodic(int t) { return t >= FVMR_BC_PERIODIC_0 && t <= FVMR_BC_PERIODIC_0 + 9; }

// ---------------------------------------------------------------------------
// Face physics (FluxGPU.cu:34-205, RoeGPU.cu, LocalLaxFriedrichsGPU.cu)
// ---------------------------------------------------------------------------
// A reconstructed or ghost state: conserved U and its primitives, computed once.
template <class C>
struct FaceState {
    C U[4];
    C r, u, v, p, ir;
};

template <class C>
__device__ __forceinline__ void set_prim(FaceState<C> &s, C gamma) {
    s.ir = frcp(s.U[0]);
    s.r = s.U[0];
    s.u = s.U[1] * s.ir;
    s.v = s.U[2] * s.ir;
    s.p = s.r * (gamma - 1) * (s.U[3] * s.ir - (s.u * s.u + s.v * s.v) * C(0.5));
}

// Where a face kernel finds the state, centroid and limited gradient of a cell: global
// records (cell kernels) or a tile's LDS (tile kernels).  has(c): c is a cell (not -1).
template <class G>
struct GlobalCells {
    const Mesh<G> &m;
    const S4 *__restrict__ U;
    const Grad<G> *__restrict__ grad;
    __device__ __forceinline__ void state(int c, double u[4]) const { ld(U, c, u); }
    template <class C> __device__ __forceinline__ void state_c(int c, C u[4]) const {
        double d[4];
        ld(U, c, d);
#pragma unroll
        for (int k = 0; k < 4; k++) u[k] = (C)d[k];
    }
    __device__ __forceinline__ V2<G> centroid(int c) const {
        const V4<G> g = m.cgeo[c];
        return V2<G>{g.x, g.y};
    }
    __device__ __forceinline__ Grad<G> gradient(int c) const { return grad[c]; }
    __device__ __forceinline__ bool has(int c) const { return !(c < 0 || c >= m.nc); }
};

// The records of a face's two cells loaded together before the flux uses them (R = L where the
// face has no cell across), so the R loads do not wait behind the L arithmetic; answers for
// those two cells only.
template <class G>
struct PairCells {
    int L, nc;
    double uL[4], uR[4];
    V2<G> cL, cR;
    Grad<G> gL, gR;
    __device__ __forceinline__ PairCells(const Mesh<G> &m, const S4 *__restrict__ U, const Grad<G> *__restrict__ grad,
                                         int Lc, int Rc) : L(Lc), nc(m.nc) {
        const int Rs = (Rc < 0 || Rc >= m.nc) ? Lc : Rc;
        ld(U, Lc, uL);
        ld(U, Rs, uR);
        const V4<G> a = m.cgeo[Lc], b = m.cgeo[Rs];
        gL = grad[Lc];
        gR = grad[Rs];
        cL = V2<G>{a.x, a.y};
        cR = V2<G>{b.x, b.y};
    }
    __device__ __forceinline__ void state(int c, double u[4]) const {
#pragma unroll
        for (int k = 0; k < 4; k++) u[k] = (c == L) ? uL[k] : uR[k];
    }
    template <class C> __device__ __forceinline__ void state_c(int c, C u[4]) const {
#pragma unroll
        for (int k = 0; k < 4; k++) u[k] = (C)((c == L) ? uL[k] : uR[k]);
    }
    __device__ __forceinline__ V2<G> centroid(int c) const { return (c == L) ? cL : cR; }
    __device__ __forceinline__ Grad<G> gradient(int c) const { return (c == L) ? gL : gR; }
    __device__ __forceinline__ bool has(int c) const { return !(c < 0 || c >= nc); }
};

// get_Uvector_reconstruct_2_device (FluxGPU.cu:47-61): linear (3002) or constant
// reconstruction; back to the cell value when rho or p leave their range.
template <class C, class G, class Cells>
__device__ __forceinline__ void reconstruct(const Cells &cells, const Phys &ph, int c, C x, C y, FaceState<C> &s) {
    const C gamma = (C)ph.gamma;
    C Uc[4];
    cells.state_c(c, Uc);
    if (ph.reconstruct == FVMR_REC_LINEAR) {
        const V2<G> g = cells.centroid(c);
        const Grad<G> gr = cells.gradient(c);
        const C dx = x - (C)g.x, dy = y - (C)g.y;
#pragma unroll
        for (int i = 0; i < 4; i++) s.U[i] = Uc[i] + (C)gr.x[i] * dx + (C)gr.y[i] * dy;
    } else {
#pragma unroll
        for (int i = 0; i < 4; i++) s.U[i] = Uc[i];
    }
    set_prim(s, gamma);
    if (out_of_range(s.r, s.p)) {
#pragma unroll
        for (int i = 0; i < 4; i++) s.U[i] = Uc[i];
        set_prim(s, gamma);
    }
}

// get_UR_wallNonViscous_device (FluxGPU.cu:63-77)
template <class C>
__device__ __forceinline__ void ghost_slip(const FaceState<C> &L, FaceState<C> &R, C nx, C ny, C gamma) {
    const C unL = L.u * nx + L.v * ny;
    const C utL = L.u * ny - L.v * nx;
    const C unR = -unL, utR = utL;
    const C uxR = unR * nx + utR * ny;
    const C uyR = unR * ny - utR * nx;
    R.U[0] = L.U[0];
    R.U[1] = R.U[0] * uxR;
    R.U[2] = R.U[0] * uyR;
    R.U[3] = L.U[3];
    set_prim(R, gamma);
}

// get_UR_farField_device (FluxGPU.cu:87-167): Riemann-invariant ghost in four regimes
template <class C, class P>
__device__ __forceinline__ void ghost_farfield(const FaceState<C> &L, const P *ruvp_inf, FaceState<C> &R, C nx,
                                               C ny, C gamma) {
    const C ga1 = gamma - 1;
    const C rho_f = (C)ruvp_inf[0], u_f = (C)ruvp_inf[1], v_f = (C)ruvp_inf[2], p_f = (C)ruvp_inf[3];
    const C rho_e = L.r, u_e = L.u, v_e = L.v, p_e = L.p;
    const C a2_f = fdiv(gamma * p_f, rho_f);
    const C af = fsqrt(a2_f);
    const C ae = fsqrt(fdiv(gamma * p_e, rho_e));
    const C qnf = u_f * nx + v_f * ny;
    const C qne = u_e * nx + v_e * ny;
    const C rf = qnf - fdiv(C(2.0) * af, ga1);
    const C re = qne + fdiv(C(2.0) * ae, ga1);
    const C qn = C(0.5) * (re + rf);
    const C as = fdiv(ga1 * (re - rf), C(4.0));
    C r0[4];
    if (qn <= -as) {
        r0[0] = rho_f; r0[1] = u_f; r0[2] = v_f; r0[3] = p_f;
    } else if (qn < as) {
        // subsonic inflow (qn < 0) takes entropy and tangential velocity from the freestream,
        // subsonic outflow from the interior
        const bool in = qn < 0;
        const C s_p = in ? p_f : p_e, s_r = in ? rho_f : rho_e;
        const C s_u = in ? u_f : u_e, s_v = in ? v_f : v_e, s_qn = in ? qnf : qne;
        const C son = fdiv(s_p, pow(s_r, gamma));
        const C qtx = s_u - s_qn * nx, qty = s_v - s_qn * ny;
        r0[0] = pow(fdiv(fdiv(as * as, son), gamma), frcp(ga1));
        r0[1] = qtx + qn * nx; r0[2] = qty + qn * ny;
        r0[3] = fdiv(as * as * r0[0], gamma);
    } else {
        r0[0] = rho_e; r0[1] = u_e; r0[2] = v_e; r0[3] = p_e;
    }
    ruvp2u(r0, R.U, gamma);
    set_prim(R, gamma);
}

// New: no-slip isothermal wall ghost (u_R = -u_L, p_R = p_L, T_R = 2 T_w - T_L, floored).
template <class C>
__device__ __forceinline__ void ghost_isothermal(const FaceState<C> &L, FaceState<C> &R, C Tw, C Rg, C gamma) {
    const C TL = fdiv(L.p, L.r * Rg);
    C TR = C(2.0) * Tw - TL;
    if (TR < C(0.5) * Tw) TR = C(0.5) * Tw;
    const C g[4] = {fdiv(L.p, Rg * TR), -L.u, -L.v, L.p};
    ruvp2u(g, R.U, gamma);
    set_prim(R, gamma);
}

// LocalLaxFriedrichs2d (LocalLaxFriedrichsGPU.cu:4-49) on precomputed primitives
template <class C>
__device__ __forceinline__ void flux_llf(const FaceState<C> &L, const FaceState<C> &R, C nx, C ny, C length,
                                         C flux[4], C gamma, C g_ga1) {
    const C unL = nx * L.u + ny * L.v, unR = nx * R.u + ny * R.v;
    C FnL[4], FnR[4];
    FnL[0] = L.r * unL; FnL[1] = L.r * unL * L.u + nx * L.p; FnL[2] = L.r * unL * L.v + ny * L.p;
    const C hl = L.p * g_ga1 * L.ir + C(0.5) * (L.u * L.u + L.v * L.v);
    FnL[3] = L.r * unL * hl;
    FnR[0] = R.r * unR; FnR[1] = R.r * unR * R.u + nx * R.p; FnR[2] = R.r * unR * R.v + ny * R.p;
    const C hr = R.p * g_ga1 * R.ir + C(0.5) * (R.u * R.u + R.v * R.v);
    FnR[3] = R.r * unR * hr;
    const C lam = tmax(fabs(unL) + fsqrt(gamma * L.p * L.ir), fabs(unR) + fsqrt(gamma * R.p * R.ir));
#pragma unroll
    for (int i = 0; i < 4; i++) flux[i] = C(0.5) * (FnL[i] + FnR[i] - lam * (R.U[i] - L.U[i])) * length;
}

// ConvectRoeCommon2d + RoeDissapationTerm2d (RoeGPU.cu:4-159): no entropy fix, the
// reference energy jump Etm; divisions folded into reciprocals
template <class C>
__device__ __forceinline__ void flux_roe(const FaceState<C> &L, const FaceState<C> &R, C nx, C ny, C area, C F[4],
                                         C gamma, C g_ga1, C inv_g) {
    const C rL = L.r, uL = L.u, vL = L.v, pL = L.p;
    const C rR = R.r, uR = R.u, vR = R.v, pR = R.p;
    const C ul = nx * uL + ny * vL, ur = nx * uR + ny * vR;
    const C rul = rL * ul, rur = rR * ur;
    const C V2L = uL * uL + vL * vL, V2R = uR * uR + vR * vR;
    const C ga1 = gamma - C(1.0);
    const C HL = pL * g_ga1 * L.ir + C(0.5) * V2L;
    const C HR = pR * g_ga1 * R.ir + C(0.5) * V2R;
    const C psum = pL + pR;
    F[0] = C(0.5) * (rul + rur) * area;
    F[1] = C(0.5) * (rul * uL + rur * uR + nx * psum) * area;
    F[2] = C(0.5) * (rul * vL + rur * vR + ny * psum) * area;
    F[3] = C(0.5) * (rul * HL + rur * HR) * area;
    const C rrorl = fsqrt(rR * L.ir);
    const C inv_p1 = frcp(rrorl + C(1.0));
    const C rm = rL * rrorl;
    const C um = (uL + uR * rrorl) * inv_p1;
    const C vm = (vL + vR * rrorl) * inv_p1;
    const C Hm = (HL + HR * rrorl) * inv_p1;
    const C Vm2 = um * um + vm * vm;
    const C am2 = ga1 * (Hm - C(0.5) * Vm2);
    const C am = fsqrt(am2);
    const C inv_am = frcp(am);
    const C unm = um * nx + vm * ny;
    const C eig1 = fabs(unm);
    const C eig2 = fabs(unm + am);
    const C eig3 = fabs(unm - am);
    const C ER = HR - pR * R.ir, EL = HL - pL * L.ir;
    const C Etm = Hm * inv_g + ga1 * inv_g * C(0.5) * Vm2;
    const C dW0 = rR - rL;
    const C dW1 = rm * (uR - uL) + dW0 * um;
    const C dW2 = rm * (vR - vL) + dW0 * vm;
    const C dW3 = rm * (ER - EL) + dW0 * Etm;
    const C astar = (eig2 + eig3) * C(0.5);
    const C Mstar = (eig2 - eig3) * C(0.5) * inv_am;
    const C dun = nx * (uR - uL) + ny * (vR - vL);
    const C dUroe = Mstar * dun + fdiv((astar - eig1) * (pR - pL), rm) * (inv_am * inv_am);
    const C dProe = Mstar * (pR - pL) + (astar - eig1) * rm * dun;
    F[0] -= C(0.5) * (eig1 * dW0 + dUroe * rm) * area;
    F[1] -= C(0.5) * (eig1 * dW1 + dUroe * rm * um + dProe * nx) * area;
    F[2] -= C(0.5) * (eig1 * dW2 + dUroe * rm * vm + dProe * ny) * area;
    F[3] -= C(0.5) * (eig1 * dW3 + dUroe * rm * Hm + dProe * unm) * area;
}

__device__ __forceinline__ double mu_of_T(const Phys &ph, double T) {
    if (ph.mu_model == FVMR_MU_CONSTANT) return ph.mu0;
    // get_mu_using_Sutherland_air_host_device (ViscousFluxGPU.h:30-32)
    return ph.sutherland_C1 * pow(T, 1.5) / (T + 110.4);
}

// viscous_flux_device_kernel body (ViscousFluxGPU.cu:139-229); returns F_v (to be subtracted)
template <class C, class G, class Cells>
__device__ __forceinline__ void flux_viscous(const Cells &cells, const Phys &ph, int L, int R, C nx, C ny, C area,
                                             C F[4]) {
    const C ga1 = (C)ph.gamma - C(1.0);
    const C Cv = fdiv((C)ph.R, ga1);
    const C Cp = Cv * (C)ph.gamma;
    const bool hasR = cells.has(R);
    C gx[4], gy[4], Um[4];
    {
        const Grad<G> gL = cells.gradient(L);
        C uL[4];
        cells.state_c(L, uL);
        if (hasR) {
            const Grad<G> gR = cells.gradient(R);
            C uR[4];
            cells.state_c(R, uR);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                gx[k] = C(0.5) * ((C)gL.x[k] + (C)gR.x[k]);
                gy[k] = C(0.5) * ((C)gL.y[k] + (C)gR.y[k]);
                Um[k] = C(0.5) * ((C)uL[k] + (C)uR[k]);
            }
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                gx[k] = (C)gL.x[k];
                gy[k] = (C)gL.y[k];
                Um[k] = (C)uL[k];
            }
        }
    }
    const C one_on_rho = frcp(Um[0]);
    const C u = Um[1] * one_on_rho, v = Um[2] * one_on_rho, E = Um[3] * one_on_rho;
    const C dudx = one_on_rho * (gx[1] - u * gx[0]);
    const C dvdx = one_on_rho * (gx[2] - v * gx[0]);
    const C dudy = one_on_rho * (gy[1] - u * gy[0]);
    const C dvdy = one_on_rho * (gy[2] - v * gy[0]);
    const C dEdx = one_on_rho * (gx[3] - E * gx[0]);
    const C dEdy = one_on_rho * (gy[3] - E * gy[0]);
    const C one_on_Cv = frcp(Cv);
    const C V2 = u * u + v * v;
    const C T = one_on_Cv * (E - C(0.5) * V2);
    const C dTdx = (dEdx - u * dudx - v * dvdx) * one_on_Cv;
    const C dTdy = (dEdy - u * dudy - v * dvdy) * one_on_Cv;
    const C mu = (C)mu_of_T(ph, (double)T);
    const C NMD = C(-2.0 / 3.0);
    const C div = dudx + dvdy;
    const C txx = mu * (NMD * div + C(2.0) * dudx);
    const C tyy = mu * (NMD * div + C(2.0) * dvdy);
    const C txy = mu * (dudy + dvdx);
    const C akmu = Cp * fdiv(mu, (C)ph.Pr_visc);
    const C Phix = akmu * dTdx + u * txx + v * txy;
    const C Phiy = akmu * dTdy + u * txy + v * tyy;
    F[0] = C(0.0);
    F[1] = (txx * nx + txy * ny) * area;
    F[2] = (txy * nx + tyy * ny) * area;
    F[3] = (Phix * nx + Phiy * ny) * area;
}

// The boundary state a farfield-type face uses: inlet and outlet their own, the double Mach
// reflection boundary inlet or outlet by the shock position at t_rk, the others the
// freestream (FluxGPU.cu:236-276).
__device__ __forceinline__ const double *bc_state(const Phys &ph, const CaseBC &bc, int bType, double xm, double ym,
                                                  double t_rk) {
    if (bType == FVMR_BC_INLET) return bc.inlet;
    if (bType == FVMR_BC_OUTLET) return bc.outlet;
    if (bType == FVMR_BC_DSR) {
        const double right = ph.dsr_x + (ym + ph.dsr_speed * t_rk) / sqrt(3.0);
        return (xm < right) ? bc.inlet : bc.outlet;
    }
    return bc.inf;
}

// Ghost state R of a boundary face of type bType from the reconstructed interior state L
// (FluxGPU.cu:63-167): slip walls and symmetry reflect the normal velocity, the adiabatic
// wall keeps rho and rhoE with zero momenta, the isothermal wall is new, every other type
// is a Riemann-invariant farfield against ruvp.
template <class C, class P>
__device__ __forceinline__ void ghost_state(const Phys &ph, int bType, const P *ruvp, const FaceState<C> &L,
                                            FaceState<C> &R, C nx, C ny, C gamma) {
    if (bType == FVMR_BC_SYMMETRY || bType == FVMR_BC_WALL_NONVISCOUS) {
        ghost_slip(L, R, nx, ny, gamma);
    } else if (bType == FVMR_BC_WALL_ADIABAT) {
        // FluxGPU.cu:79-85: momenta of a zeroed U_R negated
        R.U[0] = L.U[0]; R.U[1] = -C(0); R.U[2] = -C(0); R.U[3] = L.U[3];
        set_prim(R, gamma);
    } else if (bType == FVMR_BC_WALL_ISOTHERMAL) {
        ghost_isothermal(L, R, (C)ph.wall_T, (C)ph.R, gamma);
    } else {
        ghost_farfield(L, ruvp, R, nx, ny, gamma);
    }
}

// get_flux_RiemannSolve_device (FluxGPU.cu:190-205): LLF or Roe on the two face states.
template <class C>
__device__ __forceinline__ void riemann_flux(const Phys &ph, const FaceState<C> &L, const FaceState<C> &R, C nx, C ny,
                                             C len, C F[4]) {
    const C gamma = (C)ph.gamma, g_ga1 = (C)ph.g_ga1;
    if (ph.scheme == FVMR_SOL_LLF) flux_llf(L, R, nx, ny, len, F, gamma, g_ga1);
    else flux_roe(L, R, nx, ny, len, F, gamma, g_ga1, (C)ph.inv_g);
}

// edge_convection_flux_device_kernel body (FluxGPU.cu:207-283) (+ viscous part for NS):
// the canonical flux of a face, oriented L -> R.  bType: BC code of the face's set (-1
// interior); (pmx, pmy): midpoint of the periodic pair face, where R is reconstructed.
template <class C, class G, bool NS, class Cells>
__device__ __forceinline__ void face_flux(const Cells &cells, const Phys &ph, const CaseBC &bc, double t_rk, int Lc,
                                          int Rc, int bType, bool has_pair, const V4<G> eg, const V2<G> em,
                                          const V2<G> pm, C F[4]) {
    const C nx = (C)eg.x, ny = (C)eg.y, len = (C)eg.z;
    const C xe = (C)em.x, ye = (C)em.y;
    const C gamma = (C)ph.gamma;
    FaceState<C> L, R;
    reconstruct<C, G>(cells, ph, Lc, xe, ye, L);
    if (bType == -1 || is_periodic(bType)) {
        // interior and periodic edges (get_UR_inner_and_periodic_device, FluxGPU.cu:169-188)
        if (Rc != -1) {
            C xr = xe, yr = ye;
            if (bType != -1 && has_pair) {
                xr = (C)pm.x;
                yr = (C)pm.y;
            }
            reconstruct<C, G>(cells, ph, Rc, xr, yr, R);
        } else {
#pragma unroll
            for (int i = 0; i < 4; i++) R.U[i] = C(0);
            set_prim(R, gamma);
        }
    } else {
        ghost_state(ph, bType, bc_state(ph, bc, bType, (double)em.x, (double)em.y, t_rk), L, R, nx, ny, gamma);
    }
    riemann_flux(ph, L, R, nx, ny, len, F);
    if (NS) {
        C Fv[4];
        flux_viscous<C, G>(cells, ph, Lc, Rc, nx, ny, len, Fv);
#pragma unroll
        for (int k = 0; k < 4; k++) F[k] -= Fv[k];
    }
}

// The flux of edge e from the global records; the BC type and pair are read only for
// boundary-set edges (bset).  PAIR: the two cells' records prefetched together (PairCells).
template <class C, class G, bool NS, bool PAIR = false>
__device__ __forceinline__ void edge_flux(const Mesh<G> &m, const Phys &ph, const S4 *__restrict__ U,
                                          const Grad<G> *__restrict__ grad, double t_rk, int e, int Lc, int Rc,
                                          bool bset, C F[4]) {
    const V4<G> eg = m.egeo[e];
    const V2<G> em = m.emid[e];
    int bType = -1, pe = -1;
    if (bset) {
        const int4 ei = m.eint[e];
        bType = ei.z;
        pe = ei.w;
    }
    const bool has_pair = pe >= 0 && pe < m.nf;
    V2<G> pm = em;
    if (bType != -1 && has_pair) pm = m.emid[pe];
    if (PAIR) {
        const PairCells<G> cells(m, U, grad, Lc, Rc);
        face_flux<C, G, NS>(cells, ph, ph.bc, t_rk, Lc, Rc, bType, has_pair, eg, em, pm, F);
    } else {
        const GlobalCells<G> cells{m, U, grad};
        face_flux<C, G, NS>(cells, ph, ph.bc, t_rk, Lc, Rc, bType, has_pair, eg, em, pm, F);
    }
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Spectral radius of one edge for the cells a and b (has_b false: boundary, one-sided)
// (update_alphaC_host, CalculateDt.cpp:24-62).  ra/rb: primitives of a and b.  Returns the
// inviscid part in *inv and the viscous part divided by V_c in *vis (0 for Euler).  T: the
// arithmetic (FP64; FP32 in mixed precision, whose dt minimum stays FP64); the viscous part is
// formed in FP64.
template <class T>
__device__ __forceinline__ void edge_alpha(const Phys &ph, const T ra[4], const T rb[4], bool has_b, T nx, T ny, T len,
                                           T vol_c, T *inv, T *vis) {
    const T gamma = (T)ph.gamma;
    T rho = ra[0], u = ra[1], v = ra[2], p = ra[3];
    if (has_b) {
        rho = T(0.5) * (ra[0] + rb[0]);
        u = T(0.5) * (ra[1] + rb[1]);
        v = T(0.5) * (ra[2] + rb[2]);
        p = T(0.5) * (ra[3] + rb[3]);
    }
    const T dx = nx * len, dy = ny * len;
    if (std::is_same<T, double>::value)
        *inv = fabs(u * dx + v * dy) + sqrt(fdiv(gamma * p, rho)) * sqrt(len);
    else
        *inv = fabs(u * dx + v * dy) + fsqrt(fdiv(gamma * p, rho)) * fsqrt(len);
    *vis = T(0);
    if (ph.equation == FVMR_EQ_NS) {
        const double value_1 = fmax(4.0 / 3.0, ph.gamma / ph.Pr);
        const double Tt = (double)p / ph.R / (double)rho;
        *vis = (T)(2.0 * (double)len * value_1 * mu_of_T(ph, Tt) / (double)rho / (double)vol_c);
    }
}

// Gradient and limiter of one cell from its face-slot data (GradientGreenGaussKernel,
// GradientGPU.cu:223-425, or the least-squares gradient, Gradient.cpp:160-246, with the
// correct 2x2 inverse unless compat; then LimiterBarth_device_kernel, GradientGPU.cu:453-502,
// or Venkatakrishnan).  Per slot i: fl bit 0 = a cell across, bit 1 = a non-periodic
// neighbour; A = (gC, Sx s, Sy s) for Green-Gauss or (dx, dy) for least squares; Uo = the
// state across (own state when none); N = vertex i relative to the centroid.  NV: the number of
// conserved variables formed (they do not mix; the adjoint forms one at a time).
template <class C, class Slots, int NV = 4, class V = C>
__device__ __forceinline__ void grad_limit(const Phys &ph, const C Uc[NV], const Slots &S, const int fl[3], V vol,
                                           C gx[NV], C gy[NV]) {
    // the slot coefficients and vertex offsets keep the accessor's type (the adjoint's are
    // plain doubles beside dual states)
    using A_t = typename std::decay<decltype(S.A(0, 0))>::type;
    using N_t = typename std::decay<decltype(S.N(0, 0))>::type;
    if (ph.gradient == FVMR_GRA_GREENGAUSS) {
        const V inv_vol = frcp(vol);
#pragma unroll
        for (int k = 0; k < NV; k++) {
            C s0 = C(0), s1 = C(0);
#pragma unroll
            for (int i = 0; i < 3; i++) {
                // a slot with no cell across has g = 1/2 and the cell's own state, for which the
                // blend gives U_C exactly, so the plain types blend every slot
                const A_t g = S.A(i, 0);
                const bool blend = std::is_floating_point<C>::value || (fl[i] & 1);
                const C phif = blend ? g * Uc[k] + (A_t(1.0) - g) * S.Uo(i, k) : Uc[k];
                s0 += phif * S.A(i, 1);
                s1 += phif * S.A(i, 2);
            }
            gx[k] = inv_vol * s0;
            gy[k] = inv_vol * s1;
        }
    } else {
        A_t M[4];   // A^T A
#pragma unroll
        for (int x = 0; x < 2; x++)
#pragma unroll
            for (int y = 0; y < 2; y++) {
                A_t acc = A_t(0);
#pragma unroll
                for (int i = 0; i < 3; i++)
                    if (fl[i] & 2) acc += S.A(i, x) * S.A(i, y);
                M[x * 2 + y] = acc;
            }
        A_t det = M[0] * M[3] - M[1] * M[2];
        const A_t eps = A_t(fvmr::kEPS);
        if (ph.compat_inv) {  // MatrixKernel.h:48-49 as written
            if (det >= A_t(0) && det < eps) det = eps;
            else if (det > -eps) det = -eps;
        } else {
            if (det >= A_t(0) && det < eps) det = eps;
            else if (det < A_t(0) && det > -eps) det = -eps;
        }
        const A_t idet = frcp(det);
        const A_t i00 = M[3] * idet, i01 = -M[1] * idet, i10 = -M[2] * idet, i11 = M[0] * idet;
        const bool any = (fl[0] | fl[1] | fl[2]) & 2;
#pragma unroll
        for (int k = 0; k < NV; k++) {
            C a = C(0), b = C(0);
#pragma unroll
            for (int i = 0; i < 3; i++)
                if (fl[i] & 2) a += (i00 * S.A(i, 0) + i01 * S.A(i, 1)) * (S.Uo(i, k) - Uc[k]);
#pragma unroll
            for (int i = 0; i < 3; i++)
                if (fl[i] & 2) b += (i10 * S.A(i, 0) + i11 * S.A(i, 1)) * (S.Uo(i, k) - Uc[k]);
            gx[k] = any ? a : C(0);
            gy[k] = any ? b : C(0);
        }
    }
    if (ph.limiter != FVMR_LIM_NONE) {
        C eps2 = C(0);
        if (ph.limiter == FVMR_LIM_VENKATAKRISHNAN) {
            const V Kh = (V)ph.venkat_K * fsqrt(vol);
            eps2 = Kh * Kh * Kh;
        }
#pragma unroll
        for (int k = 0; k < NV; k++) {
            // bounds over the cell and its non-periodic neighbours (fmin/fmax: a NaN bound
            // already makes the Green-Gauss or least-squares gradient NaN).  Without periodic
            // faces a slot with no neighbour holds the cell's own state, which leaves both
            // bounds as they are, so the plain types skip the per-slot test
            C lo = Uc[k], hi = Uc[k];
            if (std::is_floating_point<C>::value && !ph.has_periodic) {
#pragma unroll
                for (int i = 0; i < 3; i++) { lo = fmin(lo, S.Uo(i, k)); hi = fmax(hi, S.Uo(i, k)); }
            } else {
#pragma unroll
                for (int i = 0; i < 3; i++)
                    if (fl[i] & 2) { lo = fmin(lo, S.Uo(i, k)); hi = fmax(hi, S.Uo(i, k)); }
            }
            C phi = C(1.0);
            const C hrel = hi - Uc[k], lrel = lo - Uc[k];
            if (ph.limiter == FVMR_LIM_BARTH && !isnan(hrel) && !isnan(lrel)) {
                // min over vertices of min(1, hrel / (rel + eps)) (rel > 0) and of
                // min(1, lrel / (rel - eps)) (rel < 0): division rounds monotonically, so each
                // minimum is one division by the extreme denominator, needed only when the
                // quotient can fall below 1
                C dpos = C(0), dneg = C(0);
                if (std::is_floating_point<C>::value) {
                    // rounding is monotonic, so the largest rel + eps is the largest rel plus
                    // eps: the extremes first, eps once (the same values, no per-vertex selects)
                    C mp = C(0), mn = C(0);
#pragma unroll
                    for (int jn = 0; jn < 3; jn++) {
                        const C Un = Uc[k] + gx[k] * S.N(jn, 0) + gy[k] * S.N(jn, 1);
                        const C rel = Un - Uc[k];
                        mp = fmax(mp, rel);
                        mn = fmin(mn, rel);
                    }
                    if (mp > C(0)) dpos = mp + C(fvmr::kEPS);
                    if (mn < C(0)) dneg = mn - C(fvmr::kEPS);
                } else {
#pragma unroll
                    for (int jn = 0; jn < 3; jn++) {
                        const C Un = Uc[k] + gx[k] * S.N(jn, 0) + gy[k] * S.N(jn, 1);
                        const C rel = Un - Uc[k];
                        if (rel > C(0)) dpos = fmax(dpos, rel + C(fvmr::kEPS));
                        else if (rel < C(0)) dneg = fmin(dneg, rel - C(fvmr::kEPS));
                    }
                }
                if (dpos > C(0) && hrel < dpos) phi = tmin(phi, fdiv(hrel, dpos));
                if (dneg < C(0) && lrel > dneg) phi = tmin(phi, fdiv(lrel, dneg));
            } else if (ph.limiter == FVMR_LIM_BARTH) {
                // NaN bounds: the vertex loop as the reference writes it
#pragma unroll
                for (int jn = 0; jn < 3; jn++) {
                    const C Un = Uc[k] + gx[k] * S.N(jn, 0) + gy[k] * S.N(jn, 1);
                    C rel = Un - Uc[k];
                    C pn = C(1.0);
                    if (rel > C(0)) { rel += C(fvmr::kEPS); pn = tmin(C(1), fdiv(hrel, rel)); }
                    else if (rel < C(0)) { rel -= C(fvmr::kEPS); pn = tmin(C(1), fdiv(lrel, rel)); }
                    phi = tmin(phi, pn);
                }
            } else {
#pragma unroll
                for (int jn = 0; jn < 3; jn++) {
                    const N_t ndx = S.N(jn, 0), ndy = S.N(jn, 1);
                    C pn = C(1.0);
                    const C d2 = gx[k] * ndx + gy[k] * ndy;
                    if (d2 != C(0)) {
                        const C d1 = (d2 > C(0)) ? (hi - Uc[k]) : (lo - Uc[k]);
                        const C num = (d1 * d1 + eps2) * d2 + C(2.0) * d2 * d2 * d1;
                        const C den = d1 * d1 + C(2.0) * d2 * d2 + d1 * d2 + eps2;
                        pn = fdiv(num, den * d2);
                    }
                    phi = tmin(phi, pn);
                }
            }
            gx[k] *= phi;
            gy[k] *= phi;
        }
    }
}

// Gradient_2 (GradientGPU.cu:513-551): GG (GradientGreenGaussKernel :223-425) or LSQ
// (Gradient.cpp:160-246, correct inverse unless compat), then Barth (:453-502) or
// Venkatakrishnan, one thread per cell: its three slots' neighbour states, face geometry and
// vertices are gathered into registers (all loads issued in two rounds before any is used).
// With DT the kernel also forms alphaC of the cell from the same loads, summed in the
// reference edge order, and writes the block minimum of CFL V / alphaC (get_global_dt_device,
// CalculateDtGPU.cu:97-103).
template <class C, class G, bool DT>
__global__ __launch_bounds__(kBlock) FVMR_GRAD_ATTR void k_gradient(Mesh<G> m, Phys ph, const S4 *__restrict__ U,
                                                     Grad<G> *__restrict__ grad, double *__restrict__ partial) {
    __shared__ double sMin[kBlock / 64];
    const int blk = xcd_block();
    const int c = blk * kBlock + threadIdx.x;
    const int nc = m.nc;
    double dtc = 1e10;
    if (c < nc) {
        // the loads in two rounds: the cell's own records, then what they point to
        const V4<G> cg = m.cgeo[c];
        int2 sl[3];
        int cn[3];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            sl[j] = m.cslot[j * nc + c];
            cn[j] = m.cnode[j * nc + c];
        }
        double Ud[4];
        ld(U, c, Ud);
        double Uo[3][4];
        V2<G> nd[3], em[3];
        V4<G> eg[3], og[3];
        int pe[3];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int e = sl[j].x & kSlotEdgeMask, out = sl[j].y;
            const int oc = out != -1 ? out : c;
            ld(U, oc, Uo[j]);
            nd[j] = m.nxy[cn[j]];
            eg[j] = m.egeo[e];
            em[j] = m.emid[e];
            og[j] = m.cgeo[oc];
            pe[j] = (sl[j].x & kSlotPeriodic) ? m.eint[e].w : -1;
        }
        const C xc = (C)cg.x, yc = (C)cg.y;
        struct {
            C a[3][3], u[3][4], n[3][2];
            __device__ __forceinline__ C A(int i, int q) const { return a[i][q]; }
            __device__ __forceinline__ C Uo(int i, int k) const { return u[i][k]; }
            __device__ __forceinline__ C N(int i, int q) const { return n[i][q]; }
        } S;
        int fl[3];
        double al[DT ? 3 : 1][4];
        double rc[4];
        if (DT) prim(Ud, rc, ph.gamma);
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int out = sl[j].y;
            const bool periodic = (sl[j].x & kSlotPeriodic) != 0;
            const C sgn = (sl[j].x & kSlotCellIsR) ? C(-1) : C(1);
#pragma unroll
            for (int k = 0; k < 4; k++) S.u[j][k] = (C)Uo[j][k];
            S.n[j][0] = (C)nd[j].x - xc;
            S.n[j][1] = (C)nd[j].y - yc;
            if (ph.gradient == FVMR_GRA_GREENGAUSS) {
                const C len = (C)eg[j].z;
                C gC = C(0.5);
                if (out != -1) {
                    const C fx = (C)em[j].x, fy = (C)em[j].y;
                    C Fx = (C)og[j].x, Fy = (C)og[j].y;
                    if (periodic && pe[j] >= 0) {
                        const V2<G> pm = m.emid[pe[j]];
                        Fx += fx - (C)pm.x;
                        Fy += fy - (C)pm.y;
                    }
                    const C ax = fx - Fx, ay = fy - Fy, bx = xc - Fx, by = yc - Fy;
                    const C dFf = fsqrt(ax * ax + ay * ay);
                    C dFC = fsqrt(bx * bx + by * by);
                    if (dFC < C(fvmr::kEPS)) dFC = C(fvmr::kEPS);
                    gC = fdiv(dFf, dFC);
                }
                S.a[j][0] = gC;
                S.a[j][1] = (C)eg[j].x * len * sgn;
                S.a[j][2] = (C)eg[j].y * len * sgn;
            } else {
                const int nbr = periodic ? -1 : out;   // element neighbours exclude periodic faces
                C dxn = C(0), dyn = C(0);
                if (nbr != -1) {
                    dxn = (C)og[j].x - xc;
                    dyn = (C)og[j].y - yc;
                }
                S.a[j][0] = dxn;
                S.a[j][1] = dyn;
                S.a[j][2] = C(0);
            }
            fl[j] = ((out != -1) ? 1 : 0) | ((out != -1 && !periodic) ? 2 : 0);
            if (DT) {
                double ro[4];
                prim(Uo[j], ro, ph.gamma);
                const double vol_c = (double)cg.z;
                double a0, v0, a1 = 0.0, v1 = 0.0;
                edge_alpha(ph, rc, ro, out != -1, (double)eg[j].x, (double)eg[j].y, (double)eg[j].z, vol_c, &a0, &v0);
                if (pe[j] >= 0) {
                    const V4<G> pg = m.egeo[pe[j]];
                    edge_alpha(ph, rc, ro, true, (double)pg.x, (double)pg.y, (double)pg.z, vol_c, &a1, &v1);
                }
                al[DT ? j : 0][0] = a0;
                al[DT ? j : 0][1] = v0;
                al[DT ? j : 0][2] = a1;
                al[DT ? j : 0][3] = v1;
            }
        }
        C Uc[4];
#pragma unroll
        for (int k = 0; k < 4; k++) Uc[k] = (C)Ud[k];
        C gx[4], gy[4];
        grad_limit<C>(ph, Uc, S, fl, (C)cg.z, gx, gy);
        Grad<G> o;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            o.x[k] = (G)gx[k];
            o.y[k] = (G)gy[k];
        }
        grad[c] = o;
        if (DT) {
            // alphaC in the reference edge order: count (3 bits) + 6 x (slot 2 bits, pair 1 bit)
            const int code = m.dtcode[c];
            const int cnt = code & 7;
            double alpha_c = 0.0;
            for (int q = 0; q < cnt; q++) {
                const int ent = (code >> (3 + 3 * q)) & 7;
                const int slot = ent & 3, pr = ent >> 2;
                double av = 0.0, vv = 0.0;
#pragma unroll
                for (int j = 0; j < 3; j++)
                    if (slot == j) {
                        av = al[DT ? j : 0][2 * pr];
                        vv = al[DT ? j : 0][2 * pr + 1];
                    }
                alpha_c += av;
                if (ph.equation == FVMR_EQ_NS) alpha_c += vv;
            }
            dtc = fdiv(ph.bc.CFL * (double)cg.z, alpha_c);
        }
    }
    if (DT) {
        const double wm = wave_min(dtc);
        if ((threadIdx.x & 63) == 0) sMin[threadIdx.x >> 6] = wm;
        __syncthreads();
        if (threadIdx.x == 0) {
            double b = sMin[0];
            for (int w = 1; w < kBlock / 64; w++) b = fmin(b, sMin[w]);
            partial[blk] = b;
        }
    }
}

// One RK stage: rhs = -(1/V) sum_j s_j F_{e_j} (add_edge_flux_to_element_flux_kernel, FluxGPU.cu:285-323,
// vector_weighted_divide_kernel, BasicAlgorithmGPU.cu:4-9), then
//   stage 0:    Uout = Uin + dt rhs                      (vector_weighted_add_kernel)
//   stage 1, 2: Uout = a Un + b (Uin + dt rhs)           (SSP-RK3)
// Wave j evaluates the canonical flux of face slot j of 64 cells; wave 0 sums the three
// signed fluxes in slot order and updates.  In the step's last stage (last != 0) cells whose
// new state is NaN or out of range are appended to the restriction list.
template <class C, class G, bool NS>
__global__ __launch_bounds__(kGroup3) FVMR_STAGE_ATTR void k_stage(Mesh<G> m, Phys ph, const S4 *__restrict__ Uin,
                                                   const S4 *__restrict__ Un, S4 *__restrict__ Uout,
                                                   const Grad<G> *__restrict__ grad, S4 *__restrict__ rhs_out,
                                                   Scalars *__restrict__ sc, int *__restrict__ bad_list, double a,
                                                   double b, int first, int last) {
    __shared__ double sF[3][4][kCellsPerGroup];
    const int lane = threadIdx.x & 63, j = threadIdx.x >> 6;
    const int c = xcd_block() * kCellsPerGroup + lane;
    const int nc = m.nc;
    if (c < nc) {
        const double t_rk = sc->t_rk;
        const int2 sl = m.cslot[j * nc + c];
        const bool cR = (sl.x & kSlotCellIsR) != 0;
        C F[4];
        // FP64 Euler prefetches both cells' records (the FP32 and viscous variants would lose occupancy)
        edge_flux<C, G, NS, !NS && std::is_same<C, double>::value>(m, ph, Uin, grad, t_rk, sl.x & kSlotEdgeMask,
                                                                   cR ? sl.y : c, cR ? c : sl.y,
                                                                   (sl.x & kSlotBoundary) != 0, F);
        const double s = cR ? -1.0 : 1.0;   // +1 when the cell is not the edge's R
#pragma unroll
        for (int k = 0; k < 4; k++) sF[j][k][lane] = s * (double)F[k];
    }
    __syncthreads();
    if (j != 0 || c >= nc) return;
    const double dt = sc->dt;
    const double w = -1.0 * (double)m.cgeo[c].z;
    double Ui[4], Ub[4] = {0, 0, 0, 0}, r[4], Unew[4];
    ld(Uin, c, Ui);
    if (!first) ld(Un, c, Ub);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        double acc = 0.0;
        acc += sF[0][k][lane];
        acc += sF[1][k][lane];
        acc += sF[2][k][lane];
        r[k] = fdiv(acc, w);
        const double s = Ui[k] + dt * r[k];
        Unew[k] = first ? s : a * Ub[k] + b * s;
    }
    if (rhs_out) st(rhs_out, c, r);
    if (std::is_same<C, double>::value) st_stream(Uout, c, Unew);
    else st(Uout, c, Unew);
    if (!last) return;
    double rv[4];
    prim(Unew, rv, ph.gamma);
    bool bad = out_of_range(rv[0], rv[3]);
#pragma unroll
    for (int k = 0; k < 4; k++) bad = bad || isnan(rv[k]);
    if (bad) {
        const int slot = atomicAdd(&sc->bad_count, 1);
        if (slot < nc) {
            bad_list[slot] = c;
            bad_list[nc + slot] = 0;   // the case of the cell (the cell family runs one case)
        }
    }
}

// ---------------------------------------------------------------------------
// Tile kernels: one work-group per tile of consecutive cells (tile records built by
// build_tiles on the host).  The work-group stages the states of the tile and two halo
// rings, the nodes and the face records in LDS, recomputes centroids and face geometry from
// the node coordinates with the host's arithmetic, forms the limited gradients of the tile
// and the first ring in LDS, evaluates the flux of every face of the tile once, and sums
// the fluxes per cell in slot order.  Gradients never go to HBM.
// ---------------------------------------------------------------------------
// Work-group shapes of the tile kernels per arithmetic precision.  The default (S = 0) is sized
// so that the LDS of a work-group lets as many work-groups share a CU as the registers allow
// (FP64: about 104 VGPRs -> 4 waves per SIMD -> 4 work-groups of 4 waves with <= 40 KB LDS each;
// FP32: about 74 VGPRs -> 6 waves per SIMD -> 3 work-groups of 8 waves with <= 53 KB each).
// S = 1 halves the work-group (smaller tiles, more work-groups per CU), S = 2 doubles it (FP64
// only: the FP32 shape's local ids would not fit their 10-bit fields).  A handle uses one shape
// (fvmr_params.tile_threads); the tuners pick among them.
template <class C, int S = 0> struct TileCfg;
template <> struct TileCfg<double, 0> {
    static constexpr int kThreads = 256;
    static constexpr int kMaxG = 200;    // cells with a gradient (tile + first ring)
    static constexpr int kMaxL = 256;    // cells with a state (+ second ring)
    static constexpr int kMaxN = 176;    // nodes
    static constexpr int kMaxF = 352;    // faces
    static constexpr int kMaxFF = kThreads;   // faces of tile cells: one per thread
    static constexpr int kCells = 150;   // default tile size
};
template <> struct TileCfg<double, 1> {
    static constexpr int kThreads = 128;
    static constexpr int kMaxG = 100, kMaxL = 128, kMaxN = 96, kMaxF = 176, kMaxFF = kThreads;
    static constexpr int kCells = 72;
};
template <> struct TileCfg<double, 2> {
    static constexpr int kThreads = 512;
    static constexpr int kMaxG = 400, kMaxL = 512, kMaxN = 352, kMaxF = 704, kMaxFF = kThreads;
    static constexpr int kCells = 300;
};
template <> struct TileCfg<float, 0> {
    static constexpr int kThreads = 512;
    static constexpr int kMaxG = 400;
    static constexpr int kMaxL = 480;
    static constexpr int kMaxN = 320;
    static constexpr int kMaxF = 672;
    static constexpr int kMaxFF = kThreads;
    static constexpr int kCells = 300;
};
template <> struct TileCfg<float, 1> {
    static constexpr int kThreads = 256;
    static constexpr int kMaxG = 200, kMaxL = 240, kMaxN = 160, kMaxF = 336, kMaxFF = kThreads;
    static constexpr int kCells = 150;
};
constexpr int kTileMaxCells = 400;   // largest tile_cells accepted (the largest shapes' kMaxG)
constexpr int kNumShapes[2] = {3, 2};   // shapes per precision: FP64, FP32-mixed
template <class Cfg> constexpr bool tile_cfg_fits() {
    return Cfg::kMaxL <= fvmr::kLocalNone && Cfg::kMaxN <= fvmr::kLocalNone && Cfg::kMaxF <= fvmr::kLocalNone &&
           Cfg::kMaxG <= kTileMaxCells && Cfg::kMaxL <= Cfg::kThreads && Cfg::kMaxN <= Cfg::kThreads &&
           Cfg::kMaxF <= 2 * Cfg::kThreads && Cfg::kMaxG <= Cfg::kMaxL;
}
static_assert(tile_cfg_fits<TileCfg<double, 0>>() && tile_cfg_fits<TileCfg<double, 1>>() &&
              tile_cfg_fits<TileCfg<double, 2>>() && tile_cfg_fits<TileCfg<float, 0>>() &&
              tile_cfg_fits<TileCfg<float, 1>>(), "tile shapes: 10-bit local ids, one cell and two faces per thread");
// threads of shape S of a precision (mixed: FP32)
inline int shape_threads(bool mixed, int S) {
    if (mixed) return S == 0 ? TileCfg<float, 0>::kThreads : TileCfg<float, 1>::kThreads;
    return S == 0 ? TileCfg<double, 0>::kThreads : (S == 1 ? TileCfg<double, 1>::kThreads : TileCfg<double, 2>::kThreads);
}
inline int shape_cells(bool mixed, int S) {
    if (mixed) return S == 0 ? TileCfg<float, 0>::kCells : TileCfg<float, 1>::kCells;
    return S == 0 ? TileCfg<double, 0>::kCells : (S == 1 ? TileCfg<double, 1>::kCells : TileCfg<double, 2>::kCells);
}
// the shape of a precision with `threads` work-items (0: the default), -1 if none
inline int shape_of_threads(bool mixed, int threads) {
    if (threads == 0) return 0;
    for (int S = 0; S < kNumShapes[mixed ? 1 : 0]; S++)
        if (shape_threads(mixed, S) == threads) return S;
    return -1;
}

// Per cell, for the time-step kernel: alphaC code, the 3 slot faces (local ids in the
// cell's tile), volume.
struct DtRec {
    int code;
    int faces;
    double vol;
};

template <class G>
struct TileArgs {
    const int *info;          // fvmr::kTileInfo ints per tile
    const int *halo;          // device cell ids of halo cells
    const V2<double> *nxy;    // node coordinates per tile
    const int *lnodes;        // per local cell: 3 local node ids
    const int4 *gslot;        // per gradient cell: 3 slot words + dtcode
    const V4<G> *gcoef;       // per gradient cell: Green-Gauss weights of the 3 slots, volume
    const int4 *face;         // per face: nodes | BC code, L | R | pair face, float length, edge
    const DtRec *dtrec;       // per device cell
    const V2<G> *lcxy;        // per local cell: centroid (FVMR_TILE_GEOM_LOADED)
    const V2<G> *fnrm;        // per face: normal (FVMR_TILE_GEOM_LOADED)
    const CaseBC *cbc;        // per case of the batch (one case: the handle's)
    double *dtl;              // per-cell steps of a steady solve (k_tile_dt / k_tile with LDT), else NULL
};

__device__ __forceinline__ int lfield(int w, int k) { return (w >> (fvmr::kLocalBits * k)) & fvmr::kLocalMask; }

// Centroid of a cell from its nodes as the host computes it ((0 + x0) + x1 + x2) / 3 (the
// quotient by fdiv, which rounds as IEEE division does).
__device__ __forceinline__ V2<double> tile_centroid(const V2<double> *sN, int w) {
#pragma clang fp contract(off)
    const V2<double> a = sN[lfield(w, 0)], b = sN[lfield(w, 1)], c = sN[lfield(w, 2)];
    double sx = 0.0, sy = 0.0;
    sx += a.x; sx += b.x; sx += c.x;
    sy += a.y; sy += b.y; sy += c.y;
    return V2<double>{fdiv(sx, 3.0), fdiv(sy, 3.0)};
}

// Face geometry from a face record as the host computes it: normal (dy, -dx) / L with the
// stored (float) length L, midpoint (xa + xb) / 2; rounded to G like the global records.
template <class G>
__device__ __forceinline__ void tile_face_geom(const V2<double> *sN, int4 fr, V4<G> &eg, V2<G> &em) {
#pragma clang fp contract(off)
    const V2<double> a = sN[lfield(fr.x, 0)], b = sN[lfield(fr.x, 1)];
    const double len = (double)__int_as_float(fr.z);
    const double dx = b.x - a.x, dy = b.y - a.y;
    eg = V4<G>{(G)fdiv(dy, len), (G)fdiv(-dx, len), (G)len, G(0)};
    em = V2<G>{(G)((a.x + b.x) / 2.0), (G)((a.y + b.y) / 2.0)};
}
// Face geometry with the normal of face f already formed in sNrm.
template <class G>
__device__ __forceinline__ void tile_face_geom_n(const V2<double> *sN, const V2<G> *sNrm, int f, int4 fr, V4<G> &eg,
                                                 V2<G> &em) {
#pragma clang fp contract(off)
    const V2<double> a = sN[lfield(fr.x, 0)], b = sN[lfield(fr.x, 1)];
    const V2<G> n = sNrm[f];
    eg = V4<G>{n.x, n.y, (G)(double)__int_as_float(fr.z), G(0)};
    em = V2<G>{(G)((a.x + b.x) / 2.0), (G)((a.y + b.y) / 2.0)};
}
template <class G>
__device__ __forceinline__ V2<G> tile_face_mid(const V2<double> *sN, int4 fr) {
#pragma clang fp contract(off)
    const V2<double> a = sN[lfield(fr.x, 0)], b = sN[lfield(fr.x, 1)];
    return V2<G>{(G)((a.x + b.x) / 2.0), (G)((a.y + b.y) / 2.0)};
}
__device__ __forceinline__ int tile_btype(int4 fr) {
    const int bc = (fr.x >> fvmr::kBcShift) & 127;
    return bc ? bc + fvmr::kBcBase : -1;
}

// The staged state record of a tile kernel computing in C: FP64 in the plain path; in mixed
// precision the FP32 values the arithmetic uses (converted once when staged; the RK update
// takes each tile cell's FP64 state from the staging registers).
template <class C> using TileState = typename std::conditional<std::is_same<C, float>::value, V4<float>, S4>::type;

// Records in LDS as 16-byte columns: a table of n records of 32 or 64 bytes is 2 or 4 arrays
// of n 16-byte parts, part k of record i at column k, entry i.  Entry i of a column starts at
// bank 4i mod 32 (4-byte banks), so the random gathers of a wave spread over all bank groups; a
// 32-byte record starts at bank 8i mod 32 and a 64-byte one at 16i mod 32, half and a quarter
// of them.
template <class T> __device__ __forceinline__ T *lds_col(void *base, int n, int k) {
    return reinterpret_cast<T *>(base) + k * n;
}
template <class T> __device__ __forceinline__ const T *lds_col(const void *base, int n, int k) {
    return reinterpret_cast<const T *>(base) + k * n;
}
__device__ __forceinline__ void lds_put(void *b, int n, int i, const V4<double> &v) {
    lds_col<V2<double>>(b, n, 0)[i] = V2<double>{v.x, v.y};
    lds_col<V2<double>>(b, n, 1)[i] = V2<double>{v.z, v.w};
}
__device__ __forceinline__ void lds_put(void *b, int n, int i, const V4<float> &v) { lds_col<V4<float>>(b, n, 0)[i] = v; }
__device__ __forceinline__ void lds_put(void *b, int n, int i, const Grad<double> &g) {
    lds_col<V2<double>>(b, n, 0)[i] = V2<double>{g.x[0], g.x[1]};
    lds_col<V2<double>>(b, n, 1)[i] = V2<double>{g.x[2], g.x[3]};
    lds_col<V2<double>>(b, n, 2)[i] = V2<double>{g.y[0], g.y[1]};
    lds_col<V2<double>>(b, n, 3)[i] = V2<double>{g.y[2], g.y[3]};
}
__device__ __forceinline__ void lds_put(void *b, int n, int i, const Grad<float> &g) {
    lds_col<V4<float>>(b, n, 0)[i] = V4<float>{g.x[0], g.x[1], g.x[2], g.x[3]};
    lds_col<V4<float>>(b, n, 1)[i] = V4<float>{g.y[0], g.y[1], g.y[2], g.y[3]};
}
template <class T> __device__ __forceinline__ T lds_get(const void *b, int n, int i);
template <> __device__ __forceinline__ V4<double> lds_get<V4<double>>(const void *b, int n, int i) {
    const V2<double> p = lds_col<V2<double>>(b, n, 0)[i], q = lds_col<V2<double>>(b, n, 1)[i];
    return V4<double>{p.x, p.y, q.x, q.y};
}
template <> __device__ __forceinline__ V4<float> lds_get<V4<float>>(const void *b, int n, int i) {
    return lds_col<V4<float>>(b, n, 0)[i];
}
template <> __device__ __forceinline__ Grad<double> lds_get<Grad<double>>(const void *b, int n, int i) {
    const V2<double> a = lds_col<V2<double>>(b, n, 0)[i], c = lds_col<V2<double>>(b, n, 1)[i];
    const V2<double> d = lds_col<V2<double>>(b, n, 2)[i], e = lds_col<V2<double>>(b, n, 3)[i];
    Grad<double> g;
    g.x[0] = a.x; g.x[1] = a.y; g.x[2] = c.x; g.x[3] = c.y;
    g.y[0] = d.x; g.y[1] = d.y; g.y[2] = e.x; g.y[3] = e.y;
    return g;
}
template <> __device__ __forceinline__ Grad<float> lds_get<Grad<float>>(const void *b, int n, int i) {
    const V4<float> a = lds_col<V4<float>>(b, n, 0)[i], c = lds_col<V4<float>>(b, n, 1)[i];
    Grad<float> g;
    g.x[0] = a.x; g.x[1] = a.y; g.x[2] = a.z; g.x[3] = a.w;
    g.y[0] = c.x; g.y[1] = c.y; g.y[2] = c.z; g.y[3] = c.w;
    return g;
}

__device__ __forceinline__ void lds_state(const void *b, int n, int i, double u[4]) {
    const V4<double> v = lds_get<V4<double>>(b, n, i);
    u[0] = v.x; u[1] = v.y; u[2] = v.z; u[3] = v.w;
}

// The cells of a tile in LDS (states and gradients in 16-byte columns of nU and nGr entries).
template <class G, class SU = S4>
struct TileCells {
    const void *sU;
    int nU;
    const V2<G> *sC;
    const void *sGr;
    int nGr;
    __device__ __forceinline__ void state(int c, double u[4]) const {
        const SU v = lds_get<SU>(sU, nU, c);
        u[0] = v.x; u[1] = v.y; u[2] = v.z; u[3] = v.w;
    }
    template <class C> __device__ __forceinline__ void state_c(int c, C u[4]) const {
        const SU v = lds_get<SU>(sU, nU, c);
        u[0] = (C)v.x; u[1] = (C)v.y; u[2] = (C)v.z; u[3] = (C)v.w;
    }
    __device__ __forceinline__ V2<G> centroid(int c) const { return sC[c]; }
    __device__ __forceinline__ Grad<G> gradient(int c) const { return lds_get<Grad<G>>(sGr, nGr, c); }
    __device__ __forceinline__ bool has(int c) const { return c >= 0; }
};

// Scheme of a k_tile instantiation: read from Phys at run time, or fixed at compile time for
// the common Green-Gauss + Barth + Roe + linear reconstruction (the compiler then drops the
// other branches: 116 -> 99 VGPRs and about 3% of the FP64 kernel's time).
enum TileScheme { kTileAny = 0, kTileGgBarthRoe = 1 };
inline bool tile_scheme_fixed(const Phys &ph) {
    return ph.gradient == FVMR_GRA_GREENGAUSS && ph.limiter == FVMR_LIM_BARTH && ph.scheme == FVMR_SOL_ROE &&
           ph.reconstruct == FVMR_REC_LINEAR;
}

// One RK stage over the tiles (the work of k_gradient + k_stage, see there).
template <class C, class G, bool NS, bool LDT = false, int SCHEME = kTileAny, int SHAPE = 0>
__global__ __launch_bounds__((TileCfg<C, SHAPE>::kThreads)) void k_tile(TileArgs<G> ta, Phys ph, const S4 *__restrict__ Uin,
                                                        const S4 *__restrict__ Un, S4 *__restrict__ Uout,
                                                        S4 *__restrict__ rhs_out, Scalars *__restrict__ sc,
                                                        int *__restrict__ bad_list, int nc, double a, double b,
                                                        int first, int last) {
    if (SCHEME == kTileGgBarthRoe) {
        ph.gradient = FVMR_GRA_GREENGAUSS;
        ph.limiter = FVMR_LIM_BARTH;
        ph.scheme = FVMR_SOL_ROE;
        ph.reconstruct = FVMR_REC_LINEAR;
    }
    using Cfg = TileCfg<C, SHAPE>;
    constexpr int kTileThreads = Cfg::kThreads, kTileMaxG = Cfg::kMaxG, kTileMaxL = Cfg::kMaxL;
    constexpr int kTileMaxN = Cfg::kMaxN, kTileMaxF = Cfg::kMaxF, kTileMaxFF = Cfg::kMaxFF;
    static_assert(kTileMaxL <= kTileThreads && kTileMaxFF <= kTileThreads, "one local cell and one face per thread");
    using SU = TileState<C>;
    __shared__ __align__(16) char sU[kTileMaxL * sizeof(SU)];   // states, 16-byte columns of kTileMaxL
    __shared__ V2<double> sN[kTileMaxN];
    __shared__ V2<G> sC[kTileMaxL];
    __shared__ int4 sFace[kTileMaxF];
    __shared__ V2<G> sNrm[kTileMaxF];   // face normals
    // the face fluxes (16-byte columns of kTileMaxFF) replace the gradients (columns of
    // kTileMaxG) once every flux is formed
    static_assert(kTileMaxFF * sizeof(V4<C>) <= kTileMaxG * sizeof(Grad<G>), "flux buffer overlays the gradients");
    __shared__ __align__(16) char sGrFl[kTileMaxG * sizeof(Grad<G>)];
    const int tid = threadIdx.x;
    const int *ti = ta.info + (size_t)xcd_block() * fvmr::kTileInfo;
    const int c0 = ti[fvmr::kTiC0], nT = ti[fvmr::kTiNT], nG = ti[fvmr::kTiNG], nL = ti[fvmr::kTiNL];
    const int nN = ti[fvmr::kTiNN], nF = ti[fvmr::kTiNF], nFF = ti[fvmr::kTiNFF];
    const int offH = ti[fvmr::kTiOffH], offN = ti[fvmr::kTiOffN], offF = ti[fvmr::kTiOffF];
    const int offL = ti[fvmr::kTiOffL], offG = ti[fvmr::kTiOffG];
    const int cs = ti[fvmr::kTiCase];   // case of the batch; sc[cs] holds its dt and times
    // thread i owns local cell i (phases 2, 3, 5) and face i (phase 4); its records are
    // fetched with the staging loads
    const int i = tid;
    const int nw = (i < nL) ? ta.lnodes[offL + i] : 0;
    const int4 gs = (i < nG) ? ta.gslot[offG + i] : make_int4(0, 0, 0, 0);
    const V4<G> gco = (i < nG) ? ta.gcoef[offG + i] : V4<G>{G(0.5), G(0.5), G(0.5), G(1)};
    const G gvol = gco.w;
    // 1. stage nodes, states and face records: every global load is issued before the first
    // LDS store (a store between loads would make each wave wait for the loads before it)
    static_assert(kTileMaxN <= kTileThreads && kTileMaxF <= 2 * kTileThreads, "one node and two faces per thread");
    const int src = (i < nL) ? (i < nT ? c0 + i : ta.halo[offH + i - nT]) : -1;
    const S4 uin = (src >= 0) ? Uin[src] : S4{0.0, 0.0, 0.0, 0.0};
    const V2<double> nd = (tid < nN) ? ta.nxy[offN + tid] : V2<double>{0.0, 0.0};
    const int4 fr0 = (tid < nF) ? ta.face[offF + tid] : make_int4(0, 0, 0, 0);
    const int4 fr1 = (tid + kTileThreads < nF) ? ta.face[offF + tid + kTileThreads] : make_int4(0, 0, 0, 0);
    const S4 un = (!first && i < nT) ? Un[c0 + i] : S4{0.0, 0.0, 0.0, 0.0};
#if FVMR_TILE_GEOM_LOADED
    const V2<G> cc = (i < nL) ? ta.lcxy[offL + i] : V2<G>{G(0), G(0)};
    const V2<G> nr0 = (tid < nF) ? ta.fnrm[offF + tid] : V2<G>{G(0), G(0)};
    const V2<G> nr1 = (tid + kTileThreads < nF) ? ta.fnrm[offF + tid + kTileThreads] : V2<G>{G(0), G(0)};
#endif
    if (tid < nN) sN[tid] = nd;
    if (tid < nF) sFace[tid] = fr0;
    if (tid + kTileThreads < nF) sFace[tid + kTileThreads] = fr1;
    if (i < nL) lds_put(sU, kTileMaxL, i, SU{(C)uin.x, (C)uin.y, (C)uin.z, (C)uin.w});
#if FVMR_TILE_GEOM_LOADED
    if (i < nL) sC[i] = cc;
    if (tid < nF) sNrm[tid] = nr0;
    if (tid + kTileThreads < nF) sNrm[tid + kTileThreads] = nr1;
    __syncthreads();
#else
    __syncthreads();
    // 2. centroids and face normals
    if (i < nL) {
        const V2<double> ct = tile_centroid(sN, nw);
        sC[i] = V2<G>{(G)ct.x, (G)ct.y};
    }
    for (int k = tid; k < nF; k += kTileThreads) {
        V4<G> eg;
        V2<G> em;
        tile_face_geom<G>(sN, sFace[k], eg, em);
        sNrm[k] = V2<G>{eg.x, eg.y};
    }
    __syncthreads();
#endif
    // 3. limited gradients of the tile and the first ring
    if (i < nG) {
        const int sw[3] = {gs.x, gs.y, gs.z};
        const C vol = (C)gvol;
        C Uc[4];
        {
            const SU v = lds_get<SU>(sU, kTileMaxL, i);
            Uc[0] = (C)v.x; Uc[1] = (C)v.y; Uc[2] = (C)v.z; Uc[3] = (C)v.w;
        }
        const C xc = (C)sC[i].x, yc = (C)sC[i].y;
        struct {
            C a[3][3], u[3][4], n[3][2];
            __device__ __forceinline__ C A(int j, int q) const { return a[j][q]; }
            __device__ __forceinline__ C Uo(int j, int k) const { return u[j][k]; }
            __device__ __forceinline__ C N(int j, int q) const { return n[j][q]; }
        } sl;
        int fl[3];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int f = sw[j] & fvmr::kLocalMask;
            const int out = (sw[j] >> fvmr::kGsOutShift) & fvmr::kLocalMask;
            const bool has_out = out != fvmr::kLocalNone;
            const bool periodic = (sw[j] & fvmr::kGsPeriodic) != 0;
            const C sgn = (sw[j] & fvmr::kGsIsR) ? C(-1) : C(1);
            const int4 fr = sFace[f];
            V4<G> eg;
            V2<G> em;
            tile_face_geom_n<G>(sN, sNrm, f, fr, eg, em);
            {
                const SU v = lds_get<SU>(sU, kTileMaxL, has_out ? out : i);
                sl.u[j][0] = (C)v.x; sl.u[j][1] = (C)v.y; sl.u[j][2] = (C)v.z; sl.u[j][3] = (C)v.w;
            }
            const V2<double> nd = sN[lfield(nw, j)];
            sl.n[j][0] = (C)(G)nd.x - xc;
            sl.n[j][1] = (C)(G)nd.y - yc;
            if (ph.gradient == FVMR_GRA_GREENGAUSS) {
                // weight g_C of the slot precomputed by the host (the device formula in FP64)
                const C len = (C)eg.z;
                sl.a[j][0] = (C)(j == 0 ? gco.x : (j == 1 ? gco.y : gco.z));
                sl.a[j][1] = (C)eg.x * len * sgn;
                sl.a[j][2] = (C)eg.y * len * sgn;
            } else {
                C dxn = C(0), dyn = C(0);
                if (has_out && !periodic) {   // element neighbours exclude periodic faces
                    dxn = (C)sC[out].x - xc;
                    dyn = (C)sC[out].y - yc;
                }
                sl.a[j][0] = dxn;
                sl.a[j][1] = dyn;
                sl.a[j][2] = C(0);
            }
            fl[j] = (has_out ? 1 : 0) | ((has_out && !periodic) ? 2 : 0);
        }
        C gx[4], gy[4];
        grad_limit<C>(ph, Uc, sl, fl, vol, gx, gy);
        Grad<G> g;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            g.x[k] = (G)gx[k];
            g.y[k] = (G)gy[k];
        }
        lds_put(sGrFl, kTileMaxG, i, g);
    }
    __syncthreads();
    // 4. one canonical flux per face of the tile (nFF <= threads), written over the gradients
    const double t_rk = sc[cs].t_rk;
    const CaseBC &bc = ta.cbc[cs];
    const TileCells<G, SU> cells{sU, kTileMaxL, sC, sGrFl, kTileMaxG};
    C F[4] = {C(0), C(0), C(0), C(0)};
    const int f = tid;
    if (f < nFF) {
        const int4 fr = sFace[f];
        V4<G> eg;
        V2<G> em;
        tile_face_geom_n<G>(sN, sNrm, f, fr, eg, em);
        const int L = lfield(fr.y, 0), R = lfield(fr.y, 1), pl = lfield(fr.y, 2);
        const int bType = tile_btype(fr);
        const bool has_pair = pl != fvmr::kLocalNone;
        const V2<G> pm = (bType != -1 && has_pair) ? tile_face_mid<G>(sN, sFace[pl]) : em;
        face_flux<C, G, NS>(cells, ph, bc, t_rk, L, R == fvmr::kLocalNone ? -1 : R, bType, has_pair, eg, em, pm, F);
    }
    __syncthreads();
    if (f < nFF) lds_put(sGrFl, kTileMaxFF, f, V4<C>{F[0], F[1], F[2], F[3]});
    __syncthreads();
    // 5. cell-owned sum in slot order, /(-V), RK update (as k_stage); LDT: the cell's own dt
    double dt = sc[cs].dt;
    if (i < nT) {
        const int c = c0 + i;
        if (LDT) dt = ta.dtl[c];
        const int sw[3] = {gs.x, gs.y, gs.z};
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const V4<C> fv = lds_get<V4<C>>(sGrFl, kTileMaxFF, sw[j] & fvmr::kLocalMask);
            const double s = (sw[j] & fvmr::kGsIsR) ? -1.0 : 1.0;
            acc[0] += s * (double)fv.x;
            acc[1] += s * (double)fv.y;
            acc[2] += s * (double)fv.z;
            acc[3] += s * (double)fv.w;
        }
        const double w = -1.0 * (double)gvol;
        // the cell's own FP64 state: from LDS in the plain path, from this thread's staging
        // registers in mixed precision (its LDS copy is FP32)
        double Ui[4];
        if (std::is_same<SU, S4>::value) {
            const SU v = lds_get<SU>(sU, kTileMaxL, i);
            Ui[0] = v.x; Ui[1] = v.y; Ui[2] = v.z; Ui[3] = v.w;
        } else {
            Ui[0] = uin.x; Ui[1] = uin.y; Ui[2] = uin.z; Ui[3] = uin.w;
        }
        double Ub[4] = {0, 0, 0, 0}, r[4], Unew[4];
        if (!first) { Ub[0] = un.x; Ub[1] = un.y; Ub[2] = un.z; Ub[3] = un.w; }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            r[k] = fdiv(acc[k], w);
            const double s = Ui[k] + dt * r[k];
            Unew[k] = first ? s : a * Ub[k] + b * s;
        }
        if (rhs_out) st(rhs_out, c, r);
        if (std::is_same<C, double>::value) st_stream(Uout, c, Unew);
        else st(Uout, c, Unew);
        if (!last) return;
        double rv[4];
        prim(Unew, rv, ph.gamma);
        bool bad = out_of_range(rv[0], rv[3]);
#pragma unroll
        for (int k = 0; k < 4; k++) bad = bad || isnan(rv[k]);
        if (bad) {
            const int slot = atomicAdd(&sc->bad_count, 1);   // the handle's list: sc[0]
            if (slot < nc) {
                bad_list[slot] = c;
                bad_list[nc + slot] = cs;
            }
        }
    }
}

// Block minima of CFL V / alphaC over the tiles (the dt part of k_gradient<DT>): the
// spectral radius of every face of the tile once (inviscid part, and the viscous part up
// to the division by the cell volume), then per cell the sum in the reference edge order.
// LDT: the per-cell pseudo-time step of a steady solve instead, written to ta.dtl (the
// consistent spectral radius (|u.n| + a) L of the face-mean primitives summed over the cell's
// three faces, with k_local_dt's operation order).
template <class G, bool NS, bool LDT = false, int SHAPE = 0>
__global__ __launch_bounds__((TileCfg<G, SHAPE>::kThreads)) void k_tile_dt(TileArgs<G> ta, Phys ph,
                                                                          const S4 *__restrict__ U,
                                                                          double *__restrict__ partial) {
    if (!NS) ph.equation = FVMR_EQ_EULER;   // lets the compiler drop the viscous terms
    using Cfg = TileCfg<G, SHAPE>;
    constexpr int kTileThreads = Cfg::kThreads, kTileMaxG = Cfg::kMaxG, kTileMaxN = Cfg::kMaxN;
    constexpr int kTileMaxFF = Cfg::kMaxFF;
    __shared__ __align__(16) char sU[kTileMaxG * sizeof(S4)];   // states, 16-byte columns (lds_put)
    __shared__ V2<double> sN[kTileMaxN];
    __shared__ int4 sFace[kTileMaxFF];   // faces of the tile cells
    __shared__ double sAi[kTileMaxFF], sAv[kTileMaxFF];
    __shared__ double sMin[kTileThreads / 64];
    static_assert(kTileMaxG <= kTileThreads, "one cell per thread");
    const int tid = threadIdx.x;
    const int blk = xcd_block();
    const int *ti = ta.info + (size_t)blk * fvmr::kTileInfo;
    const int c0 = ti[fvmr::kTiC0], nT = ti[fvmr::kTiNT], nG = ti[fvmr::kTiNG];
    const int nN = ti[fvmr::kTiNN], nFF = ti[fvmr::kTiNFF];
    const int offH = ti[fvmr::kTiOffH], offN = ti[fvmr::kTiOffN], offF = ti[fvmr::kTiOffF], offG = ti[fvmr::kTiOffG];
    // every global load is issued before the first LDS store
    static_assert(kTileMaxN <= kTileThreads, "one node per thread");
    const DtRec dr = (tid < nT) ? ta.dtrec[c0 + tid] : DtRec{0, 0, 1.0};
    const V2<double> nd = (tid < nN) ? ta.nxy[offN + tid] : V2<double>{0.0, 0.0};
    const int4 fr = (tid < nFF) ? ta.face[offF + tid] : make_int4(0, 0, 0, 0);
    const int src = (tid < nG) ? (tid < nT ? c0 + tid : ta.halo[offH + tid - nT]) : -1;
    const S4 u = (src >= 0) ? U[src] : S4{0.0, 0.0, 0.0, 0.0};
    if (tid < nN) sN[tid] = nd;
    if (tid < nFF) sFace[tid] = fr;
    if (tid < nG) lds_put(sU, kTileMaxG, tid, u);
    __syncthreads();
    // faces of the tile cells (one per thread): L and R (when present) are tile or first-ring cells
    static_assert(kTileMaxFF <= kTileThreads, "one face per thread");
    if (tid < nFF) {
        const int f = tid;
        const int L = lfield(fr.y, 0), R = lfield(fr.y, 1);
        V4<G> eg;
        V2<G> em;
        tile_face_geom<G>(sN, fr, eg, em);
        double uL[4], rL[4], uR[4], rR[4];
        lds_state(sU, kTileMaxG, L, uL);
        const bool has_R = R != fvmr::kLocalNone;
        lds_state(sU, kTileMaxG, has_R ? R : L, uR);
        double av, vv;
        if (!LDT && !std::is_same<G, double>::value) {
            // mixed precision: the face's primitives and spectral radius in FP32 (the per-cell
            // sums and the minimum in FP64)
            G ua[4], ub[4], ra[4], rb[4], a32, v32;
#pragma unroll
            for (int k = 0; k < 4; k++) { ua[k] = (G)uL[k]; ub[k] = (G)uR[k]; }
            prim(ua, ra, (G)ph.gamma);
            prim(ub, rb, (G)ph.gamma);
            edge_alpha<G>(ph, ra, rb, has_R, eg.x, eg.y, eg.z, G(1), &a32, &v32);
            av = (double)a32;
            vv = (double)v32;
        } else if (LDT) {
            prim(uL, rL, ph.gamma);
            prim(uR, rR, ph.gamma);
            double r[4];
#pragma unroll
            for (int k = 0; k < 4; k++) r[k] = has_R ? 0.5 * (rL[k] + rR[k]) : rL[k];
            const double Lf = (double)eg.z;
            av = (fabs(r[1] * (double)eg.x + r[2] * (double)eg.y) + sqrt(ph.gamma * r[3] / r[0])) * Lf;
            vv = 0.0;
            if (ph.equation == FVMR_EQ_NS) {
                const double T = r[3] / ph.R / r[0];
                vv = fmax(4.0 / 3.0, ph.gamma / ph.Pr) * mu_of_T(ph, T) / r[0] * Lf * Lf;
            }
        } else {
            prim(uL, rL, ph.gamma);
            prim(uR, rR, ph.gamma);
            edge_alpha<double>(ph, rL, rR, has_R, (double)eg.x, (double)eg.y, (double)eg.z, 1.0, &av, &vv);
        }
        sAi[f] = av;
        sAv[f] = vv;
    }
    __syncthreads();
    if (LDT) {
        if (tid < nT) {
            const double vol_c = dr.vol;
            double alpha_c = 0.0;
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const int f = lfield(dr.faces, j);
                alpha_c += sAi[f];
                if (ph.equation == FVMR_EQ_NS) alpha_c += sAv[f] / vol_c;
            }
            ta.dtl[c0 + tid] = ta.cbc[ti[fvmr::kTiCase]].CFL * vol_c / alpha_c;
        }
        return;
    }
    double dtc = 1e10;
    if (tid < nT) {
        const int i = tid;
        const int code = dr.code;
        const double vol_c = dr.vol;
        // alphaC in the reference edge order: count (3 bits) + 6 x (slot 2 bits, pair 1 bit)
        const int cnt = code & 7;
        double alpha_c = 0.0;
        for (int q = 0; q < cnt; q++) {
            const int ent = (code >> (3 + 3 * q)) & 7;
            const int f = lfield(dr.faces, ent & 3);
            double av, vv;
            if (!(ent >> 2)) {
                av = sAi[f];
                vv = sAv[f] / vol_c;
            } else {
                // the periodic pair of face f, seen from this cell (the face's L) and its R
                const int4 ff = sFace[f];
                const int out = lfield(ff.y, 1);
                const int pf = lfield(ff.y, 2);
                const int4 fr = ta.face[offF + pf];
                V4<G> eg;
                V2<G> em;
                tile_face_geom<G>(sN, fr, eg, em);
                double Uc[4], rc[4], Uo[4], ro[4];
                lds_state(sU, kTileMaxG, i, Uc);
                prim(Uc, rc, ph.gamma);
                lds_state(sU, kTileMaxG, out, Uo);
                prim(Uo, ro, ph.gamma);
                edge_alpha(ph, rc, ro, true, (double)eg.x, (double)eg.y, (double)eg.z, vol_c, &av, &vv);
            }
            alpha_c += av;
            if (ph.equation == FVMR_EQ_NS) alpha_c += vv;
        }
        dtc = fmin(dtc, fdiv(ta.cbc[ti[fvmr::kTiCase]].CFL * vol_c, alpha_c));
    }
    const double wm = wave_min(dtc);
    if ((tid & 63) == 0) sMin[tid >> 6] = wm;
    __syncthreads();
    if (tid == 0) {
        double m = sMin[0];
        for (int w = 1; w < kTileThreads / 64; w++) m = fmin(m, sMin[w]);
        partial[blk] = m;
    }
}

// update_nodeField_ruvp + update_nodeField_other_variables (FieldWriter.cpp:371-452): one
// thread per node, the cells of the node in the caller's element order, U2ruvp as the
// reference writes it (PhysicalKernelGPU.h:11-16), no contraction.  out: ruvp (4 nv), Cp, Ma.
__device__ __forceinline__ void node_mean_ruvp(int v, const int *__restrict__ off, const int *__restrict__ cells,
                                               const S4 *__restrict__ U, double gamma, double sum[4]) {
#pragma clang fp contract(off)
    sum[0] = sum[1] = sum[2] = sum[3] = 0.0;
    const int k0 = off[v], k1 = off[v + 1];
    for (int k = k0; k < k1; k++) {
        const S4 u = U[cells[k]];
        double r[4];
        r[0] = u.x;
        r[1] = u.y / u.x;
        r[2] = u.z / u.x;
        r[3] = r[0] * (gamma - 1) * (u.w / u.x - (r[1] * r[1] + r[2] * r[2]) * 0.5);
        for (int q = 0; q < 4; q++) sum[q] += r[q];
    }
    const int n = k1 - k0;
    if (n > 0)
        for (int q = 0; q < 4; q++) sum[q] /= n;
}
// Pressure coefficient of a node mean pressure (update_nodeField_other_variables, FieldWriter.cpp:420-452).
__device__ __forceinline__ double node_cp(double p, double p_inf, double p_dyn) {
#pragma clang fp contract(off)
    return (p - p_inf) / p_dyn;
}

__global__ __launch_bounds__(kBlock) void k_node_field(int nv, const int *__restrict__ off,
                                                        const int *__restrict__ cells, const S4 *__restrict__ U,
                                                        double gamma, double p_inf, double p_dyn,
                                                        double *__restrict__ out) {
#pragma clang fp contract(off)
    const int v = blockIdx.x * kBlock + threadIdx.x;
    if (v >= nv) return;
    double sum[4];
    node_mean_ruvp(v, off, cells, U, gamma, sum);
    for (int q = 0; q < 4; q++) out[(size_t)q * nv + v] = sum[q];
    const double rho = sum[0], u = sum[1], w = sum[2], p = sum[3];
    out[(size_t)4 * nv + v] = node_cp(p, p_inf, p_dyn);
    const double V2 = u * u + w * w;
    const double a2 = gamma * p / rho;
    out[(size_t)5 * nv + v] = sqrt(V2 / a2);
}

// Colour-map control points of the GUI kernels: {r, g, b, position}, positions increasing.
constexpr int kMaxColorPoints = 64;
struct ColorMapArg {
    int n;
    float4 p[kMaxColorPoints];
};

// getColor_device (CudaGrid_kernels.cu:22-65): clamp to [0, 1], the interval of the control
// points that holds the input, the cubic blend of its end colours with their mean as both
// inner control points; purple (252, 3, 253) / 255 when no interval holds it.
__device__ __forceinline__ float3 colormap_color(const ColorMapArg &cm, float x) {
#pragma clang fp contract(off)
    if (x < 0.0f) x = 0.0f;
    if (x > 1.0f) x = 1.0f;
    for (int i = 0; i + 1 < cm.n; ++i) {
        const float4 a = cm.p[i], b = cm.p[i + 1];
        if (x >= a.w && x <= b.w) {
            const float mx = 0.5f * (a.x + b.x), my = 0.5f * (a.y + b.y), mz = 0.5f * (a.z + b.z);
            const float n = (x - a.w) / (b.w - a.w);
            const float omn = 1 - n, omn2 = omn * omn, omn3 = omn2 * omn, n2 = n * n, n3 = n2 * n;
            const float w1 = 3 * omn2 * n, w2 = 3 * omn * n2;
            return float3{omn3 * a.x + w1 * mx + w2 * mx + n3 * b.x, omn3 * a.y + w1 * my + w2 * my + n3 * b.y,
                          omn3 * a.z + w1 * mz + w2 * mz + n3 * b.z};
        }
    }
    return float3{252.0f / 255.0f, 3.0f / 255.0f, 253.0f / 255.0f};
}

// GPU::setVerticesColor (CudaGrid_kernels.cu:185-232): per node the largest of q_c * n over its
// n cells, from 0 (accumulateElementDataToNodeData_k, :77-95, without its lost updates), times
// (float)(1.0 / n) (divideNodeData_k, :97-104), normalised by [lo, hi] and coloured
// (nodeDataToVertexColor_k, :106-123) into vertices[6 v + 3 .. 5].  q = primitive var of U as
// float (the reference renders its float ruvp).
__global__ __launch_bounds__(kBlock) void k_vertex_colors(int nv, const int *__restrict__ off,
                                                           const int *__restrict__ cells, const S4 *__restrict__ U,
                                                           double gamma, int var, float lo, float hi, ColorMapArg cm,
                                                           float *__restrict__ vertices) {
#pragma clang fp contract(off)
    const int v = blockIdx.x * kBlock + threadIdx.x;
    if (v >= nv) return;
    const int k0 = off[v], k1 = off[v + 1], cnt = k1 - k0;
    float node = 0.0f;
    for (int k = k0; k < k1; k++) {
        const S4 u = U[cells[k]];
        double q;
        if (var == 0) q = u.x;
        else if (var == 1) q = u.y / u.x;
        else if (var == 2) q = u.z / u.x;
        else {
            const double uu = u.y / u.x, vv = u.z / u.x;
            q = u.x * (gamma - 1) * (u.w / u.x - (uu * uu + vv * vv) * 0.5);
        }
        float in = (float)q;
        in *= (float)cnt;
        node = (node > in) ? node : in;
    }
    float div = (float)(1.0 / cnt);
    if (div > 1e5f) div = 1.0f;
    node *= div;
    const float len = fabsf(hi - lo);
    const float x = (len == 0.0f) ? 0.0f : (node - lo) / len;
    const float3 c = colormap_color(cm, x);
    vertices[(size_t)6 * v + 3] = c.x;
    vertices[(size_t)6 * v + 4] = c.y;
    vertices[(size_t)6 * v + 5] = c.z;
}

// ReferenceData::calculate_edgeSet_force (FieldWriter.cpp:174-205): one thread per set, its
// edges in set order.  edge e: nodes en[2e], en[2e+1], geometry eg[3e..3e+2] = nx, ny, L.
__global__ void k_surface_force(int nsets, const int *__restrict__ off, const int *__restrict__ en,
                                const double *__restrict__ eg, const double *__restrict__ Cp, double S_ref,
                                double *__restrict__ out) {
#pragma clang fp contract(off)
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsets) return;
    const double one_on_Sref = 1 / S_ref;
    double cd = 0.0, cl = 0.0;
    for (int k = off[s]; k < off[s + 1]; k++) {
        const double cp = 0.5 * (Cp[en[2 * k]] + Cp[en[2 * k + 1]]);
        const double dC = cp * one_on_Sref * eg[3 * k + 2];
        cd += dC * eg[3 * k];
        cl += dC * eg[3 * k + 1];
    }
    out[2 * s] = cd;
    out[2 * s + 1] = cl;
}

// k_surface_force without the node-field buffer: the caller's edge ids of set s,
// edges[off[s] .. off[s+1]-1], are looked up in the handle's boundary-edge table (bref
// ascending, nb entries; every id was checked on the host to be there) and the Cp of each
// edge node is formed where it is used, with the arithmetic of k_node_field, so the sums
// equal k_surface_force's bit for bit.  Results: cx[s], cy[s].
__global__ void k_surface_force_direct(int nsets, const int *__restrict__ off, const int *__restrict__ edges, int nb,
                                       const int *__restrict__ bref, const int2 *__restrict__ bnodes,
                                       const V4<double> *__restrict__ bgeo, const int *__restrict__ noff,
                                       const int *__restrict__ ncells, const S4 *__restrict__ U, double gamma,
                                       double p_inf, double p_dyn, double S_ref, double *__restrict__ cx,
                                       double *__restrict__ cy) {
#pragma clang fp contract(off)
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsets) return;
    const double one_on_Sref = 1 / S_ref;
    double cd = 0.0, cl = 0.0;
    for (int k = off[s]; k < off[s + 1]; k++) {
        const int e = edges[k];
        int lo = 0, hi = nb - 1;   // bref[lo] <= e <= bref[hi] holds for every listed id
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (bref[mid] < e) lo = mid + 1;
            else hi = mid;
        }
        const int2 nd = bnodes[lo];
        const V4<double> g = bgeo[lo];
        double m0[4], m1[4];
        node_mean_ruvp(nd.x, noff, ncells, U, gamma, m0);
        node_mean_ruvp(nd.y, noff, ncells, U, gamma, m1);
        const double cp = 0.5 * (node_cp(m0[3], p_inf, p_dyn) + node_cp(m1[3], p_inf, p_dyn));
        const double dC = cp * one_on_Sref * g.z;
        cd += dC * g.x;
        cl += dC * g.y;
    }
    cx[s] = cd;
    cy[s] = cl;
}

// Device order -> the caller's order (out: 4 x nc, component-major) and back, for the nc cells
// of one case whose perm values are base .. base+nc-1 (the whole handle: base 0).
__global__ __launch_bounds__(kBlock) void k_to_caller(int nc, int base, const int *__restrict__ perm,
                                                      const S4 *__restrict__ U, double *__restrict__ out) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= nc) return;
    const size_t r = (size_t)(perm[i] - base), n = (size_t)nc;
    const S4 u = U[i];
    out[r] = u.x;
    out[n + r] = u.y;
    out[2 * n + r] = u.z;
    out[3 * n + r] = u.w;
}
__global__ __launch_bounds__(kBlock) void k_from_caller(int nc, int base, const int *__restrict__ perm,
                                                        const double *__restrict__ in, S4 *__restrict__ U) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= nc) return;
    const size_t r = (size_t)(perm[i] - base), n = (size_t)nc;
    U[i] = S4{in[r], in[n + r], in[2 * n + r], in[3 * n + r]};
}

// modifyElementFieldU2d_device_kernel (RestrictGPU.cu:6-84) with Jacobi reads: one work-group,
// phase 1 computes all repairs from the unrepaired field, phase 2 writes them.
// The body of k_restrict (every thread of the work-group calls it).
template <class G>
__device__ __forceinline__ void restrict_cells(const Mesh<G> &m, S4 *__restrict__ U, Scalars *__restrict__ sc,
                                               const int *__restrict__ bad_list, S4 *__restrict__ fix) {
    const int nc = m.nc;
    const int count = min(sc->bad_count, nc);
    if (count == 0) return;   // the usual case: nothing flagged, nothing to write
    for (int i = threadIdx.x; i < count; i += blockDim.x) {
        const int c = bad_list[i];
        int nvalid = 0;
        double Usum[4] = {0, 0, 0, 0}, vs = 0.0;
        for (int j = 0; j < 3; j++) {
            const int2 sl = m.cslot[j * nc + c];
            const int nb = (sl.x & kSlotPeriodic) ? -1 : sl.y;   // periodic neighbours are not element neighbours
            if (nb == -1) continue;
            nvalid++;
            const double v = (double)m.cgeo[nb].z;
            double Unb[4];
            ld(U, nb, Unb);
            for (int k = 0; k < 4; k++) {
                Usum[k] += Unb[k] * v;
                vs += v;
            }
        }
        double res[4];
        if (nvalid == 0) ld(U, c, res);
        else for (int k = 0; k < 4; k++) res[k] = Usum[k] / vs;
        st(fix, i, res);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < count; i += blockDim.x) {
        U[bad_list[i]] = fix[i];
        atomicAdd(&sc[bad_list[nc + i]].n_restricted, 1);   // counted per case of a batch
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        sc->bad_count = 0;
    }
}
template <class G>
__global__ __launch_bounds__(1024) void k_restrict(Mesh<G> m, S4 *__restrict__ U, Scalars *__restrict__ sc,
                                                   const int *__restrict__ bad_list, S4 *__restrict__ fix) {
    restrict_cells<G>(m, U, sc, bad_list, fix);
}

// Minimum of the block partials of k_gradient<DT>; then the clamp to T - t and the t
// update of CIteration::iteration_device (CIteration.cpp:134-139).
__global__ __launch_bounds__(1024) void k_dt_finish(int nparts, const double *__restrict__ partial,
                                                    Scalars *__restrict__ sc, int advance) {
    __shared__ double s[16];
    double b = 1e10;
    double v[8];   // the first 8 partials of each thread as independent loads
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int i = (int)threadIdx.x + j * (int)blockDim.x;
        v[j] = i < nparts ? partial[i] : 1e10;
    }
#pragma unroll
    for (int j = 0; j < 8; j++) b = fmin(b, v[j]);
    for (int i = threadIdx.x + 8 * blockDim.x; i < nparts; i += blockDim.x) b = fmin(b, partial[i]);
    b = wave_min(b);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = b;
    __syncthreads();
    if (threadIdx.x == 0) {
        double d = s[0];
        for (int i = 1; i < (int)(blockDim.x >> 6); i++) d = fmin(d, s[i]);
        if (advance) {
            const double t = sc->t;
            if (t + d > sc->T) d = sc->T - t;
            sc->t_rk = t;
            sc->t = t + d;
            sc->n_restricted = 0;
        }
        sc->dt = d;
    }
}

// k_dt_finish per case of a batch: block c takes the minimum over the partials of its tiles
// tile_begin[c] .. tile_begin[c+1]-1 and clamps / advances the times of sc[c].
__global__ __launch_bounds__(1024) void k_dt_finish_cases(const int *__restrict__ tile_begin,
                                                          const double *__restrict__ partial,
                                                          Scalars *__restrict__ sc, int advance) {
    __shared__ double s[16];
    const int c = blockIdx.x;
    const int beg = tile_begin[c], end = tile_begin[c + 1];
    double b = 1e10;
    // the first 8 partials of each thread as independent loads (one latency instead of eight)
    double v[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int i = beg + (int)threadIdx.x + j * 1024;
        v[j] = i < end ? partial[i] : 1e10;
    }
#pragma unroll
    for (int j = 0; j < 8; j++) b = fmin(b, v[j]);
    for (int i = beg + (int)threadIdx.x + 8 * 1024; i < end; i += 1024) b = fmin(b, partial[i]);
    b = wave_min(b);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = b;
    __syncthreads();
    if (threadIdx.x == 0) {
        double d = s[0];
        for (int i = 1; i < 16; i++) d = fmin(d, s[i]);
        if (advance) {
            const double t = sc[c].t;
            if (t + d > sc[c].T) d = sc[c].T - t;
            sc[c].t_rk = t;
            sc[c].t = t + d;
            sc[c].n_restricted = 0;
        }
        sc[c].dt = d;
    }
}

// get_residual_host (Residual.cpp:4-43) on the device: 1 = L1, 2 = L2, 3 = Linf, per
// variable; block partials, then one work-group finishes (k_norm_finish).
__global__ __launch_bounds__(kBlock) void k_norm(int nc, const S4 *__restrict__ rhs, int norm,
                                                 double *__restrict__ partial) {
    const int blk = blockIdx.x;
    const int c = blk * kBlock + threadIdx.x;
    __shared__ double s[4][kBlock / 64];
    double r[4] = {0, 0, 0, 0};
    if (c < nc) ld(rhs, c, r);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const double f = fabs(r[k]);
        const double v = (norm == FVMR_NORM_INF) ? wave_max(f) : wave_sum(norm == FVMR_NORM_2 ? f * f : f);
        if ((threadIdx.x & 63) == 0) s[k][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        const int k = threadIdx.x;
        double b = s[k][0];
        for (int i = 1; i < kBlock / 64; i++) b = (norm == FVMR_NORM_INF) ? fmax(b, s[k][i]) : b + s[k][i];
        partial[4 * blk + k] = b;
    }
}
__global__ __launch_bounds__(kBlock) void k_norm_finish(int nparts, int norm, const double *__restrict__ partial,
                                                        double *__restrict__ out) {
    // per variable: thread t combines partials t, t + kBlock, ... in order, then a fixed-shape
    // tree over the threads (the same order on every run)
    __shared__ double s[kBlock];
    for (int k = 0; k < 4; k++) {
        double b = 0.0;
        for (int i = threadIdx.x; i < nparts; i += kBlock) {
            const double v = partial[4 * i + k];
            b = (norm == FVMR_NORM_INF) ? fmax(b, v) : b + v;
        }
        s[threadIdx.x] = b;
        __syncthreads();
        for (int w = kBlock / 2; w > 0; w >>= 1) {
            if (threadIdx.x < w) s[threadIdx.x] = (norm == FVMR_NORM_INF) ? fmax(s[threadIdx.x], s[threadIdx.x + w])
                                                                          : s[threadIdx.x] + s[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[k] = (norm == FVMR_NORM_2) ? sqrt(s[0]) : s[0];
        __syncthreads();
    }
}

// The FP64 quotient and reciprocal of the face, gradient and update arithmetic, on n operand
// pairs (a parity hook for those forms).
__global__ __launch_bounds__(kBlock) void k_division(int n, const double *__restrict__ a, const double *__restrict__ b,
                                                     double *__restrict__ q, double *__restrict__ r) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    q[i] = fdiv(a[i], b[i]);
    r[i] = frcp(b[i]);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
template <class T>
int upload(T **dst, const std::vector<T> &src, hipStream_t s) {
    HIP_OK(hipMalloc((void **)dst, std::max<size_t>(1, src.size()) * sizeof(T)));
    if (!src.empty()) HIP_OK(hipMemcpyAsync(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, s));
    return FVMR_OK;
}

// Reverse Cuthill-McKee over the cell adjacency (faces, periodic faces included),
// starting each component from a pseudo-peripheral cell; ties by lower index.
std::vector<int> rcm_order(int nc, const std::vector<int> &adj_off, const std::vector<int> &adj) {
    std::vector<int> order;
    order.reserve(nc);
    std::vector<char> seen(nc, 0);
    auto deg = [&](int c) { return adj_off[c + 1] - adj_off[c]; };
    std::vector<int> level(nc, -1);
    auto bfs_last = [&](int s, int &ecc) {
        std::vector<int> q{s};
        level[s] = 0;
        size_t h = 0;
        int last = s;
        while (h < q.size()) {
            int c = q[h++];
            last = c;
            for (int k = adj_off[c]; k < adj_off[c + 1]; k++) {
                int d = adj[k];
                if (level[d] < 0) { level[d] = level[c] + 1; q.push_back(d); }
            }
        }
        // among the last level pick the minimum degree (then lowest index)
        ecc = level[last];
        int best = last;
        for (int c : q) if (level[c] == ecc && (deg(c) < deg(best) || (deg(c) == deg(best) && c < best))) best = c;
        for (int c : q) level[c] = -1;
        return best;
    };
    for (int s0 = 0; s0 < nc; s0++) {
        if (seen[s0]) continue;
        int start = s0, ecc = 0;
        for (int it = 0; it < 4; it++) {
            int e2;
            int far = bfs_last(start, e2);
            if (it > 0 && e2 <= ecc) break;
            ecc = e2;
            start = far;
        }
        size_t h = order.size();
        order.push_back(start);
        seen[start] = 1;
        std::vector<int> nb;
        while (h < order.size()) {
            int c = order[h++];
            nb.clear();
            for (int k = adj_off[c]; k < adj_off[c + 1]; k++) if (!seen[adj[k]]) nb.push_back(adj[k]);
            std::sort(nb.begin(), nb.end(), [&](int a, int b) { return deg(a) != deg(b) ? deg(a) < deg(b) : a < b; });
            for (int d : nb) { seen[d] = 1; order.push_back(d); }
        }
    }
    std::reverse(order.begin(), order.end());
    return order;
}

struct Prof {
    double total_ms = 0;
    int count = 0;
};

}  // namespace

// Staging of fvmr_surface_force_noalloc, which allocates nothing: per call the set offsets,
// the caller's edge ids and the per-set results.  One copy per device, held until the process
// ends; every call on a device shares it (calls that overlap on one device race).  Host code
// never uses this name as an address: it uses the device address taken once per device at
// fvmr_create (scratch_ptr).
constexpr int kNoAllocMaxSets = 256;
constexpr int kNoAllocMaxEdges = 131072;
constexpr size_t kScratchOffInts = 0;                                   // int[kNoAllocMaxSets + 1]
constexpr size_t kScratchEdgeInts = 512;                                // int[kNoAllocMaxEdges]
constexpr size_t kScratchOutDoubles = (kScratchEdgeInts + kNoAllocMaxEdges) / 2;   // double[2 kNoAllocMaxSets]
constexpr size_t kScratchDoubles = kScratchOutDoubles + 2 * kNoAllocMaxSets;
static_assert(kNoAllocMaxSets + 1 <= kScratchEdgeInts && (kScratchEdgeInts + kNoAllocMaxEdges) % 2 == 0,
              "scratch regions overlap");
static __device__ __align__(256) double scratch[kScratchDoubles];   // 1 MiB + 8 KiB

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, taken once per device.
int scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return set_error(FVMR_ERR_UNSUPPORTED, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return FVMR_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return set_error(FVMR_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = (double *)p;
    return FVMR_OK;
}
double *scratch_ptr(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[device];
}
}  // namespace

struct fvmr_solver {
    int nc = 0, nf = 0, nv = 0, nb = 0;
    fvmr_params p{};
    Phys ph{};
    hipStream_t stream = nullptr;
    int device = 0;
    bool mixed = false;
    std::vector<int> perm, iperm;   // perm[new] = reference cell; iperm[reference] = new
    // device storage
    S4 *Ubuf[3] = {nullptr, nullptr, nullptr};
    int cur = 0;                    // index of the buffer holding the current state
    void *grad = nullptr;           // Grad<G>[nc]
    S4 *rhs = nullptr, *fix = nullptr;
    void *cgeo = nullptr, *nxy = nullptr, *egeo = nullptr, *emid = nullptr;
    int *cnode = nullptr, *dtcode = nullptr;
    int2 *cslot = nullptr;
    int4 *eint = nullptr;
    double *partial = nullptr, *norm_out = nullptr;
    int *bad_list = nullptr;
    Scalars *sc = nullptr;
    // steps replayed from captured graphs: three steps bring the rotating state buffers back,
    // so one graph of three steps per starting buffer (captured on first use)
    bool use_graphs = false;
    hipGraph_t graph[3] = {nullptr, nullptr, nullptr};
    hipGraphExec_t gexec[3] = {nullptr, nullptr, nullptr};
    // cases of a batch (one for fvmr_create): cells case_cell_begin[c] .. [c+1]-1 in the caller's
    // numbering; per case a Scalars entry (sc[c]) and boundary states / CFL (cbc[c])
    int ncases = 1;
    std::vector<int> case_cell_begin;
    CaseBC *cbc = nullptr;
    int *tile_case_begin = nullptr;       // tiles of case c: [c] .. [c+1]-1
    int *perm_dev = nullptr;              // perm on the device (state transfers, first use)
    double *io_soa = nullptr;             // 4 x nc doubles in the caller's order (state transfers)
    bool io_ready = false;                // perm_dev holds perm (set by ensure_io)
    // tile kernels (fvmr_tiles.h)
    bool tiles = false;
    int ntiles = 0;
    int shape = 0;                        // work-group shape of the tile kernels (TileCfg<., shape>)
    int *t_info = nullptr, *t_halo = nullptr, *t_lnodes = nullptr;
    V2<double> *t_nxy = nullptr;
    int4 *t_gslot = nullptr, *t_face = nullptr;
    void *t_gcoef = nullptr;
    void *t_lcxy = nullptr, *t_fnrm = nullptr;   // V2<G> per local cell / per face (FVMR_TILE_GEOM_LOADED)
    DtRec *t_dtrec = nullptr;
    // output quantities: node -> cells (caller's element order), node buffers, boundary edges
    int *nc_off = nullptr, *nc_cells = nullptr;
    double *node_buf = nullptr;               // 6 x nv: ruvp, Cp, Ma
    std::vector<int> bedge_ref;               // caller ids of boundary-set edges, ascending
    std::vector<int> bedge_nodes;             // 2 per boundary edge
    std::vector<double> bedge_geo;            // nx, ny, L per boundary edge
    int *bedge_ref_dev = nullptr;             // the same three tables on the device
    int2 *bedge_nodes_dev = nullptr;
    V4<double> *bedge_geo_dev = nullptr;
    std::vector<int> enodes;                  // 2 per device edge: its nodes in the caller's order
    int bc_flags = 0;                         // boundary kinds present (periodic, DSR, isothermal)
    void *adj = nullptr;                      // steady-solve / adjoint workspace
    bool rhs_valid = false;
    bool profiling = false;
    Prof prof[kNumProf];
    std::vector<hipEvent_t> pending[kNumProf];
};

namespace {

template <class G>
Mesh<G> mesh_args(const fvmr_solver *h) {
    Mesh<G> m;
    m.nc = h->nc; m.nf = h->nf;
    m.cgeo = (const V4<G> *)h->cgeo;
    m.cslot = h->cslot; m.cnode = h->cnode;
    m.nxy = (const V2<G> *)h->nxy;
    m.egeo = (const V4<G> *)h->egeo;
    m.emid = (const V2<G> *)h->emid;
    m.eint = h->eint;
    m.dtcode = h->dtcode;
    return m;
}

template <class G>
TileArgs<G> tile_args(const fvmr_solver *h) {
    TileArgs<G> t;
    t.info = h->t_info; t.halo = h->t_halo; t.nxy = h->t_nxy; t.lnodes = h->t_lnodes;
    t.gslot = h->t_gslot; t.gcoef = (const V4<G> *)h->t_gcoef; t.face = h->t_face; t.dtrec = h->t_dtrec;
    t.lcxy = (const V2<G> *)h->t_lcxy; t.fnrm = (const V2<G> *)h->t_fnrm;
    t.cbc = h->cbc;
    t.dtl = nullptr;
    return t;
}

template <class G, class F>
int with_shape(const fvmr_solver *h, F &&f) {
    if (h->shape == 1) return f(std::integral_constant<int, 1>{});
    if constexpr (std::is_same<G, double>::value) {
        if (h->shape == 2) return f(std::integral_constant<int, 2>{});
    }
    return f(std::integral_constant<int, 0>{});
}

int prof_begin(fvmr_solver *h, int k, hipEvent_t *out) {
    if (!h->profiling) return FVMR_OK;
    HIP_OK(hipEventCreate(out));
    HIP_OK(hipEventRecord(*out, h->stream));
    return FVMR_OK;
}
int prof_end(fvmr_solver *h, int k, hipEvent_t start) {
    if (!h->profiling) return FVMR_OK;
    hipEvent_t e;
    HIP_OK(hipEventCreate(&e));
    HIP_OK(hipEventRecord(e, h->stream));
    h->pending[k].push_back(start);
    h->pending[k].push_back(e);
    return FVMR_OK;
}
int prof_collect(fvmr_solver *h) {
    for (int k = 0; k < kNumProf; k++) {
        auto &v = h->pending[k];
        for (size_t i = 0; i + 1 < v.size(); i += 2) {
            HIP_OK(hipEventSynchronize(v[i + 1]));
            float ms = 0;
            HIP_OK(hipEventElapsedTime(&ms, v[i], v[i + 1]));
            h->prof[k].total_ms += ms;
            h->prof[k].count++;
            HIP_OK(hipEventDestroy(v[i]));
            HIP_OK(hipEventDestroy(v[i + 1]));
        }
        v.clear();
    }
    return FVMR_OK;
}

int grid_of(int n) { return (n + kBlock - 1) / kBlock; }
int grid3_of(int n) { return (n + kCellsPerGroup - 1) / kCellsPerGroup; }

// Gradient + limiter; with dt_mode >= 0 also the global dt (dt_mode 1: clamp and advance t).
template <class C, class G>
int launch_gradient(fvmr_solver *h, const S4 *U, int dt_mode) {
    hipEvent_t e0{};
    int rc = prof_begin(h, 1, &e0);
    if (rc) return rc;
    if (dt_mode >= 0)
        k_gradient<C, G, true><<<grid_of(h->nc), kBlock, 0, h->stream>>>(mesh_args<G>(h), h->ph, U,
                                                                          (Grad<G> *)h->grad, h->partial);
    else
        k_gradient<C, G, false><<<grid_of(h->nc), kBlock, 0, h->stream>>>(mesh_args<G>(h), h->ph, U,
                                                                           (Grad<G> *)h->grad, h->partial);
    HIP_OK(hipGetLastError());
    if ((rc = prof_end(h, 1, e0))) return rc;
    if (dt_mode < 0) return FVMR_OK;
    if ((rc = prof_begin(h, 2, &e0))) return rc;
    k_dt_finish<<<1, 1024, 0, h->stream>>>(grid_of(h->nc), h->partial, h->sc, dt_mode);
    HIP_OK(hipGetLastError());
    return prof_end(h, 2, e0);
}

// The restriction after the step's last stage (CIteration.cpp:79 and :152 restrict once per
// iteration, after the evolve).
template <class G>
int launch_restrict(fvmr_solver *h, S4 *U) {
    hipEvent_t e0{};
    int rc = prof_begin(h, 3, &e0);
    if (rc) return rc;
    k_restrict<G><<<1, 1024, 0, h->stream>>>(mesh_args<G>(h), U, h->sc, h->bad_list, h->fix);
    HIP_OK(hipGetLastError());
    return prof_end(h, 3, e0);
}

template <class C, class G>
int launch_stage(fvmr_solver *h, const S4 *Uin, const S4 *Un, S4 *Uout, double a, double b, int first, int last,
                 S4 *rhs_out) {
    hipEvent_t e0{};
    int rc = prof_begin(h, 0, &e0);
    if (rc) return rc;
    if (h->ph.equation == FVMR_EQ_NS)
        k_stage<C, G, true><<<grid3_of(h->nc), kGroup3, 0, h->stream>>>(mesh_args<G>(h), h->ph, Uin, Un, Uout,
            (const Grad<G> *)h->grad, rhs_out, h->sc, h->bad_list, a, b, first, last);
    else
        k_stage<C, G, false><<<grid3_of(h->nc), kGroup3, 0, h->stream>>>(mesh_args<G>(h), h->ph, Uin, Un, Uout,
            (const Grad<G> *)h->grad, rhs_out, h->sc, h->bad_list, a, b, first, last);
    HIP_OK(hipGetLastError());
    if ((rc = prof_end(h, 0, e0))) return rc;
    return last ? launch_restrict<G>(h, Uout) : FVMR_OK;
}

// Global dt over the tiles (dt_mode as launch_gradient).
template <class C, class G>
int launch_tile_dt(fvmr_solver *h, const S4 *U, int dt_mode) {
    hipEvent_t e0{};
    int rc = prof_begin(h, 1, &e0);
    if (rc) return rc;
    with_shape<G>(h, [&](auto sh) {
        constexpr int S = decltype(sh)::value;
        if (h->ph.equation == FVMR_EQ_NS)
            k_tile_dt<G, true, false, S><<<h->ntiles, TileCfg<G, S>::kThreads, 0, h->stream>>>(tile_args<G>(h), h->ph, U,
                                                                                               h->partial);
        else
            k_tile_dt<G, false, false, S><<<h->ntiles, TileCfg<G, S>::kThreads, 0, h->stream>>>(tile_args<G>(h), h->ph, U,
                                                                                                h->partial);
        return FVMR_OK;
    });
    HIP_OK(hipGetLastError());
    if ((rc = prof_end(h, 1, e0))) return rc;
    if ((rc = prof_begin(h, 2, &e0))) return rc;
    k_dt_finish_cases<<<h->ncases, 1024, 0, h->stream>>>(h->tile_case_begin, h->partial, h->sc, dt_mode);
    HIP_OK(hipGetLastError());
    return prof_end(h, 2, e0);
}

template <class C, class G>
int launch_tile_stage(fvmr_solver *h, const S4 *Uin, const S4 *Un, S4 *Uout, double a, double b, int first, int last,
                      S4 *rhs_out) {
    hipEvent_t e0{};
    int rc = prof_begin(h, 0, &e0);
    if (rc) return rc;
    const bool ns = h->ph.equation == FVMR_EQ_NS, fixed = tile_scheme_fixed(h->ph);
    with_shape<C>(h, [&](auto sh) {
        constexpr int S = decltype(sh)::value;
        auto *kern = ns ? (fixed ? k_tile<C, G, true, false, kTileGgBarthRoe, S> : k_tile<C, G, true, false, kTileAny, S>)
                        : (fixed ? k_tile<C, G, false, false, kTileGgBarthRoe, S> : k_tile<C, G, false, false, kTileAny, S>);
        kern<<<h->ntiles, TileCfg<C, S>::kThreads, 0, h->stream>>>(tile_args<G>(h), h->ph, Uin, Un, Uout, rhs_out, h->sc,
                                                                   h->bad_list, h->nc, a, b, first, last);
        return FVMR_OK;
    });
    HIP_OK(hipGetLastError());
    if ((rc = prof_end(h, 0, e0))) return rc;
    return last ? launch_restrict<G>(h, Uout) : FVMR_OK;
}

// One step on the device with the tile kernels.
template <class C, class G>
int device_step_tiles(fvmr_solver *h, bool store_rhs) {
    int rc;
    const int a = h->cur, b = (a + 1) % 3, c = (a + 2) % 3;
    S4 *Ua = h->Ubuf[a], *Ub = h->Ubuf[b], *Uc = h->Ubuf[c];
    if ((rc = launch_tile_dt<C, G>(h, Ua, 1))) return rc;
    const int rk3 = h->p.time_advance == FVMR_EVO_RK3;
    if ((rc = launch_tile_stage<C, G>(h, Ua, Ua, Ub, 0.0, 1.0, 1, !rk3, store_rhs ? h->rhs : nullptr))) return rc;
    if (rk3) {
        if ((rc = launch_tile_stage<C, G>(h, Ub, Ua, Uc, 0.75, 0.25, 0, 0, nullptr))) return rc;
        if ((rc = launch_tile_stage<C, G>(h, Uc, Ua, Ub, 1.0 / 3.0, 2.0 / 3.0, 0, 1, nullptr))) return rc;
    }
    h->cur = b;
    if (store_rhs) h->rhs_valid = true;
    return FVMR_OK;
}

// One step on the device (no host synchronisation).  store_rhs: keep stage-1 rhs.
template <class C, class G>
int device_step(fvmr_solver *h, bool store_rhs) {
    int rc;
    const int a = h->cur, b = (a + 1) % 3, c = (a + 2) % 3;
    S4 *Ua = h->Ubuf[a], *Ub = h->Ubuf[b], *Uc = h->Ubuf[c];
    if ((rc = launch_gradient<C, G>(h, Ua, 1))) return rc;
    const int rk3 = h->p.time_advance == FVMR_EVO_RK3;
    if ((rc = launch_stage<C, G>(h, Ua, Ua, Ub, 0.0, 1.0, 1, !rk3, store_rhs ? h->rhs : nullptr))) return rc;
    if (rk3) {
        // SSP-RK3 (Shu-Osher): U1 = U + dt R(U); U2 = 3/4 U + 1/4 (U1 + dt R(U1));
        // U^{n+1} = 1/3 U + 2/3 (U2 + dt R(U2)); the restriction follows the last stage
        if ((rc = launch_gradient<C, G>(h, Ub, -1))) return rc;
        if ((rc = launch_stage<C, G>(h, Ub, Ua, Uc, 0.75, 0.25, 0, 0, nullptr))) return rc;
        if ((rc = launch_gradient<C, G>(h, Uc, -1))) return rc;
        if ((rc = launch_stage<C, G>(h, Uc, Ua, Ub, 1.0 / 3.0, 2.0 / 3.0, 0, 1, nullptr))) return rc;
    }
    h->cur = b;
    if (store_rhs) h->rhs_valid = true;
    return FVMR_OK;
}

int run_step(fvmr_solver *h, bool store_rhs) {
    if (h->tiles) return h->mixed ? device_step_tiles<float, float>(h, store_rhs) : device_step_tiles<double, double>(h, store_rhs);
    if (h->mixed) return device_step<float, float>(h, store_rhs);
    return device_step<double, double>(h, store_rhs);
}

// n steps; the last one keeps its stage-1 rhs when store_rhs_last.  Without profiling, runs of
// three steps replay a captured graph (the same launches with the same arguments).
int run_steps(fvmr_solver *h, int n, bool store_rhs_last) {
    int rc, done = 0;
    if (h->use_graphs && !h->profiling) {
        const int eager_tail = store_rhs_last ? 1 : 0;
        while (n - done - eager_tail >= 3) {
            const int a = h->cur;
            if (!h->gexec[a]) {
                HIP_OK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
                rc = FVMR_OK;
                for (int k = 0; k < 3 && !rc; k++) rc = run_step(h, false);
                hipGraph_t g = nullptr;
                const hipError_t e = hipStreamEndCapture(h->stream, &g);
                h->cur = a;   // capture launches nothing; the buffers rotate when the graph runs
                if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
                if (e != hipSuccess) return set_error(FVMR_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
                hipGraphExec_t x = nullptr;
                const hipError_t e2 = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
                if (e2 != hipSuccess) {
                    (void)hipGraphDestroy(g);
                    return set_error(FVMR_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e2));
                }
                h->graph[a] = g;
                h->gexec[a] = x;
            }
            HIP_OK(hipGraphLaunch(h->gexec[a], h->stream));
            done += 3;
        }
    }
    for (; done < n; done++)
        if ((rc = run_step(h, store_rhs_last && done == n - 1))) return rc;
    return FVMR_OK;
}

template <class G>
int build_geometry(fvmr_solver *h, const fvmr_mesh_soa *m, const std::vector<int> &eperm) {
    // eperm[new edge] = reference edge
    const int nc = h->nc, nf = h->nf, nv = h->nv;
    std::vector<V4<G>> cg(nc), eg(nf);
    std::vector<V2<G>> nx(nv), em(nf);
    for (int i = 0; i < nc; i++) {
        const int r = h->perm[i];
        cg[i] = V4<G>{(G)m->element_xy[0][r], (G)m->element_xy[1][r], (G)m->element_volume[r], G(0)};
    }
    for (int v = 0; v < nv; v++) nx[v] = V2<G>{(G)m->node_xy[0][v], (G)m->node_xy[1][v]};
    for (int e = 0; e < nf; e++) {
        const int r = eperm[e];
        eg[e] = V4<G>{(G)m->edge_normal[0][r], (G)m->edge_normal[1][r], (G)m->edge_length[r], G(0)};
        em[e] = V2<G>{(G)m->edge_xy[0][r], (G)m->edge_xy[1][r]};
    }
    int rc;
    if ((rc = upload((V4<G> **)&h->cgeo, cg, h->stream)) || (rc = upload((V2<G> **)&h->nxy, nx, h->stream)) ||
        (rc = upload((V4<G> **)&h->egeo, eg, h->stream)) || (rc = upload((V2<G> **)&h->emid, em, h->stream)))
        return rc;
    HIP_OK(hipStreamSynchronize(h->stream));  // the host vectors go out of scope
    return FVMR_OK;
}

void free_graphs(fvmr_solver *h) {
    for (int k = 0; k < 3; k++) {
        if (h->gexec[k]) (void)hipGraphExecDestroy(h->gexec[k]);
        if (h->graph[k]) (void)hipGraphDestroy(h->graph[k]);
        h->gexec[k] = nullptr;
        h->graph[k] = nullptr;
    }
}

void adj_free(fvmr_solver *h);

void free_all(fvmr_solver *h) {
    free_graphs(h);
    adj_free(h);
    for (auto &b : h->Ubuf) if (b) (void)hipFree(b);
    void *ptrs[] = {h->grad, h->rhs, h->fix, h->cgeo, h->nxy, h->egeo, h->emid, h->cslot, h->cnode, h->dtcode,
                    h->eint, h->partial, h->norm_out, h->bad_list, h->sc, h->t_info, h->t_halo, h->t_lnodes,
                    h->t_nxy, h->t_gslot, h->t_face, h->t_gcoef, h->t_dtrec, h->t_lcxy, h->t_fnrm,
                    h->nc_off, h->nc_cells, h->node_buf, h->cbc, h->tile_case_begin, h->perm_dev, h->io_soa,
                    h->bedge_ref_dev, h->bedge_nodes_dev, h->bedge_geo_dev};
    for (void *q : ptrs) if (q) (void)hipFree(q);
    for (int k = 0; k < kNumProf; k++)
        for (auto e : h->pending[k]) (void)hipEventDestroy(e);
    if (h->stream) (void)hipStreamDestroy(h->stream);
}

int validate_params(const fvmr_params *p) {
    if (p->equation != FVMR_EQ_EULER && p->equation != FVMR_EQ_NS) return set_error(FVMR_ERR_INVALID_ARG, "invalid equation");
    if (p->flux_scheme != FVMR_SOL_LLF && p->flux_scheme != FVMR_SOL_ROE) return set_error(FVMR_ERR_INVALID_ARG, "invalid scheme @get_flux_RiemannSolve");
    if (p->reconstruct != FVMR_REC_CONSTANT && p->reconstruct != FVMR_REC_LINEAR && p->reconstruct != FVMR_REC_MUSCL)
        return set_error(FVMR_ERR_INVALID_ARG, "invalid flag_reconstruct");
    if (p->gradient != FVMR_GRA_GREENGAUSS && p->gradient != FVMR_GRA_LEASTSQUARE) return set_error(FVMR_ERR_INVALID_ARG, "invalid graident type");
    if (p->limiter != FVMR_LIM_NONE && p->limiter != FVMR_LIM_BARTH && p->limiter != FVMR_LIM_VENKATAKRISHNAN)
        return set_error(FVMR_ERR_INVALID_ARG, "invalid limiter type");
    if (p->time_advance != FVMR_EVO_EXPLICIT && p->time_advance != FVMR_EVO_RK3)
        return set_error(FVMR_ERR_UNSUPPORTED, "invalid temporal_choice");
    if (p->precision != FVMR_PREC_FP64 && p->precision != FVMR_PREC_MIXED) return set_error(FVMR_ERR_INVALID_ARG, "invalid precision");
    if (!(p->gamma > 1.0) || !(p->CFL > 0.0)) return set_error(FVMR_ERR_INVALID_ARG, "invalid gamma or CFL");
    if (p->renumber < FVMR_RENUMBER_NONE || p->renumber > FVMR_RENUMBER_HILBERT)
        return set_error(FVMR_ERR_INVALID_ARG, "invalid renumber");
    if (p->kernels < FVMR_KERNELS_AUTO || p->kernels > FVMR_KERNELS_TILE) return set_error(FVMR_ERR_INVALID_ARG, "invalid kernels");
    if (p->tile_cells < 0 || p->tile_cells > kTileMaxCells) return set_error(FVMR_ERR_INVALID_ARG, "tile_cells out of [0, 400]");
    if (shape_of_threads(p->precision == FVMR_PREC_MIXED, p->tile_threads) < 0)
        return set_error(FVMR_ERR_INVALID_ARG, p->precision == FVMR_PREC_MIXED
                                                   ? "tile_threads must be 0, 512 or 256 in mixed precision"
                                                   : "tile_threads must be 0, 256, 128 or 512 in FP64");
    return FVMR_OK;
}

int validate_mesh(const fvmr_mesh_soa *m) {
    if (!m || m->num_element <= 0 || m->num_edge <= 0 || m->num_node <= 0) return set_error(FVMR_ERR_INVALID_ARG, "empty mesh");
    const int nc = m->num_element, nf = m->num_edge, nv = m->num_node;
    for (int e = 0; e < nf; e++) {
        const int L = m->edge_elementL[e], R = m->edge_elementR[e], s = m->edge_setID[e], pe = m->edge_periodicPair[e];
        if (L < 0 || L >= nc || R < -1 || R >= nc) return set_error(FVMR_ERR_MESH, "edge owner out of range");
        if (s != -1 && (s < 1 || s > m->num_boundary_set)) return set_error(FVMR_ERR_MESH, "edge setID out of range");
        if (pe < -1 || pe >= nf) return set_error(FVMR_ERR_MESH, "periodic pair out of range");
        if (R == -1 && s == -1) return set_error(FVMR_ERR_MESH, "boundary edge without a boundary set");
    }
    for (int c = 0; c < nc; c++) {
        for (int j = 0; j < 3; j++) {
            const int e = m->element_edges[j][c], v = m->element_nodes[j][c];
            if (e < 0 || e >= nf) return set_error(FVMR_ERR_MESH, "element edge index out of range (triangles only)");
            if (v < 0 || v >= nv) return set_error(FVMR_ERR_MESH, "element node index out of range");
            const int L = m->edge_elementL[e], R = m->edge_elementR[e];
            if (L != c && R != c) return set_error(FVMR_ERR_MESH, "element edge does not own the element");
            const int out = (c != R) ? R : L;
            const int expect = (m->edge_periodicPair[e] >= 0) ? -1 : out;
            if (m->element_neighbors[j][c] != expect)
                return set_error(FVMR_ERR_MESH, "element neighbours disagree with the edge owners");
        }
    }
    return FVMR_OK;
}

}  // namespace

extern "C" {

}  // extern "C"

// fvmr_create for one case or a batch: m is the caller's SoA (for a batch, the cases' SoAs
// one after the other), case_begin[c] .. case_begin[c+1]-1 the cells of case c, case_params[c]
// its parameters (NULL: p for every case).
static int create_impl(const fvmr_mesh_soa *m, const fvmr_params *p, int ncases, const int *case_begin,
                       const fvmr_params *case_params, fvmr_handle **out) {
    if (!m || !p || !out) return set_error(FVMR_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    int rc = validate_params(p);
    if (rc) return rc;
    if ((rc = validate_mesh(m))) return rc;
    int ndev = 0;
    HIP_OK(hipGetDeviceCount(&ndev));
    if (ndev <= 0) return set_error(FVMR_ERR_HIP, "no HIP device");
    fvmr_solver *h = new fvmr_solver();
    h->nc = m->num_element; h->nf = m->num_edge; h->nv = m->num_node; h->nb = m->num_boundary_set;
    h->p = *p;
    h->mixed = (p->precision == FVMR_PREC_MIXED);
    (void)hipGetDevice(&h->device);
    auto bail = [&](int code) { free_all(h); delete h; return code; };
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) return bail(set_error(FVMR_ERR_HIP, "hipStreamCreate failed"));
    const int nc = h->nc, nf = h->nf;
    // physics constants
    Phys &ph = h->ph;
    ph.gamma = p->gamma; ph.R = p->R; ph.Pr = p->Pr; ph.Pr_visc = p->Pr_viscous; ph.mu0 = p->mu0;
    ph.sutherland_C1 = p->mu0 / std::pow(p->T0, 1.5) * (p->T0 + 110.4);  // ViscousFluxGPU.h:26-28
    ph.bc.CFL = p->CFL; ph.wall_T = p->wall_temperature; ph.venkat_K = p->venkat_K;
    ph.g_ga1 = p->gamma / (p->gamma - 1.0); ph.inv_g = 1.0 / p->gamma;
    ph.dsr_x = p->dsr_shock_x; ph.dsr_y = p->dsr_shock_y; ph.dsr_speed = p->dsr_shock_speed;
    for (int k = 0; k < 4; k++) { ph.bc.inf[k] = p->inf_ruvp[k]; ph.bc.inlet[k] = p->inlet_ruvp[k]; ph.bc.outlet[k] = p->outlet_ruvp[k]; }
    h->ncases = ncases;
    h->case_cell_begin.assign(case_begin, case_begin + ncases + 1);
    ph.equation = p->equation; ph.scheme = p->flux_scheme;
    ph.reconstruct = (p->reconstruct == FVMR_REC_LINEAR) ? FVMR_REC_LINEAR : FVMR_REC_CONSTANT;
    ph.gradient = p->gradient; ph.limiter = p->limiter; ph.mu_model = p->viscosity_model; ph.compat_inv = p->compat_inv2x2;
    // cell permutation, within each case (case c keeps cells case_begin[c] .. [c+1]-1)
    h->perm.resize(nc);
    for (int cs = 0; cs < ncases; cs++) {
        const int b = case_begin[cs], n = case_begin[cs + 1] - b;
        std::vector<int> local;
        if (p->renumber == FVMR_RENUMBER_HILBERT) {
            local = fvmr::hilbert_order(n, m->element_xy[0] + b, m->element_xy[1] + b);
        } else if (p->renumber == FVMR_RENUMBER_RCM) {
            std::vector<int> off(n + 1, 0), adj;
            adj.reserve(3 * (size_t)n);
            for (int c = 0; c < n; c++) {
                for (int j = 0; j < 3; j++) {
                    const int e = m->element_edges[j][b + c];
                    const int L = m->edge_elementL[e], R = m->edge_elementR[e];
                    const int o = (b + c != R) ? R : L;
                    if (o >= 0 && o != b + c) adj.push_back(o - b);
                }
                off[c + 1] = (int)adj.size();
            }
            local = rcm_order(n, off, adj);
        } else {
            local.resize(n);
            std::iota(local.begin(), local.end(), 0);
        }
        for (int i = 0; i < n; i++) h->perm[b + i] = b + local[i];
    }
    h->iperm.assign(nc, -1);
    for (int i = 0; i < nc; i++) h->iperm[h->perm[i]] = i;
    // edge order: first appearance over the new cell order
    std::vector<int> enew(nf, -1), eperm;
    eperm.reserve(nf);
    for (int i = 0; i < nc; i++)
        for (int j = 0; j < 3; j++) {
            const int e = m->element_edges[j][h->perm[i]];
            if (enew[e] < 0) { enew[e] = (int)eperm.size(); eperm.push_back(e); }
        }
    for (int e = 0; e < nf; e++) if (enew[e] < 0) { enew[e] = (int)eperm.size(); eperm.push_back(e); }
    h->enodes.resize(2 * (size_t)nf);
    for (int e = 0; e < nf; e++) {
        const int r = eperm[e];
        h->enodes[2 * (size_t)e] = m->edge_nodes[0][r];
        h->enodes[2 * (size_t)e + 1] = m->edge_nodes[1][r];
        const int s = m->edge_setID[r];
        const int t = (s == -1) ? -1 : m->boundary_type[s - 1];
        if (m->edge_periodicPair[r] >= 0 || (t >= FVMR_BC_PERIODIC_0 && t <= FVMR_BC_PERIODIC_0 + 9)) h->bc_flags |= 1;
        if (t == FVMR_BC_DSR) h->bc_flags |= 2;
        if (t == FVMR_BC_WALL_ISOTHERMAL) h->bc_flags |= 4;
    }
    ph.has_periodic = (h->bc_flags & 1) ? 1 : 0;
    // connectivity in the new numbering
    if (nf > kSlotEdgeMask) return bail(set_error(FVMR_ERR_UNSUPPORTED, "too many edges for the face-slot encoding"));
    std::vector<int2> cslot(3 * (size_t)nc);
    std::vector<int> cnode(3 * (size_t)nc);
    std::vector<int4> eint(nf);
    for (int e = 0; e < nf; e++) {
        const int r = eperm[e];
        const int L = m->edge_elementL[r], R = m->edge_elementR[r], s = m->edge_setID[r];
        eint[e] = make_int4(h->iperm[L], R >= 0 ? h->iperm[R] : -1, (s == -1) ? -1 : m->boundary_type[s - 1],
                            m->edge_periodicPair[r] >= 0 ? enew[m->edge_periodicPair[r]] : -1);
    }
    for (int i = 0; i < nc; i++) {
        const int r = h->perm[i];
        for (int j = 0; j < 3; j++) {
            const int er = m->element_edges[j][r];
            const int e = enew[er];
            const int4 ei = eint[e];
            // reference: the cell across is R unless the cell is R itself (GradientGPU.cu:264-271)
            const bool cR = (i == ei.y);
            int flags = cR ? kSlotCellIsR : 0;
            if (m->edge_setID[er] != -1) flags |= kSlotBoundary;
            if (m->edge_periodicPair[er] >= 0) flags |= kSlotPeriodic;
            cslot[(size_t)j * nc + i] = make_int2(e | flags, cR ? ei.x : ei.y);
            cnode[(size_t)j * nc + i] = m->element_nodes[j][r];
        }
    }
    // alphaC contributions in the reference edge order (CalculateDt.cpp:22-63), encoded per
    // cell as count (3 bits) + 6 x (face slot 2 bits, pair-edge flag 1 bit)
    std::vector<int> dtcode(nc, 0);
    {
        std::vector<int> cnt(nc, 0);
        for (int r = 0; r < nf; r++) {
            const int cells[2] = {m->edge_elementL[r], m->edge_elementR[r]};
            for (int q = 0; q < 2; q++) {
                if (cells[q] < 0) continue;
                const int ref = cells[q], ci = h->iperm[ref];
                int ent = -1;
                for (int j = 0; j < 3 && ent < 0; j++)
                    if (m->element_edges[j][ref] == r) ent = j;
                for (int j = 0; j < 3 && ent < 0; j++)
                    if (m->edge_periodicPair[m->element_edges[j][ref]] == r) ent = 4 | j;
                if (ent < 0) return bail(set_error(FVMR_ERR_MESH, "edge owner does not list the edge"));
                if (cnt[ci] >= kMaxDtSlots) return bail(set_error(FVMR_ERR_MESH, "too many edges touch one cell"));
                dtcode[ci] |= ent << (3 + 3 * cnt[ci]);
                cnt[ci]++;
            }
        }
        for (int i = 0; i < nc; i++) dtcode[i] |= cnt[i];
    }
    // node -> cells in the caller's element order; boundary-set edges for the surface force
    {
        const int nv = h->nv;
        std::vector<int> off(nv + 1, 0), cells(3 * (size_t)nc);
        for (int r = 0; r < nc; r++)
            for (int j = 0; j < 3; j++) off[m->element_nodes[j][r] + 1]++;
        for (int v = 0; v < nv; v++) off[v + 1] += off[v];
        std::vector<int> fill(off.begin(), off.end() - 1);
        for (int r = 0; r < nc; r++)
            for (int j = 0; j < 3; j++) cells[fill[m->element_nodes[j][r]]++] = h->iperm[r];
        if ((rc = upload(&h->nc_off, off, h->stream)) || (rc = upload(&h->nc_cells, cells, h->stream))) return bail(rc);
        if (hipStreamSynchronize(h->stream) != hipSuccess) return bail(set_error(FVMR_ERR_HIP, "node table upload failed"));
        for (int e = 0; e < nf; e++) {
            if (m->edge_setID[e] == -1) continue;
            h->bedge_ref.push_back(e);
            h->bedge_nodes.push_back(m->edge_nodes[0][e]);
            h->bedge_nodes.push_back(m->edge_nodes[1][e]);
            h->bedge_geo.push_back(m->edge_normal[0][e]);
            h->bedge_geo.push_back(m->edge_normal[1][e]);
            h->bedge_geo.push_back(m->edge_length[e]);
        }
        const size_t nbe = h->bedge_ref.size();
        std::vector<int2> bn(nbe);
        std::vector<V4<double>> bg(nbe);
        for (size_t b = 0; b < nbe; b++) {
            bn[b] = make_int2(h->bedge_nodes[2 * b], h->bedge_nodes[2 * b + 1]);
            bg[b] = V4<double>{h->bedge_geo[3 * b], h->bedge_geo[3 * b + 1], h->bedge_geo[3 * b + 2], 0.0};
        }
        if ((rc = upload(&h->bedge_ref_dev, h->bedge_ref, h->stream)) || (rc = upload(&h->bedge_nodes_dev, bn, h->stream)) ||
            (rc = upload(&h->bedge_geo_dev, bg, h->stream)))
            return bail(rc);
        if (hipStreamSynchronize(h->stream) != hipSuccess) return bail(set_error(FVMR_ERR_HIP, "boundary table upload failed"));
        // the allocation-free entry point's staging array on this device
        if ((rc = scratch_init(h->device))) return bail(rc);
    }
    // tile records
    fvmr::TileMesh tm;
    if (p->kernels != FVMR_KERNELS_CELL) {
        const char *why = nullptr;
        if (!fvmr::standard_geometry(m, &why)) {
            if (p->kernels == FVMR_KERNELS_TILE) return bail(set_error(FVMR_ERR_UNSUPPORTED, std::string("tile kernels: ") + why));
        } else {
            std::vector<int> sx(3 * (size_t)nc), sy(3 * (size_t)nc), eL(nf), eR(nf), et(nf), ep(nf), en0(nf), en1(nf);
            std::vector<double> elen(nf), vol(nc);
            for (size_t k = 0; k < sx.size(); k++) { sx[k] = cslot[k].x; sy[k] = cslot[k].y; }
            for (int e = 0; e < nf; e++) {
                eL[e] = eint[e].x; eR[e] = eint[e].y; et[e] = eint[e].z; ep[e] = eint[e].w;
                en0[e] = m->edge_nodes[0][eperm[e]]; en1[e] = m->edge_nodes[1][eperm[e]];
                elen[e] = m->edge_length[eperm[e]];
            }
            for (int i = 0; i < nc; i++) vol[i] = m->element_volume[h->perm[i]];
            std::vector<double> ccx(nc), ccy(nc), emx(nf), emy(nf), enx(nf), eny(nf);
            for (int i = 0; i < nc; i++) { ccx[i] = m->element_xy[0][h->perm[i]]; ccy[i] = m->element_xy[1][h->perm[i]]; }
            for (int e = 0; e < nf; e++) {
                emx[e] = m->edge_xy[0][eperm[e]]; emy[e] = m->edge_xy[1][eperm[e]];
                enx[e] = m->edge_normal[0][eperm[e]]; eny[e] = m->edge_normal[1][eperm[e]];
            }
            fvmr::TileInput ti{nc, nf, sx.data(), sy.data(), cnode.data(), eL.data(), eR.data(), et.data(), ep.data(),
                               en0.data(), en1.data(), elen.data(), m->node_xy[0], m->node_xy[1], vol.data(), dtcode.data(),
                               ccx.data(), ccy.data(), emx.data(), emy.data(), enx.data(), eny.data(), ncases, case_begin};
            auto caps_of = [&](auto cfg) {
                using Cfg = decltype(cfg);
                return fvmr::TileCaps{std::min(p->tile_cells ? p->tile_cells : Cfg::kCells, Cfg::kMaxG), Cfg::kMaxG,
                                      Cfg::kMaxL, Cfg::kMaxN, Cfg::kMaxF, Cfg::kMaxFF};
            };
            h->shape = shape_of_threads(h->mixed, p->tile_threads);
            fvmr::TileCaps caps;
            if (h->mixed) caps = (h->shape == 1) ? caps_of(TileCfg<float, 1>{}) : caps_of(TileCfg<float, 0>{});
            else caps = (h->shape == 1) ? caps_of(TileCfg<double, 1>{})
                                        : ((h->shape == 2) ? caps_of(TileCfg<double, 2>{}) : caps_of(TileCfg<double, 0>{}));
            if ((rc = fvmr::build_tiles(ti, caps, tm))) return bail(rc);
            h->tiles = true;
            h->ntiles = tm.ntiles;
            std::vector<V2<double>> txy(tm.nxy.size() / 2);
            for (size_t k = 0; k < txy.size(); k++) txy[k] = V2<double>{tm.nxy[2 * k], tm.nxy[2 * k + 1]};
            std::vector<int4> tgs(tm.gslot.size() / 4), tfc(tm.face.size() / 4);
            for (size_t k = 0; k < tgs.size(); k++)
                tgs[k] = make_int4(tm.gslot[4 * k], tm.gslot[4 * k + 1], tm.gslot[4 * k + 2], tm.gslot[4 * k + 3]);
            for (size_t k = 0; k < tfc.size(); k++)
                tfc[k] = make_int4(tm.face[4 * k], tm.face[4 * k + 1], tm.face[4 * k + 2], tm.face[4 * k + 3]);
            if ((rc = upload(&h->t_info, tm.info, h->stream)) || (rc = upload(&h->t_halo, tm.halo, h->stream)) ||
                (rc = upload(&h->t_lnodes, tm.lnodes, h->stream)) || (rc = upload(&h->t_nxy, txy, h->stream)) ||
                (rc = upload(&h->t_gslot, tgs, h->stream)) || (rc = upload(&h->t_face, tfc, h->stream)))
                return bail(rc);
            std::vector<V4<float>> gcf;
            std::vector<V4<double>> gcd(tm.gcoef.size() / 4);
            for (size_t k = 0; k < gcd.size(); k++)
                gcd[k] = V4<double>{tm.gcoef[4 * k], tm.gcoef[4 * k + 1], tm.gcoef[4 * k + 2], tm.gcoef[4 * k + 3]};
            if (h->mixed) {
                gcf.resize(gcd.size());
                for (size_t k = 0; k < gcd.size(); k++)
                    gcf[k] = V4<float>{(float)gcd[k].x, (float)gcd[k].y, (float)gcd[k].z, (float)gcd[k].w};
                V4<float> *d = nullptr;
                if ((rc = upload(&d, gcf, h->stream))) return bail(rc);
                h->t_gcoef = d;
            } else {
                V4<double> *d = nullptr;
                if ((rc = upload(&d, gcd, h->stream))) return bail(rc);
                h->t_gcoef = d;
            }
            if (FVMR_TILE_GEOM_LOADED) {
                // centroids and normals as the kernels would recompute them (standard geometry:
                // the caller's values), rounded to the kernels' geometry type
                auto pack = [&](const std::vector<double> &xy, void **dst) -> int {
                    if (h->mixed) {
                        std::vector<V2<float>> v(xy.size() / 2);
                        for (size_t k = 0; k < v.size(); k++) v[k] = V2<float>{(float)xy[2 * k], (float)xy[2 * k + 1]};
                        V2<float> *d = nullptr;
                        int r = upload(&d, v, h->stream);
                        *dst = d;
                        if (!r && hipStreamSynchronize(h->stream) != hipSuccess) r = set_error(FVMR_ERR_HIP, "tile upload failed");
                        return r;
                    }
                    std::vector<V2<double>> v(xy.size() / 2);
                    for (size_t k = 0; k < v.size(); k++) v[k] = V2<double>{xy[2 * k], xy[2 * k + 1]};
                    V2<double> *d = nullptr;
                    int r = upload(&d, v, h->stream);
                    *dst = d;
                    if (!r && hipStreamSynchronize(h->stream) != hipSuccess) r = set_error(FVMR_ERR_HIP, "tile upload failed");
                    return r;
                };
                if ((rc = pack(tm.lcxy, &h->t_lcxy)) || (rc = pack(tm.fnrm, &h->t_fnrm))) return bail(rc);
            }
            std::vector<DtRec> dtr(nc);
            for (int i = 0; i < nc; i++)
                dtr[i] = DtRec{dtcode[i], tm.dtface[i], h->mixed ? (double)(float)vol[i] : vol[i]};
            if ((rc = upload(&h->t_dtrec, dtr, h->stream))) return bail(rc);
            if ((rc = upload(&h->tile_case_begin, tm.case_tile_begin, h->stream))) return bail(rc);
            // the host vectors of these uploads end with this block
            if (hipStreamSynchronize(h->stream) != hipSuccess) return bail(set_error(FVMR_ERR_HIP, "tile upload failed"));
        }
    }
    if (ncases > 1 && !h->tiles)
        return bail(set_error(FVMR_ERR_UNSUPPORTED, "a batch needs the tile kernels (standard geometry, kernels != CELL)"));
    {
        std::vector<CaseBC> cb(ncases);
        for (int cs = 0; cs < ncases; cs++) {
            const fvmr_params *q = case_params ? &case_params[cs] : p;
            CaseBC &b = cb[cs];
            for (int k = 0; k < 4; k++) { b.inf[k] = q->inf_ruvp[k]; b.inlet[k] = q->inlet_ruvp[k]; b.outlet[k] = q->outlet_ruvp[k]; }
            b.CFL = q->CFL;
            b.pad[0] = b.pad[1] = b.pad[2] = 0.0;
        }
        if ((rc = upload(&h->cbc, cb, h->stream))) return bail(rc);
        if (hipStreamSynchronize(h->stream) != hipSuccess) return bail(set_error(FVMR_ERR_HIP, "case upload failed"));
    }
    // device memory
    const size_t gsz = h->mixed ? sizeof(Grad<float>) : sizeof(Grad<double>);
    for (int k = 0; k < 3; k++)
        if (hipMalloc((void **)&h->Ubuf[k], (size_t)nc * sizeof(S4)) != hipSuccess) return bail(set_error(FVMR_ERR_HIP, "hipMalloc U"));
    const int nparts = std::max(std::max(grid3_of(nc), 4 * grid_of(nc)), h->ntiles);
    if (hipMalloc(&h->grad, (size_t)nc * gsz) != hipSuccess ||
        hipMalloc((void **)&h->rhs, (size_t)nc * sizeof(S4)) != hipSuccess ||
        hipMalloc((void **)&h->fix, (size_t)nc * sizeof(S4)) != hipSuccess ||
        hipMalloc((void **)&h->bad_list, 2 * (size_t)nc * sizeof(int)) != hipSuccess ||
        hipMalloc((void **)&h->partial, (size_t)nparts * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&h->norm_out, 4 * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&h->sc, ncases * sizeof(Scalars)) != hipSuccess)
        return bail(set_error(FVMR_ERR_HIP, "hipMalloc failed"));
    if ((rc = upload(&h->cslot, cslot, h->stream)) || (rc = upload(&h->cnode, cnode, h->stream)) ||
        (rc = upload(&h->dtcode, dtcode, h->stream)) || (rc = upload(&h->eint, eint, h->stream)))
        return bail(rc);
    rc = h->mixed ? build_geometry<float>(h, m, eperm) : build_geometry<double>(h, m, eperm);
    if (rc) return bail(rc);
    std::vector<Scalars> s0(ncases);
    if (hipMemsetAsync(h->Ubuf[0], 0, (size_t)nc * sizeof(S4), h->stream) != hipSuccess ||
        hipMemcpyAsync(h->sc, s0.data(), ncases * sizeof(Scalars), hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
        return bail(set_error(FVMR_ERR_HIP, "initial copy failed"));
    *out = h;
    return FVMR_OK;
}

extern "C" {

int fvmr_create(const fvmr_mesh_soa *m, const fvmr_params *p, fvmr_handle **out) {
    if (!m) return set_error(FVMR_ERR_INVALID_ARG, "null argument");
    const int cb[2] = {0, m->num_element};
    return create_impl(m, p, 1, cb, nullptr, out);
}

// Settings every case of a batch shares (all of fvmr_params but the boundary states and CFL).
static bool same_shared_params(const fvmr_params &a, const fvmr_params &b) {
    return a.gamma == b.gamma && a.R == b.R && a.Pr == b.Pr && a.Pr_viscous == b.Pr_viscous && a.T0 == b.T0 &&
           a.mu0 == b.mu0 && a.equation == b.equation && a.flux_scheme == b.flux_scheme &&
           a.reconstruct == b.reconstruct && a.gradient == b.gradient && a.limiter == b.limiter &&
           a.time_advance == b.time_advance && a.viscosity_model == b.viscosity_model && a.precision == b.precision &&
           a.renumber == b.renumber && a.compat_inv2x2 == b.compat_inv2x2 && a.kernels == b.kernels &&
           a.tile_cells == b.tile_cells && a.tile_threads == b.tile_threads && a.wall_temperature == b.wall_temperature &&
           a.venkat_K == b.venkat_K &&
           a.dsr_shock_x == b.dsr_shock_x && a.dsr_shock_y == b.dsr_shock_y && a.dsr_shock_speed == b.dsr_shock_speed;
}

int fvmr_create_batch(int ncases, const fvmr_mesh_soa *const *meshes, const fvmr_params *params, fvmr_handle **out) {
    if (ncases < 1 || !meshes || !params || !out) return set_error(FVMR_ERR_INVALID_ARG, "invalid batch arguments");
    *out = nullptr;
    for (int c = 0; c < ncases; c++) {
        if (!meshes[c]) return set_error(FVMR_ERR_INVALID_ARG, "null mesh in batch");
        if (!same_shared_params(params[0], params[c]))
            return set_error(FVMR_ERR_INVALID_ARG, "batch cases differ in a shared setting (only the boundary states and CFL may differ)");
        int rc = validate_mesh(meshes[c]);
        if (rc) return rc;
    }
    // the cases' SoAs one after the other, indices offset
    int64_t tn = 0, tc = 0, tf = 0, tb = 0;
    for (int c = 0; c < ncases; c++) {
        tn += meshes[c]->num_node; tc += meshes[c]->num_element; tf += meshes[c]->num_edge; tb += meshes[c]->num_boundary_set;
    }
    if (tn > INT32_MAX || tc > INT32_MAX || tf > INT32_MAX) return set_error(FVMR_ERR_UNSUPPORTED, "batch too large");
    std::vector<double> nxy[2], exy[2], evol, elen, gxy[2], gnrm[2];
    std::vector<int> en[4], ee[4], enb[4], gn[2], gset, gpair, gL, gR, btype;
    std::vector<int> case_begin{0};
    int noff = 0, coff = 0, foff = 0, soff = 0;
    auto add = [](int v, int off) { return v < 0 ? v : v + off; };
    for (int c = 0; c < ncases; c++) {
        const fvmr_mesh_soa &q = *meshes[c];
        for (int k = 0; k < 2; k++) nxy[k].insert(nxy[k].end(), q.node_xy[k], q.node_xy[k] + q.num_node);
        for (int i = 0; i < q.num_element; i++) {
            for (int j = 0; j < 4; j++) {
                en[j].push_back(add(q.element_nodes[j][i], noff));
                ee[j].push_back(add(q.element_edges[j][i], foff));
                enb[j].push_back(add(q.element_neighbors[j][i], coff));
            }
            for (int k = 0; k < 2; k++) exy[k].push_back(q.element_xy[k][i]);
            evol.push_back(q.element_volume[i]);
        }
        for (int e = 0; e < q.num_edge; e++) {
            for (int k = 0; k < 2; k++) {
                gn[k].push_back(q.edge_nodes[k][e] + noff);
                gxy[k].push_back(q.edge_xy[k][e]);
                gnrm[k].push_back(q.edge_normal[k][e]);
            }
            gset.push_back(add(q.edge_setID[e], soff));
            gpair.push_back(add(q.edge_periodicPair[e], foff));
            gL.push_back(q.edge_elementL[e] + coff);
            gR.push_back(add(q.edge_elementR[e], coff));
            elen.push_back(q.edge_length[e]);
        }
        btype.insert(btype.end(), q.boundary_type, q.boundary_type + q.num_boundary_set);
        noff += q.num_node; coff += q.num_element; foff += q.num_edge; soff += q.num_boundary_set;
        case_begin.push_back(coff);
    }
    fvmr_mesh_soa m{};
    m.num_node = noff; m.num_element = coff; m.num_edge = foff; m.num_boundary_set = soff;
    for (int k = 0; k < 2; k++) {
        m.node_xy[k] = nxy[k].data(); m.element_xy[k] = exy[k].data(); m.edge_nodes[k] = gn[k].data();
        m.edge_xy[k] = gxy[k].data(); m.edge_normal[k] = gnrm[k].data();
    }
    for (int j = 0; j < 4; j++) { m.element_nodes[j] = en[j].data(); m.element_edges[j] = ee[j].data(); m.element_neighbors[j] = enb[j].data(); }
    m.element_volume = evol.data(); m.edge_length = elen.data(); m.edge_setID = gset.data();
    m.edge_periodicPair = gpair.data(); m.edge_elementL = gL.data(); m.edge_elementR = gR.data();
    m.boundary_type = btype.data();
    return create_impl(&m, &params[0], ncases, case_begin.data(), params, out);
}

int fvmr_batch_cases(fvmr_handle *h, int *ncases, int *case_cell_begin) {
    if (!h || !ncases) return set_error(FVMR_ERR_INVALID_ARG, "null argument");
    *ncases = h->ncases;
    if (case_cell_begin) std::copy(h->case_cell_begin.begin(), h->case_cell_begin.end(), case_cell_begin);
    return FVMR_OK;
}

int fvmr_destroy(fvmr_handle *h) {
    if (!h) return FVMR_OK;
    (void)hipStreamSynchronize(h->stream);
    free_all(h);
    delete h;
    return FVMR_OK;
}

// The state crosses the boundary as four host arrays in the caller's order: one contiguous copy
// per component into (out of) a device scratch in that order, and the permutation to (from) the
// device order as a gather (scatter) on the device.  perm is a permutation of 0 .. nc-1 (built
// by fvmr_create), so every index stays inside the nc-long arrays.
static int ensure_io(fvmr_handle *h) {
    if (h->io_ready) return FVMR_OK;
    const size_t nc = (size_t)h->nc;
    auto release = [&]() {
        if (h->perm_dev) (void)hipFree(h->perm_dev);
        if (h->io_soa) (void)hipFree(h->io_soa);
        h->perm_dev = nullptr;
        h->io_soa = nullptr;
    };
    if (hipMalloc((void **)&h->perm_dev, nc * sizeof(int)) != hipSuccess ||
        hipMalloc((void **)&h->io_soa, 4 * nc * sizeof(double)) != hipSuccess) {
        release();
        return set_error(FVMR_ERR_HIP, "hipMalloc state transfer buffers");
    }
    // ready only once the permutation is on the device: the transfer kernels index with it
    const hipError_t e = hipMemcpy(h->perm_dev, h->perm.data(), nc * sizeof(int), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        release();
        return set_error(FVMR_ERR_HIP, std::string("permutation upload: ") + hipGetErrorString(e));
    }
    h->io_ready = true;
    return FVMR_OK;
}

int fvmr_set_state(fvmr_handle *h, const double *const U[4]) {
    if (!h || !U) return set_error(FVMR_ERR_INVALID_ARG, "null argument");
    for (int k = 0; k < 4; k++)
        if (!U[k]) return set_error(FVMR_ERR_INVALID_ARG, "null state component");
    int rc = ensure_io(h);
    if (rc) return rc;
    const int nc = h->nc;
    for (int k = 0; k < 4; k++)
        HIP_OK(hipMemcpyAsync(h->io_soa + (size_t)k * nc, U[k], (size_t)nc * sizeof(double), hipMemcpyHostToDevice,
                              h->stream));
    k_from_caller<<<grid_of(nc), kBlock, 0, h->stream>>>(nc, 0, h->perm_dev, h->io_soa, h->Ubuf[h->cur]);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(h->stream));
    h->rhs_valid = false;
    return FVMR_OK;
}

static int read_back(fvmr_handle *h, const S4 *dev, double *const out[4]) {
    for (int k = 0; k < 4; k++)
        if (!out[k]) return set_error(FVMR_ERR_INVALID_ARG, "null output component");
    int rc = ensure_io(h);
    if (rc) return rc;
    const int nc = h->nc;
    k_to_caller<<<grid_of(nc), kBlock, 0, h->stream>>>(nc, 0, h->perm_dev, dev, h->io_soa);
    HIP_OK(hipGetLastError());
    for (int k = 0; k < 4; k++)
        HIP_OK(hipMemcpyAsync(out[k], h->io_soa + (size_t)k * nc, (size_t)nc * sizeof(double), hipMemcpyDeviceToHost,
                              h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    return FVMR_OK;
}

int fvmr_get_state(fvmr_handle *h, double *const U[4]) {
    if (!h || !U) return set_error(FVMR_ERR_INVALID_ARG, "null argument");
    return read_back(h, h->Ubuf[h->cur], U);
}

static int read_scalars(fvmr_handle *h, Scalars &s) {
    HIP_OK(hipMemcpyAsync(&s, h->sc, sizeof(Scalars), hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    return FVMR_OK;
}

int fvmr_advance(fvmr_handle *h, int nsteps, double *t, double T, fvmr_step_info *info) {
    if (!h || !t || nsteps < 0) return set_error(FVMR_ERR_INVALID_ARG, "invalid argument");
    Scalars s{};
    int rc = read_scalars(h, s);
    if (rc) return rc;
    s.t = *t;
    s.T = T;
    s.counter = 0; s.bad_count = 0; s.n_restricted = 0;
    HIP_OK(hipMemcpyAsync(h->sc, &s, sizeof(Scalars), hipMemcpyHostToDevice, h->stream));
    if ((rc = run_steps(h, nsteps, true))) return rc;
    if ((rc = read_scalars(h, s))) return rc;
    if (info) {
        info->t = s.t;
        info->dt = s.dt;
        info->n_restricted = s.n_restricted;
        info->steps_done = nsteps;
        if (info->residual_norm && nsteps > 0 && (rc = fvmr_residual(h, info->residual_norm, info->residual))) return rc;
    }
    if (h->profiling && (rc = prof_collect(h))) return rc;
    *t = s.t;
    return FVMR_OK;
}

int fvmr_step(fvmr_handle *h, double *t, double T, fvmr_step_info *info) { return fvmr_advance(h, 1, t, T, info); }

// The device rows of case c are c's own cells: [case_cell_begin[c], case_cell_begin[c+1]).
int fvmr_batch_set_state(fvmr_handle *h, int c, const double *const U[4]) {
    if (!h || !U || c < 0 || c >= h->ncases) return set_error(FVMR_ERR_INVALID_ARG, "invalid case or argument");
    for (int k = 0; k < 4; k++)
        if (!U[k]) return set_error(FVMR_ERR_INVALID_ARG, "null state component");
    const int b = h->case_cell_begin[c], n = h->case_cell_begin[c + 1] - b;
    int rc = ensure_io(h);
    if (rc) return rc;
    if (n > 0) {   // the case's cells through the transfer scratch (perm of case c maps into b .. b+n-1)
        for (int k = 0; k < 4; k++)
            HIP_OK(hipMemcpyAsync(h->io_soa + (size_t)k * n, U[k], (size_t)n * sizeof(double), hipMemcpyHostToDevice,
                                  h->stream));
        k_from_caller<<<grid_of(n), kBlock, 0, h->stream>>>(n, b, h->perm_dev + b, h->io_soa, h->Ubuf[h->cur] + b);
        HIP_OK(hipGetLastError());
    }
    HIP_OK(hipStreamSynchronize(h->stream));
    h->rhs_valid = false;
    return FVMR_OK;
}

int fvmr_batch_get_state(fvmr_handle *h, int c, double *const U[4]) {
    if (!h || !U || c < 0 || c >= h->ncases) return set_error(FVMR_ERR_INVALID_ARG, "invalid case or argument");
    for (int k = 0; k < 4; k++)
        if (!U[k]) return set_error(FVMR_ERR_INVALID_ARG, "null output component");
    const int b = h->case_cell_begin[c], n = h->case_cell_begin[c + 1] - b;
    int rc = ensure_io(h);
    if (rc) return rc;
    if (n > 0) {
        k_to_caller<<<grid_of(n), kBlock, 0, h->stream>>>(n, b, h->perm_dev + b, h->Ubuf[h->cur] + b, h->io_soa);
        HIP_OK(hipGetLastError());
        for (int k = 0; k < 4; k++)
            HIP_OK(hipMemcpyAsync(U[k], h->io_soa + (size_t)k * n, (size_t)n * sizeof(double), hipMemcpyDeviceToHost,
                                  h->stream));
    }
    HIP_OK(hipStreamSynchronize(h->stream));
    return FVMR_OK;
}

int fvmr_batch_advance(fvmr_handle *h, int nsteps, double *t, const double *T) {
    if (!h || !t || !T || nsteps < 0) return set_error(FVMR_ERR_INVALID_ARG, "invalid argument");
    if (!h->tiles) return set_error(FVMR_ERR_UNSUPPORTED, "fvmr_batch_advance needs the tile kernels");
    std::vector<Scalars> s(h->ncases);
    HIP_OK(hipMemcpyAsync(s.data(), h->sc, s.size() * sizeof(Scalars), hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    for (int c = 0; c < h->ncases; c++) { s[c].t = t[c]; s[c].T = T[c]; }
    s[0].counter = 0; s[0].bad_count = 0; s[0].n_restricted = 0;
    HIP_OK(hipMemcpyAsync(h->sc, s.data(), s.size() * sizeof(Scalars), hipMemcpyHostToDevice, h->stream));
    int rc;
    if ((rc = run_steps(h, nsteps, false))) return rc;
    HIP_OK(hipMemcpyAsync(s.data(), h->sc, s.size() * sizeof(Scalars), hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    if (h->profiling && (rc = prof_collect(h))) return rc;
    for (int c = 0; c < h->ncases; c++) t[c] = s[c].t;
    return FVMR_OK;
}

int fvmr_batch_step_info(fvmr_handle *h, int c, fvmr_step_info *info) {
    if (!h || !info || c < 0 || c >= h->ncases) return set_error(FVMR_ERR_INVALID_ARG, "invalid case or argument");
    Scalars s{};
    HIP_OK(hipMemcpyAsync(&s, h->sc + c, sizeof(Scalars), hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    info->t = s.t;
    info->dt = s.dt;
    info->n_restricted = s.n_restricted;
    return FVMR_OK;
}

int fvmr_device_division(int n, const double *a, const double *b, double *quotient, double *reciprocal) {
    if (n < 0 || (n > 0 && (!a || !b || !quotient || !reciprocal))) return set_error(FVMR_ERR_INVALID_ARG, "invalid argument");
    if (n == 0) return FVMR_OK;
    const size_t bytes = (size_t)n * sizeof(double);
    double *d = nullptr;
    if (hipMalloc((void **)&d, 4 * bytes) != hipSuccess) return set_error(FVMR_ERR_HIP, "hipMalloc division operands");
    hipError_t e = hipMemcpy(d, a, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + n, b, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        k_division<<<grid_of(n), kBlock>>>(n, d, d + n, d + 2 * (size_t)n, d + 3 * (size_t)n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(quotient, d + 2 * (size_t)n, bytes, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(reciprocal, d + 3 * (size_t)n, bytes, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return set_error(FVMR_ERR_HIP, std::string("fvmr_device_division: ") + hipGetErrorString(e));
    return FVMR_OK;
}

int fvmr_stage_rhs(fvmr_handle *h, double *const rhs[4]) {
    if (!h || !rhs) return set_error(FVMR_ERR_INVALID_ARG, "null argument");
    // one forward-Euler stage with dt = 0 into a scratch buffer: rhs of the current state
    Scalars s{};
    int rc = read_scalars(h, s);
    if (rc) return rc;
    Scalars z = s;
    z.dt = 0.0;
    z.t_rk = s.t;
    HIP_OK(hipMemcpyAsync(h->sc, &z, sizeof(Scalars), hipMemcpyHostToDevice, h->stream));
    const int a = h->cur, b = (a + 1) % 3;
    if (h->tiles) {
        rc = h->mixed ? launch_tile_stage<float, float>(h, h->Ubuf[a], h->Ubuf[a], h->Ubuf[b], 0, 1, 1, 0, h->rhs)
                      : launch_tile_stage<double, double>(h, h->Ubuf[a], h->Ubuf[a], h->Ubuf[b], 0, 1, 1, 0, h->rhs);
        if (rc) return rc;
    } else if (h->mixed) {
        if ((rc = launch_gradient<float, float>(h, h->Ubuf[a], -1)) ||
            (rc = launch_stage<float, float>(h, h->Ubuf[a], h->Ubuf[a], h->Ubuf[b], 0, 1, 1, 0, h->rhs))) return rc;
    } else {
        if ((rc = launch_gradient<double, double>(h, h->Ubuf[a], -1)) ||
            (rc = launch_stage<double, double>(h, h->Ubuf[a], h->Ubuf[a], h->Ubuf[b], 0, 1, 1, 0, h->rhs))) return rc;
    }
    HIP_OK(hipMemcpyAsync(h->sc, &s, sizeof(Scalars), hipMemcpyHostToDevice, h->stream));
    h->rhs_valid = true;
    return read_back(h, h->rhs, rhs);
}

int fvmr_gradient(fvmr_handle *h, double *const Ux[4], double *const Uy[4]) {
    if (!h || !Ux || !Uy) return set_error(FVMR_ERR_INVALID_ARG, "null argument");
    for (int k = 0; k < 4; k++)
        if (!Ux[k] || !Uy[k]) return set_error(FVMR_ERR_INVALID_ARG, "null output component");
    int rc = h->mixed ? launch_gradient<float, float>(h, h->Ubuf[h->cur], -1)
                      : launch_gradient<double, double>(h, h->Ubuf[h->cur], -1);
    if (rc) return rc;
    const int nc = h->nc;
    auto scatter = [&](auto &buf) {
        for (int i = 0; i < nc; i++) {
            const int r = h->perm[i];
            for (int k = 0; k < 4; k++) { Ux[k][r] = buf[i].x[k]; Uy[k][r] = buf[i].y[k]; }
        }
    };
    if (h->mixed) {
        std::vector<Grad<float>> buf(nc);
        HIP_OK(hipMemcpyAsync(buf.data(), h->grad, buf.size() * sizeof(Grad<float>), hipMemcpyDeviceToHost, h->stream));
        HIP_OK(hipStreamSynchronize(h->stream));
        scatter(buf);
    } else {
        std::vector<Grad<double>> buf(nc);
        HIP_OK(hipMemcpyAsync(buf.data(), h->grad, buf.size() * sizeof(Grad<double>), hipMemcpyDeviceToHost, h->stream));
        HIP_OK(hipStreamSynchronize(h->stream));
        scatter(buf);
    }
    return FVMR_OK;
}

int fvmr_global_dt(fvmr_handle *h, double *dt) {
    if (!h || !dt) return set_error(FVMR_ERR_INVALID_ARG, "null argument");
    int rc;
    if (h->tiles)
        rc = h->mixed ? launch_tile_dt<float, float>(h, h->Ubuf[h->cur], 0)
                      : launch_tile_dt<double, double>(h, h->Ubuf[h->cur], 0);
    else
        rc = h->mixed ? launch_gradient<float, float>(h, h->Ubuf[h->cur], 0)
                      : launch_gradient<double, double>(h, h->Ubuf[h->cur], 0);
    if (rc) return rc;
    Scalars s;
    if ((rc = read_scalars(h, s))) return rc;
    *dt = s.dt;
    return FVMR_OK;
}

int fvmr_residual(fvmr_handle *h, int norm, double out[4]) {
    if (!h || !out) return set_error(FVMR_ERR_INVALID_ARG, "null argument");
    if (norm != FVMR_NORM_1 && norm != FVMR_NORM_2 && norm != FVMR_NORM_INF) return set_error(FVMR_ERR_INVALID_ARG, "invalid norm type");
    if (!h->rhs_valid) return set_error(FVMR_ERR_INVALID_ARG, "no residual available: run a step first");
    hipEvent_t e0{};
    int rc = prof_begin(h, 4, &e0);
    if (rc) return rc;
    k_norm<<<grid_of(h->nc), kBlock, 0, h->stream>>>(h->nc, h->rhs, norm, h->partial);
    HIP_OK(hipGetLastError());
    k_norm_finish<<<1, kBlock, 0, h->stream>>>(grid_of(h->nc), norm, h->partial, h->norm_out);
    HIP_OK(hipGetLastError());
    if ((rc = prof_end(h, 4, e0))) return rc;
    HIP_OK(hipMemcpyAsync(out, h->norm_out, 4 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    return FVMR_OK;
}

int fvmr_permutation(fvmr_handle *h, int *perm) {
    if (!h || !perm) return set_error(FVMR_ERR_INVALID_ARG, "null argument");
    std::copy(h->perm.begin(), h->perm.end(), perm);
    return FVMR_OK;
}

// Node fields of the current state into h->node_buf (6 x nv on the device).
static int node_field_device(fvmr_handle *h) {
    const int nv = h->nv;
    if (!h->node_buf && hipMalloc((void **)&h->node_buf, 6 * (size_t)nv * sizeof(double)) != hipSuccess)
        return set_error(FVMR_ERR_HIP, "hipMalloc node fields");
    const double *inf = h->p.inf_ruvp;
    const double V2 = inf[1] * inf[1] + inf[2] * inf[2];
    k_node_field<<<grid_of(nv), kBlock, 0, h->stream>>>(nv, h->nc_off, h->nc_cells, h->Ubuf[h->cur], h->p.gamma,
                                                        inf[3], 0.5 * inf[0] * V2, h->node_buf);
    HIP_OK(hipGetLastError());
    return FVMR_OK;
}

int fvmr_node_field(fvmr_handle *h, double *const ruvp[4], double *Cp, double *Ma) {
    if (!h) return set_error(FVMR_ERR_INVALID_ARG, "null handle");
    int rc = node_field_device(h);
    if (rc) return rc;
    const size_t nv = h->nv;
    std::vector<double> buf(6 * nv);
    HIP_OK(hipMemcpyAsync(buf.data(), h->node_buf, buf.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    if (ruvp)
        for (int k = 0; k < 4; k++)
            if (ruvp[k]) std::copy(buf.begin() + k * nv, buf.begin() + (k + 1) * nv, ruvp[k]);
    if (Cp) std::copy(buf.begin() + 4 * nv, buf.begin() + 5 * nv, Cp);
    if (Ma) std::copy(buf.begin() + 5 * nv, buf.begin() + 6 * nv, Ma);
    return FVMR_OK;
}

int fvmr_vertex_colors(fvmr_handle *h, int var, float lo, float hi, int npts, const float *colormap, float *vertices,
                       int on_device) {
    if (!h || !colormap || !vertices || var < 0 || var > 3 || npts < 1 || npts > kMaxColorPoints)
        return set_error(FVMR_ERR_INVALID_ARG, "invalid argument");
    ColorMapArg cm{};
    cm.n = npts;
    for (int i = 0; i < npts; i++)
        cm.p[i] = make_float4(colormap[4 * i], colormap[4 * i + 1], colormap[4 * i + 2], colormap[4 * i + 3]);
    const size_t nv = h->nv;
    float *dv = vertices;
    if (!on_device) {
        if (hipMalloc((void **)&dv, 6 * nv * sizeof(float)) != hipSuccess) return set_error(FVMR_ERR_HIP, "hipMalloc vertices");
        if (hipMemcpyAsync(dv, vertices, 6 * nv * sizeof(float), hipMemcpyHostToDevice, h->stream) != hipSuccess) {
            (void)hipFree(dv);
            return set_error(FVMR_ERR_HIP, "vertex upload");
        }
    }
    k_vertex_colors<<<grid_of((int)nv), kBlock, 0, h->stream>>>((int)nv, h->nc_off, h->nc_cells, h->Ubuf[h->cur],
                                                               h->p.gamma, var, lo, hi, cm, dv);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && !on_device) e = hipMemcpyAsync(vertices, dv, 6 * nv * sizeof(float), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (!on_device) (void)hipFree(dv);
    if (e != hipSuccess) return set_error(FVMR_ERR_HIP, std::string("fvmr_vertex_colors: ") + hipGetErrorString(e));
    return FVMR_OK;
}

int fvmr_surface_force(fvmr_handle *h, int nsets, const int *set_offsets, const int *set_edges, double S_ref,
                       double *Cx, double *Cy) {
    if (!h || nsets < 0 || (nsets > 0 && (!set_offsets || !Cx || !Cy))) return set_error(FVMR_ERR_INVALID_ARG, "null argument");
    if (nsets == 0) return FVMR_OK;
    if (set_offsets[0] != 0) return set_error(FVMR_ERR_INVALID_ARG, "set_offsets[0] must be 0");
    for (int s = 0; s < nsets; s++)
        if (set_offsets[s + 1] < set_offsets[s]) return set_error(FVMR_ERR_INVALID_ARG, "set_offsets must not decrease");
    const int ne = set_offsets[nsets];
    if (ne > 0 && !set_edges) return set_error(FVMR_ERR_INVALID_ARG, "null set_edges");
    std::vector<int> en(2 * (size_t)std::max(ne, 1));
    std::vector<double> eg(3 * (size_t)std::max(ne, 1));
    for (int k = 0; k < ne; k++) {
        const auto it = std::lower_bound(h->bedge_ref.begin(), h->bedge_ref.end(), set_edges[k]);
        if (it == h->bedge_ref.end() || *it != set_edges[k])
            return set_error(FVMR_ERR_INVALID_ARG, "surface force: edge " + std::to_string(set_edges[k]) + " is not in a boundary set");
        const size_t b = it - h->bedge_ref.begin();
        en[2 * k] = h->bedge_nodes[2 * b];
        en[2 * k + 1] = h->bedge_nodes[2 * b + 1];
        for (int q = 0; q < 3; q++) eg[3 * k + q] = h->bedge_geo[3 * b + q];
    }
    int rc = node_field_device(h);
    if (rc) return rc;
    std::vector<int> off(set_offsets, set_offsets + nsets + 1);
    int *d_off = nullptr, *d_en = nullptr;
    double *d_eg = nullptr, *d_out = nullptr;
    auto release = [&]() {
        if (d_off) (void)hipFree(d_off);
        if (d_en) (void)hipFree(d_en);
        if (d_eg) (void)hipFree(d_eg);
        if (d_out) (void)hipFree(d_out);
    };
    std::vector<double> out(2 * (size_t)nsets);
    if ((rc = upload(&d_off, off, h->stream)) || (rc = upload(&d_en, en, h->stream)) || (rc = upload(&d_eg, eg, h->stream)) ||
        hipMalloc((void **)&d_out, out.size() * sizeof(double)) != hipSuccess) {
        release();
        return rc ? rc : set_error(FVMR_ERR_HIP, "hipMalloc surface force");
    }
    k_surface_force<<<(nsets + 63) / 64, 64, 0, h->stream>>>(nsets, d_off, d_en, d_eg, h->node_buf + 4 * (size_t)h->nv,
                                                           S_ref, d_out);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out.data(), d_out, out.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    release();
    if (e != hipSuccess) return set_error(FVMR_ERR_HIP, std::string("surface force: ") + hipGetErrorString(e));
    for (int s = 0; s < nsets; s++) { Cx[s] = out[2 * s]; Cy[s] = out[2 * s + 1]; }
    return FVMR_OK;
}

// fvmr_surface_force with no allocation: the caller's set offsets and edge ids are copied into
// the device's `scratch` array (through scratch_ptr, taken at fvmr_create), the node Cp values
// are formed inside the kernel, and the results come back from the same array.
int fvmr_surface_force_noalloc(fvmr_handle *h, int nsets, const int *set_offsets, const int *set_edges,
                               double S_ref, double *Cx, double *Cy) {
    if (!h || nsets < 0 || (nsets > 0 && (!set_offsets || !Cx || !Cy))) return set_error(FVMR_ERR_INVALID_ARG, "null argument");
    if (nsets == 0) return FVMR_OK;
    if (nsets > kNoAllocMaxSets)
        return set_error(FVMR_ERR_INVALID_ARG, "fvmr_surface_force_noalloc: at most " + std::to_string(kNoAllocMaxSets) +
                                                   " sets per call (asked for " + std::to_string(nsets) + ")");
    if (set_offsets[0] != 0) return set_error(FVMR_ERR_INVALID_ARG, "set_offsets[0] must be 0");
    for (int s = 0; s < nsets; s++)
        if (set_offsets[s + 1] < set_offsets[s]) return set_error(FVMR_ERR_INVALID_ARG, "set_offsets must not decrease");
    const int ne = set_offsets[nsets];
    if (ne > kNoAllocMaxEdges)
        return set_error(FVMR_ERR_INVALID_ARG, "fvmr_surface_force_noalloc: at most " + std::to_string(kNoAllocMaxEdges) +
                                                   " edges per call (asked for " + std::to_string(ne) + ")");
    if (ne > 0 && !set_edges) return set_error(FVMR_ERR_INVALID_ARG, "null set_edges");
    for (int k = 0; k < ne; k++)
        if (!std::binary_search(h->bedge_ref.begin(), h->bedge_ref.end(), set_edges[k]))
            return set_error(FVMR_ERR_INVALID_ARG, "surface force: edge " + std::to_string(set_edges[k]) + " is not in a boundary set");
    double *base = scratch_ptr(h->device);
    if (!base) return set_error(FVMR_ERR_HIP, "fvmr_surface_force_noalloc: scratch of this device not initialised");
    int *d_off = reinterpret_cast<int *>(base) + kScratchOffInts;
    int *d_edges = reinterpret_cast<int *>(base) + kScratchEdgeInts;
    double *d_cx = base + kScratchOutDoubles, *d_cy = d_cx + kNoAllocMaxSets;
    HIP_OK(hipMemcpyAsync(d_off, set_offsets, (size_t)(nsets + 1) * sizeof(int), hipMemcpyHostToDevice, h->stream));
    if (ne > 0) HIP_OK(hipMemcpyAsync(d_edges, set_edges, (size_t)ne * sizeof(int), hipMemcpyHostToDevice, h->stream));
    const double *inf = h->p.inf_ruvp;
    const double V2 = inf[1] * inf[1] + inf[2] * inf[2];
    k_surface_force_direct<<<(nsets + 63) / 64, 64, 0, h->stream>>>(
        nsets, d_off, d_edges, (int)h->bedge_ref.size(), h->bedge_ref_dev, h->bedge_nodes_dev, h->bedge_geo_dev,
        h->nc_off, h->nc_cells, h->Ubuf[h->cur], h->p.gamma, inf[3], 0.5 * inf[0] * V2, S_ref, d_cx, d_cy);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(Cx, d_cx, (size_t)nsets * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipMemcpyAsync(Cy, d_cy, (size_t)nsets * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    return FVMR_OK;
}

}  // extern "C"

// Candidates of the tuners for a precision: every work-group shape, each with its default tile
// size and 0.8 and 1.15 of it (within the shape's LDS capacity).
struct TileCand {
    int threads, cells;
};
static std::vector<TileCand> tile_candidates(bool mixed) {
    std::vector<TileCand> cand;
    for (int S = 0; S < kNumShapes[mixed ? 1 : 0]; S++) {
        const int dflt = shape_cells(mixed, S), th = shape_threads(mixed, S);
        for (double f : {1.0, 0.8, 1.15}) {
            const int c = std::min(kTileMaxCells, std::max(8, (int)std::lround(f * dflt)));
            cand.push_back(TileCand{th, c});
        }
    }
    return cand;
}

// Times `steps` steps of the uniform freestream of each case on a handle made by `make` for
// every candidate (all handles at once, after a warm-up; the faster of two timed runs) and
// returns the fastest; timings: the ms per step of every candidate, in candidate order (may be
// NULL).
template <class Make>
static int autotune_tiles(int ncases, const fvmr_mesh_soa *const *meshes, const fvmr_params *params, int steps,
                          Make make, int *best_tile_threads, int *best_tile_cells, double *best_ms, double *timings) {
    const bool mixed = params[0].precision == FVMR_PREC_MIXED;
    const std::vector<TileCand> cand = tile_candidates(mixed);
    // the cases' uniform freestream states
    std::vector<std::vector<double>> U0(4 * (size_t)ncases);
    for (int c = 0; c < ncases; c++) {
        double *rows[4];
        for (int k = 0; k < 4; k++) {
            U0[4 * c + k].resize(meshes[c]->num_element);
            rows[k] = U0[4 * c + k].data();
        }
        int rc = fvmr_field_init(meshes[c], &params[c], FVMR_INIT_UNIFORM, nullptr, rows);
        if (rc) return rc;
    }
    // every candidate's handle first (their set-up is host work, during which the GPU's clocks
    // drop), then a warm-up that keeps the GPU busy for at least 40 ms (the clocks rise over
    // tens of milliseconds of load), then two timed passes, forward and backward, the faster of
    // each candidate's two runs counting
    const size_t nq = cand.size();
    std::vector<fvmr_handle *> hs(nq, nullptr);
    std::vector<std::vector<double>> tq(nq, std::vector<double>(ncases, 0.0));
    const std::vector<double> T(ncases, 1e300);
    auto release = [&]() {
        for (auto *h : hs)
            if (h) fvmr_destroy(h);
    };
    int rc = FVMR_OK;
    for (size_t q = 0; q < nq && !rc; q++) {
        rc = make(cand[q], &hs[q]);
        for (int c = 0; c < ncases && !rc; c++) {
            const double *rows[4] = {U0[4 * c].data(), U0[4 * c + 1].data(), U0[4 * c + 2].data(), U0[4 * c + 3].data()};
            rc = (ncases == 1) ? fvmr_set_state(hs[q], rows) : fvmr_batch_set_state(hs[q], c, rows);
        }
    }
    auto timed = [&](size_t q, float *ms) -> int {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        int r = FVMR_OK;
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
            r = set_error(FVMR_ERR_HIP, "hipEventCreate failed");
        if (!r && hipEventRecord(e0, hs[q]->stream) != hipSuccess) r = set_error(FVMR_ERR_HIP, "hipEventRecord failed");
        if (!r) r = fvmr_batch_advance(hs[q], steps, tq[q].data(), T.data());
        if (!r && (hipEventRecord(e1, hs[q]->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                   hipEventElapsedTime(ms, e0, e1) != hipSuccess))
            r = set_error(FVMR_ERR_HIP, "event timing failed");
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        return r;
    };
    float warm = 0.0f;
    for (int k = 0; k < 1000 && !rc && warm < 40.0f; k++) {
        float ms = 0.0f;
        rc = timed((size_t)k % nq, &ms);
        warm += ms;
    }
    std::vector<double> tbest(nq, 1e300);
    for (int pass = 0; pass < 2 && !rc; pass++)
        for (size_t k = 0; k < nq && !rc; k++) {
            const size_t q = pass == 0 ? k : nq - 1 - k;
            float ms = 0.0f;
            rc = timed(q, &ms);
            tbest[q] = std::min(tbest[q], (double)ms);
        }
    release();
    if (rc) return rc;
    double best = 1e300;
    TileCand best_c = cand[0];
    for (size_t q = 0; q < nq; q++) {
        if (timings) timings[q] = tbest[q] / steps;
        if (tbest[q] < best) { best = tbest[q]; best_c = cand[q]; }
    }
    if (best_tile_threads) *best_tile_threads = best_c.threads;
    *best_tile_cells = best_c.cells;
    if (best_ms) *best_ms = best / steps;
    return FVMR_OK;
}

extern "C" {

int fvmr_autotune(const fvmr_mesh_soa *m, const fvmr_params *p, int steps, int *best_tile_threads,
                  int *best_tile_cells, double *best_ms) {
    if (!m || !p || !best_tile_cells || steps < 1) return set_error(FVMR_ERR_INVALID_ARG, "invalid argument");
    const fvmr_mesh_soa *ms[1] = {m};
    return autotune_tiles(1, ms, p, steps, [&](TileCand tc, fvmr_handle **h) {
        fvmr_params q = *p;
        q.tile_threads = tc.threads;
        q.tile_cells = tc.cells;
        q.kernels = FVMR_KERNELS_TILE;
        return fvmr_create(m, &q, h);
    }, best_tile_threads, best_tile_cells, best_ms, nullptr);
}

int fvmr_autotune_batch(int ncases, const fvmr_mesh_soa *const *meshes, const fvmr_params *params, int steps,
                        int *best_tile_threads, int *best_tile_cells, double *best_ms, double *timings) {
    if (ncases < 1 || !meshes || !params || !best_tile_cells || steps < 1)
        return set_error(FVMR_ERR_INVALID_ARG, "invalid argument");
    std::vector<fvmr_params> q(params, params + ncases);
    return autotune_tiles(ncases, meshes, params, steps, [&](TileCand tc, fvmr_handle **h) {
        for (auto &x : q) { x.tile_threads = tc.threads; x.tile_cells = tc.cells; x.kernels = FVMR_KERNELS_TILE; }
        return fvmr_create_batch(ncases, meshes, q.data(), h);
    }, best_tile_threads, best_tile_cells, best_ms, timings);
}

int fvmr_autotune_candidates(int precision, int *n, int *tile_threads, int *tile_cells) {
    if (!n || (precision != FVMR_PREC_FP64 && precision != FVMR_PREC_MIXED))
        return set_error(FVMR_ERR_INVALID_ARG, "invalid argument");
    const std::vector<TileCand> cand = tile_candidates(precision == FVMR_PREC_MIXED);
    *n = (int)cand.size();
    for (size_t k = 0; k < cand.size(); k++) {
        if (tile_threads) tile_threads[k] = cand[k].threads;
        if (tile_cells) tile_cells[k] = cand[k].cells;
    }
    return FVMR_OK;
}

int fvmr_use_graphs(fvmr_handle *h, int on) {
    if (!h) return set_error(FVMR_ERR_INVALID_ARG, "null handle");
    h->use_graphs = (on != 0);
    return FVMR_OK;
}

int fvmr_profile_enable(fvmr_handle *h, int on) {
    if (!h) return set_error(FVMR_ERR_INVALID_ARG, "null handle");
    h->profiling = (on != 0);
    return FVMR_OK;
}

int fvmr_profile_query(fvmr_handle *h, int kernel, double *avg_ms, int *count) {
    if (!h || kernel < 0 || kernel >= kNumProf) return set_error(FVMR_ERR_INVALID_ARG, "invalid kernel class");
    int rc = prof_collect(h);
    if (rc) return rc;
    if (avg_ms) *avg_ms = h->prof[kernel].count ? h->prof[kernel].total_ms / h->prof[kernel].count : 0.0;
    if (count) *count = h->prof[kernel].count;
    return FVMR_OK;
}

int fvmr_profile_reset(fvmr_handle *h) {
    if (!h) return set_error(FVMR_ERR_INVALID_ARG, "null handle");
    int rc = prof_collect(h);
    if (rc) return rc;
    for (int k = 0; k < kNumProf; k++) { h->prof[k].total_ms = 0; h->prof[k].count = 0; }
    return FVMR_OK;
}

int fvmr_synchronize(fvmr_handle *h) {
    if (!h) return set_error(FVMR_ERR_INVALID_ARG, "null handle");
    HIP_OK(hipStreamSynchronize(h->stream));
    return FVMR_OK;
}

}  // extern "C"

#include "fvmr_adjoint.inl"
