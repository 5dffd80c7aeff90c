// Steady solve and discrete adjoint of the residual (no reference counterpart: the reference
// has neither).  Included at the end of fvmr_solver.hip, so it uses the solver's device
// physics directly.
//
// Residual.  r_c(U) = -(1/V_c) sum_j s_j F_{e_j}, the rhs of one stage (k_stage), with F the
// face flux of reconstruct -> ghost_state -> riemann_flux and the reconstruction fed by the
// limited gradient of grad_limit.  Steady states solve r(U) = 0; fvmr_steady_solve iterates
// U <- U + D r(U) (the handle's forward Euler or SSP-RK3, D = per-cell pseudo-time steps).
//
// Adjoint.  For J(U, alpha, X) with r(U*) = 0: psi solves (dr/dU)^T psi = -dJ/dU, then
// dJ/dalpha = dJ/dalpha|_U + psi . dr/dalpha and dJ/dX = dJ/dX|_U + psi . dr/dX.  psi comes
// from the same pseudo-time iteration psi <- psi + D ((dr/dU)^T psi + dJ/dU), whose iteration
// matrix I + D (dr/dU)^T has the eigenvalues of the primal's linearised I + D dr/dU.
//
// Vector-Jacobian product.  psi . r = sum_e mu_e . F_e with mu_e = lambda_L - lambda_R and
// lambda_c = -psi_c / V_c (0 outside the domain).  Per face (k_adj_faces): the reverse mode of
// the face physics (primitives, ghost state, Riemann flux, written out by hand) contracted with
// mu_e gives the adjoints of the face's two reconstructed states.  Per cell (k_adj_vjp):
// reconstruction passes its three face-state adjoints to U_c and to c's limited gradient; the
// limited gradient's derivatives in U_c and in the three states across (forward mode: the
// solver's grad_limit on Dual<4>, one conserved variable at a time, the branch taken by the
// values) send that on to U_c and to the neighbours: the share of neighbour n goes to a
// per-(edge, side) slot, which n's gather (k_adj_update) adds in slot order, so the product is
// deterministic.  The limited gradients of U* (a primal quantity) are formed once per solve;
// the derivatives of flux, gradient and limiter are recomputed in every product, and no
// Jacobian is stored.
//
// Scope: FP64 handles of one case, Euler, no periodic faces and no double Mach reflection
// boundary (a time-dependent state).  dJ/dX assumes the standard geometry relations (centroid =
// vertex mean, volume = triangle area, normal = (dy, -dx)/L from the edge's first node, midpoint
// = node mean); the reference rounds lengths to float, which is constant between float values,
// and the derivative is that of the unrounded length.
#include "fvmr_dual.h"

namespace {

using fvmr::Dual;
constexpr int kAdjBlock = 128;
constexpr int kAdjCheck = 50;       // iterations between residual checks
constexpr int kXRec = 18;           // node entries per cell of the coordinate product

constexpr int kBcFlagPeriodic = 1, kBcFlagDsr = 2, kBcFlagIsothermal = 4;

// Device workspace of the steady solve and the adjoint (allocated on first use).
struct AdjWork {
    S4 *psi[4] = {nullptr, nullptr, nullptr, nullptr};   // 0, 1 iterates; 2, 3 stages
    S4 *base = nullptr, *wedge = nullptr, *fadj = nullptr, *g = nullptr, *res = nullptr;
    double *dtl = nullptr, *wnode = nullptr, *partial = nullptr;
    int2 *enode = nullptr;
    Grad<double> *gbar = nullptr;
    double *xval = nullptr;
    int *xnode = nullptr;
    int cur = 0;
    bool psi_valid = false;
    bool grad_ready = false;   // h->grad holds the limited gradients of the state being differentiated
};

void adj_free(fvmr_solver *h) {
    AdjWork *w = (AdjWork *)h->adj;
    if (!w) return;
    void *ptrs[] = {w->psi[0], w->psi[1], w->psi[2], w->psi[3], w->base, w->wedge, w->fadj, w->g, w->res, w->dtl,
                    w->wnode, w->partial, w->enode, w->gbar, w->xval, w->xnode};
    for (void *q : ptrs) if (q) (void)hipFree(q);
    delete w;
    h->adj = nullptr;
}

// ---------------------------------------------------------------------------
// Device
// ---------------------------------------------------------------------------
// The face state of a cell at (xm, ym) with reconstruct's arithmetic; returns whether the
// linear reconstruction is kept (false: constant reconstruction or the out-of-range fallback).
__device__ __forceinline__ bool recon_state(const Phys &ph, const double Uc[4], const Grad<double> &g, double cx,
                                            double cy, double xm, double ym, double Uf[4]) {
    if (ph.reconstruct == FVMR_REC_LINEAR) {
        const double dx = xm - cx, dy = ym - cy;
        FaceState<double> s;
#pragma unroll
        for (int i = 0; i < 4; i++) s.U[i] = Uc[i] + g.x[i] * dx + g.y[i] * dy;
        set_prim(s, ph.gamma);
        if (!out_of_range(s.r, s.p)) {
#pragma unroll
            for (int i = 0; i < 4; i++) Uf[i] = s.U[i];
            return true;
        }
    }
#pragma unroll
    for (int i = 0; i < 4; i++) Uf[i] = Uc[i];
    return false;
}

// Gradient coefficients of one face slot as k_gradient forms them: Green-Gauss (gC, Sx s, Sy s)
// or least squares (dx, dy) to the cell across.  (xc, yc) centroid, (fx, fy) face midpoint,
// (Fx, Fy) centroid across, (sx, sy) = normal * length * sign.
template <class T>
__device__ __forceinline__ void slot_coef(const Phys &ph, bool has_out, T xc, T yc, T fx, T fy, T Fx, T Fy, T sx,
                                          T sy, T a[3]) {
    if (ph.gradient == FVMR_GRA_GREENGAUSS) {
        T gC = T(0.5);
        if (has_out) {
            const T ax = fx - Fx, ay = fy - Fy, bx = xc - Fx, by = yc - Fy;
            const T dFf = fsqrt(ax * ax + ay * ay);
            T dFC = fsqrt(bx * bx + by * by);
            if (dFC < T(fvmr::kEPS)) dFC = T(fvmr::kEPS);
            gC = fdiv(dFf, dFC);
        }
        a[0] = gC;
        a[1] = sx;
        a[2] = sy;
    } else {
        a[0] = has_out ? Fx - xc : T(0);
        a[1] = has_out ? Fy - yc : T(0);
        a[2] = T(0);
    }
}

// Per-cell pseudo-time step dt_c = CFL V_c / sum over c's faces of (|u.n| + a) L (+ for NS
// max(4/3, gamma/Pr) mu / rho L^2 / V_c), u, a from the mean of the two cells' primitives (the
// cell's own at a boundary).  The reference's global dt (CalculateDt.cpp:24-62) scales the sound
// speed with sqrt(L); that is safe under its global minimum but too large for cells with L > 1
// as a per-cell step, so the pseudo-time step uses the consistent spectral radius.  The step
// does not change the steady state.
__global__ __launch_bounds__(kBlock) void k_local_dt(Mesh<double> m, Phys ph, const S4 *__restrict__ U,
                                                      double *__restrict__ dtl) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    const int nc = m.nc;
    if (c >= nc) return;
    double Uc[4], rc[4];
    ld(U, c, Uc);
    prim(Uc, rc, ph.gamma);
    const V4<double> cg = m.cgeo[c];
    double alpha_c = 0.0;
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const int2 sl = m.cslot[j * nc + c];
        const int e = sl.x & kSlotEdgeMask, out = sl.y;
        double r[4] = {rc[0], rc[1], rc[2], rc[3]};
        if (out != -1) {
            double Uo[4], ro[4];
            ld(U, out, Uo);
            prim(Uo, ro, ph.gamma);
#pragma unroll
            for (int k = 0; k < 4; k++) r[k] = 0.5 * (rc[k] + ro[k]);
        }
        const V4<double> eg = m.egeo[e];
        const double L = eg.z;
        alpha_c += (fabs(r[1] * eg.x + r[2] * eg.y) + sqrt(ph.gamma * r[3] / r[0])) * L;
        if (ph.equation == FVMR_EQ_NS) {
            const double T = r[3] / ph.R / r[0];
            alpha_c += fmax(4.0 / 3.0, ph.gamma / ph.Pr) * mu_of_T(ph, T) / r[0] * L * L / cg.z;
        }
    }
    dtl[c] = ph.bc.CFL * cg.z / alpha_c;
}

// One pseudo-time stage with per-cell dt (k_stage's residual, one thread per cell):
// r = -(1/V) sum_j s_j F_j; Uout = Uin + dt_c r (first) or a Un + b (Uin + dt_c r).
// Uout NULL: residual only.
template <bool NS>
__global__ __launch_bounds__(kBlock) void k_steady_stage(Mesh<double> m, Phys ph, const S4 *__restrict__ Uin,
                                                          const S4 *__restrict__ Un, S4 *__restrict__ Uout,
                                                          const Grad<double> *__restrict__ grad,
                                                          const double *__restrict__ dtl, double a, double b,
                                                          int first, S4 *__restrict__ rhs_out) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    const int nc = m.nc;
    if (c >= nc) return;
    double f[3][4];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const int2 sl = m.cslot[j * nc + c];
        const bool cR = (sl.x & kSlotCellIsR) != 0;
        double F[4];
        edge_flux<double, double, NS>(m, ph, Uin, grad, 0.0, sl.x & kSlotEdgeMask, cR ? sl.y : c, cR ? c : sl.y,
                                      (sl.x & kSlotBoundary) != 0, F);
        const double s = cR ? -1.0 : 1.0;
#pragma unroll
        for (int k = 0; k < 4; k++) f[j][k] = s * F[k];
    }
    const double w = -1.0 * m.cgeo[c].z;
    double r[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        double acc = 0.0;
        acc += f[0][k];
        acc += f[1][k];
        acc += f[2][k];
        r[k] = fdiv(acc, w);
    }
    if (rhs_out) st(rhs_out, c, r);
    if (!Uout) return;
    double Ui[4], Ub[4] = {0, 0, 0, 0}, Unew[4];
    ld(Uin, c, Ui);
    if (!first) ld(Un, c, Ub);
    const double dt = dtl[c];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const double s = Ui[k] + dt * r[k];
        Unew[k] = first ? s : a * Ub[k] + b * s;
    }
    st_stream(Uout, c, Unew);
}

// ---------------------------------------------------------------------------
// Reverse mode of the face physics (double values).  Each function recomputes the forward
// values of its template counterpart and accumulates the adjoints of its inputs from the
// adjoints of its outputs, statement by statement in reverse order; |x| takes the sign the
// forward-mode fabs uses (+1 at 0), branches follow the values.
// ---------------------------------------------------------------------------
// Adjoints of a FaceState's primitives (set_prim's outputs).
struct PrimBar {
    double r, u, v, p, ir;
};

__device__ __forceinline__ double sgn1(double x) { return x < 0.0 ? -1.0 : 1.0; }

// set_prim: ir = 1/U0, r = U0, u = U1 ir, v = U2 ir, p = r (gamma-1) (U3 ir - (u u + v v) / 2)
__device__ __forceinline__ void set_prim_adj(const FaceState<double> &s, double gamma, PrimBar b, double Ub[4]) {
    const double ga1 = gamma - 1.0;
    const double q = s.U[3] * s.ir - (s.u * s.u + s.v * s.v) * 0.5;
    b.r += b.p * ga1 * q;
    const double qb = b.p * s.r * ga1;
    Ub[3] += qb * s.ir;
    b.ir += qb * s.U[3];
    b.u -= qb * s.u;
    b.v -= qb * s.v;
    Ub[1] += b.u * s.ir;
    b.ir += b.u * s.U[1];
    Ub[2] += b.v * s.ir;
    b.ir += b.v * s.U[2];
    Ub[0] += b.r - b.ir * s.ir * s.ir;
}

// flux_roe, reverse: mu = adjoint of the four flux components
__device__ void roe_adj(const FaceState<double> &L, const FaceState<double> &R, double nx, double ny, double area,
                        double gamma, double g_ga1, double inv_g, const double mu[4], PrimBar &Lb, PrimBar &Rb) {
    const double rL = L.r, uL = L.u, vL = L.v, pL = L.p, rR = R.r, uR = R.u, vR = R.v, pR = R.p;
    const double ul = nx * uL + ny * vL, ur = nx * uR + ny * vR;
    const double rul = rL * ul, rur = rR * ur;
    const double V2L = uL * uL + vL * vL, V2R = uR * uR + vR * vR;
    const double ga1 = gamma - 1.0;
    const double HL = pL * g_ga1 * L.ir + 0.5 * V2L, HR = pR * g_ga1 * R.ir + 0.5 * V2R;
    const double rrorl = sqrt(rR * L.ir);
    const double inv_p1 = 1.0 / (rrorl + 1.0);
    const double rm = rL * rrorl;
    const double um = (uL + uR * rrorl) * inv_p1, vm = (vL + vR * rrorl) * inv_p1, Hm = (HL + HR * rrorl) * inv_p1;
    const double Vm2 = um * um + vm * vm;
    const double am = sqrt(ga1 * (Hm - 0.5 * Vm2));
    const double inv_am = 1.0 / am;
    const double unm = um * nx + vm * ny;
    const double eig1 = fabs(unm), eig2 = fabs(unm + am), eig3 = fabs(unm - am);
    const double ER = HR - pR * R.ir, EL = HL - pL * L.ir;
    const double Etm = Hm * inv_g + ga1 * inv_g * 0.5 * Vm2;
    const double dW0 = rR - rL;
    const double dW1 = rm * (uR - uL) + dW0 * um, dW2 = rm * (vR - vL) + dW0 * vm, dW3 = rm * (ER - EL) + dW0 * Etm;
    const double astar = (eig2 + eig3) * 0.5;
    const double Mstar = (eig2 - eig3) * 0.5 * inv_am;
    const double dun = nx * (uR - uL) + ny * (vR - vL);
    const double dp = pR - pL;
    const double B = (astar - eig1) * dp / rm, Cq = inv_am * inv_am;
    const double dUroe = Mstar * dun + B * Cq;
    const double dProe = Mstar * dp + (astar - eig1) * rm * dun;
    // central part 0.5 (F_L + F_R) area
    const double h = 0.5 * area;
    const double rulb = h * (mu[0] + mu[1] * uL + mu[2] * vL + mu[3] * HL);
    const double rurb = h * (mu[0] + mu[1] * uR + mu[2] * vR + mu[3] * HR);
    double uLb = h * mu[1] * rul, vLb = h * mu[2] * rul, HLb = h * mu[3] * rul;
    double uRb = h * mu[1] * rur, vRb = h * mu[2] * rur, HRb = h * mu[3] * rur;
    const double psb = h * (mu[1] * nx + mu[2] * ny);
    double pLb = psb, pRb = psb, rLb = 0.0, rRb = 0.0, irLb = 0.0, irRb = 0.0;
    // dissipation -0.5 (...) area
    const double m0 = -h * mu[0], m1 = -h * mu[1], m2 = -h * mu[2], m3 = -h * mu[3];
    double eig1b = m0 * dW0 + m1 * dW1 + m2 * dW2 + m3 * dW3;
    double dW0b = m0 * eig1;
    const double dW1b = m1 * eig1, dW2b = m2 * eig1, dW3b = m3 * eig1;
    const double dUroeb = (m0 + m1 * um + m2 * vm + m3 * Hm) * rm;
    double rmb = (m0 + m1 * um + m2 * vm + m3 * Hm) * dUroe;
    double umb = m1 * dUroe * rm, vmb = m2 * dUroe * rm, Hmb = m3 * dUroe * rm;
    const double dProeb = m1 * nx + m2 * ny + m3 * unm;
    double unmb = m3 * dProe;
    // dProe = Mstar dp + (astar - eig1) rm dun
    double Mstarb = dProeb * dp, dpb = dProeb * Mstar, astarb = dProeb * rm * dun;
    eig1b -= dProeb * rm * dun;
    rmb += dProeb * (astar - eig1) * dun;
    double dunb = dProeb * (astar - eig1) * rm;
    // dUroe = Mstar dun + B Cq, B = (astar - eig1) dp / rm, Cq = inv_am^2
    Mstarb += dUroeb * dun;
    dunb += dUroeb * Mstar;
    const double Bb = dUroeb * Cq;
    double inv_amb = dUroeb * B * 2.0 * inv_am;
    const double Ab = Bb / rm;
    rmb -= Bb * B / rm;
    astarb += Ab * dp;
    eig1b -= Ab * dp;
    dpb += Ab * (astar - eig1);
    pRb += dpb;
    pLb -= dpb;
    uRb += dunb * nx; uLb -= dunb * nx; vRb += dunb * ny; vLb -= dunb * ny;
    // Mstar, astar
    const double eig2b = 0.5 * inv_am * Mstarb + 0.5 * astarb;
    const double eig3b = -0.5 * inv_am * Mstarb + 0.5 * astarb;
    inv_amb += Mstarb * (eig2 - eig3) * 0.5;
    // dW3, dW2, dW1, dW0
    rmb += dW3b * (ER - EL) + dW2b * (vR - vL) + dW1b * (uR - uL);
    const double ERb = dW3b * rm, ELb = -dW3b * rm;
    dW0b += dW3b * Etm + dW2b * vm + dW1b * um;
    const double Etmb = dW3b * dW0;
    vRb += dW2b * rm; vLb -= dW2b * rm; vmb += dW2b * dW0;
    uRb += dW1b * rm; uLb -= dW1b * rm; umb += dW1b * dW0;
    rRb += dW0b;
    rLb -= dW0b;
    // Etm, ER, EL
    Hmb += Etmb * inv_g;
    double Vm2b = Etmb * ga1 * inv_g * 0.5;
    HRb += ERb; pRb -= ERb * R.ir; irRb -= ERb * pR;
    HLb += ELb; pLb -= ELb * L.ir; irLb -= ELb * pL;
    // eigenvalues, unm
    unmb += eig1b * sgn1(unm) + eig2b * sgn1(unm + am) + eig3b * sgn1(unm - am);
    double amb = eig2b * sgn1(unm + am) - eig3b * sgn1(unm - am);
    umb += unmb * nx;
    vmb += unmb * ny;
    // inv_am, am = sqrt(ga1 (Hm - Vm2 / 2))
    amb -= inv_amb * inv_am * inv_am;
    const double am2b = amb * 0.5 / am;
    Hmb += am2b * ga1;
    Vm2b -= am2b * ga1 * 0.5;
    umb += Vm2b * 2.0 * um;
    vmb += Vm2b * 2.0 * vm;
    // Roe averages
    double rrorlb = 0.0, inv_p1b = 0.0;
    uLb += umb * inv_p1; uRb += umb * rrorl * inv_p1; rrorlb += umb * uR * inv_p1; inv_p1b += umb * (uL + uR * rrorl);
    vLb += vmb * inv_p1; vRb += vmb * rrorl * inv_p1; rrorlb += vmb * vR * inv_p1; inv_p1b += vmb * (vL + vR * rrorl);
    HLb += Hmb * inv_p1; HRb += Hmb * rrorl * inv_p1; rrorlb += Hmb * HR * inv_p1; inv_p1b += Hmb * (HL + HR * rrorl);
    rLb += rmb * rrorl;
    rrorlb += rmb * rL;
    rrorlb -= inv_p1b * inv_p1 * inv_p1;
    const double sb = rrorlb * 0.5 / rrorl;
    rRb += sb * L.ir;
    irLb += sb * rR;
    // total enthalpies
    pLb += HLb * g_ga1 * L.ir; irLb += HLb * pL * g_ga1;
    pRb += HRb * g_ga1 * R.ir; irRb += HRb * pR * g_ga1;
    uLb += HLb * uL; vLb += HLb * vL;
    uRb += HRb * uR; vRb += HRb * vR;
    // mass fluxes
    rLb += rulb * ul;
    rRb += rurb * ur;
    const double ulb = rulb * rL, urb = rurb * rR;
    uLb += ulb * nx; vLb += ulb * ny;
    uRb += urb * nx; vRb += urb * ny;
    Lb.r += rLb; Lb.u += uLb; Lb.v += vLb; Lb.p += pLb; Lb.ir += irLb;
    Rb.r += rRb; Rb.u += uRb; Rb.v += vRb; Rb.p += pRb; Rb.ir += irRb;
}

// flux_llf, reverse; the dissipation's conserved jump goes to ULc / URc directly
__device__ void llf_adj(const FaceState<double> &L, const FaceState<double> &R, double nx, double ny, double len,
                        double gamma, double g_ga1, const double mu[4], PrimBar &Lb, PrimBar &Rb, double ULc[4],
                        double URc[4]) {
    const double unL = nx * L.u + ny * L.v, unR = nx * R.u + ny * R.v;
    const double cL = sqrt(gamma * L.p * L.ir), cR = sqrt(gamma * R.p * R.ir);
    const double aL = fabs(unL) + cL, aR = fabs(unR) + cR;
    const double lam = (aL > aR) ? aL : aR;
    const double h = 0.5 * len;
    double lamb = 0.0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        lamb -= h * mu[k] * (R.U[k] - L.U[k]);
        URc[k] -= h * mu[k] * lam;
        ULc[k] += h * mu[k] * lam;
    }
    const double aLb = (aL > aR) ? lamb : 0.0, aRb = (aL > aR) ? 0.0 : lamb;
    auto side = [&](const FaceState<double> &S, double un, double c, double ab, PrimBar &Sb) {
        const double hs = S.p * g_ga1 * S.ir + 0.5 * (S.u * S.u + S.v * S.v);
        const double t = S.r * un;
        const double tb = h * (mu[0] + mu[1] * S.u + mu[2] * S.v + mu[3] * hs);
        const double hsb = h * mu[3] * t;
        double unb = tb * S.r + ab * sgn1(un);
        Sb.r += tb * un;
        Sb.u += h * mu[1] * t + hsb * S.u;
        Sb.v += h * mu[2] * t + hsb * S.v;
        Sb.p += h * (mu[1] * nx + mu[2] * ny) + hsb * g_ga1 * S.ir;
        Sb.ir += hsb * S.p * g_ga1;
        const double wb = ab * 0.5 / c;
        Sb.p += wb * gamma * S.ir;
        Sb.ir += wb * gamma * S.p;
        Sb.u += unb * nx;
        Sb.v += unb * ny;
    };
    side(L, unL, cL, aLb, Lb);
    side(R, unR, cR, aRb, Rb);
}

// ghost_farfield, reverse: Rc = adjoint of the ghost's conserved state; into L's primitives
__device__ void farfield_adj(const FaceState<double> &L, const double *ruvp_inf, double nx, double ny, double gamma,
                             const double Rc[4], PrimBar &Lb) {
    const double ga1 = gamma - 1;
    const double rho_f = ruvp_inf[0], u_f = ruvp_inf[1], v_f = ruvp_inf[2], p_f = ruvp_inf[3];
    const double rho_e = L.r, u_e = L.u, v_e = L.v, p_e = L.p;
    const double af = sqrt(gamma * p_f / rho_f);
    const double ae = sqrt(gamma * p_e / rho_e);
    const double qnf = u_f * nx + v_f * ny, qne = u_e * nx + v_e * ny;
    const double rf = qnf - 2.0 * af / ga1, re = qne + 2.0 * ae / ga1;
    const double qn = 0.5 * (re + rf), as = ga1 * (re - rf) / 4.0;
    if (qn <= -as) return;   // freestream state: no dependence on L
    double r0[4];
    const bool sub = qn < as, in = qn < 0;
    double s_p = 0, s_r = 0, son = 0, X = 0;
    if (sub) {
        s_p = in ? p_f : p_e;
        s_r = in ? rho_f : rho_e;
        const double s_u = in ? u_f : u_e, s_v = in ? v_f : v_e, s_qn = in ? qnf : qne;
        son = s_p / pow(s_r, gamma);
        X = as * as / son / gamma;
        r0[0] = pow(X, 1.0 / ga1);
        r0[1] = s_u - s_qn * nx + qn * nx;
        r0[2] = s_v - s_qn * ny + qn * ny;
        r0[3] = as * as * r0[0] / gamma;
    } else {
        r0[0] = rho_e; r0[1] = u_e; r0[2] = v_e; r0[3] = p_e;
    }
    // ruvp2u, reverse
    double r0b[4];
    r0b[0] = Rc[0] + Rc[1] * r0[1] + Rc[2] * r0[2] + Rc[3] * 0.5 * (r0[1] * r0[1] + r0[2] * r0[2]);
    r0b[1] = Rc[1] * r0[0] + Rc[3] * r0[0] * r0[1];
    r0b[2] = Rc[2] * r0[0] + Rc[3] * r0[0] * r0[2];
    r0b[3] = Rc[3] / ga1;
    if (!sub) {
        Lb.r += r0b[0]; Lb.u += r0b[1]; Lb.v += r0b[2]; Lb.p += r0b[3];
        return;
    }
    const double r00b = r0b[0] + r0b[3] * as * as / gamma;
    double asb = r0b[3] * 2.0 * as * r0[0] / gamma;
    double qnb = r0b[1] * nx + r0b[2] * ny;
    const double s_ub = r0b[1], s_vb = r0b[2], s_qnb = -r0b[1] * nx - r0b[2] * ny;
    const double Xb = r00b * (1.0 / ga1) * pow(X, 1.0 / ga1 - 1.0);
    asb += Xb * 2.0 * as / son / gamma;
    const double sonb = -Xb * as * as / (son * son) / gamma;
    double qneb = 0.0;
    if (!in) {
        const double srg = pow(s_r, gamma);
        Lb.p += sonb / srg;
        Lb.r -= sonb * s_p * gamma * pow(s_r, gamma - 1.0) / (srg * srg);
        Lb.u += s_ub;
        Lb.v += s_vb;
        qneb += s_qnb;
    }
    const double reb = 0.5 * qnb + asb * ga1 / 4.0;
    qneb += reb;
    const double wb = reb * 2.0 / ga1 * 0.5 / ae;
    Lb.p += wb * gamma / rho_e;
    Lb.r -= wb * gamma * p_e / (rho_e * rho_e);
    Lb.u += qneb * nx;
    Lb.v += qneb * ny;
}

// The adjoints of the two reconstructed face states of one face for the flux cotangent mu:
// ULb = (dF/dU_L^f)^T mu and URb = (dF/dU_R^f)^T mu (a boundary face's ghost folded into ULb).
__device__ void face_vjp(const Phys &ph, int bType, const double *ruvp, const double ULf[4], const double URf[4],
                         double nx, double ny, double len, const double mu[4], double ULb[4], double URb[4]) {
    const double gamma = ph.gamma;
    FaceState<double> L, R;
#pragma unroll
    for (int i = 0; i < 4; i++) L.U[i] = ULf[i];
    set_prim(L, gamma);
    if (bType == -1) {
#pragma unroll
        for (int i = 0; i < 4; i++) R.U[i] = URf[i];
        set_prim(R, gamma);
    } else {
        ghost_state(ph, bType, ruvp, L, R, nx, ny, gamma);
    }
    PrimBar Lb{0, 0, 0, 0, 0}, Rb{0, 0, 0, 0, 0};
    double ULc[4] = {0, 0, 0, 0}, URc[4] = {0, 0, 0, 0};
    if (ph.scheme == FVMR_SOL_LLF) llf_adj(L, R, nx, ny, len, gamma, ph.g_ga1, mu, Lb, Rb, ULc, URc);
    else roe_adj(L, R, nx, ny, len, gamma, ph.g_ga1, ph.inv_g, mu, Lb, Rb);
    set_prim_adj(R, gamma, Rb, URc);
    if (bType == -1) {
#pragma unroll
        for (int i = 0; i < 4; i++) URb[i] = URc[i];
    } else if (bType == FVMR_BC_SYMMETRY || bType == FVMR_BC_WALL_NONVISCOUS) {
        // R.U0 = L.U0, R.U3 = L.U3, R's velocity the reflection of L's
        const double unL = L.u * nx + L.v * ny, utL = L.u * ny - L.v * nx;
        const double unR = -unL, utR = utL;
        const double uxR = unR * nx + utR * ny, uyR = unR * ny - utR * nx;
        ULc[0] += URc[0] + URc[1] * uxR + URc[2] * uyR;
        ULc[3] += URc[3];
        const double uxRb = URc[1] * L.U[0], uyRb = URc[2] * L.U[0];
        const double unRb = uxRb * nx + uyRb * ny, utRb = uxRb * ny - uyRb * nx;
        Lb.u += -unRb * nx + utRb * ny;
        Lb.v += -unRb * ny - utRb * nx;
    } else if (bType == FVMR_BC_WALL_ADIABAT) {
        ULc[0] += URc[0];
        ULc[3] += URc[3];
    } else {
        farfield_adj(L, ruvp, nx, ny, gamma, URc, Lb);
    }
    set_prim_adj(L, gamma, Lb, ULc);
#pragma unroll
    for (int i = 0; i < 4; i++) ULb[i] = ULc[i];
}

// Part 0 of the product (one thread per face): the adjoints of the face's two reconstructed
// states, fadj[2 e] for L and fadj[2 e + 1] for R, from mu_e = lambda_L - lambda_R.
__global__ __launch_bounds__(kBlock) void k_adj_faces(Mesh<double> m, Phys ph, const S4 *__restrict__ U,
                                                       const Grad<double> *__restrict__ grad,
                                                       const S4 *__restrict__ psi, S4 *__restrict__ fadj) {
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= m.nf) return;
    const int4 ei = m.eint[e];
    const int Lc = ei.x, Rc = ei.y, bType = ei.z;
    const V4<double> eg = m.egeo[e];
    const V2<double> em = m.emid[e];
    const bool interior = bType == -1;
    // both sides are loaded before either is used (a boundary face reads its L cell twice), so
    // the R loads do not wait behind the L arithmetic
    const int Rs = interior ? Rc : Lc;
    const V4<double> gL = m.cgeo[Lc], gR = m.cgeo[Rs];
    const Grad<double> grL = grad[Lc], grR = grad[Rs];
    double UL[4], UR[4], pL[4], pR[4], ULf[4], URf[4] = {0, 0, 0, 0}, mu[4];
    ld(U, Lc, UL);
    ld(U, Rs, UR);
    ld(psi, Lc, pL);
    ld(psi, Rs, pR);
    recon_state(ph, UL, grL, gL.x, gL.y, em.x, em.y, ULf);
#pragma unroll
    for (int k = 0; k < 4; k++) mu[k] = -pL[k] / gL.z;
    if (interior) {
        recon_state(ph, UR, grR, gR.x, gR.y, em.x, em.y, URf);
#pragma unroll
        for (int k = 0; k < 4; k++) mu[k] += pR[k] / gR.z;
    }
    double ULb[4], URb[4];
    face_vjp(ph, bType, bc_state(ph, ph.bc, bType, em.x, em.y, 0.0), ULf, URf, eg.x, eg.y, eg.z, mu, ULb, URb);
    st(fadj, 2 * e, ULb);
    if (interior) st(fadj, 2 * e + 1, URb);
}

// Gradient slots of grad_limit for one conserved variable with its four derivative directions
// (0 = the cell's own state, 1 + i = the state across slot i).
struct SlotsU1 {
    double a[3][3];
    Dual<4> uo[3];
    double n[3][2];
    __device__ __forceinline__ double A(int i, int q) const { return a[i][q]; }
    __device__ __forceinline__ Dual<4> Uo(int i, int) const { return uo[i]; }
    __device__ __forceinline__ double N(int i, int q) const { return n[i][q]; }
};

// The product with (dr/dU)^T, part 1 (one thread per cell c): the adjoints of c's three face
// states (k_adj_faces) pass through the reconstruction to U_c and to c's limited gradient, whose
// derivatives in U_c and the three states across (grad_limit on Dual<4>, one conserved variable
// at a time: the variables do not mix) send them on: base[c] = the terms that land on U_c,
// wedge[2 e + side] = the terms c sends to the cell across its edge e (side 1 when c is the
// edge's R).  With gbar, also the adjoint of c's limited gradient.
#ifndef FVMR_ADJ_WAVES
#define FVMR_ADJ_WAVES 2   // minimum waves per SIMD k_adj_vjp is compiled for (3 spills: slower)
#endif
template <bool GBAR>
__global__ __launch_bounds__(kAdjBlock, FVMR_ADJ_WAVES) void k_adj_vjp(Mesh<double> m, Phys ph, const S4 *__restrict__ U,
                                                        const Grad<double> *__restrict__ grad,
                                                        const S4 *__restrict__ fadj, S4 *__restrict__ base,
                                                        S4 *__restrict__ wedge, Grad<double> *__restrict__ gbar) {
    using D = Dual<4>;
    const int c = blockIdx.x * kAdjBlock + threadIdx.x;
    const int nc = m.nc;
    if (c >= nc) return;
    const V4<double> cg = m.cgeo[c];
    double Uc[4];
    ld(U, c, Uc);
    const Grad<double> gc = grad[c];
    double ub[4] = {0, 0, 0, 0}, gbx[4] = {0, 0, 0, 0}, gby[4] = {0, 0, 0, 0};
    int outs[3], edges[3];
    bool isR[3];
    // the three slots' records first, then their midpoints and face adjoints, before any is used
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const int2 sl = m.cslot[j * nc + c];
        outs[j] = sl.y;
        edges[j] = sl.x & kSlotEdgeMask;
        isR[j] = (sl.x & kSlotCellIsR) != 0;
    }
    V2<double> ems[3];
    double u4s[3][4];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        ems[j] = m.emid[edges[j]];
        ld(fadj, 2 * edges[j] + (isR[j] ? 1 : 0), u4s[j]);   // the adjoint of this cell's face state on e
    }
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const V2<double> em = ems[j];
        double Uf[4];
        const double *u4 = u4s[j];
        const bool lin = recon_state(ph, Uc, gc, cg.x, cg.y, em.x, em.y, Uf);
        const double dx = em.x - cg.x, dy = em.y - cg.y;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            ub[i] += u4[i];
            if (lin) {
                gbx[i] += u4[i] * dx;
                gby[i] += u4[i] * dy;
            }
        }
    }
    double self[4] = {0, 0, 0, 0}, w[3][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
    if (ph.reconstruct == FVMR_REC_LINEAR) {
        double Uo[3][4];
        SlotsU1 S;
        int fl[3];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int out = outs[j];
            const V4<double> eg = m.egeo[edges[j]];
            const V2<double> em = m.emid[edges[j]];
            ld(U, out >= 0 ? out : c, Uo[j]);
            const V2<double> nd = m.nxy[m.cnode[j * nc + c]];
            S.n[j][0] = nd.x - cg.x;
            S.n[j][1] = nd.y - cg.y;
            double Fx = 0.0, Fy = 0.0;
            if (out >= 0) {
                const V4<double> og = m.cgeo[out];
                Fx = og.x;
                Fy = og.y;
            }
            const double sgn = isR[j] ? -1.0 : 1.0;
            slot_coef<double>(ph, out >= 0, cg.x, cg.y, em.x, em.y, Fx, Fy, eg.x * eg.z * sgn, eg.y * eg.z * sgn, S.a[j]);
            fl[j] = (out >= 0) ? 3 : 0;
        }
#pragma unroll 1
        for (int k = 0; k < 4; k++) {
#pragma unroll
            for (int j = 0; j < 3; j++) S.uo[j] = D::seed(Uo[j][k], 1 + j);
            const D Ud[1] = {D::seed(Uc[k], 0)};
            D gx[1], gy[1];
            grad_limit<D, SlotsU1, 1, double>(ph, Ud, S, fl, cg.z, gx, gy);
            self[k] = gbx[k] * gx[0].d[0] + gby[k] * gy[0].d[0];
#pragma unroll
            for (int j = 0; j < 3; j++) w[j][k] = gbx[k] * gx[0].d[1 + j] + gby[k] * gy[0].d[1 + j];
        }
    }
    base[c] = S4{ub[0] + self[0], ub[1] + self[1], ub[2] + self[2], ub[3] + self[3]};
#pragma unroll
    for (int j = 0; j < 3; j++)
        if (outs[j] >= 0) wedge[2 * edges[j] + (isR[j] ? 1 : 0)] = S4{w[j][0], w[j][1], w[j][2], w[j][3]};
    if (GBAR) {
        Grad<double> o;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            o.x[k] = gbx[k];
            o.y[k] = gby[k];
        }
        gbar[c] = o;
    }
}

// Part 2 (one thread per cell): v = base[c] + the wedge terms of c's neighbours in slot order
// (+ g[c]).  Without dtl: out = v.  With dtl, one pseudo-time stage: s = psi_in + dt_c v,
// psi_out = s (first) or a psi_n + b s.  res (optional) = v.
__global__ __launch_bounds__(kBlock) void k_adj_update(Mesh<double> m, const S4 *__restrict__ base,
                                                        const S4 *__restrict__ wedge, const S4 *__restrict__ g,
                                                        const S4 *__restrict__ pin, const S4 *__restrict__ pn,
                                                        S4 *__restrict__ pout, const double *__restrict__ dtl,
                                                        double a, double b, int first, S4 *__restrict__ res) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    const int nc = m.nc;
    if (c >= nc) return;
    // the three shares are loaded before any is used (every slot has an edge, so the index is in
    // range; a share counts only where a cell lies across): no load waits for another
    int2 sl[3];
#pragma unroll
    for (int j = 0; j < 3; j++) sl[j] = m.cslot[j * nc + c];
    double q[3][4];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const int e = sl[j].x & kSlotEdgeMask;
        const bool cR = (sl[j].x & kSlotCellIsR) != 0;
        ld(wedge, 2 * e + (cR ? 0 : 1), q[j]);
    }
    double v[4];
    ld(base, c, v);
#pragma unroll
    for (int j = 0; j < 3; j++)
        if (sl[j].y >= 0) {
#pragma unroll
            for (int k = 0; k < 4; k++) v[k] += q[j][k];
        }
    if (g) {
        double q[4];
        ld(g, c, q);
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] += q[k];
    }
    if (res) st(res, c, v);
    if (!dtl) {
        st(pout, c, v);
        return;
    }
    const double dt = dtl[c];
    double pi[4], pb[4] = {0, 0, 0, 0}, o[4];
    ld(pin, c, pi);
    if (!first) ld(pn, c, pb);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const double s = pi[k] + dt * v[k];
        o[k] = first ? s : a * pb[k] + b * s;
    }
    st(pout, c, o);
}

// dJ/dU of J = sum_v W_v Cp_v with Cp_v = (p_v - p_inf) / p_dyn and p_v the mean of the
// pressures of the node's cells (k_node_field): g_c = dp/dU_c sum over c's nodes W_v / n_v.
// wnode[v] = W_v / p_dyn.
__global__ __launch_bounds__(kBlock) void k_adj_dJdU(int nc, const int *__restrict__ cnode,
                                                      const int *__restrict__ node_off,
                                                      const double *__restrict__ wnode, const S4 *__restrict__ U,
                                                      double gamma, S4 *__restrict__ g) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= nc) return;
    double ws = 0.0;
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const int v = cnode[j * nc + c];
        const double w = wnode[v];
        if (w != 0.0) ws += w / (double)(node_off[v + 1] - node_off[v]);
    }
    double u[4];
    ld(U, c, u);
    const double ga1 = gamma - 1.0, vx = u[1] / u[0], vy = u[2] / u[0];
    const double o[4] = {ws * ga1 * 0.5 * (vx * vx + vy * vy), -ws * ga1 * vx, -ws * ga1 * vy, ws * ga1};
    st(g, c, o);
}

__device__ __forceinline__ bool uses_freestream(int bType) {
    return bType != -1 && !is_periodic(bType) && bType != FVMR_BC_SYMMETRY && bType != FVMR_BC_WALL_NONVISCOUS &&
           bType != FVMR_BC_WALL_ADIABAT && bType != FVMR_BC_WALL_ISOTHERMAL && bType != FVMR_BC_INLET &&
           bType != FVMR_BC_OUTLET && bType != FVMR_BC_DSR;
}

// psi . dr/dalpha: the freestream velocity turns with alpha, d(u_inf, v_inf)/dalpha =
// (-v_inf, u_inf); only faces against the freestream depend on it.  Block sums in partial.
__global__ __launch_bounds__(kAdjBlock) void k_adj_alpha(Mesh<double> m, Phys ph, const S4 *__restrict__ U,
                                                          const Grad<double> *__restrict__ grad,
                                                          const S4 *__restrict__ psi, double *__restrict__ partial) {
    using D = Dual<1>;
    __shared__ double sred[kAdjBlock / 64];
    const int c = blockIdx.x * kAdjBlock + threadIdx.x;
    const int nc = m.nc;
    double acc = 0.0;
    if (c < nc) {
        const V4<double> cg = m.cgeo[c];
        double Uc[4], lc[4];
        ld(U, c, Uc);
        ld(psi, c, lc);
        const Grad<double> gc = grad[c];
        const D gam(ph.gamma);
        D inf[4] = {D(ph.bc.inf[0]), D(ph.bc.inf[1]), D(ph.bc.inf[2]), D(ph.bc.inf[3])};
        inf[1].d[0] = -ph.bc.inf[2];
        inf[2].d[0] = ph.bc.inf[1];
        for (int j = 0; j < 3; j++) {
            const int2 sl = m.cslot[j * nc + c];
            if (!(sl.x & kSlotBoundary)) continue;
            const int e = sl.x & kSlotEdgeMask;
            const int bType = m.eint[e].z;
            if (!uses_freestream(bType)) continue;
            const V4<double> eg = m.egeo[e];
            const V2<double> em = m.emid[e];
            double Uf[4];
            recon_state(ph, Uc, gc, cg.x, cg.y, em.x, em.y, Uf);
            FaceState<D> L, R;
#pragma unroll
            for (int i = 0; i < 4; i++) L.U[i] = D(Uf[i]);
            set_prim(L, gam);
            ghost_state(ph, bType, inf, L, R, D(eg.x), D(eg.y), gam);
            D F[4];
            riemann_flux(ph, L, R, D(eg.x), D(eg.y), D(eg.z), F);
#pragma unroll
            for (int k = 0; k < 4; k++) acc += (-lc[k] / cg.z) * F[k].d[0];
        }
    }
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) sred[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = sred[0];
        for (int i = 1; i < kAdjBlock / 64; i++) s += sred[i];
        partial[blockIdx.x] = s;
    }
}

// Gradient slots with coordinate derivatives (geometry as duals, states constant).
template <int K>
struct SlotsX {
    Dual<K> a[3][3];
    double uo[3][4];
    Dual<K> n[3][2];
    __device__ __forceinline__ Dual<K> A(int i, int q) const { return a[i][q]; }
    __device__ __forceinline__ Dual<K> Uo(int i, int k) const { return Dual<K>(uo[i][k]); }
    __device__ __forceinline__ Dual<K> N(int i, int q) const { return n[i][q]; }
};

// psi . dr/dX per cell, as (node, d/dx, d/dy) entries: (1) the faces c owns as L (their normal,
// length and midpoint and the centroids in the reconstruction offsets: nodes a, b of the edge and
// the third nodes of L and R; Dual<8>); (2) c's limited gradient, contracted with its adjoint
// gbar (c's vertices and the third nodes of its neighbours; Dual<12>), with the volume term
// -(psi_c . r_c) / V_c dV_c/dX of the 1/V_c scaling.
__global__ __launch_bounds__(64) void k_adj_coords(Mesh<double> m, Phys ph, const S4 *__restrict__ U,
                                                   const Grad<double> *__restrict__ grad,
                                                   const S4 *__restrict__ psi, const Grad<double> *__restrict__ gbar,
                                                   const S4 *__restrict__ rhs, const int2 *__restrict__ enode,
                                                   int *__restrict__ xnode, double *__restrict__ xval) {
    const int c = blockIdx.x * 64 + threadIdx.x;
    const int nc = m.nc;
    if (c >= nc) return;
    int *rn = xnode + (size_t)kXRec * c;
    double *rv = xval + (size_t)2 * kXRec * c;
#pragma unroll 1
    for (int q = 0; q < kXRec; q++) {
        rn[q] = -1;
        rv[2 * q] = rv[2 * q + 1] = 0.0;
    }
    const V4<double> cg = m.cgeo[c];
    double Uc[4], lc[4];
    ld(U, c, Uc);
    ld(psi, c, lc);
    const double pr_c = lc[0], pr_1 = lc[1], pr_2 = lc[2], pr_3 = lc[3];
#pragma unroll
    for (int k = 0; k < 4; k++) lc[k] = -lc[k] / cg.z;
    const Grad<double> gc = grad[c];
    int vid[3];
#pragma unroll
    for (int j = 0; j < 3; j++) vid[j] = m.cnode[j * nc + c];
    // third node of the cell across slot j (-1 none)
    int opp[3];
    int2 sls[3];
#pragma unroll 1
    for (int j = 0; j < 3; j++) {
        sls[j] = m.cslot[j * nc + c];
        const int out = sls[j].y;
        opp[j] = -1;
        if (out >= 0) {
            const int2 en = enode[sls[j].x & kSlotEdgeMask];
            for (int q = 0; q < 3; q++) {
                const int v = m.cnode[q * nc + out];
                if (v != en.x && v != en.y) opp[j] = v;
            }
        }
    }
    // (1) faces of c as L
    {
        using D = Dual<8>;
        const D gam(ph.gamma);
#pragma unroll 1
        for (int j = 0; j < 3; j++) {
            const int2 sl = sls[j];
            if (sl.x & kSlotCellIsR) continue;
            const int e = sl.x & kSlotEdgeMask, out = sl.y;
            const int bType = (sl.x & kSlotBoundary) ? m.eint[e].z : -1;
            const int2 en = enode[e];
            int oL = -1;
            for (int q = 0; q < 3; q++)
                if (vid[q] != en.x && vid[q] != en.y) oL = vid[q];
            const int oR = (out >= 0) ? opp[j] : -1;
            if (oL < 0 || (out >= 0 && oR < 0)) continue;   // not a triangle of the edge's nodes
            const V2<double> pa = m.nxy[en.x], pb = m.nxy[en.y], pL = m.nxy[oL];
            const V2<double> pR = (oR >= 0) ? m.nxy[oR] : V2<double>{0.0, 0.0};
            const D xa = D::seed(pa.x, 0), ya = D::seed(pa.y, 1), xb = D::seed(pb.x, 2), yb = D::seed(pb.y, 3);
            const D xo = D::seed(pL.x, 4), yo = D::seed(pL.y, 5), xr = D::seed(pR.x, 6), yr = D::seed(pR.y, 7);
            const D ddx = xb - xa, ddy = yb - ya;
            const D len = sqrt(ddx * ddx + ddy * ddy);
            const D nx = ddy / len, ny = -ddx / len;
            const D xm = (xa + xb) * 0.5, ym = (ya + yb) * 0.5;
            const D cxL = ((xa + xb) + xo) / 3.0, cyL = ((ya + yb) + yo) / 3.0;
            const V2<double> emv = m.emid[e];
            double Uf[4];
            const bool lin = recon_state(ph, Uc, gc, cg.x, cg.y, emv.x, emv.y, Uf);
            FaceState<D> L, R;
            const D dxL = xm - cxL, dyL = ym - cyL;
#pragma unroll
            for (int i = 0; i < 4; i++) L.U[i] = lin ? Uc[i] + gc.x[i] * dxL + gc.y[i] * dyL : D(Uc[i]);
            set_prim(L, gam);
            double mu[4];
            if (bType == -1) {
                double Uo[4], Ufo[4], po[4];
                ld(U, out, Uo);
                ld(psi, out, po);
                const V4<double> og = m.cgeo[out];
                const Grad<double> go = grad[out];
                const bool lo = recon_state(ph, Uo, go, og.x, og.y, emv.x, emv.y, Ufo);
                const D cxR = ((xa + xb) + xr) / 3.0, cyR = ((ya + yb) + yr) / 3.0;
                const D dxR = xm - cxR, dyR = ym - cyR;
#pragma unroll
                for (int i = 0; i < 4; i++) R.U[i] = lo ? Uo[i] + go.x[i] * dxR + go.y[i] * dyR : D(Uo[i]);
                set_prim(R, gam);
#pragma unroll
                for (int k = 0; k < 4; k++) mu[k] = lc[k] + po[k] / og.z;
            } else {
                ghost_state(ph, bType, bc_state(ph, ph.bc, bType, emv.x, emv.y, 0.0), L, R, nx, ny, gam);
#pragma unroll
                for (int k = 0; k < 4; k++) mu[k] = lc[k];
            }
            D F[4];
            riemann_flux(ph, L, R, nx, ny, len, F);
            double d[8];
#pragma unroll
            for (int s = 0; s < 8; s++) d[s] = mu[0] * F[0].d[s] + mu[1] * F[1].d[s] + mu[2] * F[2].d[s] + mu[3] * F[3].d[s];
            const int ids[4] = {en.x, en.y, oL, oR};
#pragma unroll
            for (int q = 0; q < 4; q++) {
                rn[6 + 4 * j + q] = ids[q];
                rv[2 * (6 + 4 * j + q)] = d[2 * q];
                rv[2 * (6 + 4 * j + q) + 1] = d[2 * q + 1];
            }
        }
    }
    // (2) the limited gradient of c and the volume scaling
    {
        using D = Dual<12>;
        D vx[3], vy[3], ox[3], oy[3];
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const V2<double> p = m.nxy[vid[q]];
            vx[q] = D::seed(p.x, 2 * q);
            vy[q] = D::seed(p.y, 2 * q + 1);
            const V2<double> o = (opp[q] >= 0) ? m.nxy[opp[q]] : V2<double>{0.0, 0.0};
            ox[q] = D::seed(o.x, 6 + 2 * q);
            oy[q] = D::seed(o.y, 7 + 2 * q);
        }
        const D xc = ((vx[0] + vx[1]) + vx[2]) / 3.0, yc = ((vy[0] + vy[1]) + vy[2]) / 3.0;
        const D vol = fabs(0.5 * ((vx[1] - vx[0]) * (vy[2] - vy[0]) - (vx[2] - vx[0]) * (vy[1] - vy[0])));
        double dv[12];
#pragma unroll
        for (int s = 0; s < 12; s++) dv[s] = 0.0;
        if (ph.reconstruct == FVMR_REC_LINEAR) {
            const Grad<double> gb = gbar[c];
            SlotsX<12> S;
            int fl[3];
#pragma unroll 1
            for (int j = 0; j < 3; j++) {
                const int out = sls[j].y;
                const int e = sls[j].x & kSlotEdgeMask;
                const int2 en = enode[e];
                int ia = 0, ib = 0;
                for (int q = 0; q < 3; q++) {
                    if (vid[q] == en.x) ia = q;
                    if (vid[q] == en.y) ib = q;
                }
                const D ddx = vx[ib] - vx[ia], ddy = vy[ib] - vy[ia];
                const D len = sqrt(ddx * ddx + ddy * ddy);
                const D nx = ddy / len, ny = -ddx / len;
                const D fx = (vx[ia] + vx[ib]) * 0.5, fy = (vy[ia] + vy[ib]) * 0.5;
                D Fx(0.0), Fy(0.0);
                if (out >= 0) {
                    Fx = ((vx[ia] + vx[ib]) + ox[j]) / 3.0;
                    Fy = ((vy[ia] + vy[ib]) + oy[j]) / 3.0;
                }
                const double sgn = (sls[j].x & kSlotCellIsR) ? -1.0 : 1.0;
                slot_coef<D>(ph, out >= 0, xc, yc, fx, fy, Fx, Fy, nx * len * sgn, ny * len * sgn, S.a[j]);
                double Uo[4];
                ld(U, out >= 0 ? out : c, Uo);
#pragma unroll
                for (int k = 0; k < 4; k++) S.uo[j][k] = Uo[k];
                S.n[j][0] = vx[j] - xc;
                S.n[j][1] = vy[j] - yc;
                fl[j] = (out >= 0) ? 3 : 0;
            }
            D Ud[4], gx[4], gy[4];
#pragma unroll
            for (int k = 0; k < 4; k++) Ud[k] = D(Uc[k]);
            grad_limit<D>(ph, Ud, S, fl, vol, gx, gy);
#pragma unroll 1
            for (int s = 0; s < 12; s++) {
                double acc = 0.0;
#pragma unroll
                for (int k = 0; k < 4; k++) acc += gb.x[k] * gx[k].d[s] + gb.y[k] * gy[k].d[s];
                dv[s] = acc;
            }
        }
        double r[4];
        ld(rhs, c, r);
        const double pv = -(pr_c * r[0] + pr_1 * r[1] + pr_2 * r[2] + pr_3 * r[3]) / cg.z;
#pragma unroll
        for (int s = 0; s < 6; s++) dv[s] += pv * vol.d[s];
#pragma unroll
        for (int q = 0; q < 3; q++) {
            rn[q] = vid[q];
            rv[2 * q] = dv[2 * q];
            rv[2 * q + 1] = dv[2 * q + 1];
            rn[3 + q] = opp[q];
            rv[2 * (3 + q)] = dv[6 + 2 * q];
            rv[2 * (3 + q) + 1] = dv[7 + 2 * q];
        }
    }
}

// ---------------------------------------------------------------------------
// Host
// ---------------------------------------------------------------------------
int adj_check(const fvmr_solver *h, bool adjoint) {
    if (h->mixed) return set_error(FVMR_ERR_UNSUPPORTED, "steady solve and adjoint need an FP64 handle");
    if (h->ncases != 1) return set_error(FVMR_ERR_UNSUPPORTED, "steady solve and adjoint need a one-case handle");
    if (!adjoint) return FVMR_OK;
    if (h->p.equation != FVMR_EQ_EULER) return set_error(FVMR_ERR_UNSUPPORTED, "adjoint: Euler only");
    if (h->bc_flags & kBcFlagPeriodic) return set_error(FVMR_ERR_UNSUPPORTED, "adjoint: periodic boundaries");
    if (h->bc_flags & kBcFlagDsr) return set_error(FVMR_ERR_UNSUPPORTED, "adjoint: double Mach reflection boundary");
    if (h->bc_flags & kBcFlagIsothermal) return set_error(FVMR_ERR_UNSUPPORTED, "adjoint: isothermal wall");
    return FVMR_OK;
}

int adj_alloc(fvmr_solver *h) {
    if (h->adj) return FVMR_OK;
    AdjWork *w = new AdjWork();
    h->adj = w;
    const size_t nc = h->nc, nf = h->nf, nv = h->nv;
    bool ok = true;
    for (auto &q : w->psi) ok = ok && hipMalloc((void **)&q, nc * sizeof(S4)) == hipSuccess;
    ok = ok && hipMalloc((void **)&w->base, nc * sizeof(S4)) == hipSuccess &&
         hipMalloc((void **)&w->wedge, 2 * nf * sizeof(S4)) == hipSuccess &&
         hipMalloc((void **)&w->fadj, 2 * nf * sizeof(S4)) == hipSuccess &&
         hipMalloc((void **)&w->g, nc * sizeof(S4)) == hipSuccess &&
         hipMalloc((void **)&w->res, nc * sizeof(S4)) == hipSuccess &&
         hipMalloc((void **)&w->dtl, nc * sizeof(double)) == hipSuccess &&
         hipMalloc((void **)&w->wnode, nv * sizeof(double)) == hipSuccess &&
         hipMalloc((void **)&w->partial, ((nc + kAdjBlock - 1) / kAdjBlock) * sizeof(double)) == hipSuccess;
    if (!ok) {
        adj_free(h);
        return set_error(FVMR_ERR_HIP, "hipMalloc adjoint workspace");
    }
    std::vector<int2> en(nf);
    for (size_t e = 0; e < nf; e++) en[e] = make_int2(h->enodes[2 * e], h->enodes[2 * e + 1]);
    int rc = upload(&w->enode, en, h->stream);
    if (!rc && hipStreamSynchronize(h->stream) != hipSuccess) rc = set_error(FVMR_ERR_HIP, "upload edge nodes");
    if (rc) adj_free(h);
    return rc;
}

int adj_grid(int n) { return (n + kAdjBlock - 1) / kAdjBlock; }

// Per-cell pseudo-time steps of U into the workspace (the tile kernel's form when the handle
// has tiles; both give the same values).
int local_dt(fvmr_solver *h, const S4 *U) {
    AdjWork *w = (AdjWork *)h->adj;
    if (h->tiles && h->shape == 0) {   // (other work-group shapes use the cell kernels here)
        TileArgs<double> t = tile_args<double>(h);
        t.dtl = w->dtl;
        if (h->ph.equation == FVMR_EQ_NS)
            k_tile_dt<double, true, true><<<h->ntiles, TileCfg<double>::kThreads, 0, h->stream>>>(t, h->ph, U, nullptr);
        else
            k_tile_dt<double, false, true><<<h->ntiles, TileCfg<double>::kThreads, 0, h->stream>>>(t, h->ph, U, nullptr);
    } else {
        k_local_dt<<<grid_of(h->nc), kBlock, 0, h->stream>>>(mesh_args<double>(h), h->ph, U, w->dtl);
    }
    HIP_OK(hipGetLastError());
    return FVMR_OK;
}

// Residual-only evaluation at U into rhs (gradient first).
int steady_rhs(fvmr_solver *h, const S4 *U, S4 *rhs) {
    int rc = launch_gradient<double, double>(h, U, -1);
    if (rc) return rc;
    const Mesh<double> m = mesh_args<double>(h);
    if (h->p.equation == FVMR_EQ_NS)
        k_steady_stage<true><<<grid_of(h->nc), kBlock, 0, h->stream>>>(m, h->ph, U, U, nullptr, (const Grad<double> *)h->grad,
                                                                      nullptr, 0, 0, 1, rhs);
    else
        k_steady_stage<false><<<grid_of(h->nc), kBlock, 0, h->stream>>>(m, h->ph, U, U, nullptr, (const Grad<double> *)h->grad,
                                                                       nullptr, 0, 0, 1, rhs);
    HIP_OK(hipGetLastError());
    return FVMR_OK;
}

int steady_stage(fvmr_solver *h, const S4 *Uin, const S4 *Un, S4 *Uout, double a, double b, int first, S4 *rhs) {
    const AdjWork *w = (const AdjWork *)h->adj;
    if (h->tiles && h->shape == 0) {   // the tile kernel with per-cell dt (the same residual as the cell kernels)
        TileArgs<double> t = tile_args<double>(h);
        t.dtl = w->dtl;
        if (h->ph.equation == FVMR_EQ_NS)
            k_tile<double, double, true, true><<<h->ntiles, TileCfg<double>::kThreads, 0, h->stream>>>(
                t, h->ph, Uin, Un, Uout, rhs, h->sc, h->bad_list, h->nc, a, b, first, 0);
        else
            k_tile<double, double, false, true><<<h->ntiles, TileCfg<double>::kThreads, 0, h->stream>>>(
                t, h->ph, Uin, Un, Uout, rhs, h->sc, h->bad_list, h->nc, a, b, first, 0);
        HIP_OK(hipGetLastError());
        return FVMR_OK;
    }
    int rc = launch_gradient<double, double>(h, Uin, -1);
    if (rc) return rc;
    const Mesh<double> m = mesh_args<double>(h);
    if (h->p.equation == FVMR_EQ_NS)
        k_steady_stage<true><<<grid_of(h->nc), kBlock, 0, h->stream>>>(m, h->ph, Uin, Un, Uout, (const Grad<double> *)h->grad,
                                                                      w->dtl, a, b, first, rhs);
    else
        k_steady_stage<false><<<grid_of(h->nc), kBlock, 0, h->stream>>>(m, h->ph, Uin, Un, Uout, (const Grad<double> *)h->grad,
                                                                       w->dtl, a, b, first, rhs);
    HIP_OK(hipGetLastError());
    return FVMR_OK;
}

// L2 norms of the 4 components of an S4 field (device -> host).
int field_norm2(fvmr_solver *h, const S4 *f, double out[4]) {
    k_norm<<<grid_of(h->nc), kBlock, 0, h->stream>>>(h->nc, f, FVMR_NORM_2, h->partial);
    HIP_OK(hipGetLastError());
    k_norm_finish<<<1, kBlock, 0, h->stream>>>(grid_of(h->nc), FVMR_NORM_2, h->partial, h->norm_out);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(out, h->norm_out, 4 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    return FVMR_OK;
}

// One product with (dr/dU)^T at U (gradients recomputed first): base/wedge, then the gather.
int adj_product(fvmr_solver *h, const S4 *U, const S4 *psi, bool with_gbar) {
    AdjWork *w = (AdjWork *)h->adj;
    if (!w->grad_ready) {   // the limited gradients of U (a primal quantity) once per solve
        const int rc = launch_gradient<double, double>(h, U, -1);
        if (rc) return rc;
        w->grad_ready = true;
    }
    const Mesh<double> m = mesh_args<double>(h);
    k_adj_faces<<<grid_of(h->nf), kBlock, 0, h->stream>>>(m, h->ph, U, (const Grad<double> *)h->grad, psi, w->fadj);
    HIP_OK(hipGetLastError());
    if (with_gbar) {
        if (!w->gbar && hipMalloc((void **)&w->gbar, (size_t)h->nc * sizeof(Grad<double>)) != hipSuccess)
            return set_error(FVMR_ERR_HIP, "hipMalloc adjoint gradient");
        k_adj_vjp<true><<<adj_grid(h->nc), kAdjBlock, 0, h->stream>>>(m, h->ph, U, (const Grad<double> *)h->grad,
                                                                                   w->fadj, w->base, w->wedge, w->gbar);
    } else {
        k_adj_vjp<false><<<adj_grid(h->nc), kAdjBlock, 0, h->stream>>>(m, h->ph, U, (const Grad<double> *)h->grad,
                                                                                    w->fadj, w->base, w->wedge, nullptr);
    }
    HIP_OK(hipGetLastError());
    return FVMR_OK;
}

int adj_update(fvmr_solver *h, const S4 *g, const S4 *pin, const S4 *pn, S4 *pout, bool step, double a, double b,
               int first, S4 *res) {
    AdjWork *w = (AdjWork *)h->adj;
    k_adj_update<<<grid_of(h->nc), kBlock, 0, h->stream>>>(mesh_args<double>(h), w->base, w->wedge, g, pin, pn, pout,
                                                           step ? w->dtl : nullptr, a, b, first, res);
    HIP_OK(hipGetLastError());
    return FVMR_OK;
}

// psi . dr/dalpha (block sums summed on the host in block order).
int adj_alpha(fvmr_solver *h, const S4 *U, const S4 *psi, double *out) {
    int rc = launch_gradient<double, double>(h, U, -1);
    if (rc) return rc;
    AdjWork *w = (AdjWork *)h->adj;
    const int nb = adj_grid(h->nc);
    k_adj_alpha<<<nb, kAdjBlock, 0, h->stream>>>(mesh_args<double>(h), h->ph, U, (const Grad<double> *)h->grad, psi,
                                                 w->partial);
    HIP_OK(hipGetLastError());
    std::vector<double> part(nb);
    HIP_OK(hipMemcpyAsync(part.data(), w->partial, nb * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    double s = 0.0;
    for (double v : part) s += v;
    *out = s;
    return FVMR_OK;
}

// psi . dr/dX into dxy (2 nv, interleaved x, y; accumulated in cell and entry order).
int adj_coords(fvmr_solver *h, const S4 *U, const S4 *psi, double *dxy) {
    AdjWork *w = (AdjWork *)h->adj;
    w->grad_ready = false;   // steady_rhs below recomputes h->grad
    const size_t nc = h->nc;
    if (!w->xnode && (hipMalloc((void **)&w->xnode, nc * kXRec * sizeof(int)) != hipSuccess ||
                      hipMalloc((void **)&w->xval, nc * kXRec * 2 * sizeof(double)) != hipSuccess))
        return set_error(FVMR_ERR_HIP, "hipMalloc coordinate product");
    int rc;
    if ((rc = steady_rhs(h, U, w->res))) return rc;      // r at U (volume term)
    if ((rc = adj_product(h, U, psi, true))) return rc;  // gbar; leaves the gradient of U in h->grad
    k_adj_coords<<<(int)((nc + 63) / 64), 64, 0, h->stream>>>(mesh_args<double>(h), h->ph, U, (const Grad<double> *)h->grad,
                                                              psi, w->gbar, w->res, w->enode, w->xnode, w->xval);
    HIP_OK(hipGetLastError());
    std::vector<int> nodes(nc * kXRec);
    std::vector<double> vals(nc * kXRec * 2);
    HIP_OK(hipMemcpyAsync(nodes.data(), w->xnode, nodes.size() * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipMemcpyAsync(vals.data(), w->xval, vals.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    std::fill(dxy, dxy + 2 * (size_t)h->nv, 0.0);
    for (size_t q = 0; q < nodes.size(); q++) {
        const int v = nodes[q];
        if (v < 0) continue;
        dxy[2 * (size_t)v] += vals[2 * q];
        dxy[2 * (size_t)v + 1] += vals[2 * q + 1];
    }
    return FVMR_OK;
}

// Host copy of a caller-order field into device order.
int upload_field(fvmr_solver *h, const double *const F[4], S4 *dev) {
    for (int k = 0; k < 4; k++)
        if (!F[k]) return set_error(FVMR_ERR_INVALID_ARG, "null field component");
    std::vector<S4> buf(h->nc);
    for (int i = 0; i < h->nc; i++) {
        const int r = h->perm[i];
        buf[i] = S4{F[0][r], F[1][r], F[2][r], F[3][r]};
    }
    HIP_OK(hipMemcpyAsync(dev, buf.data(), buf.size() * sizeof(S4), hipMemcpyHostToDevice, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    return FVMR_OK;
}

// Surface-force sets of J: boundary edges (caller ids) -> nodes and geometry, with validation.
struct JSets {
    std::vector<int> en;      // 2 per edge
    std::vector<double> eg;   // nx, ny, L per edge
};
int j_sets(const fvmr_solver *h, int nsets, const int *off, const int *edges, JSets &js) {
    if (nsets < 1 || !off || !edges) return set_error(FVMR_ERR_INVALID_ARG, "J needs at least one boundary set");
    if (off[0] != 0) return set_error(FVMR_ERR_INVALID_ARG, "set_offsets[0] must be 0");
    for (int s = 0; s < nsets; s++)
        if (off[s + 1] < off[s]) return set_error(FVMR_ERR_INVALID_ARG, "set_offsets must not decrease");
    for (int k = 0; k < off[nsets]; k++) {
        const auto it = std::lower_bound(h->bedge_ref.begin(), h->bedge_ref.end(), edges[k]);
        if (it == h->bedge_ref.end() || *it != edges[k])
            return set_error(FVMR_ERR_INVALID_ARG, "J: edge " + std::to_string(edges[k]) + " is not in a boundary set");
        const size_t b = it - h->bedge_ref.begin();
        js.en.push_back(h->bedge_nodes[2 * b]);
        js.en.push_back(h->bedge_nodes[2 * b + 1]);
        for (int q = 0; q < 3; q++) js.eg.push_back(h->bedge_geo[3 * b + q]);
    }
    return FVMR_OK;
}

// One pseudo-time iteration of the adjoint: out = Phi(x) with source g (NULL: the
// homogeneous map T x).  Stage buffers psi[2], psi[3]; res (optional) = (dr/dU)^T x + g.
int adj_iteration(fvmr_solver *h, const S4 *U, const S4 *g, const S4 *x, S4 *out, S4 *res) {
    AdjWork *w = (AdjWork *)h->adj;
    S4 *Pb = w->psi[2], *Pc = w->psi[3];
    int rc;
    if (h->p.time_advance == FVMR_EVO_RK3) {
        if ((rc = adj_product(h, U, x, false)) || (rc = adj_update(h, g, x, x, Pb, true, 0.0, 1.0, 1, res)) ||
            (rc = adj_product(h, U, Pb, false)) || (rc = adj_update(h, g, Pb, x, Pc, true, 0.75, 0.25, 0, nullptr)) ||
            (rc = adj_product(h, U, Pc, false)) ||
            (rc = adj_update(h, g, Pc, x, out, true, 1.0 / 3.0, 2.0 / 3.0, 0, nullptr)))
            return rc;
        return FVMR_OK;
    }
    if ((rc = adj_product(h, U, x, false)) || (rc = adj_update(h, g, x, x, out, true, 0.0, 1.0, 1, res))) return rc;
    return FVMR_OK;
}

}  // namespace

extern "C" {

int fvmr_set_boundary_states(fvmr_handle *h, const double inf[4], const double inlet[4], const double outlet[4]) {
    if (!h || !inf || !inlet || !outlet) return set_error(FVMR_ERR_INVALID_ARG, "null argument");
    if (h->ncases != 1) return set_error(FVMR_ERR_UNSUPPORTED, "fvmr_set_boundary_states: one-case handles only");
    for (int k = 0; k < 4; k++) {
        h->p.inf_ruvp[k] = h->ph.bc.inf[k] = inf[k];
        h->p.inlet_ruvp[k] = h->ph.bc.inlet[k] = inlet[k];
        h->p.outlet_ruvp[k] = h->ph.bc.outlet[k] = outlet[k];
    }
    HIP_OK(hipMemcpyAsync(h->cbc, &h->ph.bc, sizeof(CaseBC), hipMemcpyHostToDevice, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    free_graphs(h);
    return FVMR_OK;
}

int fvmr_steady_solve(fvmr_handle *h, int max_iters, double tol, int *iters, double res[2]) {
    if (!h || max_iters < 0 || !(tol >= 0)) return set_error(FVMR_ERR_INVALID_ARG, "invalid argument");
    int rc = adj_check(h, false);
    if (rc || (rc = adj_alloc(h))) return rc;
    AdjWork *w = (AdjWork *)h->adj;
    const Mesh<double> m = mesh_args<double>(h);
    double r0 = -1.0, r = -1.0, nrm[4];
    int it = 0;
    while (it < max_iters) {
        const int a = h->cur, b = (a + 1) % 3, c = (a + 2) % 3;
        S4 *Ua = h->Ubuf[a], *Ub = h->Ubuf[b], *Uc = h->Ubuf[c];
        const bool check = (it % kAdjCheck == 0);
        if ((rc = local_dt(h, Ua))) return rc;
        if ((rc = steady_stage(h, Ua, Ua, Ub, 0.0, 1.0, 1, check ? h->rhs : nullptr))) return rc;
        if (h->p.time_advance == FVMR_EVO_RK3) {
            if ((rc = steady_stage(h, Ub, Ua, Uc, 0.75, 0.25, 0, nullptr)) ||
                (rc = steady_stage(h, Uc, Ua, Ub, 1.0 / 3.0, 2.0 / 3.0, 0, nullptr)))
                return rc;
        }
        h->cur = b;
        it++;
        if (check) {
            if ((rc = field_norm2(h, h->rhs, nrm))) return rc;
            r = nrm[0];
            if (r0 < 0) r0 = r;
            if (!std::isfinite(r)) return set_error(FVMR_ERR_NUMERIC, "steady solve: the residual is not finite");
            if (r <= tol * r0) break;
        }
    }
    h->rhs_valid = false;
    free_graphs(h);
    if (iters) *iters = it;
    if (res) { res[0] = r0; res[1] = r; }
    return FVMR_OK;
}

int fvmr_adjoint_vjp(fvmr_handle *h, const double *const psi[4], double *const out[4], double *dalpha, double *dxy) {
    if (!h || !psi || !out) return set_error(FVMR_ERR_INVALID_ARG, "null argument");
    int rc = adj_check(h, true);
    if (rc || (rc = adj_alloc(h))) return rc;
    AdjWork *w = (AdjWork *)h->adj;
    const S4 *U = h->Ubuf[h->cur];
    w->grad_ready = false;
    if ((rc = upload_field(h, psi, w->psi[0]))) return rc;
    if ((rc = adj_product(h, U, w->psi[0], false))) return rc;
    if ((rc = adj_update(h, nullptr, nullptr, nullptr, w->psi[1], false, 0, 0, 1, nullptr))) return rc;
    if ((rc = read_back(h, w->psi[1], out))) return rc;
    if (dalpha && (rc = adj_alpha(h, U, w->psi[0], dalpha))) return rc;
    if (dxy && (rc = adj_coords(h, U, w->psi[0], dxy))) return rc;
    w->psi_valid = false;
    return FVMR_OK;
}

int fvmr_adjoint_solve(fvmr_handle *h, int nsets, const int *set_offsets, const int *set_edges, double S_ref,
                       int max_iters, double tol, double *dJ_dxy, fvmr_adjoint_result *out) {
    if (!h || !out || max_iters < 0 || !(tol >= 0) || !(S_ref != 0)) return set_error(FVMR_ERR_INVALID_ARG, "invalid argument");
    int rc = adj_check(h, true);
    if (rc || (rc = adj_alloc(h))) return rc;
    JSets js;
    if ((rc = j_sets(h, nsets, set_offsets, set_edges, js))) return rc;
    AdjWork *w = (AdjWork *)h->adj;
    const S4 *U = h->Ubuf[h->cur];
    w->grad_ready = false;
    const int nv = h->nv, ne = (int)js.en.size() / 2;
    // J and its explicit parts from the node Cp of U
    if ((rc = node_field_device(h))) return rc;
    std::vector<double> Cp(nv);
    HIP_OK(hipMemcpyAsync(Cp.data(), h->node_buf + 4 * (size_t)nv, nv * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    const double *inf = h->p.inf_ruvp;
    const double alpha = std::atan2(inf[2], inf[1]), sa = std::sin(alpha), ca = std::cos(alpha);
    const double p_dyn = 0.5 * inf[0] * (inf[1] * inf[1] + inf[2] * inf[2]);
    double Cx = 0.0, Cy = 0.0;
    std::vector<double> wn(nv, 0.0);
    for (int k = 0; k < ne; k++) {
        const int a = js.en[2 * k], b = js.en[2 * k + 1];
        const double nx = js.eg[3 * k], ny = js.eg[3 * k + 1], L = js.eg[3 * k + 2];
        const double cp = 0.5 * (Cp[a] + Cp[b]);
        Cx += cp * (1 / S_ref) * L * nx;
        Cy += cp * (1 / S_ref) * L * ny;
        const double wgt = 0.5 * (1 / S_ref) * L * (-sa * nx + ca * ny) / p_dyn;
        wn[a] += wgt;
        wn[b] += wgt;
    }
    out->J = -sa * Cx + ca * Cy;
    HIP_OK(hipMemcpyAsync(w->wnode, wn.data(), nv * sizeof(double), hipMemcpyHostToDevice, h->stream));
    k_adj_dJdU<<<grid_of(h->nc), kBlock, 0, h->stream>>>(h->nc, h->cnode, h->nc_off, w->wnode, U, h->p.gamma, w->g);
    HIP_OK(hipGetLastError());
    // pseudo-time iteration from psi = 0 with the primal's local dt at U
    k_local_dt<<<grid_of(h->nc), kBlock, 0, h->stream>>>(mesh_args<double>(h), h->ph, U, w->dtl);
    HIP_OK(hipGetLastError());
    double nrm[4], r0 = -1.0, r = -1.0;
    int it = 0, a = 0;
    HIP_OK(hipMemsetAsync(w->psi[0], 0, (size_t)h->nc * sizeof(S4), h->stream));
    while (it < max_iters) {
        const bool check = (it % kAdjCheck == 0);
        if ((rc = adj_iteration(h, U, w->g, w->psi[a], w->psi[1 - a], check ? w->res : nullptr))) return rc;
        a = 1 - a;
        it++;
        if (check) {
            if ((rc = field_norm2(h, w->res, nrm))) return rc;
            r = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2] + nrm[3] * nrm[3]);
            if (r0 < 0) r0 = r;
            if (!std::isfinite(r)) return set_error(FVMR_ERR_NUMERIC, "adjoint: the residual is not finite");
            if (r <= tol * r0) break;
        }
    }
    // the residual of the adjoint equation at the start (psi = 0: dJ/dU) and at the end
    if ((rc = field_norm2(h, w->g, nrm))) return rc;
    r0 = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2] + nrm[3] * nrm[3]);
    if ((rc = adj_product(h, U, w->psi[a], false)) ||
        (rc = adj_update(h, w->g, nullptr, nullptr, w->res, false, 0, 0, 1, nullptr)) ||
        (rc = field_norm2(h, w->res, nrm)))
        return rc;
    r = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2] + nrm[3] * nrm[3]);
    w->cur = a;
    w->psi_valid = true;
    double pa = 0.0;
    if ((rc = adj_alpha(h, U, w->psi[a], &pa))) return rc;
    out->dJ_dalpha = (-ca * Cx - sa * Cy) + pa;
    out->iters = it;
    out->res0 = r0;
    out->res = r;
    if (dJ_dxy) {
        if ((rc = adj_coords(h, U, w->psi[a], dJ_dxy))) return rc;
        // explicit part: J = sum cp/S_ref (-sin a dy - cos a dx), (dx, dy) = node b - node a
        for (int k = 0; k < ne; k++) {
            const int na = js.en[2 * k], nb = js.en[2 * k + 1];
            const double cp = 0.5 * (Cp[na] + Cp[nb]) / S_ref;
            dJ_dxy[2 * (size_t)na] += ca * cp;
            dJ_dxy[2 * (size_t)nb] -= ca * cp;
            dJ_dxy[2 * (size_t)na + 1] += sa * cp;
            dJ_dxy[2 * (size_t)nb + 1] -= sa * cp;
        }
    }
    h->rhs_valid = false;
    return FVMR_OK;
}

int fvmr_adjoint_state(fvmr_handle *h, double *const psi[4]) {
    if (!h || !psi) return set_error(FVMR_ERR_INVALID_ARG, "null argument");
    const AdjWork *w = (const AdjWork *)h->adj;
    if (!w || !w->psi_valid) return set_error(FVMR_ERR_INVALID_ARG, "no adjoint state: run fvmr_adjoint_solve first");
    return read_back(h, w->psi[w->cur], psi);
}

}  // extern "C"
