"""Shared case builders of the test-suite (meshes, parameters, initial fields)."""
import numpy as np

import cfd_solver_2d_gpu_amd as F


def vortex_case(n=64, gradient=F.GRA_GREENGAUSS, scheme=F.SOL_ROE, advance=F.EVO_RK3, limiter=F.LIM_BARTH, **kw):
    """C1 family: periodic [0,10]^2, n x n quads split, isentropic vortex chi=5 at (5,5)."""
    m = F.Mesh.generate(F.GEN_PERIODIC_SQUARE, n, n, a=0.0, b=10.0)
    p = F.default_params(inf_ruvp=[1, 1, 1, 1], CFL=0.5, gradient=gradient, flux_scheme=scheme,
                         time_advance=advance, limiter=limiter, **kw)
    U = F.field_init(m, p, F.INIT_ISENTROPIC_VORTEX, (5.0, 5.0, 5.0))
    return m, p, U


def wedge_case(ni=64, nj=32, scheme=F.SOL_LLF, advance=F.EVO_RK3, seed=7, **kw):
    """C2 family: M=2 flow over a 15 deg ramp in a [0,3]x[0,1] channel, jittered nodes."""
    m = F.Mesh.generate(F.GEN_WEDGE_CHANNEL, ni, nj, a=3.0, b=1.0, c=0.5, d=15.0, jitter=0.1, seed=seed)
    ruvp = [1.4, 2.0, 0.0, 1.0]
    p = F.default_params(inf_ruvp=ruvp, inlet_ruvp=ruvp, outlet_ruvp=ruvp, CFL=0.5, flux_scheme=scheme,
                         time_advance=advance, **kw)
    U = F.field_init(m, p, F.INIT_UNIFORM)
    return m, p, U


def read_pause(path):
    """Parses a reference continue file: (t, istep, CFL), nodes, elements (ID n0 n1 n2 U0..U3), sets."""
    lines = open(path).read().splitlines()
    t, istep, cfl = lines[1].split()
    i = lines.index("nodes: ID, x, y")
    j = lines.index("elements: ID, nodes, U")
    k = lines.index("boundaries: ID, name; edgeIDs")
    nodes = np.array([l.split() for l in lines[i + 1:j]], dtype=float)
    elems = np.array([l.split() for l in lines[j + 1:k]], dtype=float)
    sets = []
    for l in lines[k + 1:]:
        w = l.split()
        if len(w) == 2 and not w[1].lstrip("-").isdigit():
            sets.append((w[1], []))
        elif w:
            sets[-1][1].extend(int(x) for x in w)
    return (float(t), int(istep), float(cfl)), nodes, elems, sets


def cylinder_case(golden_dir):
    """quick_start 2D_cylinder: NS, Ma 0.5, Re 6.5e6 (mu0 from Re), farfield + adiabatic wall,
    state of the reference's own continue file at step 20000."""
    m = F.Mesh.read_su2(golden_dir / "cylinder_allfar_50.su2", equation=F.EQ_NS)
    inf = F.ruvp_from_mach(0.5, 0.0)
    u_ref = float(np.hypot(inf[1], inf[2]))
    p = F.default_params(equation=F.EQ_NS, inf_ruvp=inf, inlet_ruvp=inf, outlet_ruvp=inf, CFL=2.2,
                         Pr=0.73, mu0=inf[0] * u_ref * 1.0 / 6.5e6)
    _, _, elems, _ = read_pause(golden_dir / "pause_cylinder_allfar_50_20000.dat")
    U = np.ascontiguousarray(elems[:, 4:8].T)
    return m, p, U
