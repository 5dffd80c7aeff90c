"""Shared case builders of the test-suite (meshes, parameters, initial fields, input.toml trees)."""
import os

import numpy as np
import tomli

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


class SoaMesh:
    """A caller-built fvmr_mesh_soa over numpy arrays (the layout soa_arrays returns), as a
    host code that fills the reference's SoA itself would pass it to fvmr_create."""

    def __init__(self, arrays):
        import ctypes as C
        self.arrays = {k: np.ascontiguousarray(v) for k, v in arrays.items()}
        a = self.arrays
        s = F.MeshSoA()
        s.num_node = a["node_xy"].shape[1]
        s.num_element = a["element_volume"].shape[0]
        s.num_edge = a["edge_length"].shape[0]
        s.num_boundary_set = a["boundary_type"].shape[0]

        def dp(x):
            return x.ctypes.data_as(C.POINTER(C.c_double))

        def ip(x):
            return x.ctypes.data_as(C.POINTER(C.c_int))

        for k in range(2):
            s.node_xy[k] = dp(a["node_xy"][k]); s.element_xy[k] = dp(a["element_xy"][k])
            s.edge_nodes[k] = ip(a["edge_nodes"][k]); s.edge_xy[k] = dp(a["edge_xy"][k])
            s.edge_normal[k] = dp(a["edge_normal"][k])
        for k in range(4):
            s.element_nodes[k] = ip(a["element_nodes"][k]); s.element_edges[k] = ip(a["element_edges"][k])
            s.element_neighbors[k] = ip(a["element_neighbors"][k])
        s.element_volume = dp(a["element_volume"]); s.edge_setID = ip(a["edge_setID"])
        s.edge_periodicPair = ip(a["edge_periodicPair"]); s.edge_elementL = ip(a["edge_elementL"])
        s.edge_elementR = ip(a["edge_elementR"]); s.edge_length = dp(a["edge_length"])
        s.boundary_type = ip(a["boundary_type"])
        self.soa = s
        self.num_element = s.num_element


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def toml_tree(name):
    """The parsed tree of tests/golden/<name> (a quick_start input.toml)."""
    with open(os.path.join(GOLDEN, name), "rb") as f:
        return tomli.load(f)


def tree_set(tree, key, value):
    """Set (or, with None, remove) a dotted key of a parsed TOML tree."""
    node = tree
    parts = key.split(".")
    for p in parts[:-1]:
        node = node.setdefault(p, {})
    if value is None:
        node.pop(parts[-1], None)
    else:
        node[parts[-1]] = value


def write_toml(path, tree):
    """Write a tree of tables, strings, booleans and numbers as TOML."""
    lines = []

    def walk(prefix, node):
        scalars = {k: v for k, v in node.items() if not isinstance(v, dict)}
        if scalars and prefix:
            lines.append(f"[{'.'.join(prefix)}]")
            for k, v in scalars.items():
                if isinstance(v, bool):
                    lines.append(f"{k} = {'true' if v else 'false'}")
                elif isinstance(v, str):
                    lines.append(f'{k} = "{v}"')
                else:
                    lines.append(f"{k} = {v!r}")
        for k, v in node.items():
            if isinstance(v, dict):
                walk(prefix + [k], v)

    walk([], tree)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
