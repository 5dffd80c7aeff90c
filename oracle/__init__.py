"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of the CPU oracle (oracle/fvmr_oracle.cpp).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
package, and only as the checker (or the timed CPU baseline), never as the product.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from cfd_solver_2d_gpu_amd._abi import MeshSoA, Params, StepInfo  # struct layouts of include/fvmr.h

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libfvmr_oracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_SIG = {
    "orc_last_error": (C.c_char_p, []),
    "orc_mesh_read_su2": (C.c_void_p, [C.c_char_p, C.c_int, C.c_double]),
    "orc_mesh_from_lists": (C.c_void_p, [C.c_int, _dp, _dp, _ip, C.c_int, _ip, _ip, C.c_int,
                                         C.POINTER(C.c_char_p), _ip, _ip, C.c_int]),
    "orc_mesh_free": (None, [C.c_void_p]),
    "orc_mesh_soa": (None, [C.c_void_p, C.POINTER(MeshSoA)]),
    "orc_mesh_element_node_ids": (_ip, [C.c_void_p]),
    "orc_mesh_num_sets": (C.c_int, [C.c_void_p]),
    "orc_mesh_set_name": (C.c_char_p, [C.c_void_p, C.c_int]),
    "orc_mesh_set_type": (C.c_int, [C.c_void_p, C.c_int]),
    "orc_mesh_set_size": (C.c_int, [C.c_void_p, C.c_int]),
    "orc_mesh_set_edges": (_ip, [C.c_void_p, C.c_int]),
    "orc_field_init": (C.c_int, [C.POINTER(MeshSoA), C.POINTER(Params), C.c_int, _dp, _dp * 4]),
    "orc_solver_create": (C.c_void_p, [C.POINTER(MeshSoA), C.POINTER(Params)]),
    "orc_solver_free": (None, [C.c_void_p]),
    "orc_set_threads": (C.c_int, [C.c_int]),
    "orc_set_state": (None, [C.c_void_p, _dp * 4]),
    "orc_get_state": (None, [C.c_void_p, _dp * 4]),
    "orc_gradient": (C.c_int, [C.c_void_p, _dp * 4, _dp * 4]),
    "orc_stage_rhs": (C.c_int, [C.c_void_p, C.c_double, _dp * 4]),
    "orc_edge_flux": (None, [C.c_void_p, _dp * 4]),
    "orc_global_dt": (C.c_double, [C.c_void_p]),
    "orc_step": (C.c_int, [C.c_void_p, _dp, C.c_double, C.POINTER(StepInfo)]),
    "orc_residual": (None, [C.c_void_p, C.c_int, _dp]),
    "orc_restrict": (C.c_int, [C.c_void_p]),
    "orc_roe": (None, [_dp, _dp, C.c_double, C.c_double, C.c_double, C.c_double, _dp]),
    "orc_llf": (None, [_dp, _dp, C.c_double, C.c_double, C.c_double, C.c_double, _dp]),
    "orc_farfield": (None, [_dp, _dp, C.c_double, C.c_double, C.c_double, _dp]),
    "orc_slip": (None, [_dp, C.c_double, C.c_double, _dp]),
    "orc_isothermal": (None, [_dp, C.c_double, C.c_double, C.c_double, _dp]),
    "orc_u2ruvp": (None, [_dp, C.c_double, _dp]),
    "orc_ruvp2u": (None, [_dp, C.c_double, _dp]),
    "orc_inv2x2": (None, [_dp, C.c_int]),
    "orc_node_field": (None, [C.POINTER(MeshSoA), _dp * 4, C.c_double, _dp, _dp * 4, _dp, _dp]),
    "orc_surface_force": (None, [C.POINTER(MeshSoA), _dp, C.c_int, _ip, _ip, C.c_double, _dp, _dp]),
    "orc_vertex_colors": (None, [C.POINTER(MeshSoA), _dp * 4, C.c_double, C.c_int, C.c_float, C.c_float, C.c_int,
                                 C.POINTER(C.c_float), C.POINTER(C.c_float)]),
}
_libs = {}


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _cpu_tag():
    """A short tag of this host's CPU (model and flags): the fast build is -march=native."""
    import hashlib
    info = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith(("model name", "flags")):
                    info += line
                if line.strip() == "" and info:
                    break
    except OSError:
        pass
    return hashlib.sha1(info.encode()).hexdigest()[:12]


def fast_lib_path():
    return os.path.join(HERE, "_build", "fast_" + _cpu_tag(), "libfvmr_oracle_fast.so")


def build_fast():
    """The -O3 -march=native -fopenmp build for this host's CPU (the benchmark's CPU baseline)."""
    subprocess.run(["make", "-s", "-C", HERE, "fast", "FAST_LIB=" + fast_lib_path()], check=True)


def _load(path):
    L = C.CDLL(path)
    for k, (r, a) in _SIG.items():
        f = getattr(L, k)
        f.restype = r
        f.argtypes = a
    return L


def lib():
    """The plain build (one thread), which the tests use."""
    if "plain" not in _libs:
        build()   # make: rebuilds only when the source or the shared header is newer
        _libs["plain"] = _load(LIB_PATH)
    return _libs["plain"]


def fast_lib(threads=None):
    """The OpenMP build for this host (built on first use); threads: OpenMP threads to use."""
    if "fast" not in _libs:
        path = fast_lib_path()
        build_fast()
        _libs["fast"] = _load(path)
    if threads is not None:
        _libs["fast"].orc_set_threads(int(threads))
    return _libs["fast"]


def _err(what, L=None):
    raise RuntimeError(f"oracle {what} failed: {(L or lib()).orc_last_error().decode()}")


def _dptr(a):
    return a.ctypes.data_as(_dp)


def _iptr(a):
    return a.ctypes.data_as(_ip)


def _rows(U):
    U = np.ascontiguousarray(U, dtype=np.float64)
    return U, (_dp * 4)(*[_dptr(U[k]) for k in range(4)])


def vec(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, _dptr(a)


class OracleMesh:
    """Connectivity built by the reference algorithm (linear-search edge lookup)."""

    def __init__(self, h):
        if not h:
            _err("mesh build")
        self._h = C.c_void_p(h)
        self.soa = MeshSoA()
        lib().orc_mesh_soa(self._h, C.byref(self.soa))

    @classmethod
    def read_su2(cls, path, equation=1, scale=1.0):
        return cls(lib().orc_mesh_read_su2(str(path).encode(), equation, scale))

    @classmethod
    def from_lists(cls, x, y, node_ids, elem_types, elem_conn, markers, equation=1):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        ids = np.ascontiguousarray(node_ids, dtype=np.int32)
        et = np.ascontiguousarray(elem_types, dtype=np.int32)
        ec = np.full((len(et), 4), -1, dtype=np.int32)
        conn = np.asarray(elem_conn)
        ec[:, :conn.shape[1]] = conn
        names = (C.c_char_p * max(1, len(markers)))(*[m[0].encode() for m in markers])
        offs = np.zeros(len(markers) + 1, dtype=np.int32)
        pairs = []
        for k, (_, e) in enumerate(markers):
            pairs.extend(e)
            offs[k + 1] = len(pairs)
        en = np.ascontiguousarray(np.array(pairs, dtype=np.int32).reshape(-1, 2)) if pairs else np.zeros((1, 2), np.int32)
        return cls(lib().orc_mesh_from_lists(len(x), _dptr(x), _dptr(y), _iptr(ids), len(et), _iptr(et), _iptr(ec),
                                             len(markers), names, _iptr(offs), _iptr(en), equation))

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().orc_mesh_free(self._h)
            self._h = C.c_void_p()

    @property
    def num_element(self):
        return self.soa.num_element

    def arrays(self):
        from cfd_solver_2d_gpu_amd import soa_arrays
        return soa_arrays(self.soa)

    def element_node_ids(self):
        p = lib().orc_mesh_element_node_ids(self._h)
        return np.ctypeslib.as_array(p, shape=(self.num_element * 3,)).reshape(-1, 3).copy()

    def boundary_sets(self):
        out = []
        for s in range(lib().orc_mesh_num_sets(self._h)):
            n = lib().orc_mesh_set_size(self._h, s)
            e = np.ctypeslib.as_array(lib().orc_mesh_set_edges(self._h, s), shape=(n,)).copy() if n else np.zeros(0, np.int32)
            out.append((lib().orc_mesh_set_name(self._h, s).decode(), lib().orc_mesh_set_type(self._h, s), e))
        return out


def field_init(soa, params, init_type, args=()):
    n = soa.num_element
    U, rows = _rows(np.zeros((4, n)))
    a = np.zeros(max(4, len(args)))
    a[:len(args)] = args
    if lib().orc_field_init(C.byref(soa), C.byref(params), init_type, _dptr(a), rows):
        _err("field_init")
    return U


class OracleSolver:
    """FP64 CPU restatement of CIteration::iteration_host (forward Euler) plus SSP-RK3."""

    def __init__(self, mesh, params, fast=False):
        """fast: the OpenMP build (fast_lib), with the same results."""
        self.mesh = mesh
        self.n = mesh.soa.num_element
        self.nf = mesh.soa.num_edge
        self._L = fast_lib() if fast else lib()
        self._h = C.c_void_p(self._L.orc_solver_create(C.byref(mesh.soa), C.byref(params)))

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            self._L.orc_solver_free(self._h)
            self._h = C.c_void_p()

    def set_state(self, U):
        U, rows = _rows(U)
        self._L.orc_set_state(self._h, rows)

    def get_state(self):
        U, rows = _rows(np.zeros((4, self.n)))
        self._L.orc_get_state(self._h, rows)
        return U

    def gradient(self):
        gx, rx = _rows(np.zeros((4, self.n)))
        gy, ry = _rows(np.zeros((4, self.n)))
        if self._L.orc_gradient(self._h, rx, ry):
            _err("gradient", self._L)
        return gx, gy

    def stage_rhs(self, t_rk=0.0):
        R, rows = _rows(np.zeros((4, self.n)))
        if self._L.orc_stage_rhs(self._h, t_rk, rows):
            _err("stage_rhs", self._L)
        return R

    def edge_flux(self):
        F, rows = _rows(np.zeros((4, self.nf)))
        self._L.orc_edge_flux(self._h, rows)
        return F

    def global_dt(self):
        return self._L.orc_global_dt(self._h)

    def step(self, t, T, residual_norm=0):
        tt = C.c_double(t)
        info = StepInfo(residual_norm=residual_norm)
        if self._L.orc_step(self._h, C.byref(tt), T, C.byref(info)):
            _err("step", self._L)
        return tt.value, info

    def residual(self, norm=3):
        out = (C.c_double * 4)()
        self._L.orc_residual(self._h, norm, out)
        return np.array(out[:])

    def restrict(self):
        return self._L.orc_restrict(self._h)


def roe(UL, UR, nx, ny, length, gamma=1.4):
    UL, pl = vec(UL)
    UR, pr = vec(UR)
    F, pf = vec(np.zeros(4))
    lib().orc_roe(pl, pr, nx, ny, length, gamma, pf)
    return F


def llf(UL, UR, nx, ny, length, gamma=1.4):
    UL, pl = vec(UL)
    UR, pr = vec(UR)
    F, pf = vec(np.zeros(4))
    lib().orc_llf(pl, pr, nx, ny, length, gamma, pf)
    return F


def farfield(UL, ruvp_inf, nx, ny, gamma=1.4):
    UL, pl = vec(UL)
    ri, pi = vec(ruvp_inf)
    UR, pr = vec(np.zeros(4))
    lib().orc_farfield(pl, pi, nx, ny, gamma, pr)
    return UR


def slip(UL, nx, ny):
    UL, pl = vec(UL)
    UR, pr = vec(np.zeros(4))
    lib().orc_slip(pl, nx, ny, pr)
    return UR


def isothermal(UL, Tw, R=287.06, gamma=1.4):
    """Ghost state of the no-slip isothermal wall at temperature Tw."""
    UL, pl = vec(UL)
    UR, pr = vec(np.zeros(4))
    lib().orc_isothermal(pl, Tw, R, gamma, pr)
    return UR


def u2ruvp(U, gamma=1.4):
    U, pu = vec(U)
    r, pr = vec(np.zeros(4))
    lib().orc_u2ruvp(pu, gamma, pr)
    return r


def ruvp2u(r, gamma=1.4):
    r, pr = vec(r)
    U, pu = vec(np.zeros(4))
    lib().orc_ruvp2u(pr, gamma, pu)
    return U


def inv2x2(m, compat=False):
    m, pm = vec(np.array(m, dtype=np.float64).reshape(4))
    lib().orc_inv2x2(pm, 1 if compat else 0)
    return m


def node_field(soa, U, gamma, ruvp_inf):
    """Node ruvp (4 x num_node), Cp and Ma as the reference's FieldWriter forms them."""
    nv = soa.num_node
    U = np.ascontiguousarray(U, dtype=np.float64)
    R = np.zeros((4, nv))
    cp = np.zeros(nv)
    ma = np.zeros(nv)
    inf, pinf = vec(np.asarray(ruvp_inf, dtype=np.float64))
    lib().orc_node_field(C.byref(soa), (_dp * 4)(*[U[k].ctypes.data_as(_dp) for k in range(4)]), gamma, pinf,
                         (_dp * 4)(*[R[k].ctypes.data_as(_dp) for k in range(4)]), cp.ctypes.data_as(_dp),
                         ma.ctypes.data_as(_dp))
    return R, cp, ma


def surface_force(soa, Cp, sets, S_ref):
    """Per boundary set (a list of edge-index lists, set order) the pressure force
    coefficients (CD, CL) of FieldWriter::ReferenceData::calculate_edgeSet_force."""
    off = np.zeros(len(sets) + 1, dtype=np.int32)
    off[1:] = np.cumsum([len(e) for e in sets])
    edges = np.ascontiguousarray(np.concatenate([np.asarray(e, dtype=np.int32) for e in sets])
                                 if sets else np.zeros(0, np.int32))
    Cp = np.ascontiguousarray(Cp, dtype=np.float64)
    cd = np.zeros(len(sets))
    cl = np.zeros(len(sets))
    lib().orc_surface_force(C.byref(soa), Cp.ctypes.data_as(_dp), len(sets), off.ctypes.data_as(_ip),
                            edges.ctypes.data_as(_ip), S_ref, cd.ctypes.data_as(_dp), cl.ctypes.data_as(_dp))
    return cd, cl


def vertex_colors(soa, U, gamma, var, lo, hi, colormap):
    """r, g, b per node (num_node x 3, float32) as GPU::setVerticesColor forms them."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    cmap = np.ascontiguousarray(np.asarray(colormap, dtype=np.float32).reshape(-1, 4))
    vert = np.zeros((soa.num_node, 6), np.float32)
    fp = C.POINTER(C.c_float)
    lib().orc_vertex_colors(C.byref(soa), (_dp * 4)(*[U[k].ctypes.data_as(_dp) for k in range(4)]), gamma, int(var),
                            float(lo), float(hi), cmap.shape[0], cmap.ctypes.data_as(fp), vert.ctypes.data_as(fp))
    return vert[:, 3:6].copy()
