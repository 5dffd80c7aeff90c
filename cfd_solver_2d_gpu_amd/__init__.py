"""MI355X (gfx950) explicit finite-volume residual-and-update path of FVMR.

Host-side mirror of the reference solver surface over the C ABI of libfvmr.so:
  Mesh          SU2MeshReader + FVM_2D + OldDataConverter (FVMR/src/input/SU2MeshReader.cpp,
                FVMR/src/legacy/FVM_2D.cpp, FVMR/src/solvers/OldDataConverter.cpp)
  field_init    FieldInitializer (FVMR/src/input/FieldInitializer.cpp:22-229)
  GPUSolver2    GPU::GPUSolver2 + CIteration::iteration_device
                (FVMR/src/solvers/GPUSolver2.h:32-89, FVMR/src/time/CIteration.cpp:89-161)
  driver        input.toml reader (TomlFileManager, FVMR/src/input/TomlFileManager.cpp:110-265), the
                CDriver::start loop and the Tecplot/hist/continue writers (FVMR/src/drivers/CDriver.cpp:31-257,
                FVMR/src/output/FieldWriter.cpp:44-262); `python -m cfd_solver_2d_gpu_amd.driver input.toml`
"""
import ctypes as C

import numpy as np

from . import _abi
from ._abi import *  # noqa: F401,F403  (codes)
from ._abi import AdjointResult, FvmrError, GenParams, MeshSoA, Params, StepInfo, check, lib

__all__ = ["Mesh", "GPUSolver2", "BatchSolver", "autotune", "autotune_batch", "autotune_candidates", "field_init", "default_params", "ruvp_from_mach", "FvmrError",
           "Params", "StepInfo", "MeshSoA", "lib"]


def default_params(**kw):
    """GlobalPara defaults (FVMR/src/global/GlobalPara.cpp:4-90) with overrides."""
    p = Params()
    check(lib().fvmr_params_default(C.byref(p)), "fvmr_params_default")
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError(f"fvmr_params has no field {k}")
        if k.endswith("_ruvp"):
            arr = getattr(p, k)
            for i in range(4):
                arr[i] = float(v[i])
        else:
            setattr(p, k, v)
    return p


def ruvp_from_mach(Ma, AoA, T0=288.16, p0=101325.0, gamma=1.4, R=287.06):
    out = (C.c_double * 4)()
    check(lib().fvmr_ruvp_from_mach(Ma, AoA, T0, p0, gamma, R, out), "fvmr_ruvp_from_mach")
    return np.array(out[:])


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class Mesh:
    """Host mesh in the reference SoA layout (GPU::NodeSoA/ElementSoA/EdgeSoA/BoundarySetMap)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)
        self.soa = MeshSoA()
        check(lib().fvmr_mesh_soa_view(self._h, C.byref(self.soa)), "fvmr_mesh_soa_view")

    @classmethod
    def read_su2(cls, path, equation=EQ_EULER, scale=1.0):
        h = C.c_void_p()
        check(lib().fvmr_mesh_read_su2(str(path).encode(), equation, scale, C.byref(h)), "fvmr_mesh_read_su2")
        return cls(h.value)

    @classmethod
    def read_continue(cls, path, equation=EQ_EULER):
        """ContinueFileReader (FVMR/src/input/ContinueFileReader.cpp:19-158,457-494): returns
        (mesh, U 4 x num_element, t, istep, CFL) of a pause_*.dat continue file."""
        h = C.c_void_p()
        t, cfl = C.c_double(), C.c_double()
        istep = C.c_int()
        check(lib().fvmr_mesh_read_continue(str(path).encode(), equation, C.byref(h), C.byref(t), C.byref(istep),
                                            C.byref(cfl)), "fvmr_mesh_read_continue")
        m = cls(h.value)
        U = np.zeros((4, m.num_element))
        U, rows = _rows(U)
        check(lib().fvmr_mesh_continue_state(m._h, rows), "fvmr_mesh_continue_state")
        return m, U, t.value, istep.value, cfl.value

    def write_continue(self, path, U, t, istep, CFL):
        """FieldWriter::writeContinueFile_1 layout with 17 significant digits."""
        U, rows = _rows(U, self.num_element)
        check(lib().fvmr_write_continue(self._h, str(path).encode(), float(t), int(istep), float(CFL), rows),
              "fvmr_write_continue")

    @classmethod
    def generate(cls, family, ni, nj, a=0.0, b=0.0, c=0.0, d=0.0, jitter=0.0, seed=0, naca_digits=12,
                 equation=EQ_EULER, isothermal_wall=False):
        g = GenParams(ni=ni, nj=nj, a=a, b=b, c=c, d=d, jitter=jitter, seed=seed, naca_digits=naca_digits,
                      reserved=1 if isothermal_wall else 0)
        h = C.c_void_p()
        check(lib().fvmr_mesh_generate(family, C.byref(g), equation, C.byref(h)), "fvmr_mesh_generate")
        return cls(h.value)

    @classmethod
    def from_lists(cls, x, y, node_ids, elem_types, elem_conn, markers, equation=EQ_EULER):
        """markers: list of (name, [(n0, n1), ...]) with 0-based file node ids."""
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
        h = C.c_void_p()
        check(lib().fvmr_mesh_from_lists(len(x), _dptr(x), _dptr(y), _iptr(ids), len(et), _iptr(et), _iptr(ec),
                                         len(markers), names, _iptr(offs), _iptr(en), equation, C.byref(h)),
              "fvmr_mesh_from_lists")
        return cls(h.value)

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            try:
                lib().fvmr_mesh_free(self._h)
            except Exception:
                pass
            self._h = C.c_void_p()

    @property
    def num_element(self):
        return self.soa.num_element

    @property
    def num_edge(self):
        return self.soa.num_edge

    @property
    def num_node(self):
        return self.soa.num_node

    def write_su2(self, path):
        check(lib().fvmr_mesh_write_su2(self._h, str(path).encode()), "fvmr_mesh_write_su2")

    def arrays(self):
        """numpy copies of every SoA array (keys follow the reference member names)."""
        return soa_arrays(self.soa)

    def element_node_ids(self):
        p = C.POINTER(C.c_int)()
        check(lib().fvmr_mesh_element_node_ids(self._h, C.byref(p)), "fvmr_mesh_element_node_ids")
        return np.ctypeslib.as_array(p, shape=(self.num_element * 3,)).reshape(-1, 3).copy()

    def boundary_sets(self):
        out = []
        for s in range(self.soa.num_boundary_set):
            name = C.c_char_p()
            typ = C.c_int()
            edges = C.POINTER(C.c_int)()
            n = C.c_int()
            check(lib().fvmr_mesh_boundary_set(self._h, s, C.byref(name), C.byref(typ), C.byref(edges), C.byref(n)),
                  "fvmr_mesh_boundary_set")
            e = np.ctypeslib.as_array(edges, shape=(n.value,)).copy() if n.value else np.zeros(0, np.int32)
            out.append((name.value.decode(), typ.value, e))
        return out


def soa_arrays(soa):
    """numpy copies of the arrays of an fvmr_mesh_soa view."""
    nv, ne, nf, nb = soa.num_node, soa.num_element, soa.num_edge, soa.num_boundary_set

    def d(p, n):
        return np.ctypeslib.as_array(p, shape=(n,)).copy()

    return {
        "node_xy": np.stack([d(soa.node_xy[k], nv) for k in range(2)]),
        "element_nodes": np.stack([d(soa.element_nodes[k], ne) for k in range(4)]),
        "element_edges": np.stack([d(soa.element_edges[k], ne) for k in range(4)]),
        "element_neighbors": np.stack([d(soa.element_neighbors[k], ne) for k in range(4)]),
        "element_xy": np.stack([d(soa.element_xy[k], ne) for k in range(2)]),
        "element_volume": d(soa.element_volume, ne),
        "edge_nodes": np.stack([d(soa.edge_nodes[k], nf) for k in range(2)]),
        "edge_setID": d(soa.edge_setID, nf),
        "edge_periodicPair": d(soa.edge_periodicPair, nf),
        "edge_elementL": d(soa.edge_elementL, nf),
        "edge_elementR": d(soa.edge_elementR, nf),
        "edge_length": d(soa.edge_length, nf),
        "edge_xy": np.stack([d(soa.edge_xy[k], nf) for k in range(2)]),
        "edge_normal": np.stack([d(soa.edge_normal[k], nf) for k in range(2)]),
        "boundary_type": d(soa.boundary_type, nb) if nb else np.zeros(0, np.int32),
    }


def _rows(U, n=None):
    """U as a contiguous float64 array and its four row pointers; with n, U must be 4 x n (the
    library copies n doubles from each row)."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    if n is not None and U.shape != (4, n):
        raise ValueError(f"state must be 4 x {n} (rho, rho u, rho v, rho E per cell), got shape {U.shape}")
    return U, (C.POINTER(C.c_double) * 4)(*[_dptr(U[k]) for k in range(4)])


def autotune(mesh, params, steps=20):
    """fvmr_autotune: (best tile_threads, best tile_cells, ms per step) over the candidate
    work-group shapes and tile sizes on this mesh."""
    th, best = C.c_int(), C.c_int()
    ms = C.c_double()
    check(lib().fvmr_autotune(C.byref(mesh.soa), C.byref(params), steps, C.byref(th), C.byref(best), C.byref(ms)),
          "fvmr_autotune")
    return th.value, best.value, ms.value


def autotune_candidates(precision=PREC_FP64):
    """The (tile_threads, tile_cells) pairs the tuners try for a precision."""
    n = C.c_int()
    th = np.zeros(16, dtype=np.int32)
    cells = np.zeros(16, dtype=np.int32)
    check(lib().fvmr_autotune_candidates(precision, C.byref(n), _iptr(th), _iptr(cells)), "fvmr_autotune_candidates")
    return [(int(a), int(b)) for a, b in zip(th[:n.value], cells[:n.value])]


def autotune_batch(meshes, params, steps=10):
    """fvmr_autotune_batch: (best tile_threads, best tile_cells, its ms per step,
    {(tile_threads, tile_cells): ms per step}) for a batch of these cases."""
    n = len(meshes)
    soas = (C.POINTER(MeshSoA) * n)(*[C.pointer(m.soa) for m in meshes])
    ps = (Params * n)(*params)
    th, best = C.c_int(), C.c_int()
    ms = C.c_double()
    cand = autotune_candidates(params[0].precision)
    times = np.zeros(16)
    check(lib().fvmr_autotune_batch(n, soas, ps, steps, C.byref(th), C.byref(best), C.byref(ms), _dptr(times)),
          "fvmr_autotune_batch")
    return th.value, best.value, ms.value, {c: float(t) for c, t in zip(cand, times)}


def device_division(a, b):
    """(a / b, 1 / b) as the device kernels form them in FP64 (fvmr_device_division)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    if a.shape != b.shape or a.ndim != 1:
        raise ValueError("a and b must be 1-D arrays of one length")
    q = np.zeros_like(a)
    r = np.zeros_like(a)
    check(lib().fvmr_device_division(len(a), _dptr(a), _dptr(b), _dptr(q), _dptr(r)), "fvmr_device_division")
    return q, r


def field_init(mesh, params, init_type=INIT_UNIFORM, args=()):
    """FieldInitializer::setInitialAndBoundaryCondition: U (4 x num_element)."""
    U = np.zeros((4, mesh.num_element))
    a = np.zeros(max(4, len(args)))
    a[:len(args)] = args
    U, rows = _rows(U)
    check(lib().fvmr_field_init(C.byref(mesh.soa), C.byref(params), init_type, _dptr(a), rows), "fvmr_field_init")
    return U


class GPUSolver2:
    """Device solver: GPU::GPUSolver2 (allocate/initialize/host_to_device) + CIteration."""

    PROF_STAGE, PROF_GRADIENT, PROF_DT, PROF_RESTRICT, PROF_RESIDUAL = range(5)

    def __init__(self, mesh, params):
        self.mesh = mesh  # keeps the SoA storage alive
        self.params = params
        self._h = C.c_void_p()
        check(lib().fvmr_create(C.byref(mesh.soa), C.byref(params), C.byref(self._h)), "fvmr_create")
        self.n = mesh.num_element

    def close(self):
        if self._h and self._h.value:
            check(lib().fvmr_destroy(self._h), "fvmr_destroy")
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_state(self, U):
        U, rows = _rows(U, self.n)
        check(lib().fvmr_set_state(self._h, rows), "fvmr_set_state")

    def get_state(self):
        U = np.zeros((4, self.n))
        U, rows = _rows(U)
        check(lib().fvmr_get_state(self._h, rows), "fvmr_get_state")
        return U

    def iteration_device(self, t, T, residual_norm=0):
        """CIteration::iteration_device: one step; returns (t, StepInfo)."""
        tt = C.c_double(t)
        info = StepInfo(residual_norm=residual_norm)
        check(lib().fvmr_step(self._h, C.byref(tt), T, C.byref(info)), "fvmr_step")
        return tt.value, info

    def advance(self, nsteps, t, T, residual_norm=0):
        tt = C.c_double(t)
        info = StepInfo(residual_norm=residual_norm)
        check(lib().fvmr_advance(self._h, nsteps, C.byref(tt), T, C.byref(info)), "fvmr_advance")
        return tt.value, info

    def stage_rhs(self):
        R = np.zeros((4, self.n))
        R, rows = _rows(R)
        check(lib().fvmr_stage_rhs(self._h, rows), "fvmr_stage_rhs")
        return R

    def gradient(self):
        gx, rx = _rows(np.zeros((4, self.n)))
        gy, ry = _rows(np.zeros((4, self.n)))
        check(lib().fvmr_gradient(self._h, rx, ry), "fvmr_gradient")
        return gx, gy

    def global_dt(self):
        dt = C.c_double()
        check(lib().fvmr_global_dt(self._h, C.byref(dt)), "fvmr_global_dt")
        return dt.value

    def updateResidualVector(self, norm=NORM_INF):
        out = (C.c_double * 4)()
        check(lib().fvmr_residual(self._h, norm, out), "fvmr_residual")
        return np.array(out[:])

    def permutation(self):
        p = np.zeros(self.n, dtype=np.int32)
        check(lib().fvmr_permutation(self._h, _iptr(p)), "fvmr_permutation")
        return p

    def use_graphs(self, on=True):
        """Replay runs of three steps from captured HIP graphs (default off)."""
        check(lib().fvmr_use_graphs(self._h, 1 if on else 0), "fvmr_use_graphs")

    def node_field(self):
        """FieldWriter::update_nodeField on the device: (ruvp 4 x num_node, Cp, Ma) of the
        current state (FVMR/src/output/FieldWriter.cpp:371-452)."""
        nv = self.mesh.soa.num_node
        R = np.zeros((4, nv))
        cp = np.zeros(nv)
        ma = np.zeros(nv)
        R, rows = _rows(R)
        check(lib().fvmr_node_field(self._h, rows, _dptr(cp), _dptr(ma)), "fvmr_node_field")
        return R, cp, ma

    def surface_force(self, S_ref=1.0, sets=None):
        """ReferenceData::calculate_edgeSet_force (FieldWriter.cpp:174-205) on the device: per
        boundary set (default: the mesh's sets, in set order) the pressure force coefficients
        (Cx, Cy) = the reference's (CD, CL)."""
        if sets is None:
            sets = [s[2] for s in self.mesh.boundary_sets()]
        off = np.zeros(len(sets) + 1, dtype=np.int32)
        off[1:] = np.cumsum([len(e) for e in sets])
        edges = np.ascontiguousarray(np.concatenate([np.asarray(e, dtype=np.int32) for e in sets])
                                     if len(sets) else np.zeros(1, np.int32))
        cx = np.zeros(max(1, len(sets)))
        cy = np.zeros(max(1, len(sets)))
        check(lib().fvmr_surface_force(self._h, len(sets), _iptr(off), _iptr(edges), float(S_ref), _dptr(cx),
                                       _dptr(cy)), "fvmr_surface_force")
        return cx[:len(sets)], cy[:len(sets)]

    def surface_force_noalloc(self, S_ref=1.0, sets=None):
        """surface_force through fvmr_surface_force_noalloc: the same coefficients, bit for bit,
        with no device allocation (at most 256 sets and 131072 edges per call)."""
        if sets is None:
            sets = [s[2] for s in self.mesh.boundary_sets()]
        off, edges = self._sets(sets)
        cx = np.zeros(max(1, len(sets)))
        cy = np.zeros(max(1, len(sets)))
        check(lib().fvmr_surface_force_noalloc(self._h, len(sets), _iptr(off), _iptr(edges), float(S_ref),
                                               _dptr(cx), _dptr(cy)), "fvmr_surface_force_noalloc")
        return cx[:len(sets)], cy[:len(sets)]

    # CudaGrid::CudaGrid's colour map (FVMR/src/gui/cuda/CudaGrid.cpp:33-39): {r, g, b, position}
    DEFAULT_COLORMAP = [(1.0, 0.0, 0.0, 0.00), (1.0, 1.0, 0.0, 0.25), (0.0, 1.0, 0.0, 0.50),
                        (0.0, 0.0, 1.0, 0.75), (1.0, 0.0, 1.0, 1.00)]

    def vertex_colors(self, var=0, lo=0.0, hi=1.0, colormap=None, vertices=None):
        """GPU::setVerticesColor: r, g, b of every node (num_node x 3, float32) for the primitive
        var of the current state normalised by [lo, hi]; vertices (num_node x 6 float32, x y z r g b)
        is updated in place when given."""
        cmap = np.ascontiguousarray(np.asarray(colormap if colormap is not None else self.DEFAULT_COLORMAP,
                                               dtype=np.float32).reshape(-1, 4))
        nv = self.mesh.soa.num_node
        vert = np.zeros((nv, 6), np.float32) if vertices is None else vertices
        assert vert.dtype == np.float32 and vert.shape == (nv, 6) and vert.flags.c_contiguous
        fp = C.POINTER(C.c_float)
        check(lib().fvmr_vertex_colors(self._h, int(var), float(lo), float(hi), cmap.shape[0], cmap.ctypes.data_as(fp),
                                       vert.ctypes.data_as(fp), 0), "fvmr_vertex_colors")
        return vert[:, 3:6].copy()

    def _sets(self, sets):
        off = np.zeros(len(sets) + 1, dtype=np.int32)
        off[1:] = np.cumsum([len(e) for e in sets])
        edges = np.ascontiguousarray(np.concatenate([np.asarray(e, dtype=np.int32) for e in sets])
                                     if len(sets) else np.zeros(1, np.int32))
        return off, edges

    def set_boundary_states(self, inf, inlet=None, outlet=None):
        """Replace the freestream / inlet / outlet (rho, u, v, p) of the handle."""
        p = self.params
        arr = [np.ascontiguousarray(np.asarray(v if v is not None else d, dtype=np.float64))
               for v, d in ((inf, p.inf_ruvp), (inlet, p.inlet_ruvp), (outlet, p.outlet_ruvp))]
        check(lib().fvmr_set_boundary_states(self._h, *[_dptr(a) for a in arr]), "fvmr_set_boundary_states")
        for k in range(4):
            p.inf_ruvp[k], p.inlet_ruvp[k], p.outlet_ruvp[k] = arr[0][k], arr[1][k], arr[2][k]

    def steady_solve(self, max_iters, tol):
        """Pseudo-time iteration with per-cell dt to r(U) = 0: (iterations, first, last L2 of
        the density residual)."""
        it = C.c_int()
        res = np.zeros(2)
        check(lib().fvmr_steady_solve(self._h, int(max_iters), float(tol), C.byref(it), _dptr(res)),
              "fvmr_steady_solve")
        return it.value, res[0], res[1]

    def adjoint_vjp(self, psi, alpha=False, coords=False):
        """(dr/dU)^T psi at the current state, plus psi . dr/dalpha and psi . dr/dX (2 x num_node)
        when asked."""
        P, prow = _rows(np.array(psi, dtype=np.float64, copy=True), self.n)
        out, orow = _rows(np.zeros((4, self.n)))
        da = C.c_double()
        dxy = np.zeros(2 * self.mesh.soa.num_node)
        check(lib().fvmr_adjoint_vjp(self._h, prow, orow, C.byref(da) if alpha else None,
                                     _dptr(dxy) if coords else None), "fvmr_adjoint_vjp")
        res = [out]
        if alpha:
            res.append(da.value)
        if coords:
            res.append(dxy.reshape(-1, 2).T.copy())
        return res[0] if len(res) == 1 else tuple(res)

    def adjoint_solve(self, sets, S_ref=1.0, max_iters=10000, tol=1e-10, coords=False):
        """Discrete adjoint of J = sum over sets of (-sin a Cx + cos a Cy): returns
        (AdjointResult, dJ/dX as 2 x num_node or None)."""
        off, edges = self._sets(sets)
        out = AdjointResult()
        dxy = np.zeros(2 * self.mesh.soa.num_node) if coords else None
        check(lib().fvmr_adjoint_solve(self._h, len(sets), _iptr(off), _iptr(edges), float(S_ref), int(max_iters),
                                       float(tol), _dptr(dxy) if coords else None, C.byref(out)),
              "fvmr_adjoint_solve")
        return out, (dxy.reshape(-1, 2).T.copy() if coords else None)

    def adjoint_state(self):
        P, rows = _rows(np.zeros((4, self.n)))
        check(lib().fvmr_adjoint_state(self._h, rows), "fvmr_adjoint_state")
        return P

    def profile(self, on=True):
        check(lib().fvmr_profile_enable(self._h, 1 if on else 0), "fvmr_profile_enable")

    def profile_query(self, kernel):
        ms = C.c_double()
        n = C.c_int()
        check(lib().fvmr_profile_query(self._h, kernel, C.byref(ms), C.byref(n)), "fvmr_profile_query")
        return ms.value, n.value

    def profile_reset(self):
        check(lib().fvmr_profile_reset(self._h), "fvmr_profile_reset")

    def synchronize(self):
        check(lib().fvmr_synchronize(self._h), "fvmr_synchronize")


class BatchSolver:
    """Many independent cases on one device handle (fvmr_create_batch): the cases share the
    scheme settings; boundary states, CFL, dt and time are per case."""

    def __init__(self, meshes, params):
        if len(meshes) != len(params) or not meshes:
            raise ValueError("one params per mesh")
        self.meshes = list(meshes)   # keep the SoA storage alive
        n = len(meshes)
        soas = (C.POINTER(MeshSoA) * n)(*[C.pointer(m.soa) for m in meshes])
        self._params = (Params * n)(*params)
        self._h = C.c_void_p()
        check(lib().fvmr_create_batch(n, soas, self._params, C.byref(self._h)), "fvmr_create_batch")
        self.n = n

    def close(self):
        if self._h and self._h.value:
            check(lib().fvmr_destroy(self._h), "fvmr_destroy")
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def cases(self):
        nc = C.c_int()
        b = np.zeros(self.n + 1, dtype=np.int32)
        check(lib().fvmr_batch_cases(self._h, C.byref(nc), _iptr(b)), "fvmr_batch_cases")
        return b

    def set_state(self, c, U):
        if not 0 <= c < self.n:
            raise ValueError(f"case {c} out of range [0, {self.n})")
        U, rows = _rows(U, self.meshes[c].num_element)
        check(lib().fvmr_batch_set_state(self._h, c, rows), "fvmr_batch_set_state")

    def get_state(self, c):
        U = np.zeros((4, self.meshes[c].num_element))
        U, rows = _rows(U)
        check(lib().fvmr_batch_get_state(self._h, c, rows), "fvmr_batch_get_state")
        return U

    def advance(self, nsteps, t, T):
        """nsteps steps of every case; t, T: per-case arrays; returns the new t."""
        t = np.ascontiguousarray(t, dtype=np.float64).copy()
        T = np.ascontiguousarray(T, dtype=np.float64)
        check(lib().fvmr_batch_advance(self._h, nsteps, _dptr(t), _dptr(T)), "fvmr_batch_advance")
        return t

    def synchronize(self):
        check(lib().fvmr_synchronize(self._h), "fvmr_synchronize")

    def step_info(self, c):
        """Case c after the last advance: StepInfo with t, dt and n_restricted of its last step."""
        info = StepInfo()
        check(lib().fvmr_batch_step_info(self._h, int(c), C.byref(info)), "fvmr_batch_step_info")
        return info

    def use_graphs(self, on=True):
        check(lib().fvmr_use_graphs(self._h, 1 if on else 0), "fvmr_use_graphs")

    def profile(self, on=True):
        check(lib().fvmr_profile_enable(self._h, 1 if on else 0), "fvmr_profile_enable")

    def profile_reset(self):
        check(lib().fvmr_profile_reset(self._h), "fvmr_profile_reset")

    def profile_query(self, kernel):
        ms, n = C.c_double(), C.c_int()
        check(lib().fvmr_profile_query(self._h, kernel, C.byref(ms), C.byref(n)), "fvmr_profile_query")
        return ms.value, n.value
