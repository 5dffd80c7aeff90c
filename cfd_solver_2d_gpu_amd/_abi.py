"""ctypes mirror of include/fvmr.h.

The library is the in-tree libfvmr.so built by cfd_solver_2d_gpu_amd.build; loading
fails loudly when it is missing (there is no CPU fallback of the device path).
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# the in-tree build; FVMR_LIBRARY may name another build of the same sources (a variant built
# with other compile-time settings by cfd_solver_2d_gpu_amd.build.build_library)
LIB_PATH = os.environ.get("FVMR_LIBRARY") or os.path.join(HERE, "libfvmr.so")

# status codes
FVMR_OK = 0

# codes of FVMR/src/global/Constexpr.h
EQ_EULER, EQ_NS = 1, 2
SOL_LLF, SOL_ROE = 2001, 2002
REC_CONSTANT, REC_LINEAR, REC_MUSCL = 3001, 3002, 3003
GRA_LEASTSQUARE, GRA_GREENGAUSS = 1, 2
EVO_EXPLICIT, EVO_RK3 = 3101, 3103
LIM_NONE, LIM_BARTH, LIM_VENKATAKRISHNAN = 0, 1, 4
BC_WALL_NONVISCOUS, BC_WALL_ADIABAT, BC_WALL_ISOTHERMAL = 6001, 6002, 6003
BC_INLET, BC_OUTLET, BC_INF, BC_DSR, BC_SYMMETRY, BC_PERIODIC_0 = 6011, 6012, 6013, 6014, 6021, 6100
PREC_FP64, PREC_MIXED = 0, 1
RENUMBER_NONE, RENUMBER_RCM, RENUMBER_HILBERT = 0, 1, 2
KERNELS_AUTO, KERNELS_CELL, KERNELS_TILE = 0, 1, 2
MU_SUTHERLAND, MU_CONSTANT = 0, 1
NORM_1, NORM_2, NORM_INF = 1, 2, 3
INIT_UNIFORM, INIT_ISENTROPIC_VORTEX, INIT_SHOCK_TUBE, INIT_DSR = 1, 2, 3, 4
GEN_PERIODIC_SQUARE, GEN_WEDGE_CHANNEL, GEN_CYLINDER_OGRID, GEN_NACA_OGRID, GEN_RECT_CHANNEL = 1, 2, 3, 4, 5

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class MeshSoA(C.Structure):
    _fields_ = [
        ("num_node", C.c_int), ("num_element", C.c_int), ("num_edge", C.c_int), ("num_boundary_set", C.c_int),
        ("node_xy", _dp * 2),
        ("element_nodes", _ip * 4), ("element_edges", _ip * 4), ("element_neighbors", _ip * 4),
        ("element_xy", _dp * 2), ("element_volume", _dp),
        ("edge_nodes", _ip * 2), ("edge_setID", _ip), ("edge_periodicPair", _ip),
        ("edge_elementL", _ip), ("edge_elementR", _ip), ("edge_length", _dp),
        ("edge_xy", _dp * 2), ("edge_normal", _dp * 2),
        ("boundary_type", _ip),
    ]


class Params(C.Structure):
    _fields_ = [
        ("gamma", C.c_double), ("R", C.c_double), ("Pr", C.c_double), ("Pr_viscous", C.c_double),
        ("T0", C.c_double), ("mu0", C.c_double), ("CFL", C.c_double),
        ("equation", C.c_int), ("flux_scheme", C.c_int), ("reconstruct", C.c_int), ("gradient", C.c_int),
        ("limiter", C.c_int), ("time_advance", C.c_int), ("viscosity_model", C.c_int), ("precision", C.c_int),
        ("renumber", C.c_int), ("compat_inv2x2", C.c_int),
        ("kernels", C.c_int), ("tile_cells", C.c_int), ("tile_threads", C.c_int), ("reserved", C.c_int),
        ("inf_ruvp", C.c_double * 4), ("inlet_ruvp", C.c_double * 4), ("outlet_ruvp", C.c_double * 4),
        ("wall_temperature", C.c_double), ("venkat_K", C.c_double),
        ("dsr_shock_x", C.c_double), ("dsr_shock_y", C.c_double), ("dsr_shock_speed", C.c_double),
    ]


class StepInfo(C.Structure):
    _fields_ = [
        ("t", C.c_double), ("dt", C.c_double), ("residual", C.c_double * 4),
        ("residual_norm", C.c_int), ("n_restricted", C.c_int), ("steps_done", C.c_int), ("reserved", C.c_int),
    ]


class AdjointResult(C.Structure):
    _fields_ = [
        ("J", C.c_double), ("dJ_dalpha", C.c_double), ("res0", C.c_double), ("res", C.c_double),
        ("iters", C.c_int), ("reserved", C.c_int),
    ]


class GenParams(C.Structure):
    _fields_ = [
        ("ni", C.c_int), ("nj", C.c_int),
        ("a", C.c_double), ("b", C.c_double), ("c", C.c_double), ("d", C.c_double),
        ("jitter", C.c_double), ("seed", C.c_ulonglong), ("naca_digits", C.c_int), ("reserved", C.c_int),
    ]


# name -> (restype, argtypes); every symbol of include/fvmr.h
SIGNATURES = {
    "fvmr_last_error": (C.c_char_p, []),
    "fvmr_params_default": (C.c_int, [C.POINTER(Params)]),
    "fvmr_ruvp_from_mach": (C.c_int, [C.c_double] * 6 + [_dp]),
    "fvmr_mesh_read_su2": (C.c_int, [C.c_char_p, C.c_int, C.c_double, C.POINTER(C.c_void_p)]),
    "fvmr_mesh_from_lists": (C.c_int, [C.c_int, _dp, _dp, _ip, C.c_int, _ip, _ip, C.c_int,
                                       C.POINTER(C.c_char_p), _ip, _ip, C.c_int, C.POINTER(C.c_void_p)]),
    "fvmr_mesh_generate": (C.c_int, [C.c_int, C.POINTER(GenParams), C.c_int, C.POINTER(C.c_void_p)]),
    "fvmr_mesh_write_su2": (C.c_int, [C.c_void_p, C.c_char_p]),
    "fvmr_mesh_read_continue": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_void_p), _dp, _ip, _dp]),
    "fvmr_mesh_continue_state": (C.c_int, [C.c_void_p, _dp * 4]),
    "fvmr_write_continue": (C.c_int, [C.c_void_p, C.c_char_p, C.c_double, C.c_int, C.c_double, _dp * 4]),
    "fvmr_mesh_soa_view": (C.c_int, [C.c_void_p, C.POINTER(MeshSoA)]),
    "fvmr_mesh_boundary_set": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_char_p), _ip,
                                         C.POINTER(_ip), _ip]),
    "fvmr_mesh_element_node_ids": (C.c_int, [C.c_void_p, C.POINTER(_ip)]),
    "fvmr_mesh_free": (None, [C.c_void_p]),
    "fvmr_field_init": (C.c_int, [C.POINTER(MeshSoA), C.POINTER(Params), C.c_int, _dp, _dp * 4]),
    "fvmr_create": (C.c_int, [C.POINTER(MeshSoA), C.POINTER(Params), C.POINTER(C.c_void_p)]),
    "fvmr_destroy": (C.c_int, [C.c_void_p]),
    "fvmr_set_state": (C.c_int, [C.c_void_p, _dp * 4]),
    "fvmr_get_state": (C.c_int, [C.c_void_p, _dp * 4]),
    "fvmr_step": (C.c_int, [C.c_void_p, _dp, C.c_double, C.POINTER(StepInfo)]),
    "fvmr_advance": (C.c_int, [C.c_void_p, C.c_int, _dp, C.c_double, C.POINTER(StepInfo)]),
    "fvmr_stage_rhs": (C.c_int, [C.c_void_p, _dp * 4]),
    "fvmr_gradient": (C.c_int, [C.c_void_p, _dp * 4, _dp * 4]),
    "fvmr_global_dt": (C.c_int, [C.c_void_p, _dp]),
    "fvmr_residual": (C.c_int, [C.c_void_p, C.c_int, _dp]),
    "fvmr_permutation": (C.c_int, [C.c_void_p, _ip]),
    "fvmr_device_division": (C.c_int, [C.c_int, _dp, _dp, _dp, _dp]),
    "fvmr_node_field": (C.c_int, [C.c_void_p, _dp * 4, _dp, _dp]),
    "fvmr_use_graphs": (C.c_int, [C.c_void_p, C.c_int]),
    "fvmr_autotune": (C.c_int, [C.POINTER(MeshSoA), C.POINTER(Params), C.c_int, _ip, _ip, _dp]),
    "fvmr_autotune_batch": (C.c_int, [C.c_int, C.POINTER(C.POINTER(MeshSoA)), C.POINTER(Params), C.c_int, _ip, _ip,
                                      _dp, _dp]),
    "fvmr_autotune_candidates": (C.c_int, [C.c_int, _ip, _ip, _ip]),
    "fvmr_set_boundary_states": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    "fvmr_vertex_colors": (C.c_int, [C.c_void_p, C.c_int, C.c_float, C.c_float, C.c_int, C.POINTER(C.c_float),
                                     C.POINTER(C.c_float), C.c_int]),
    "fvmr_steady_solve": (C.c_int, [C.c_void_p, C.c_int, C.c_double, _ip, _dp]),
    "fvmr_adjoint_vjp": (C.c_int, [C.c_void_p, _dp * 4, _dp * 4, _dp, _dp]),
    "fvmr_adjoint_solve": (C.c_int, [C.c_void_p, C.c_int, _ip, _ip, C.c_double, C.c_int, C.c_double, _dp,
                                     C.POINTER(AdjointResult)]),
    "fvmr_adjoint_state": (C.c_int, [C.c_void_p, _dp * 4]),
    "fvmr_create_batch": (C.c_int, [C.c_int, C.POINTER(C.POINTER(MeshSoA)), C.POINTER(Params), C.POINTER(C.c_void_p)]),
    "fvmr_batch_cases": (C.c_int, [C.c_void_p, _ip, _ip]),
    "fvmr_batch_set_state": (C.c_int, [C.c_void_p, C.c_int, _dp * 4]),
    "fvmr_batch_get_state": (C.c_int, [C.c_void_p, C.c_int, _dp * 4]),
    "fvmr_batch_advance": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp]),
    "fvmr_batch_step_info": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(StepInfo)]),
    "fvmr_surface_force": (C.c_int, [C.c_void_p, C.c_int, _ip, _ip, C.c_double, _dp, _dp]),
    "fvmr_surface_force_noalloc": (C.c_int, [C.c_void_p, C.c_int, _ip, _ip, C.c_double, _dp, _dp]),
    "fvmr_profile_enable": (C.c_int, [C.c_void_p, C.c_int]),
    "fvmr_profile_query": (C.c_int, [C.c_void_p, C.c_int, _dp, _ip]),
    "fvmr_profile_reset": (C.c_int, [C.c_void_p]),
    "fvmr_synchronize": (C.c_int, [C.c_void_p]),
}

_lib = None


class FvmrError(RuntimeError):
    def __init__(self, code, func, msg):
        super().__init__(f"{func} failed (status {code}): {msg}")
        self.code = code


def lib():
    """The loaded libfvmr.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m cfd_solver_2d_gpu_amd.build` "
                "(the device path has no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(code, func):
    if code != FVMR_OK:
        msg = lib().fvmr_last_error()
        raise FvmrError(code, func, msg.decode() if msg else "")
    return code
