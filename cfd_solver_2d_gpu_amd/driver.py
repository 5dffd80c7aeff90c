"""Host driver: input.toml -> mesh and field -> device solver -> step loop with output files.

Mirrors the reference's input side and main loop:
  read_config     TomlFileManager::treeToGlobalParameter and getValue_boundaryCondition2D
                  (FVMR/src/input/TomlFileManager.cpp:110-170,200-265), read_mu0_Re_from_config
                  (:52-103), FieldInitializer::initialize_using_config (FieldInitializer.cpp:71-111,139-145),
                  FieldWriter::initialize_outputScheme_usingConfig (FieldWriter.cpp:455-502)
  Driver          U2NITS::CDriver::start with CInput::readField_1 (FVMR/src/drivers/CDriver.cpp:31-185,
                  :203-257; FVMR/src/input/CInput.cpp:84-121), ContinueFileReader's file lookup
                  (ContinueFileReader.cpp:19-62) and COutput::updateFileName (COutput.cpp:5-20)
  write_tecplot_* FieldWriter::write_tecplot_volume_file / _boundary_file / _hist_file
                  (FieldWriter.cpp:44-172,207-262,504-594)
  run_batch       several cases on one batch handle (no reference counterpart)

Keys are read with the reference's rules: a required key that is missing or has the wrong type
is an error, all errors of a file are collected and raised together (ConfigError); a float key
accepts a TOML integer, a bool key accepts an integer, an int key accepts neither a float nor a
bool.  The table [device] holds the keys the reference does not have (precision, renumbering,
kernel family, tile size, HIP graphs, Venkatakrishnan K, isothermal wall temperature).

Inputs are read from, and outputs written to, the case directory (the directory of input.toml),
as the reference does with its empty input/output folder names (FilePathManager.h:12-13).
"""
import math
import os
import re
import time

import numpy as np
import tomli

from . import _abi as F
from . import BatchSolver, GPUSolver2, Mesh, autotune, default_params, field_init, ruvp_from_mach

EPSILON = 1e-10          # U2NITS::Math::EPSILON (FVMR/src/math/CommonValue.h:21): steady-state stop
SNAPSHOT_SECONDS = 1800.0  # CDriver::m_snapshot_cold_time (FVMR/src/drivers/CDriver.h:34)

# FieldWriter::OutputSchemePresetEnum (FVMR/src/output/FieldWriter.h:49-53)
PRESET_RUVP, PRESET_RUVP_CP, PRESET_RUVP_MA = 0, 1, 2
_PRESET_VARS = {PRESET_RUVP: ("rho", "u", "v", "p"), PRESET_RUVP_CP: ("rho", "u", "v", "p", "Cp"),
                PRESET_RUVP_MA: ("rho", "u", "v", "p", "Ma")}

_WALLS = (F.BC_WALL_ADIABAT, F.BC_WALL_ISOTHERMAL, F.BC_WALL_NONVISCOUS)

_REQUIRED = object()


class ConfigError(ValueError):
    """Invalid or missing input.toml keys (the reference logs them and exits)."""


class _Tree:
    """Typed access to a parsed TOML tree with dotted keys (cpptoml's get_qualified_as)."""

    def __init__(self, tree):
        self.tree = tree
        self.errors = []

    def has(self, key):
        node = self.tree
        for part in key.split("."):
            if not isinstance(node, dict) or part not in node:
                return False
            node = node[part]
        return True

    def _raw(self, key):
        node = self.tree
        for part in key.split("."):
            node = node[part]
        return node

    def get(self, key, typ, default=_REQUIRED):
        """The value of key as typ; default when absent (an error when default is _REQUIRED)."""
        if not self.has(key):
            if default is _REQUIRED:
                self.errors.append(f'Required parameter "{key}" doesn\'t exist.')
                return None
            return default
        v = self._raw(key)
        ok = None
        if typ is bool:
            if isinstance(v, bool):
                ok = v
            elif type(v) is int:
                ok = bool(v)
        elif typ is int:
            if type(v) is int:
                ok = v
        elif typ is float:
            if type(v) is float:
                ok = v
            elif type(v) is int:
                ok = float(v)
        elif typ is str:
            if isinstance(v, str):
                ok = v
        if ok is None:
            self.errors.append(f'Parameter "{key}" gets a wrong value type.')
            return None if default is _REQUIRED else default
        return ok


class Config:
    """The GlobalPara values an input.toml sets (FVMR/src/global/GlobalPara.cpp:4-90 defaults)."""

    def __init__(self):
        # basic
        self.continue_ = True
        self.dimension = 2
        self.filename = "default_project"
        self.meshFileType = "inp"
        self.useGPU = 0
        self.isDebugMode = False
        # render (CDriver members and GlobalPara::render)
        self.show_gui = True
        self.range_1, self.range_2 = 0.0, 1.0
        self.enable_print = True
        self.enable_write_file = True
        # constant
        self.R = 287.06
        self.T0 = 288.16
        self.p0 = 101325.0
        self.Pr = 0.73
        self.gamma = 1.4
        self.Re = 1.0e8
        self.mu0 = 17.9e-6
        self.referenceArea = 1.0
        self.mesh_scale_factor = 1.0
        self.calculate_mu0_by_Re = False
        # boundaryCondition.2D.* as (rho, u, v, p); zero when the table is absent
        self.inf_ruvp = [0.0] * 4
        self.inlet_ruvp = [0.0] * 4
        self.outlet_ruvp = [0.0] * 4
        # initialCondition
        self.initial_type = F.INIT_UNIFORM
        self.shock_tube = [0.0, 0.0, 1.0, 0.0]         # shock_x, shock_y, shock_normal_x, shock_normal_y
        self.dsr = [0.0, 0.0, 60.0]                    # shock_x, shock_y, shock_angle_degree
        self.vortex = None                             # vortex_x, vortex_y, vortex_strength (type 2)
        # output
        self.step_per_print = 50
        self.step_per_output_field = 50
        self.step_per_output_hist = 50
        self.start_output_field = 0
        self.maxIteration = 5000
        self.tolerace_residual = 1.0e-7
        self.volume_preset = PRESET_RUVP_MA
        self.boundary_preset = PRESET_RUVP_CP
        # physicsModel, time, inviscid_flux_method
        self.equation = F.EQ_EULER
        self.strategy = 0
        self.is_steady = False
        self.CFL = 0.6
        self.CFL_steady = 0.95
        self.max_physical_time = 20.0
        self.time_advance = F.EVO_EXPLICIT
        self.flag_reconstruct = F.REC_CONSTANT
        self.flag_gradient = F.GRA_LEASTSQUARE
        self.flux_conservation_scheme = F.SOL_ROE
        self.flux_limiter = 2  # _LIM_minmod
        # [device]
        self.precision = F.PREC_FP64
        self.renumber = F.RENUMBER_RCM
        self.kernels = F.KERNELS_AUTO
        self.tile_cells = 0        # 0: library default, -1: autotune at start
        self.tile_threads = 0      # work-group size of the tile kernels, 0: library default
        self.use_graphs = False
        self.venkat_K = None
        self.wall_temperature = None

    def params(self):
        """fvmr_params of this configuration (unsteady CFL; the steady driver uses CFL_steady)."""
        kw = dict(gamma=self.gamma, R=self.R, Pr=self.Pr, T0=self.T0, mu0=self.mu0, CFL=self.CFL,
                  equation=self.equation, flux_scheme=self.flux_conservation_scheme,
                  reconstruct=self.flag_reconstruct, gradient=self.flag_gradient, limiter=self.flux_limiter,
                  time_advance=self.time_advance, precision=self.precision, renumber=self.renumber,
                  kernels=self.kernels, tile_cells=max(0, self.tile_cells), tile_threads=self.tile_threads,
                  inf_ruvp=self.inf_ruvp, inlet_ruvp=self.inlet_ruvp, outlet_ruvp=self.outlet_ruvp,
                  dsr_shock_x=self.dsr[0], dsr_shock_y=self.dsr[1])
        if self.venkat_K is not None:
            kw["venkat_K"] = self.venkat_K
        if self.wall_temperature is not None:
            kw["wall_temperature"] = self.wall_temperature
        return default_params(**kw)

    def init_args(self):
        """FieldInitializer arguments of fvmr_field_init for the initial type."""
        if self.initial_type == F.INIT_ISENTROPIC_VORTEX:
            return tuple(self.vortex)
        if self.initial_type == F.INIT_SHOCK_TUBE:
            return tuple(self.shock_tube)
        if self.initial_type == F.INIT_DSR:
            return tuple(self.dsr)
        return ()


def _boundary_state(tr, cfg, parent):
    """getValue_boundaryCondition2D (TomlFileManager.cpp:200-265): (rho, u, v, p) or None when the
    table has no input_mode (the reference keeps its default)."""
    if not tr.has(parent + ".input_mode"):
        return None
    mode = tr.get(parent + ".input_mode", int, -1)
    if mode == 0:
        aoa = tr.get(parent + ".AOA", float) if tr.has(parent + ".AOA") else tr.get(parent + ".AoA", float)
        ma = tr.get(parent + ".Ma", float)
        if ma is None or aoa is None:
            return None
        return list(ruvp_from_mach(ma, aoa, cfg.T0, cfg.p0, cfg.gamma, cfg.R))
    if mode == 1:
        v = [tr.get(parent + "." + k, float) for k in ("rho", "u", "v", "p")]
        return None if None in v else v
    if mode == 2:
        U = [tr.get(parent + "." + k, float) for k in ("rho", "rhou", "rhov", "rhoE")]
        if None in U:
            return None
        # U2ruvp_host (FVMR/src/math/PhysicalKernel.h): gamma of [constant] (1.4 when absent)
        g = tr.get("constant.gamma", float, 1.4)
        rho = U[0]
        u, v = U[1] / rho, U[2] / rho
        return [rho, u, v, rho * (g - 1) * (U[3] / rho - (u * u + v * v) * 0.5)]
    if mode == 3:
        v = [tr.get(parent + "." + k, float) for k in ("rho", "u", "angle_degree", "p")]
        if None in v:
            return None
        ang = v[2] * math.pi / 180.0
        return [v[0], v[1] * math.cos(ang), v[1] * math.sin(ang), v[3]]
    tr.errors.append(f"Invalid input mode at {parent}.\n0- Ma,AoA \n1- rho,u,v,p \n2- rho, rhou, rhov, rhoE \n"
                     "3- rho, u, angle_degree, p")
    return None


def _enum(tr, key, names, default):
    """A [device] key given as an integer code or one of names (case-insensitive)."""
    if not tr.has(key):
        return default
    v = tr._raw(key)
    if isinstance(v, str):
        code = names.get(v.lower())
        if code is None:
            tr.errors.append(f'Parameter "{key}" must be one of {sorted(names)}.')
            return default
        return code
    if type(v) is int and v in names.values():
        return v
    tr.errors.append(f'Parameter "{key}" gets a wrong value type.')
    return default


def read_config(path):
    """Parse input.toml into a Config; raises ConfigError listing every failed key."""
    with open(path, "rb") as f:
        try:
            tree = tomli.load(f)
        except tomli.TOMLDecodeError as e:
            raise ConfigError(f"parse_exception: {e}") from None
    return config_from_tree(tree)


def config_from_tree(tree):
    tr = _Tree(tree)
    c = Config()
    g = tr.get
    # basic
    c.continue_ = g("basic.continue", bool)
    c.dimension = g("basic.dimension", int)
    c.filename = g("basic.filename", str)
    c.meshFileType = g("basic.meshFileType", str)
    c.useGPU = g("basic.useGPU", int)
    c.isDebugMode = g("basic.isDebugMode", bool)
    # render (optional)
    c.show_gui = g("render.show_gui", bool, c.show_gui)
    c.range_1 = g("render.range_1", float, c.range_1)
    c.range_2 = g("render.range_2", float, c.range_2)
    c.enable_print = g("render.enable_print", bool, c.enable_print)
    c.enable_write_file = g("render.enable_write_file", bool, c.enable_write_file)
    # constant; referenceArea is scaled with the mesh
    c.T0 = g("constant.T0", float)
    c.p0 = g("constant.p0", float)
    c.Pr = g("constant.Pr", float)
    c.gamma = g("constant.gamma", float)
    c.referenceArea = g("constant.referenceArea", float)
    c.mesh_scale_factor = g("constant.mesh_scale_factor", float, c.mesh_scale_factor)
    if c.referenceArea is not None and c.mesh_scale_factor is not None:
        c.referenceArea *= c.mesh_scale_factor
    if None in (c.T0, c.p0, c.gamma):
        raise ConfigError("\n".join(tr.errors))
    for name in ("inf", "inlet", "outlet"):
        r = _boundary_state(tr, c, "boundaryCondition.2D." + name)
        if r is not None:
            setattr(c, name + "_ruvp", [float(x) for x in r])
    # initial condition
    c.initial_type = g("initialCondition.type", int)
    if c.initial_type == F.INIT_SHOCK_TUBE:
        for i, k in enumerate(("shock_x", "shock_y", "shock_normal_x", "shock_normal_y")):
            c.shock_tube[i] = g("initialCondition.shockTube." + k, float, c.shock_tube[i])
    elif c.initial_type == F.INIT_DSR:
        for i, k in enumerate(("shock_x", "shock_y", "shock_angle_degree")):
            c.dsr[i] = g("initialCondition.doubleShockReflection." + k, float)
        if c.dsr[2] is not None and not 0.0 <= c.dsr[2] <= 90.0:
            tr.errors.append("shock_angle_degree out of range [0,90].")
    elif c.initial_type == F.INIT_ISENTROPIC_VORTEX:
        c.vortex = [g("initialCondition.isentropicVortex." + k, float, None)
                    for k in ("vortex_x", "vortex_y", "vortex_strength")]
    elif c.initial_type is not None and c.initial_type != F.INIT_UNIFORM:
        tr.errors.append(f"Invalid initialCondition type {c.initial_type}.")
    # mu0 and Re (read_mu0_Re_from_config): reference density, speed and length
    if c.initial_type in (F.INIT_SHOCK_TUBE, F.INIT_DSR):
        ui, uo = math.hypot(*c.inlet_ruvp[1:3]), math.hypot(*c.outlet_ruvp[1:3])
        rho_ref, U_ref = (c.inlet_ruvp[0], ui) if ui > uo else (c.outlet_ruvp[0], uo)
    else:
        rho_ref, U_ref = c.inf_ruvp[0], math.hypot(*c.inf_ruvp[1:3])
    L_ref = c.referenceArea if c.referenceArea is not None else 1.0
    c.calculate_mu0_by_Re = g("constant.calculate_mu0_by_Re", bool, False)
    if c.calculate_mu0_by_Re:
        c.Re = g("constant.Re", float)
        if c.Re is not None:
            c.mu0 = rho_ref * U_ref * L_ref / c.Re
    else:
        c.mu0 = g("constant.mu0", float)
        if c.mu0 is not None:
            c.Re = rho_ref * U_ref * L_ref / c.mu0 if c.mu0 != 0.0 else math.inf
    # output
    c.step_per_print = g("output.step_per_print", int)
    c.step_per_output_field = g("output.step_per_output_field", int)
    c.step_per_output_hist = g("output.step_per_output_hist", int)
    c.start_output_field = g("output.start_output_field", int, c.start_output_field)
    c.maxIteration = g("output.maxIteration", int)
    c.tolerace_residual = g("output.tolerace_residual", float)
    c.volume_preset = g("output.outputScheme_volume_preset", int, c.volume_preset)
    c.boundary_preset = g("output.outputScheme_boundary_preset", int, c.boundary_preset)
    for k in ("step_per_print", "step_per_output_field", "step_per_output_hist"):
        if getattr(c, k) is not None and getattr(c, k) < 1:
            tr.errors.append(f"output.{k} must be at least 1.")
    for k in ("volume_preset", "boundary_preset"):
        if getattr(c, k) not in _PRESET_VARS:
            tr.errors.append("invalid preset @FieldWriter::FieldOutputScheme::updateBoolDataByPreset.")
    # physics, time, scheme
    c.equation = g("physicsModel.equation", int)
    c.strategy = g("time.strategy", int, 0)
    c.is_steady = g("time.is_steady", bool)
    if c.is_steady:
        c.CFL_steady = g("time.CFL_steady", float)
        c.CFL = g("time.CFL", float, c.CFL)
    else:
        c.CFL_steady = g("time.CFL_steady", float, c.CFL_steady)
        c.CFL = g("time.CFL", float)
    c.max_physical_time = g("time.max_physical_time", float)
    c.time_advance = g("time.time_advance", int)
    c.flag_reconstruct = g("inviscid_flux_method.flag_reconstruct", int)
    c.flag_gradient = g("inviscid_flux_method.flag_gradient", int)
    c.flux_conservation_scheme = g("inviscid_flux_method.flux_conservation_scheme", int)
    c.flux_limiter = g("inviscid_flux_method.flux_limiter", int)
    # [device]: keys this library adds (the reference has no such table)
    c.precision = _enum(tr, "device.precision", {"fp64": F.PREC_FP64, "mixed": F.PREC_MIXED}, c.precision)
    c.renumber = _enum(tr, "device.renumber", {"none": F.RENUMBER_NONE, "rcm": F.RENUMBER_RCM,
                                               "hilbert": F.RENUMBER_HILBERT}, c.renumber)
    c.kernels = _enum(tr, "device.kernels", {"auto": F.KERNELS_AUTO, "cell": F.KERNELS_CELL,
                                             "tile": F.KERNELS_TILE}, c.kernels)
    if tr.has("device.tile_cells") and tr._raw("device.tile_cells") == "auto":
        c.tile_cells = -1
    else:
        c.tile_cells = g("device.tile_cells", int, c.tile_cells)
    c.tile_threads = g("device.tile_threads", int, c.tile_threads)
    c.use_graphs = g("device.use_graphs", bool, c.use_graphs)
    c.venkat_K = g("device.venkat_K", float, None)
    c.wall_temperature = g("device.wall_temperature", float, None)
    # what the device path cannot run
    if c.dimension not in (None, 2):
        tr.errors.append("basic.dimension must be 2.")
    if c.meshFileType is not None and c.meshFileType != "su2":
        tr.errors.append(f"Invalid mesh file type: {c.meshFileType}")
    if c.strategy == 1:
        tr.errors.append("time.strategy 1 (residual-driven CFL) is not provided; use 0.")
    if tr.errors:
        raise ConfigError("\n".join(tr.errors))
    return c


# ---------------------------------------------------------------------------------------------
# output files
# ---------------------------------------------------------------------------------------------

def _quote_list(names):
    return ", ".join(f'"{n}"' for n in names)


def _node_columns(xy, ruvp, cp, ma, preset):
    cols = [xy[0], xy[1], ruvp[0], ruvp[1], ruvp[2], ruvp[3]]
    if preset == PRESET_RUVP_CP:
        cols.append(cp)
    elif preset == PRESET_RUVP_MA:
        cols.append(ma)
    return np.stack(cols, axis=1)


def write_tecplot_volume(path, t, title, node_xy, element_nodes, ruvp, cp, ma, preset=PRESET_RUVP_MA):
    """FieldWriter::write_tecplot_volume_file (FieldWriter.cpp:44-96): point data, one
    fequadrilateral per element listing its nodes 0, 1, 2, 2 (1-based)."""
    ne = element_nodes.shape[1]
    with open(path, "w") as f:
        f.write(f'title= "{title}"\n')
        f.write("variables= " + _quote_list(("x", "y") + _PRESET_VARS[preset]) + "\n")
        f.write(f'zone t= "volume field", nodes= {node_xy.shape[1]}, elements= {ne}, datapacking= point, '
                f"zonetype= fequadrilateral, solutiontime= {t:.6e}\n")
        np.savetxt(f, _node_columns(node_xy, ruvp, cp, ma, preset), fmt="%.6e", delimiter=" ")
        en = element_nodes[:3].T.astype(np.int64) + 1
        np.savetxt(f, np.concatenate([en, en[:, 2:3]], axis=1), fmt="%d", delimiter=" ")


def write_tecplot_boundary(path, t, sets, edge_nodes, node_xy, ruvp, cp, ma, preset=PRESET_RUVP_CP):
    """FieldWriter::write_tecplot_boundary_file (FieldWriter.cpp:98-172): one zone per edge set
    ("boundary <ID> <name>", i = number of edges); an edge's first node is written only when it
    differs from the previous edge's second node."""
    vals = _node_columns(node_xy, ruvp, cp, ma, preset)
    with open(path, "w") as f:
        f.write('title= "tecplot boundary file"\n')
        f.write("variables= " + _quote_list(("x", "y") + _PRESET_VARS[preset]) + "\n")
        for set_id, name, edges in sets:
            f.write(f'zone t= "boundary {set_id} {name}", solutiontime= {t:.6e}, strandid= 0, '
                    f"i= {len(edges)}, j= 1, f= point\n")
            n0 = edge_nodes[0][edges]
            n1 = edge_nodes[1][edges]
            prev1 = np.concatenate([[-1], n1[:-1]])
            rows = []
            for a, b, p in zip(n0, n1, prev1):
                if a != p:
                    rows.append(a)
                rows.append(b)
            if rows:
                np.savetxt(f, vals[np.asarray(rows)], fmt="%.6e", delimiter=" ")


class HistWriter:
    """FieldWriter::write_tecplot_hist_file (FieldWriter.cpp:207-262,555-594): the header is
    written on the first call when the file does not exist or the run did not continue."""

    NAMES = ("Iteration", "physical_time", "R_1", "R_2", "R_3", "R_4", "C_L", "C_D")

    def __init__(self, path, continued):
        self.path = path
        self.continued = continued
        self.called = False

    def write(self, iteration, t, residual, CL, CD):
        if not self.called and (not os.path.exists(self.path) or not self.continued):
            with open(self.path, "a") as f:
                f.write('title= "tecplot hist file"\n')
                f.write("variables= " + _quote_list(self.NAMES) + "\n")
                f.write('zone t= "zone title", f= point\n')
        self.called = True
        with open(self.path, "a") as f:
            vals = [t] + [float(r) for r in residual] + [CL, CD]
            f.write(f"{int(iteration)} " + " ".join(f"{v:.6e}" for v in vals) + "\n")


def find_continue_file(directory, filename):
    """ContinueFileReader::readContinueFile_1's lookup (ContinueFileReader.cpp:26-57): the
    pause_<filename>[N].dat with the largest N > 0, or None."""
    best, best_n = None, 0
    prefix = "pause_" + filename
    try:
        names = sorted(os.listdir(directory))
    except OSError:
        return None
    for name in names:
        if not name.startswith(prefix):
            continue
        m = re.fullmatch(r"\[(-?\d+)\]\.dat", name[len(prefix):])
        if m and int(m.group(1)) > best_n:
            best, best_n = name, int(m.group(1))
    return os.path.join(directory, best) if best else None


def _step_name(istep):
    """COutput::updateFileName's "[%04d]" (COutput.cpp:13-14)."""
    return f"[{istep:04d}]"


# ---------------------------------------------------------------------------------------------
# driver
# ---------------------------------------------------------------------------------------------

class Driver:
    """CDriver::start on the device path: read the case, run steps, write the reference's files.

    Steps between output steps run as one fvmr_advance (dt and t stay on the device); an output
    step runs as one fvmr_step.  The print residual, the stop rules and the files follow the
    reference: see run()."""

    def __init__(self, config_path, log=print):
        self.config_path = os.path.abspath(config_path)
        self.dir = os.path.dirname(self.config_path)
        self.cfg = read_config(self.config_path)
        self.log = log if log is not None else (lambda *a, **k: None)
        self.files = []

    # CInput::readField_1 (CInput.cpp:102-121)
    def read_field(self):
        c = self.cfg
        self.params = c.params()
        if c.is_steady:
            self.params.CFL = c.CFL_steady
        self.t0, self.istep0 = 0.0, 0
        path = find_continue_file(self.dir, c.filename) if c.continue_ else None
        if path is not None:
            self.mesh, self.U, self.t0, self.istep0, cfl = Mesh.read_continue(path, c.equation)
            if cfl != 0.0:
                self.params.CFL = cfl   # the CFL saved in the file (ContinueFileReader.cpp:99-104)
            self.continued = True
            self.log(f"continue from: {os.path.basename(path)}")
        else:
            if c.continue_:
                self.log("failed to read continue file(pause_*.dat), will read mesh file.")
            mesh_path = os.path.join(self.dir, c.filename + "." + c.meshFileType)
            self.mesh = Mesh.read_su2(mesh_path, c.equation, c.mesh_scale_factor)
            if c.initial_type == F.INIT_ISENTROPIC_VORTEX and (c.vortex is None or None in c.vortex):
                raise ConfigError("initialCondition.isentropicVortex.{vortex_x,vortex_y,vortex_strength} are required")
            self.U = field_init(self.mesh, self.params, c.initial_type, c.init_args())
            self.continued = False

    def _flags(self, istep):
        """modifySignalPack1_output (CDriver.cpp:203-233) without the wall-time snapshot."""
        c = self.cfg
        p = istep == 1 or istep % c.step_per_print == 0
        tec = istep == 1 or (istep % c.step_per_output_field == 0 and istep >= c.start_output_field)
        hist = istep == 1 or istep % c.step_per_output_hist == 0
        return p or tec, tec, hist

    def _next_output_step(self, istep):
        """The first step after istep whose flags are set, or maxIteration."""
        c = self.cfg
        cand = [c.maxIteration]
        if istep < 1:
            cand.append(1)
        cand.append((istep // c.step_per_print + 1) * c.step_per_print)
        cand.append((istep // c.step_per_output_hist + 1) * c.step_per_output_hist)
        s = (istep // c.step_per_output_field + 1) * c.step_per_output_field
        if s < c.start_output_field:
            s = -(-c.start_output_field // c.step_per_output_field) * c.step_per_output_field
        cand.append(s)
        return max(istep + 1, min(cand))

    def _outputs(self, solver):
        """Node field and wall force coefficients of the current state (FieldWriter::update_nodeField,
        ReferenceData::calculate_edgeSet_force, HistData::set_CDCL_wall)."""
        ruvp, cp, ma = solver.node_field()
        CL = CD = 0.0
        walls = [e for _, typ, e in self.sets if typ in _WALLS]
        if walls:
            cx, cy = solver.surface_force(self.cfg.referenceArea, walls)
            CD, CL = float(np.sum(cx)), float(np.sum(cy))
        return ruvp, cp, ma, CL, CD

    def _write(self, solver, istep, t, res, out, tec, hist, cont, nan):
        c = self.cfg
        base = os.path.join(self.dir, c.filename)
        if cont:
            pre = "nan_" if nan else "pause_"
            path = os.path.join(self.dir, pre + c.filename + _step_name(istep) + ".dat")
            self.mesh.write_continue(path, solver.get_state(), t, istep, self.params.CFL)
            self.files.append(path)
        if tec or hist:
            ruvp, cp, ma, CL, CD = out
        if tec:
            path = base + _step_name(istep) + ".dat"
            write_tecplot_volume(path, t, "title", self.node_xy, self.element_nodes, ruvp, cp, ma, c.volume_preset)
            bpath = base + "_tec_boundary.dat"
            write_tecplot_boundary(bpath, t, self.zone_sets, self.edge_nodes, self.node_xy, ruvp, cp, ma,
                                   c.boundary_preset)
            self.files += [path, bpath]
        if hist:
            self.hist.write(istep, t, res, CL, CD)

    def run(self, max_steps=None):
        """Run the case to its stop signal; returns a summary dict.

        Per step, as CDriver::start: step; output flags of the step; the stop test
        (modifySignalPack2_pause: a NaN density residual, t > T or istep >= maxIteration, a density
        residual <= 1e-10) on the residual of the last print or write step; on print steps the
        infinity norm of the last rhs (updateResidualVector); on write steps the files, whose node
        fields and force coefficients are those of the state at the start of the step, as the
        reference forms them from the primitives refreshed at the start of the step.  The wall-time
        snapshot (every 1800 s) and t > T inside a run of quiet steps are tested at the end of that
        run.  max_steps (not a reference key) caps the steps of this call."""
        c = self.cfg
        self.read_field()
        if c.tile_cells == -1:
            self.params.tile_threads, self.params.tile_cells, _ = autotune(self.mesh, self.params)
        a = self.mesh.arrays()
        self.node_xy, self.element_nodes, self.edge_nodes = a["node_xy"], a["element_nodes"], a["edge_nodes"]
        self.sets = self.mesh.boundary_sets()
        set_ids = [int(a["edge_setID"][e[0]]) if len(e) else s + 1 for s, (_, _, e) in enumerate(self.sets)]
        self.zone_sets = [(sid, name, e) for sid, (name, _, e) in zip(set_ids, self.sets)]
        self.hist = HistWriter(os.path.join(self.dir, c.filename + "_hist.dat"), self.continued)
        if not c.useGPU:
            self.log("basic.useGPU = 0: the device path runs the case")
        solver = GPUSolver2(self.mesh, self.params)
        try:
            if c.use_graphs:
                solver.use_graphs(True)
            solver.set_state(self.U)
            return self._loop(solver, max_steps)
        finally:
            solver.close()

    def _loop(self, solver, max_steps):
        c = self.cfg
        T = c.max_physical_time
        istep, t = self.istep0, self.t0
        res = np.ones(4)          # GPUSolver2::residualVector starts at {1, 1, 1, 1}
        clock0 = time.perf_counter()
        last_snapshot = 0.0
        signal = None
        stop_at = None if max_steps is None else istep + int(max_steps)

        def pause_pending(r):
            return math.isnan(r[0]) or r[0] <= EPSILON

        while signal is None:
            target = istep + 1 if pause_pending(res) else self._next_output_step(istep)
            if stop_at is not None:
                target = min(target, stop_at)
            quiet = target - istep - 1
            if quiet > 0:
                t = self._quiet(solver, quiet, t, T)
                istep += quiet
                if t > T:       # only by rounding at the clamp step
                    signal = "TimeReached"
                    self._stop_write(solver, istep, t, res, nan=False)
                    break
            # the output step
            pr, tec, hist = self._flags(target)
            final = pause_pending(res) or target >= c.maxIteration
            pre = self._outputs(solver) if (c.enable_write_file and (tec or hist or final)) else None
            t = self._one(solver, t, T)
            istep = target
            now = time.perf_counter() - clock0
            cont = False
            if now - last_snapshot > SNAPSHOT_SECONDS:
                cont, last_snapshot = True, now
            nan = False
            if math.isnan(res[0]):
                signal, cont, tec, nan = "NanDetected", True, True, True
            elif t > T or istep >= c.maxIteration:
                signal, cont, tec = "TimeReached", True, True
            elif res[0] <= EPSILON:
                signal, cont, tec = "StableReached", True, True
            write = c.enable_write_file and (cont or tec or hist)
            if (c.enable_print and pr) or write:
                res = self._residual(solver)
            if c.enable_print and pr:
                rate = (istep - self.istep0) / max(now, 1e-9)
                self.log(f"step = {istep}\tt = {t:.6e}\tresidual_rho = {res[0]:.6e}\tCFL = {self.params.CFL}"
                         f"\tmean speed = {rate:.1f} step/s")
            if write:
                out = pre if (pre is not None or not (tec or hist)) else self._outputs(solver)
                self._write(solver, istep, t, res, out, tec, hist, cont, nan)
            if stop_at is not None and istep >= stop_at and signal is None:
                signal = "StepsDone"
        self.log(f"stop: {signal} at step {istep}, t = {t:.6e}")
        return {"signal": signal, "istep": istep, "t": t, "residual": res.tolist(), "files": list(self.files),
                "seconds": time.perf_counter() - clock0}

    def _stop_write(self, solver, istep, t, res, nan):
        if self.cfg.enable_write_file:
            res[:] = self._residual(solver)
            self._write(solver, istep, t, res, self._outputs(solver), True, False, True, nan)

    # one step kind per mode: unsteady (global dt, clamp to T) or steady (per-cell pseudo-time)
    def _quiet(self, solver, n, t, T):
        if self.cfg.is_steady:
            solver.steady_solve(n, 0.0)
            return t
        t, _ = solver.advance(n, t, T)
        return t

    def _one(self, solver, t, T):
        if self.cfg.is_steady:
            solver.steady_solve(1, 0.0)
            return t
        t, _ = solver.iteration_device(t, T)
        return t

    def _residual(self, solver):
        if self.cfg.is_steady:
            r = solver.stage_rhs()
            return np.max(np.abs(r), axis=1)
        return solver.updateResidualVector(F.NORM_INF)


def run_case(config_path, max_steps=None, log=print):
    """Read input.toml at config_path and run the case (CDriver::start)."""
    return Driver(config_path, log=log).run(max_steps=max_steps)


# scheme settings every case of a batch shares (fvmr_create_batch: one kernel configuration)
_BATCH_SHARED = ("gamma", "R", "Pr", "Pr_viscous", "T0", "mu0", "equation", "flux_scheme", "reconstruct", "gradient",
                 "limiter", "time_advance", "viscosity_model", "precision", "renumber", "compat_inv2x2", "kernels",
                 "tile_cells", "tile_threads", "wall_temperature", "venkat_K", "dsr_shock_x", "dsr_shock_y",
                 "dsr_shock_speed")


def run_batch(config_paths, log=print):
    """Several input.toml cases on one batch handle (fvmr_create_batch): each case keeps its
    own mesh, field (or newest continue file), boundary states, CFL, end time and maxIteration;
    the cases must agree on the scheme settings.  All cases step together; each stops at its own
    maxIteration (or when t passes its max_physical_time), where its continue file
    pause_<filename>[%04d].dat is written as the single-case driver writes it at that stop.
    Output steps (Tecplot, hist, print residuals) and the residual stop rule are the
    single-case driver's (Driver.run).  Returns one summary per case."""
    log = log if log is not None else (lambda *a, **k: None)
    drivers = [Driver(p, log=None) for p in config_paths]
    if not drivers:
        raise ConfigError("run_batch needs at least one input.toml")
    for d in drivers:
        if d.cfg.is_steady:
            raise ConfigError(f"{d.config_path}: a batch runs unsteady cases only")
        if d.cfg.tile_cells == -1:
            raise ConfigError(f"{d.config_path}: give device.tile_cells a number for a batch")
        d.read_field()
    p0 = drivers[0].params
    for d in drivers[1:]:
        diff = [k for k in _BATCH_SHARED if getattr(d.params, k) != getattr(p0, k)]
        if diff:
            raise ConfigError(f"{d.config_path}: a batch needs the same {', '.join(diff)} as {drivers[0].config_path}")
    n = len(drivers)
    bs = BatchSolver([d.mesh for d in drivers], [d.params for d in drivers])
    try:
        for c, d in enumerate(drivers):
            bs.set_state(c, d.U)
        t = np.array([d.t0 for d in drivers], dtype=np.float64)
        T = np.array([d.cfg.max_physical_time for d in drivers], dtype=np.float64)
        istep = [d.istep0 for d in drivers]
        # the reference takes at least one step before it tests istep >= maxIteration
        left = [max(1, d.cfg.maxIteration - d.istep0) for d in drivers]
        out = [None] * n
        clock0 = time.perf_counter()
        while any(o is None for o in out):
            k = min(left[c] for c in range(n) if out[c] is None)
            t_new = bs.advance(k, t, T)
            for c in range(n):
                if out[c] is not None:
                    continue
                istep[c] += k
                left[c] -= k
                t[c] = t_new[c]
                if left[c] <= 0 or t[c] > T[c]:
                    d = drivers[c]
                    path = os.path.join(d.dir, "pause_" + d.cfg.filename + _step_name(istep[c]) + ".dat")
                    d.mesh.write_continue(path, bs.get_state(c), t[c], istep[c], d.params.CFL)
                    out[c] = {"signal": "TimeReached", "istep": istep[c], "t": float(t[c]), "files": [path]}
                    log(f"case {c} ({d.cfg.filename}): stop at step {istep[c]}, t = {t[c]:.6e}")
            # finished cases ride along with dt = 0 until the last case stops
            for c in range(n):
                if out[c] is not None:
                    T[c] = t[c]
        secs = time.perf_counter() - clock0
        for o in out:
            o["seconds"] = secs
        return out
    finally:
        bs.close()


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="Run FVMR input.toml cases on the GPU (several: one batch)")
    ap.add_argument("config", nargs="+", help="input.toml paths; each case's mesh and outputs are in its directory")
    ap.add_argument("--max-steps", type=int, default=None, help="one case only: cap the steps of this run")
    a = ap.parse_args(argv)
    if len(a.config) > 1:
        run_batch(a.config)
        return 0
    s = run_case(a.config[0], a.max_steps)
    return 0 if s["signal"] in ("TimeReached", "StableReached", "StepsDone") else 1


if __name__ == "__main__":
    raise SystemExit(main())
