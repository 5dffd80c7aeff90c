"""Host driver: input.toml parsing, the step loop's schedule and the output writers (CPU only).

The loop runs against a recording stand-in for the device solver; its schedule (which steps run
as one batch, where outputs and the stop fall) is compared with a plain per-step loop of
CDriver::start's rules (FVMR/src/drivers/CDriver.cpp:64-176,203-257)."""
import math
import os
import random
import shutil

import numpy as np
import pytest

import cfd_solver_2d_gpu_amd as F
from cfd_solver_2d_gpu_amd import driver as D
from helpers import GOLDEN, toml_tree as _tree, tree_set as _set, write_toml


def _isentropic(Ma, AoA, T0=288.16, p0=101325.0, g=1.4, R=287.06):
    """Freestream of Ma and AoA from sea-level T0, p0 by the isentropic relations."""
    T = T0 / (1.0 + 0.5 * (g - 1.0) * Ma * Ma)
    p = p0 * (T / T0) ** (g / (g - 1.0))
    c = math.sqrt(g * R * T)
    a = math.radians(AoA)
    return [p / (R * T), Ma * c * math.cos(a), Ma * c * math.sin(a), p]


# ------------------------------------------------------------------------------------------
# config
# ------------------------------------------------------------------------------------------

def test_quick_start_configs():
    st = D.read_config(os.path.join(GOLDEN, "shocktube.input.toml"))
    assert st.filename == "shocktube_100" and st.initial_type == F.INIT_SHOCK_TUBE
    assert st.inlet_ruvp == [1.0, 0.0, 0.0, 1.0] and st.outlet_ruvp == [0.125, 0.0, 0.0, 0.1]
    assert st.shock_tube == [0.0, 0.0, 1.0, 0.0]      # FieldInitializer.h:23-26 defaults
    assert st.CFL == 2.5 and st.max_physical_time == 2.255788 and st.maxIteration == 1000
    assert (st.equation, st.flux_conservation_scheme, st.flag_reconstruct, st.flag_gradient, st.flux_limiter,
            st.time_advance) == (F.EQ_EULER, F.SOL_ROE, F.REC_LINEAR, F.GRA_GREENGAUSS, F.LIM_BARTH, F.EVO_EXPLICIT)
    assert st.mu0 == 0.0          # both states at rest: U_ref = 0
    assert st.continue_ is True and st.is_steady is False

    nv = D.read_config(os.path.join(GOLDEN, "neuvillette.input.toml"))
    assert nv.inlet_ruvp == [8.0, 0.0, 8.25, 116.5] and nv.shock_tube == [0.0, 0.5, 0.0, 1.0]
    assert nv.mu0 == pytest.approx(8.0 * 8.25 * 1.0 / 6.5e6, rel=1e-15)     # faster of inlet / outlet

    cy = D.read_config(os.path.join(GOLDEN, "cylinder.input.toml"))
    ref = _isentropic(0.5, 0.0)
    assert np.allclose(cy.inf_ruvp, ref, rtol=1e-12, atol=1e-12)
    assert cy.mu0 == pytest.approx(ref[0] * ref[1] * 1.0 / 6.5e6, rel=1e-12)
    assert cy.equation == F.EQ_NS and cy.CFL == 2.2 and cy.start_output_field == 30000
    assert cy.referenceArea == 1.0

    ds = D.read_config(os.path.join(GOLDEN, "dsr.input.toml"))
    assert ds.initial_type == F.INIT_DSR and ds.dsr == [0.6, 0.0, 60.0]
    a = math.radians(-30.0)
    assert np.allclose(ds.inlet_ruvp, [8.0, 8.25 * math.cos(a), 8.25 * math.sin(a), 116.5], rtol=1e-15)
    assert ds.outlet_ruvp == [1.4, 0.0, 0.0, 1.0]
    p = ds.params()
    assert p.dsr_shock_x == 0.6 and p.dsr_shock_speed == 20.0 and p.CFL == 1.6

    iv = D.read_config(os.path.join(GOLDEN, "ivortex.input.toml"))
    assert iv.initial_type == F.INIT_ISENTROPIC_VORTEX and iv.vortex == [0.0, 0.0, 5.0]
    assert iv.inf_ruvp == [1.0, 1.0, 1.0, 1.0] and iv.init_args() == (0.0, 0.0, 5.0)
    assert iv.mu0 == pytest.approx(1.0 * math.sqrt(2.0) / 6.5e6, rel=1e-15)


def test_params_mapping():
    c = D.read_config(os.path.join(GOLDEN, "cylinder.input.toml"))
    p = c.params()
    assert (p.equation, p.flux_scheme, p.reconstruct, p.gradient, p.limiter, p.time_advance) == (
        F.EQ_NS, F.SOL_ROE, F.REC_LINEAR, F.GRA_GREENGAUSS, F.LIM_BARTH, F.EVO_EXPLICIT)
    assert p.CFL == 2.2 and p.gamma == 1.4 and p.Pr == 0.73 and p.T0 == 288.16 and p.mu0 == c.mu0
    assert list(p.inf_ruvp) == c.inf_ruvp and list(p.inlet_ruvp) == [0.0] * 4
    assert p.precision == F.PREC_FP64 and p.renumber == F.RENUMBER_RCM and p.tile_cells == 0


def test_boundary_input_modes():
    t = _tree("cylinder.input.toml")
    # mode 0 with AOA (upper case) taking precedence over AoA
    _set(t, "boundaryCondition.2D.inf", {"input_mode": 0, "Ma": 0.8, "AOA": 1.25, "AoA": 9.0})
    # mode 2: conservative variables
    _set(t, "boundaryCondition.2D.inlet", {"input_mode": 2, "rho": 1.2, "rhou": 0.6, "rhov": -0.3, "rhoE": 2.5})
    # mode 3: speed and angle
    _set(t, "boundaryCondition.2D.outlet", {"input_mode": 3, "rho": 2, "u": 3, "angle_degree": 90, "p": 4})
    c = D.config_from_tree(t)
    assert np.allclose(c.inf_ruvp, _isentropic(0.8, 1.25), rtol=1e-12)
    rho, u, v = 1.2, 0.5, -0.25
    p = (1.4 - 1.0) * (2.5 - 0.5 * rho * (u * u + v * v))
    assert np.allclose(c.inlet_ruvp, [rho, u, v, p], rtol=1e-14)
    assert np.allclose(c.outlet_ruvp, [2.0, 0.0, 3.0, 4.0], atol=1e-15)


def test_scale_factor_and_mu0_modes():
    t = _tree("cylinder.input.toml")
    _set(t, "constant.mesh_scale_factor", 2.5)
    _set(t, "constant.referenceArea", 2)          # integer accepted for a float key
    c = D.config_from_tree(t)
    assert c.referenceArea == 5.0
    ref = _isentropic(0.5, 0.0)
    assert c.mu0 == pytest.approx(ref[0] * ref[1] * 5.0 / 6.5e6, rel=1e-12)
    _set(t, "constant.calculate_mu0_by_Re", False)
    _set(t, "constant.mu0", 1.8e-5)
    c = D.config_from_tree(t)
    assert c.mu0 == 1.8e-5 and c.Re == pytest.approx(ref[0] * ref[1] * 5.0 / 1.8e-5, rel=1e-12)


def test_device_table():
    t = _tree("shocktube.input.toml")
    t["device"] = {"precision": "mixed", "renumber": "hilbert", "kernels": "tile", "tile_cells": "auto",
                   "use_graphs": True, "venkat_K": 0.3}
    _set(t, "inviscid_flux_method.flux_limiter", F.LIM_VENKATAKRISHNAN)
    c = D.config_from_tree(t)
    assert (c.precision, c.renumber, c.kernels, c.tile_cells, c.use_graphs) == (
        F.PREC_MIXED, F.RENUMBER_HILBERT, F.KERNELS_TILE, -1, True)
    p = c.params()
    assert p.venkat_K == 0.3 and p.tile_cells == 0 and p.limiter == F.LIM_VENKATAKRISHNAN
    t["device"] = {"precision": 1, "tile_cells": 200}
    c = D.config_from_tree(t)
    assert c.precision == F.PREC_MIXED and c.params().tile_cells == 200
    t["device"] = {"precision": "fp16"}
    with pytest.raises(D.ConfigError, match="device.precision"):
        D.config_from_tree(t)


def test_steady_reads_cfl_steady():
    t = _tree("ivortex.input.toml")
    _set(t, "time.is_steady", 1)                  # integer accepted for a bool key
    _set(t, "time.CFL", None)
    c = D.config_from_tree(t)
    assert c.is_steady is True and c.CFL_steady == 0.05
    _set(t, "time.CFL_steady", None)
    with pytest.raises(D.ConfigError, match="time.CFL_steady"):
        D.config_from_tree(t)


@pytest.mark.parametrize("key,value,message", [
    ("basic.filename", None, 'Required parameter "basic.filename"'),
    ("output.maxIteration", 1.5, 'Parameter "output.maxIteration" gets a wrong value type'),
    ("physicsModel.equation", True, 'Parameter "physicsModel.equation" gets a wrong value type'),
    ("constant.gamma", "1.4", 'Parameter "constant.gamma" gets a wrong value type'),
    ("boundaryCondition.2D.inlet.input_mode", 7, "Invalid input mode at boundaryCondition.2D.inlet"),
    ("initialCondition.type", 9, "Invalid initialCondition type 9"),
    ("basic.meshFileType", "inp", "Invalid mesh file type: inp"),
    ("time.strategy", 1, "time.strategy 1"),
    ("output.step_per_print", 0, "output.step_per_print must be at least 1"),
    ("basic.dimension", 3, "basic.dimension must be 2"),
])
def test_config_errors(key, value, message):
    t = _tree("shocktube.input.toml")
    _set(t, key, value)
    with pytest.raises(D.ConfigError) as e:
        D.config_from_tree(t)
    assert message in str(e.value)


def test_config_errors_collected():
    t = _tree("dsr.input.toml")
    _set(t, "output.maxIteration", None)
    _set(t, "time.CFL", None)
    _set(t, "initialCondition.doubleShockReflection.shock_angle_degree", 95)
    with pytest.raises(D.ConfigError) as e:
        D.config_from_tree(t)
    s = str(e.value)
    assert '"output.maxIteration"' in s and '"time.CFL"' in s and "shock_angle_degree out of range" in s


def test_parse_error(tmp_path):
    p = tmp_path / "input.toml"
    p.write_text("[basic\nfilename = 1\n")
    with pytest.raises(D.ConfigError, match="parse_exception"):
        D.read_config(str(p))


# ------------------------------------------------------------------------------------------
# files
# ------------------------------------------------------------------------------------------

def test_find_continue_file(tmp_path):
    for n in ("pause_case[5].dat", "pause_case[0020].dat", "pause_case[0].dat", "nan_case[0040].dat",
              "pause_casex[30].dat", "pause_case[7].txt", "case[0100].dat"):
        (tmp_path / n).write_text("")
    assert D.find_continue_file(str(tmp_path), "case") == str(tmp_path / "pause_case[0020].dat")
    assert D.find_continue_file(str(tmp_path), "other") is None
    assert D.find_continue_file(str(tmp_path / "missing"), "case") is None


def test_tecplot_volume(tmp_path):
    xy = np.array([[0.0, 1.0, 0.0, 1.0], [0.0, 0.0, 1.0, 1.0]])
    en = np.array([[0, 1], [1, 3], [2, 2], [-1, -1]])
    ruvp = np.arange(16, dtype=float).reshape(4, 4) + 0.5
    cp = np.array([0.1, 0.2, 0.3, 0.4])
    ma = np.array([1.0, 2.0, 3.0, 4.0])
    p = tmp_path / "v.dat"
    D.write_tecplot_volume(str(p), 0.25, "title", xy, en, ruvp, cp, ma)
    lines = p.read_text().splitlines()
    assert lines[0] == 'title= "title"'
    assert lines[1] == 'variables= "x", "y", "rho", "u", "v", "p", "Ma"'
    assert lines[2] == ('zone t= "volume field", nodes= 4, elements= 2, datapacking= point, '
                        'zonetype= fequadrilateral, solutiontime= 2.500000e-01')
    assert lines[3] == "0.000000e+00 0.000000e+00 5.000000e-01 4.500000e+00 8.500000e+00 1.250000e+01 1.000000e+00"
    assert lines[7:] == ["1 2 3 3", "2 4 3 3"]
    D.write_tecplot_volume(str(p), 0.0, "title", xy, en, ruvp, cp, ma, D.PRESET_RUVP)
    assert p.read_text().splitlines()[1] == 'variables= "x", "y", "rho", "u", "v", "p"'


def test_tecplot_boundary(tmp_path):
    xy = np.array([[0.0, 1.0, 2.0, 3.0, 4.0], [0.0] * 5])
    ruvp = np.ones((4, 5))
    cp = np.arange(5, dtype=float)
    edge_nodes = np.array([[0, 1, 3, 4], [1, 2, 4, 0]])
    p = tmp_path / "b.dat"
    # set 1: chained edges 0-1, 1-2 then a break to 3-4 and 4-0
    D.write_tecplot_boundary(str(p), 1.0, [(1, "wall", np.array([0, 1, 2, 3])), (2, "far", np.array([], int))],
                             edge_nodes, xy, ruvp, cp, cp)
    lines = p.read_text().splitlines()
    assert lines[1] == 'variables= "x", "y", "rho", "u", "v", "p", "Cp"'
    assert lines[2] == 'zone t= "boundary 1 wall", solutiontime= 1.000000e+00, strandid= 0, i= 4, j= 1, f= point'
    xs = [float(l.split()[0]) for l in lines[3:9]]
    assert xs == [0.0, 1.0, 2.0, 3.0, 4.0, 0.0]
    assert lines[9].startswith('zone t= "boundary 2 far"') and "i= 0" in lines[9] and len(lines) == 10


def test_hist_writer(tmp_path):
    p = tmp_path / "h.dat"
    h = D.HistWriter(str(p), continued=False)
    h.write(1, 0.5, [1, 2, 3, 4], 0.25, -0.125)
    h.write(50, 1.5, [1e-3, 2e-3, 3e-3, 4e-3], 0.5, 0.0)
    lines = p.read_text().splitlines()
    assert lines[:3] == ['title= "tecplot hist file"',
                         'variables= "Iteration", "physical_time", "R_1", "R_2", "R_3", "R_4", "C_L", "C_D"',
                         'zone t= "zone title", f= point']
    assert lines[3] == "1 5.000000e-01 1.000000e+00 2.000000e+00 3.000000e+00 4.000000e+00 2.500000e-01 -1.250000e-01"
    assert lines[4].startswith("50 1.500000e+00 1.000000e-03")
    # a continued run appends to an existing file without a new header
    D.HistWriter(str(p), continued=True).write(51, 2.0, [0, 0, 0, 0], 0, 0)
    assert len(p.read_text().splitlines()) == 6
    # a fresh run writes a header even into an existing file
    D.HistWriter(str(p), continued=False).write(1, 0.0, [0, 0, 0, 0], 0, 0)
    assert p.read_text().splitlines()[6] == 'title= "tecplot hist file"'


# ------------------------------------------------------------------------------------------
# the step loop
# ------------------------------------------------------------------------------------------

class FakeSolver:
    """Records what the driver asks of the device solver.  dt is constant; the residual of a
    step is res_of(step) for steps counted from the start of the process."""
    instances = []

    def __init__(self, mesh, params, dt=0.01, res_of=None):
        self.mesh, self.params = mesh, params
        self.dt = dt
        self.res_of = res_of or (lambda s: 1.0 / (1 + s))
        self.steps = 0
        self.calls = []
        self.U = None
        FakeSolver.instances.append(self)

    def use_graphs(self, on=True):
        pass

    def set_state(self, U):
        self.U = np.array(U, dtype=float)

    def get_state(self):
        return self.U.copy()

    def _step(self, t, T):
        d = self.dt
        if t + d > T:
            d = T - t
        self.steps += 1
        self.U[0] += 1.0
        return t + d

    def advance(self, n, t, T):
        self.calls.append(("advance", n))
        for _ in range(n):
            t = self._step(t, T)
        return t, None

    def iteration_device(self, t, T):
        self.calls.append(("step", 1))
        return self._step(t, T), None

    def steady_solve(self, n, tol):
        self.calls.append(("steady", n))
        self.steps += n
        return n, 1.0, 1.0

    def stage_rhs(self):
        return np.full((4, self.U.shape[1]), self.res_of(self.steps))

    def updateResidualVector(self, norm=F.NORM_INF):
        self.calls.append(("residual", self.steps))
        return np.full(4, self.res_of(self.steps))

    def node_field(self):
        self.calls.append(("nodes", self.steps))
        nv = self.mesh.soa.num_node
        return np.zeros((4, nv)), np.zeros(nv), np.zeros(nv)

    def surface_force(self, S_ref, sets):
        self.calls.append(("force", self.steps))
        return np.zeros(len(sets)), np.zeros(len(sets))

    def close(self):
        pass


def per_step_schedule(c, istep0, t0, dt, res_of, max_steps=None):
    """CDriver::start step by step: (outputs per step {istep: (print, tec, hist, cont)}, stop step,
    signal, steps at which node fields are taken)."""
    istep, t, res = istep0, t0, 1.0
    T = c.max_physical_time
    out, nodes = {}, []
    while True:
        istep += 1
        pr = istep == 1 or istep % c.step_per_print == 0
        tec = istep == 1 or (istep % c.step_per_output_field == 0 and istep >= c.start_output_field)
        hist = istep == 1 or istep % c.step_per_output_hist == 0
        pr = pr or tec
        steps_before = istep - istep0 - 1
        d = dt if t + dt <= T else T - t
        t += d
        cont, sig = False, None
        if math.isnan(res):
            cont, tec, sig = True, True, "NanDetected"
        elif t > T or istep >= c.maxIteration:
            cont, tec, sig = True, True, "TimeReached"
        elif res <= 1e-10:
            cont, tec, sig = True, True, "StableReached"
        if pr or cont or tec or hist:
            res = res_of(istep - istep0)
        if tec or hist:
            nodes.append(steps_before)
        if pr or cont or tec or hist:
            out[istep] = (tec, hist, cont)
        if sig:
            return out, istep, sig, nodes
        if max_steps is not None and istep - istep0 >= max_steps:
            return out, istep, "StepsDone", nodes


def _case(tmp_path, golden_dir, **over):
    d = tmp_path / "case"
    d.mkdir()
    shutil.copyfile(golden_dir / "shocktube_100.su2", d / "shocktube_100.su2")
    t = _tree("shocktube.input.toml")
    _set(t, "basic.continue", False)
    for k, v in over.items():
        _set(t, k, v)
    write_toml(d / "input.toml", t)
    return d


@pytest.mark.parametrize("seed", range(16))
def test_loop_schedule_matches_per_step(tmp_path, golden_dir, monkeypatch, seed):
    rng = random.Random(seed)
    over = {"output.step_per_print": rng.choice([1, 7, 50]), "output.step_per_output_field": rng.choice([3, 40, 100]),
            "output.step_per_output_hist": rng.choice([5, 25, 60]), "output.start_output_field": rng.choice([0, 90]),
            "output.maxIteration": rng.choice([1, 57, 230]), "time.max_physical_time": rng.choice([0.8, 100.0])}
    d = _case(tmp_path, golden_dir, **over)
    stable_at = rng.choice([None, 30, 61])
    nan_at = rng.choice([None, 20, 45])

    def res_of(s):
        if nan_at is not None and s >= nan_at:
            return float("nan")
        if stable_at is not None and s >= stable_at:
            return 1e-12
        return 1.0 / (1 + s)

    FakeSolver.instances.clear()
    monkeypatch.setattr(D, "GPUSolver2", lambda m, p: FakeSolver(m, p, 0.01, res_of))
    drv = D.Driver(str(d / "input.toml"), log=None)
    s = drv.run()
    fs = FakeSolver.instances[-1]
    out, stop, sig, nodes = per_step_schedule(drv.cfg, 0, 0.0, 0.01, res_of)
    assert (s["istep"], s["signal"]) == (stop, sig)
    assert fs.steps == stop
    # node fields are taken before the steps that write Tecplot or hist files (the time-limit stop
    # inside a batch of quiet steps aside, which the schedule above does not reach: dt divides T)
    assert [n for k, n in fs.calls if k == "nodes"] == nodes
    # each output step runs alone; quiet steps run in batches between them
    steps_with_output = sorted(out)
    ran = []
    i = 0
    for k, n in fs.calls:
        if k in ("advance", "step"):
            ran.append((k, i + 1, i + n))
            i += n
    singles = [a for k, a, b in ran if k == "step"]
    assert set(steps_with_output) <= set(singles)
    assert all(k == "step" or b - a + 1 >= 1 for k, a, b in ran)
    files = sorted(os.listdir(d))
    tec = [k for k, (t_, h, c) in out.items() if t_]
    cont = [k for k, (t_, h, c) in out.items() if c]
    for k in tec:
        assert f"shocktube_100[{k:04d}].dat" in files
    for k in cont:
        pre = "nan_" if sig == "NanDetected" and k == stop else "pause_"
        assert f"{pre}shocktube_100[{k:04d}].dat" in files
    nhist = sum(1 for k, (t_, h, c) in out.items() if h)
    hist = (d / "shocktube_100_hist.dat").read_text().splitlines() if nhist else []
    assert len(hist) == (nhist + 3 if nhist else 0)


def test_loop_continue_and_max_steps(tmp_path, golden_dir, monkeypatch):
    d = _case(tmp_path, golden_dir, **{"output.maxIteration": 500, "output.step_per_output_field": 1000})
    monkeypatch.setattr(D, "GPUSolver2", lambda m, p: FakeSolver(m, p, 0.001))
    s1 = D.Driver(str(d / "input.toml"), log=None).run(max_steps=120)
    assert s1["signal"] == "StepsDone" and s1["istep"] == 120
    # no continue file yet (none at 120: not a write step); write one as the reference would
    m = F.Mesh.read_su2(str(d / "shocktube_100.su2"))
    U = np.ones((4, m.num_element))
    m.write_continue(str(d / "pause_shocktube_100[0120].dat"), U, 0.12, 120, 1.7)
    drv = D.Driver(str(d / "input.toml"), log=None)
    drv.cfg.continue_ = True
    s2 = drv.run(max_steps=30)
    assert drv.continued and drv.istep0 == 120 and drv.t0 == 0.12 and drv.params.CFL == 1.7
    assert s2["istep"] == 150 and s2["t"] == pytest.approx(0.12 + 30 * 0.001)
    # the hist file of a continued run gets no second header
    hist = (d / "shocktube_100_hist.dat").read_text().splitlines()
    assert sum(1 for l in hist if l.startswith("title=")) == 1


class FakeBatch:
    """Stand-in for the batch handle: constant dt per case, state counts steps."""

    def __init__(self, meshes, params):
        self.U = [None] * len(meshes)
        self.dt = [0.001 * (c + 1) for c in range(len(meshes))]
        self.calls = []

    def set_state(self, c, U):
        self.U[c] = np.array(U, dtype=float)

    def get_state(self, c):
        return self.U[c].copy()

    def advance(self, n, t, T):
        self.calls.append(n)
        t = np.array(t, dtype=float)
        for c in range(len(t)):
            for _ in range(n):
                d = self.dt[c] if t[c] + self.dt[c] <= T[c] else T[c] - t[c]
                t[c] += d
                self.U[c][0] += 1.0 if d > 0 else 0.0
        return t

    def close(self):
        pass


def test_run_batch_stops_each_case(tmp_path, golden_dir, monkeypatch):
    dirs = []
    for name, mx in (("a", 30), ("b", 75), ("c", 50)):
        d = tmp_path / name
        d.mkdir()
        shutil.copyfile(golden_dir / "shocktube_100.su2", d / "shocktube_100.su2")
        t = _tree("shocktube.input.toml")
        _set(t, "basic.continue", False)
        _set(t, "output.maxIteration", mx)
        if name == "b":
            _set(t, "boundaryCondition.2D.inlet.rho", 2.0)
            _set(t, "time.CFL", 1.5)
        write_toml(d / "input.toml", t)
        dirs.append(d)
    monkeypatch.setattr(D, "BatchSolver", FakeBatch)
    out = D.run_batch([str(d / "input.toml") for d in dirs], log=None)
    assert [o["istep"] for o in out] == [30, 75, 50]
    assert [o["t"] for o in out] == pytest.approx([30 * 0.001, 75 * 0.002, 50 * 0.003])
    for d, o in zip(dirs, out):
        assert os.path.basename(o["files"][0]) == f"pause_shocktube_100[{o['istep']:04d}].dat"
        _, U, t, istep, cfl = F.Mesh.read_continue(o["files"][0])
        assert istep == o["istep"] and t == pytest.approx(o["t"])
        assert cfl == (1.5 if d.name == "b" else 2.5)
        cfg = D.read_config(str(d / "input.toml"))
        m = F.Mesh.read_su2(str(d / "shocktube_100.su2"))
        U0 = F.field_init(m, cfg.params(), cfg.initial_type, cfg.init_args())
        assert np.array_equal(U[0], U0[0] + o["istep"])   # stepped exactly to its stop (density + 1 per step)


def test_run_batch_refuses_mixed_schemes(tmp_path, golden_dir):
    dirs = []
    for name, lim in (("a", 1), ("b", 0)):
        d = tmp_path / name
        d.mkdir()
        shutil.copyfile(golden_dir / "shocktube_100.su2", d / "shocktube_100.su2")
        t = _tree("shocktube.input.toml")
        _set(t, "basic.continue", False)
        _set(t, "inviscid_flux_method.flux_limiter", lim)
        write_toml(d / "input.toml", t)
        dirs.append(d)
    with pytest.raises(D.ConfigError, match="limiter"):
        D.run_batch([str(d / "input.toml") for d in dirs], log=None)
