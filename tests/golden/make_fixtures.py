"""Copies the reference's own data files used as fixtures into tests/golden (gzip).

Run in a checkout where the reference repository is at /root/reference.  The files are
data: the quick_start cases' input.toml files (copied as they are), quick_start SU2 meshes and
one continue file the reference wrote at step 20000 (FVMR/src/output/FieldWriter.cpp:264-326),
which pins the post-CCW element node lists, the reference edge numbering of the boundary edges
and a realistic FP32 state.
"""
import gzip
import os
import shutil

REF = "/root/reference/quick_start/U2NITS-10.0"
HERE = os.path.dirname(os.path.abspath(__file__))
FILES = [
    ("2D_isentropic_vortex/ivortex50.su2", "ivortex50.su2.gz"),
    ("2D_cylinder/cylinder_allfar_50.su2", "cylinder_allfar_50.su2.gz"),
    ("2D_cylinder/pause_cylinder_allfar_50[20000].dat", "pause_cylinder_allfar_50_20000.dat.gz"),
    ("1D_shocktube/shocktube_100.su2", "shocktube_100.su2.gz"),
    ("2D_Neuvillette_shock/neuvillette.su2", "neuvillette.su2.gz"),
    ("2D_double_shock_reflect/dsr131new_slippery.su2", "dsr131new_slippery.su2.gz"),
]

# the quick_start cases' input.toml files, copied as they are (config data for the driver tests)
CONFIGS = [
    ("1D_shocktube/input.toml", "shocktube.input.toml"),
    ("2D_Neuvillette_shock/input.toml", "neuvillette.input.toml"),
    ("2D_cylinder/input.toml", "cylinder.input.toml"),
    ("2D_double_shock_reflect/input.toml", "dsr.input.toml"),
    ("2D_isentropic_vortex/input.toml", "ivortex.input.toml"),
]

# the reference's log of its own GPU run of the quick_start shock tube (FP32, CFL 2.5, 1000 steps)
LOG = "/root/reference/FVMR/working_directory/2024-10-17-142244.LOG"


def log_residuals():
    """(step, residual_rho) of the log's progress lines (CDriver.cpp:115-125)."""
    import re
    out = []
    for line in open(LOG, encoding="utf-8", errors="replace"):
        m = re.search(r"step = (\d+)\tresidual_rho = (\S+)\tCFL = (\S+)", line)
        if m:
            out.append((int(m.group(1)), m.group(2), m.group(3)))
    return out


if __name__ == "__main__":
    import json
    rows = log_residuals()
    with open(os.path.join(HERE, "shocktube_100_reference_log.json"), "w") as f:
        json.dump({"source": "FVMR/working_directory/2024-10-17-142244.LOG (progress lines)",
                   "mesh": "shocktube_100.su2", "config": "shocktube.input.toml",
                   "steps": [r[0] for r in rows], "residual_rho": [r[1] for r in rows],
                   "CFL": [r[2] for r in rows]}, f, indent=1)
        f.write("\n")
    print("shocktube_100_reference_log.json", len(rows))
    for src, dst in CONFIGS:
        shutil.copyfile(os.path.join(REF, src), os.path.join(HERE, dst))
        print(dst, os.path.getsize(os.path.join(HERE, dst)))
    for src, dst in FILES:
        with open(os.path.join(REF, src), "rb") as f, open(os.path.join(HERE, dst), "wb") as raw:
            with gzip.GzipFile(fileobj=raw, mode="wb", compresslevel=9, mtime=0) as g:
                shutil.copyfileobj(f, g)
        print(dst, os.path.getsize(os.path.join(HERE, dst)))
