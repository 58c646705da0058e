"""Write the golden vectors of tests/golden/*.npz from the compiled reference.

Runs oracle/_ref/ocfd_probe (the reference's own routines driven in the order
of one NS_solver step, see oracle/ref_probe.f90) on small seeded cases and
stores, per case, the inputs (configuration, grid, initial primitives) and the
reference's outputs:

  k1_duinv  du after du_inviscous / du_inviscous_Character, stage 1
  k1_dutot  du after du_viscous, stage 1
  k1_f      f after OCFD_time_adv_RK3(1)
  n1_f      f after one whole step (filtering included)
  nN_f      f after three whole steps
  rhyb      Rhybrid (interior) after comput_Rhybrid (hybrid cases)
  met       Ajac, Akx1, Aiy1, Asz1 (halo'd) after init_mesh (curvilinear cases)
  hist      per step: Istep, tt, E, K, S (show_flow_msg), KE, enstrophy (ana_box)
  k1_T, n1_T, nN_T, negt_cells   (negt_* cases) T after stage 1, one and three
            steps, and the cells handle_NegativeT repaired (global i, j, k in
            the order the reference reported them)

Also writes ref_snapshot_w5_sw.dat: the snapshot file (save_flow_data,
OCFD_IO.f90:44-64) that the unmodified reference writes after two steps of the
"w5_sw" case, the fixture of the library's restart-file reader and writer;
and met_flat_plate.npz / met_swept_corner.npz: the halo'd metrics Ajac, Akx1,
Aiy1, Asz1 that the reference's init_mesh computes for the boundary-layer
cases (their user ghost layers, Jac_Ghost_Userdef).

Usage (needs the reference built by `make -C oracle ref probe`):
    python tests/golden/make_golden.py [case ...]
"""
from __future__ import annotations

import dataclasses
import json
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "opencfd-sc_amd"))

import pyoracle  # noqa: E402
from ocfd import cases  # noqa: E402

SHAPE = (8, 10, 14)  # (nz, ny, nx): non-cubic so a transposed index fails


def golden_cases():
    C = lambda **kw: cases.random_case(shape=SHAPE, **kw)  # noqa: E731
    F1 = [[1, 1, 1, 1, 1, 14, 1, 10, 1, 8, 1, 0.3, 0.0]]
    F2 = [[1, 1, 1, 1, 2, 13, 1, 10, 1, 8, 2, 0.5, 1e-5], [2, 1, 0, 1, 1, 14, 1, 10, 1, 8, 1, 1.0, 0.0]]
    return {
        "w5_sw": C(seed=3),
        "w7c_llf": C(seed=4, scheme="WENO7", split=2, char=1),
        "hyb_c_llf": C(seed=5, scheme="HYBRID", split=2, char=1, jump=0.6, hybrid=(0.002, 0.02)),
        "hyb_nc_sw": C(seed=6, scheme="HYBRID", split=1, char=0, jump=0.6, hybrid=(0.002, 0.02)),
        "omp6_nc": C(seed=7, scheme="OMP6", split=1, char=0, noise=0.05),
        "omp6_c_llf": C(seed=8, scheme="OMP6", split=2, char=1, noise=0.05),
        "ud7l_cd8": C(seed=9, scheme="UD7L", split=2, char=0, vis="CD8"),
        "curv_w5c": C(seed=10, scheme="WENO5", split=1, char=1, grid="curv", iperiodic=(0, 0, 1)),
        "curv_w7_cd8": C(seed=11, scheme="WENO7", split=1, char=0, grid="curv", iperiodic=(0, 1, 0), vis="CD8"),
        "curv_sb_ghost": C(seed=12, scheme="WENO5", split=2, char=0, grid="curv", iperiodic=(0, 0, 1),
                           scheme_boundary=(1, -1, 0, 1, 0, 0), ghost_cell=(1, 0, 2, 0, 0, 0),
                           bc="BC_USER_DEF", bc_para={1: 1.1}),
        "curv_wall_fo9p": C(seed=13, scheme="WENO5", split=1, char=1, grid="curv", iperiodic=(1, 0, 1),
                            scheme_boundary=(0, 0, 1, 0, 0, 0), ghost_cell=(0, 0, 2, 0, 0, 0),
                            bc="BC_USER_DEF", bc_para={1: 1.1}, filters=F1),
        "grid1d_shockfilter": C(seed=14, scheme="WENO5", split=1, char=0, iperiodic=(1, 0, 1), grid="1d",
                                filters=F2),
        "blunt_massforce": C(seed=15, scheme="OMP6", split=1, char=1, grid="curv", iperiodic=(1, 0, 1),
                             bc="BC_BLUNT2D", mass_force=(0.1, -0.2, 0.05)),
        "hyb_sensor2_curv": C(seed=16, scheme="HYBRID", split=2, char=1, jump=0.6, grid="curv",
                              iperiodic=(1, 1, 0), hybrid=(1.0, 3.0), sensor=2),
        # strong rarefactions: 3 to 6 cells with T <= 0 repaired in every stage
        "negt_w5_sw": cases.negative_t_case(),
    }


def cfg_to_json(cfg) -> str:
    d = dataclasses.asdict(cfg)
    d["bc_para"] = {str(k): v for k, v in d["bc_para"].items()}
    d["hybrid_para"] = {str(k): v for k, v in d["hybrid_para"].items()}
    return json.dumps(d)


SNAPSHOT_CASE, SNAPSHOT_STEPS = "w5_sw", 2


def snapshot(tmp):
    case = golden_cases()[SNAPSHOT_CASE]
    d = os.path.join(tmp, "snapshot")
    pyoracle.run_reference(case, d, SNAPSHOT_STEPS)
    path = os.path.join(HERE, "ref_snapshot_%s.dat" % SNAPSHOT_CASE)
    shutil.copy(os.path.join(d, "OCFD%08d.dat" % SNAPSHOT_STEPS), path)
    print("%-20s %7.1f KB" % (os.path.basename(path), os.path.getsize(path) / 1024))


def bc_metrics(tmp):
    for name, mk in (("flat_plate", cases.flat_plate), ("swept_corner", cases.swept_corner)):
        case = mk()
        ref = pyoracle.run_probe(case, os.path.join(tmp, "met_" + name), nextra=0)
        path = os.path.join(HERE, "met_%s.npz" % name)
        np.savez_compressed(path, met=np.stack([ref[k] for k in ("Ajac", "Akx1", "Aiy1", "Asz1")]))
        print("%-20s %7.1f KB" % (os.path.basename(path), os.path.getsize(path) / 1024))


def main():
    if not pyoracle.reference_available():
        raise SystemExit("build the reference first: make -C oracle ref probe")
    tmp = tempfile.mkdtemp(prefix="ocfd_golden_")
    try:
        only = set(sys.argv[1:])
        if not only or "snapshot" in only:
            snapshot(tmp)
        if not only or "bc_metrics" in only:
            bc_metrics(tmp)
        for name, case in golden_cases().items():
            if only and name not in only:
                continue
            negt = name.startswith("negt")
            ref = pyoracle.run_probe(case, os.path.join(tmp, name), nextra=2, allow_negative_t=negt)
            L = case.cfg.lap
            out = {
                "config": np.array(cfg_to_json(case.cfg)),
                "x": case.x, "y": case.y, "z": case.z, "spans": case.spans,
                "prims": np.stack(case.prims),
                "k1_duinv": ref["k1_duinv"], "k1_dutot": ref["k1_dutot"], "k1_f": ref["k1_f"],
                "n1_f": ref["n1_f"], "nN_f": ref["nN_f"], "hist": ref["hist"],
            }
            if negt:
                # the temperature after stage 1, after one step and after three
                # steps (halo'd), and the cells the reference repaired, in order
                for k in ("k1_T", "n1_T", "nN_T", "negt_cells"):
                    out[k] = ref[k]
            if "rhyb" in ref:
                out["rhyb"] = ref["rhyb"][L:-L, L:-L, L:-L]
            if case.cfg.grid_type != cases.GRID_BOX:
                out["met"] = np.stack([ref[k] for k in ("Ajac", "Akx1", "Aiy1", "Asz1")])
            path = os.path.join(HERE, "ref_%s.npz" % name)
            np.savez_compressed(path, **out)
            print("%-20s %7.1f KB" % (name, os.path.getsize(path) / 1024))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


if __name__ == "__main__":
    main()
