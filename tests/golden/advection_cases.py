"""Reads the reference's advection test fixtures (tests/golden/advection/).

Each case is the triple the reference's test harness uses
(examples/advection/tests/README.MD): X.configuration (compile-time DIM_X,
DIM_V, DEGREE, N_POINTS), X.Y.json (run parameters) and X.Y.out (the
expected diagnostics rows, compared with numdiff -a 1e-5 -r 1e-8 by
cmake/macros/macro_hyper_deal_pickup_tests.cmake:57).
"""

import glob
import json
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ADV = os.path.join(HERE, "advection")

# examples/advection/cases/hyperrectangle.h:33-40: transport direction
HYPERRECTANGLE_VELOCITY = [1.0, 0.15, -0.05, 0.0, 0.0, 0.0]
# hyperrectangle.h:223-224: the box is [-1, 1]^D
HYPERRECTANGLE_BOX = (-1.0, 1.0)

_NUM = re.compile(r"[-+]?\d\.\d+e[-+]\d+")


def parse_rows(lines):
    return np.array([[float(x) for x in _NUM.findall(l)] for l in lines
                     if _NUM.search(l)])


def format_rows(rows):
    """The reference's row format (application.h:486-492)."""
    return ["   Time:%10.3e, norm: %17.10e, error: %17.10e" % tuple(r)
            for r in rows]


def numdiff_ok(ours, golden, abs_tol=1e-5, rel_tol=1e-8):
    """numdiff semantics: numbers agree when within abs_tol OR rel_tol."""
    if ours.shape != golden.shape:
        return False
    d = np.abs(ours - golden)
    rel = d / np.maximum(np.abs(golden), 1e-300)
    return bool(np.all((d <= abs_tol) | (rel <= rel_tol)))


def _configuration(prefix):
    txt = open(os.path.join(ADV, prefix + ".configuration")).read()
    return {k: int(v) for k, v in re.findall(r"(\w+)=(\d+)", txt)}


def _truthy(s):
    return str(s).strip().lower() == "true"


def cases():
    out = []
    for js in sorted(glob.glob(os.path.join(ADV, "*.json"))):
        name = os.path.basename(js)[:-5]
        prefix = name.split(".")[0]
        cfg = _configuration(prefix)
        p = json.load(open(js))
        g, s, t = p["General"], p["SpatialDiscretization"], \
            p["TemporalDiscretization"]
        c = p["Case"]
        dim_x, dim_v = int(g["DimX"]), int(g["DimV"])
        assert dim_x == cfg["DIM_X"] and dim_v == cfg["DIM_V"]
        nsx = [int(c["NSubdivisionsX"][k]) for k in "XYZ"[:dim_x]]
        nsv = [int(c["NSubdivisionsV"][k]) for k in "XYZ"[:dim_v]]
        refx, refv = int(c["NRefinementsX"]), int(c["NRefinementsV"])
        n_cells = [n * 2 ** refx for n in nsx] + [n * 2 ** refv for n in nsv]
        out.append(dict(
            name=name,
            dim_x=dim_x, dim_v=dim_v, degree=cfg["DEGREE"],
            n_points=cfg["N_POINTS"],
            collocation=_truthy(s["DoCollocation"]),
            n_cells=n_cells,
            periodic_x=_truthy(c["PeriodicX"]),
            periodic_v=_truthy(c["PeriodicV"]),
            skew=float(p["AdvectionOperation"]["SkewFactor"]),
            use_ecl=_truthy(p["Matrixfree"]["UseECL"]),
            rk=t["RKType"], time_step=float(t["TimeStep"]),
            start_time=float(t["StartTime"]),
            final_time=float(t["FinalTime"]), cfl=float(t["CFLNumber"]),
            tick=float(t["DiagnosticsTick"]),
            golden=parse_rows(open(os.path.join(ADV, name + ".out"))),
        ))
    return out
