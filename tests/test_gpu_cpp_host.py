"""The C++ host interface (include/hyperdeal_amd.hpp) end to end: the
advection example (hyperdeal_amd/examples/advection.cpp) runs the
reference's golden cases on the GPU and its printed rows are compared with
the golden files under the reference harness's tolerance (numdiff -a 1e-5
-r 1e-8)."""

import os
import subprocess

import pytest

from tests.golden import advection_cases as AC

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "hyperdeal_amd", "examples", "advection")

CASES = [c for c in AC.cases()
         if c["n_points"] == c["degree"] + 1 or
         (c["periodic_x"] and c["periodic_v"])]


@pytest.mark.parametrize("c", CASES, ids=[c["name"] for c in CASES])
def test_cpp_example_matches_golden(c):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    assert os.path.exists(EXE), "build the example: make -C hyperdeal_amd"
    assert len(set(c["n_cells"])) == 1
    args = [EXE, str(c["dim_x"]), str(c["dim_v"]), str(c["degree"]),
            str(c["n_points"]), str(int(c["collocation"])),
            str(int(c["periodic_x"])), repr(c["skew"]), c["rk"],
            str(c["n_cells"][0]), repr(c["time_step"]),
            repr(c["final_time"]), repr(c["cfl"]), repr(c["tick"])]
    out = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rows = AC.parse_rows(out.stdout.splitlines())
    assert AC.numdiff_ok(rows, c["golden"])
