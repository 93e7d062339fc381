"""The host loop / evaluator interface (include/hyperdeal_amd_host.hpp):
hyper.deal's advection operator written against it, face-centric
(MatrixFree::loop with cell / face / boundary operations, member-pointer
form, then the inverse mass matrix through cell_loop) and element-centric
(loop_cell_centric with a std::function), against the oracle's apply of the
same problem.

Bar: max |host - oracle| <= 1e-13 max |oracle| (the same arithmetic up to
the order of the sums), over 1+1 ... 3+3, degrees 1-3, periodic and
homogeneous Dirichlet directions and skew factors 0, 0.3, 0.5, 1.
"""

import os
import subprocess

import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "host_api", "advection_host.cpp")


@pytest.fixture(scope="module")
def program(tmp_path_factory):
    O.build()
    out = str(tmp_path_factory.mktemp("host_api") / "advection_host")
    odir = os.path.join(ROOT, "oracle")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Werror",
                    "-I" + os.path.join(ROOT, "include"), "-I" + odir, "-o",
                    out, SRC, "-L" + odir, "-lhd_oracle",
                    "-Wl,-rpath," + odir], check=True)
    return out


CASES = [
    (1, 1, 3, 1, 0.0), (1, 1, 2, 1, 0.5), (1, 1, 1, 0, 1.0),
    (2, 2, 3, 1, 0.0), (2, 2, 2, 0, 0.0), (2, 2, 1, 1, 0.3),
    (1, 2, 3, 0, 1.0), (2, 1, 1, 1, 0.3), (3, 3, 2, 1, 0.0),
    (3, 3, 1, 0, 0.5),
]


@pytest.mark.parametrize("dx,dv,k,periodic,skew", CASES)
def test_host_operator_matches_oracle(program, dx, dv, k, periodic, skew):
    r = subprocess.run([program, str(dx), str(dv), str(k), str(periodic),
                        str(skew)], capture_output=True, text=True,
                       timeout=120, check=True)
    words = r.stdout.split()
    assert words[0] == "fcl" and words[2] == "ecl", r.stdout
    fcl, ecl = float(words[1]), float(words[3])
    assert fcl <= 1e-13 and ecl <= 1e-13, r.stdout
