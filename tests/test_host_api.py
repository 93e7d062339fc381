"""The host loop / evaluator interface (include/hyperdeal_amd_host.hpp):
hyper.deal's advection operator written against it in the reference's own
shape (tests/host_api/advection_op.hpp: AdvectionOperation<dim_x, dim_v,
degree, n_points, Number, VectorType, VectorizedArrayType> with evaluators
built as (data, 0, 0, 0, 0), data.loop(&This::local_apply_cell, ...,
this, dst, src, DataAccessOnFaces::values, DataAccessOnFaces::values,
timers) + cell_loop for the inverse mass matrix, and loop_cell_centric(...,
DataAccessOnFaces::values, timers); advection_operation.h:134-209), run
face-centric and element-centric (member-pointer and std::function forms),
against the oracle's apply of the same problem.

Bar: max |host - oracle| <= 1e-13 max |oracle| (the same arithmetic up to
the order of the sums), over 1+1 ... 3+3, degrees 1-3, n_points = k+1, k+2,
k+3 (Gauss) and Gauss-Lobatto collocation, periodic and homogeneous
Dirichlet directions and skew factors 0, 0.3, 0.5, 1.
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


# (dim_x, dim_v, degree, n_points, collocation, periodic, skew)
CASES = [
    (1, 1, 3, 4, 0, 1, 0.0), (1, 1, 2, 3, 0, 1, 0.5), (1, 1, 1, 2, 0, 0, 1.0),
    (2, 2, 3, 4, 0, 1, 0.0), (2, 2, 2, 3, 0, 0, 0.0), (2, 2, 1, 2, 0, 1, 0.3),
    (1, 2, 3, 4, 0, 0, 1.0), (2, 1, 1, 2, 0, 1, 0.3), (3, 3, 2, 3, 0, 1, 0.0),
    (3, 3, 1, 2, 0, 0, 0.5),
    # over-integration (basis change to n_points > degree + 1)
    (1, 1, 3, 5, 0, 0, 0.5), (2, 2, 3, 5, 0, 1, 0.0), (2, 2, 3, 5, 0, 0, 0.3),
    (1, 2, 2, 4, 0, 1, 1.0), (2, 2, 2, 4, 0, 0, 0.0), (3, 3, 1, 3, 0, 0, 0.5),
    # Gauss-Lobatto collocation
    (1, 1, 3, 4, 1, 0, 0.0), (2, 2, 3, 4, 1, 1, 0.0), (2, 2, 3, 4, 1, 0, 0.5),
    (3, 3, 1, 2, 1, 0, 0.0), (2, 1, 1, 2, 1, 1, 1.0),
]


@pytest.mark.parametrize("dx,dv,k,nq,coll,periodic,skew", CASES)
def test_host_operator_matches_oracle(program, dx, dv, k, nq, coll, periodic,
                                      skew):
    r = subprocess.run([program] + [str(x) for x in (dx, dv, k, nq, coll,
                                                       periodic, skew)],
                       capture_output=True, text=True, timeout=120,
                       check=True)
    words = r.stdout.split()
    assert words[0] == "fcl" and words[2] == "ecl" and words[4] == "fn", \
        r.stdout
    errs = [float(words[i]) for i in (1, 3, 5)]
    assert max(errs) <= 1e-13, r.stdout
