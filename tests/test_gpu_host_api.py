"""Both sides of the drop-in boundary: the advection operator written
against the host loop / evaluator interface (include/hyperdeal_amd_host.hpp)
in hyper.deal's AdvectionOperation shape, face-centric and element-centric,
against the GPU operator of the same problem through the C++ wrapper of the
C ABI (include/hyperdeal_amd.hpp).

Bar: max |gpu - host| <= 1e-12 max |host| (FP64 on both sides), for 1+1,
2+2 and 3+3, periodic and homogeneous Dirichlet directions (both through
the collapsed kernel, Dirichlet with its boundary correction), skew 0 and
0.5, Gauss points n = k+1 and
k+2 (over-integration) and Gauss-Lobatto collocation.
"""

import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "host_api", "host_vs_gpu.cpp")
LIBDIR = os.path.join(ROOT, "hyperdeal_amd")


@pytest.fixture(scope="module")
def program(tmp_path_factory):
    if not os.path.exists(os.path.join(LIBDIR, "libhyperdeal_amd.so")):
        pytest.fail("libhyperdeal_amd.so is not built")
    out = str(tmp_path_factory.mktemp("host_gpu") / "host_vs_gpu")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Werror",
                    "-I" + os.path.join(ROOT, "include"),
                    "-I" + os.path.join(ROOT, "tests", "host_api"), "-o",
                    out, SRC, "-L" + LIBDIR, "-lhyperdeal_amd",
                    "-Wl,-rpath," + LIBDIR, "-L/opt/rocm/lib",
                    "-Wl,-rpath-link,/opt/rocm/lib"], check=True)
    return out


# (dim_x, dim_v, degree, n_points, collocation, periodic, skew)
CASES = [(1, 1, 3, 4, 0, 1, 0.0), (2, 2, 3, 4, 0, 1, 0.5),
         (2, 2, 2, 3, 0, 0, 0.0), (3, 3, 2, 3, 0, 1, 0.0),
         (3, 3, 1, 2, 0, 0, 0.5),
         # over-integration: periodic and Dirichlet
         (1, 1, 3, 5, 0, 1, 0.0), (2, 2, 3, 5, 0, 1, 0.5),
         (2, 2, 3, 5, 0, 0, 0.0), (3, 3, 1, 3, 0, 0, 0.0),
         # Gauss-Lobatto collocation
         (2, 2, 3, 4, 1, 1, 0.0), (2, 2, 2, 3, 1, 1, 0.5),
         (3, 3, 1, 2, 1, 1, 0.0)]


@pytest.mark.parametrize("dx,dv,k,nq,coll,periodic,skew", CASES)
def test_host_operator_equals_gpu_operator(program, dx, dv, k, nq, coll,
                                           periodic, skew):
    print("case: host vs gpu", dx, dv, k, nq, coll, periodic, skew,
          flush=True)
    r = subprocess.run([program] + [str(x) for x in (dx, dv, k, nq, coll,
                                                       periodic, skew)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    words = r.stdout.split()
    assert words[0] == "fcl" and words[2] == "ecl", r.stdout
    assert float(words[1]) <= 1e-12 and float(words[3]) <= 1e-12, r.stdout
