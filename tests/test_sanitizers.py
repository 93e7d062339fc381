"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU
only): the oracle and the host loop / evaluator interface
(tests/host_api/advection_host.cpp with oracle/hd_oracle.cpp compiled in),
and the library's host-side table and geometry builders
(hyperdeal_amd/csrc/hd_tables.cpp, hd_geometry.cpp, driven by
tests/host_api/tables_check.cpp).  Any invalid access, leak or undefined
behaviour aborts the program; the runs must also reproduce the plain
build's results."""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-O1", "-g", "-fno-omit-frame-pointer",
       "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]
ENV = dict(os.environ, OMP_NUM_THREADS="2",
           ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None,
                                reason="no host compiler")


def _build(out, sources, includes):
    cmd = (["g++", "-std=c++17"] + SAN + ["-I" + i for i in includes] +
           ["-o", out] + sources + ["-fopenmp"])
    subprocess.run(cmd, check=True, capture_output=True, text=True)


def test_oracle_and_host_interface_under_sanitizers(tmp_path):
    out = str(tmp_path / "advection_host_san")
    _build(out, [os.path.join(ROOT, "tests", "host_api", "advection_host.cpp"),
                 os.path.join(ROOT, "oracle", "hd_oracle.cpp")],
           [os.path.join(ROOT, "include"), os.path.join(ROOT, "oracle")])
    for case in ((2, 2, 3, 4, 0, 1, 0.0), (1, 2, 2, 4, 0, 0, 0.5),
                 (2, 1, 1, 2, 1, 1, 1.0), (3, 3, 1, 2, 0, 0, 0.3)):
        r = subprocess.run([out] + [str(x) for x in case], env=ENV,
                           capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, r.stderr[-3000:]
        words = r.stdout.split()
        errs = [float(words[i]) for i in (1, 3, 5)]
        assert max(errs) <= 1e-13, r.stdout


def test_library_host_tables_under_sanitizers(tmp_path):
    out = str(tmp_path / "tables_check_san")
    csrc = os.path.join(ROOT, "hyperdeal_amd", "csrc")
    _build(out, [os.path.join(ROOT, "tests", "host_api", "tables_check.cpp"),
                 os.path.join(csrc, "hd_tables.cpp"),
                 os.path.join(csrc, "hd_geometry.cpp")], [csrc])
    r = subprocess.run([out], env=ENV, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr[-3000:]
    assert "bad 0" in r.stdout
