"""AddressSanitizer + UndefinedBehaviorSanitizer builds of the host C code (CPU only, gcc).

* The CPU drop-ins of libgade_host.so (csr, solver_common, matfree, bicgstab), compiled from their sources
  into tests/abi/matfree_check.c, run the matrix-free products and a BiCGSTAB solve on generated couplings.
* The C oracle (oracle/gade_oracle.c, test infrastructure) runs every routine the tests use from
  tests/sanitize/oracle_run.c, with two OpenMP threads and the all-cores paths.

Any out-of-bounds access, use after free, leak or undefined operation ends the run with a non-zero status
(-fno-sanitize-recover=all, leak detection on)."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT, SMALL, dt_of

SAN = ["-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
       "-fopenmp", "-std=gnu11"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
HOST = [os.path.join(ROOT, "dirac_hydrogen_code_amd", "host", f) for f in
        ("csr.c", "solver_common.c", "matfree.c", "bicgstab.c")]
CASES = [("tiny", 2), ("nondipole_small", 4)]

pytestmark = pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")


def build(tmp_path, name, srcs, incs, defs=()):
    exe = str(tmp_path / name)
    res = subprocess.run(["gcc", *SAN, *defs, *[f"-I{i}" for i in incs], *srcs, "-o", exe, "-lm"],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    return exe


def run(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True, env=ENV, timeout=600)
    assert res.returncode == 0, (res.returncode, res.stdout[-2000:], res.stderr[-4000:])
    assert "runtime error" not in res.stderr and "Sanitizer" not in res.stderr, res.stderr[-4000:]
    return res.stdout


@pytest.mark.parametrize("cfg,lmax", CASES)
def test_host_library_sources_under_sanitizers(datadir, tmp_path, cfg, lmax):
    exe = build(tmp_path, "matfree_check_san", [os.path.join(ROOT, "tests", "abi", "matfree_check.c"), *HOST],
                [os.path.join(ROOT, "include")], ["-DGADE_HOST_LIB"])
    c = SMALL[cfg]
    run([exe, datadir(cfg), str(lmax), repr(dt_of(cfg)), repr(c["intensity"]), repr(c["omega"]),
         repr(c["cycles"]), repr(0.37 * c["cycles"] * 6.283185307179586 / c["omega"]), "1e-13",
         str(tmp_path / "out.bin")])


@pytest.mark.parametrize("cfg,lmax", CASES)
def test_oracle_under_sanitizers(datadir, tmp_path, cfg, lmax):
    exe = build(tmp_path, "oracle_run", [os.path.join(ROOT, "tests", "sanitize", "oracle_run.c"),
                                         os.path.join(ROOT, "oracle", "gade_oracle.c")],
                [os.path.join(ROOT, "oracle")])
    out = run([exe, datadir(cfg), str(lmax), repr(dt_of(cfg)), "2"])
    assert out.startswith("checksum"), out
