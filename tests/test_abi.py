"""The drop-in boundary's struct layouts against the reference's own headers.

tests/abi/layout.c prints the size and every field offset of sparse_csr_t, gpu_sparse_csr_t, gpu_dense_vec_t,
gpu_lu_t, gpu_solver_workspace_t and solver_workspace_t.  It is compiled against include/ and against the
reference's tools/*.h (its GPU build: -DUSE_HIP, no MPI, hipSPARSE descriptor types from ROCm's headers); a
driver compiled against either header set passes the same bytes, so the outputs must be equal line for
line.  Needs the reference tree (the headers are read, nothing of the reference is run or linked)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_TOOLS = "/root/reference/tools"
SRC = os.path.join(ROOT, "tests", "abi", "layout.c")


def build_and_run(tmp_path, name, incs, defs=()):
    exe = str(tmp_path / name)
    cmd = ["gcc", "-std=gnu11", "-w", *defs, *[f"-I{i}" for i in incs], SRC, "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return subprocess.run([exe], check=True, capture_output=True, text=True).stdout.splitlines()


@pytest.mark.skipif(not os.path.isdir(REF_TOOLS), reason="the reference headers are not present")
def test_struct_layouts_equal_reference_headers(tmp_path):
    ours = build_and_run(tmp_path, "ours", [os.path.join(ROOT, "include")])
    ref = build_and_run(tmp_path, "ref", [REF_TOOLS, "/opt/rocm/include"], ["-DUSE_HIP", "-D__HIP_PLATFORM_AMD__"])
    assert len(ours) == len(ref) > 50
    diff = [(a, b) for a, b in zip(ours, ref) if a != b]
    assert not diff, diff
