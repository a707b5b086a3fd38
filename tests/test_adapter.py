"""The reference-side adapter (mmc_amd/csrc/adapter/mmc_run_hip.c): the file
a maintainer drops into the reference's src/ to get the
`void mmc_run_hip(mcconfig*, tetmesh*, raytracer*)` backend (the shape of
mmc_run_cl, src/mmc_cl_host.h:126).  It is compiled here against the
reference's own headers (src/mmc_utils.h, src/mmc_mesh.h) with the flags of
the reference's sse target, so every field it reads or writes exists with the
type it assumes; its static asserts hold only where float3 is 16 bytes (the
tracer->n layout it hands over as the element planes).  Skipped where the
reference sources are absent."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/src"
ADAPTER = os.path.join(ROOT, "mmc_amd", "csrc", "adapter", "mmc_run_hip.c")

pytestmark = pytest.mark.skipif(not os.path.isdir(REF), reason="reference sources not present")


def compile_adapter(*defs):
    cmd = ["gcc", "-fsyntax-only", "-Wall", "-Werror=implicit-function-declaration", "-Werror=incompatible-pointer-types",
           *defs, "-I" + REF, "-I" + os.path.join(ROOT, "include"), ADAPTER]
    return subprocess.run(cmd, capture_output=True, text=True)


def test_adapter_compiles_against_the_reference_headers():
    for defs in (["-DMMC_USE_SSE", "-DMMC_XORSHIFT"], ["-DMMC_USE_SSE", "-DMMC_XORSHIFT", "-DUSE_OPENCL"]):
        r = compile_adapter(*defs)
        assert r.returncode == 0, r.stderr
        ours = [l for l in r.stderr.splitlines() if "mmc_run_hip.c" in l and "warning" in l]
        assert not ours, ours


def test_adapter_refuses_a_12_byte_float3():
    r = compile_adapter("-DMMC_XORSHIFT")
    assert r.returncode != 0 and "mmc_hip_float3_is_16_bytes" in r.stderr
