"""The device's safe_inv (voxel.wgsl:65-71) computes 1/b as v_rcp_f32 refined
by one Newton step instead of an IEEE division.  This runs the library's own
safe_inv on the GPU for all 2^32 float bit patterns and requires every result
to equal the IEEE quotient s / max(|v|, 1e-8) bit for bit
(tests/rcp_exhaustive.hip, built with the library)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_build", "rcp_exhaustive")


def test_safe_inv_equals_ieee_division_on_every_float():
    if not os.path.isfile(BIN):
        pytest.fail(f"{BIN} is not built: run voxelrt_amd.build() (make -C voxelrt_amd/csrc)")
    print("[rcp] safe_inv over 2^32 floats", flush=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches 0" in r.stdout
