"""The C caller of tests/capi_check.c on the device: a solve through
iblgf_gpu_apply_lgf_typed(AMR2AMR) equals iblgf_gpu_apply_lgf bit for bit,
and the unbuilt fmm types are refused with a live context."""
import subprocess

import pytest

from test_capi import build_c_caller

pytestmark = pytest.mark.gpu


def test_c_caller_solves_through_the_abi(tmp_path):
    print("case: C caller, uniform 4^3 blocks of 8^3, typed vs plain apply_lgf", flush=True)
    r = subprocess.run([build_c_caller(tmp_path), "gpu"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout
