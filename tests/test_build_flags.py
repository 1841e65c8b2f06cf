"""The kernels are built without packed FP32 instructions: the build flags
disable the feature for device code, and the generated gfx950 assembly of
every kernel source is checked for v_pk_{add,mul,fma}_f32 (CPU only)."""
import os
import re
import subprocess
import tempfile

import pytest

from ettention_amd import build as B


@pytest.mark.parametrize("src", [s for s in B.SOURCES if s.endswith(".hip")])
def test_no_packed_fp32_in_device_code(src):
    assert "-packed-fp32-ops" in B.FLAGS
    with tempfile.TemporaryDirectory() as tmp:
        # the device pass alone, to gfx950 assembly
        out = os.path.join(tmp, "k.s")
        cmd = [B.hipcc()] + B.FLAGS + ["--cuda-device-only", "-S", os.path.join(B.CSRC, src), "-o", out]
        subprocess.run(cmd, cwd=tmp, check=True, capture_output=True)
        text = open(out).read()
    assert "amdgcn" in text and "gfx950" in text
    assert "_kernel" in text
    assert not re.search(r"v_pk_(add|mul|fma)_f32", text)
