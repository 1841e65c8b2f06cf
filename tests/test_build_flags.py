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
        cmd = [B.hipcc()] + B.FLAGS + ["-c", os.path.join(B.CSRC, src), "-o", os.path.join(tmp, "k.o"),
                                       "--save-temps"]
        subprocess.run(cmd, cwd=tmp, check=True, capture_output=True)
        asm = [f for f in os.listdir(tmp) if f.endswith(".s") and "gfx950" in f]
        assert asm, os.listdir(tmp)
        text = open(os.path.join(tmp, asm[0])).read()
    assert "_kernel" in text
    assert not re.search(r"v_pk_(add|mul|fma)_f32", text)
