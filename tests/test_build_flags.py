"""The kernels are built without packed FP32 instructions: the build flags
disable the feature for device code, and the generated gfx950 assembly of
every kernel translation unit is checked for v_pk_{add,mul,fma}_f32 (CPU only)."""
import concurrent.futures as cf
import os
import re
import subprocess
import tempfile

from ettention_amd import build as B


def _device_asm(unit, tmp):
    src, stem, extra = unit
    out = os.path.join(tmp, stem + ".s")
    # the device pass alone, to gfx950 assembly
    cmd = [B.hipcc()] + B.FLAGS + extra + ["--cuda-device-only", "-S", os.path.join(B.CSRC, src), "-o", out]
    subprocess.run(cmd, cwd=tmp, check=True, capture_output=True)
    return stem, open(out).read()


def test_no_packed_fp32_in_device_code():
    assert "-packed-fp32-ops" in B.FLAGS
    units = [u for u in B.units() if u[0].endswith(".hip")]
    assert sum(u[0] == "fused_inst.hip" for u in units) == len(B.FUSED_VARIANTS)
    with tempfile.TemporaryDirectory() as tmp:
        with cf.ThreadPoolExecutor(max_workers=min(8, len(units))) as ex:
            results = list(ex.map(lambda u: _device_asm(u, tmp), units))
    for stem, text in results:
        assert "amdgcn" in text and "gfx950" in text, stem
        # fused.hip holds the fused kernels' host side; fused_inst.hip their instances
        assert stem == "fused" or "_kernel" in text, stem
        assert not re.search(r"v_pk_(add|mul|fma)_f32", text), stem
