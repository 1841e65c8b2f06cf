"""Parameter files, command line and the parameter sets they resolve to
(csrc/config.cpp, csrc/framework.cpp) on the CPU.

The reference's own parameter files are resolved where the reference test data
is present (this checkout); the rest uses parameter files written here with the
reference's keys.  Expected values come from the reference's parameter-set
sources: AlgorithmParameterSet.cpp:8-53 (block sizes), VolumeParameterSet.cpp:9-38
(resolution = dimension when unset), OutputParameterSet.cpp:19-52 (float32,
xzy_order), XMLParameterSource.cpp:15-25 (paths relative to the file),
main.cpp:36-49 (the command line wins), ParameterSourceTest.cpp:100-145.
"""
import os

import pytest

from ettention_amd import io as EI
from ettention_amd._lib import InvalidArgument, Unsupported

REF_DATA = "/root/reference/ettention/test/data"
needs_reference = pytest.mark.skipif(not os.path.isdir(REF_DATA), reason="reference test data not present")


def norm(p):
    return os.path.normpath(p)


@needs_reference
@pytest.mark.parametrize("name,alg,block,iters,lam,fp,bp,sub", [
    ("sart.xml", "sart", 1, 2, 0.1, 4, 9, 4),
    ("sirt.xml", "sirt", 0xFFFFFFFF, None, None, None, None, 4),
    ("blockIterative.xml", "blockIterative", 10, 10, 1.0, 1, 1, 4),
])
def test_reference_synthetic_parameter_files(name, alg, block, iters, lam, fp, bp, sub):
    d = os.path.join(REF_DATA, "synthetic_3_parallel")
    c = EI.resolve_config(["--config", os.path.join(d, name)])
    assert c["algorithm"] == alg and c["block_size"] == block and c["subvolume_count"] == sub
    if iters is not None:
        assert c["iterations"] == iters and abs(c["lambda"] - lam) < 1e-7
        assert c["fp_samples"] == fp and c["bp_samples"] == bp
    assert c["resolution"] == [64, 64, 64] and c["dimension"] == [64, 64, 64] and c["origin"] == [0, 0, 0]
    assert norm(c["projections"]) == norm(d)
    # "..\..\work\synthetic\generated_*.mrc", relative to the parameter file
    assert norm(c["output_filename"]) == norm(os.path.join(d, "..", "..", "work", "synthetic",
                                                           "generated_" + name.replace(".xml", ".mrc")))
    assert c["output_voxel_type"] == "float32" and c["output_orientation"] == "xzy_order"
    assert c["projection_iteration"] == "random" and c["voxel_representation"] == "float"
    assert not c["use_long_object_compensation"]


@needs_reference
def test_reference_jason64_parameter_file():
    c = EI.resolve_config(["--config", os.path.join(REF_DATA, "jason_64_SART", "jason64.xml")])
    assert c["algorithm"] == "sart" and c["block_size"] == 1 and c["iterations"] == 1
    assert c["use_long_object_compensation"] and abs(c["lambda"] - 0.3) < 1e-7
    assert c["resolution"] == [128, 128, 32] and c["subvolume_count"] == 1
    # "\data\jason_64_SART\jason64.ali": backslashes become slashes, an absolute path stays
    assert c["projections"] == "/data/jason_64_SART/jason64.ali"
    assert c["tilt_angles"] == "/data/jason_64_SART/jason64.tlt"


@needs_reference
def test_reference_volume_parameter_file():
    c = EI.resolve_config(["--config", os.path.join(REF_DATA, "volume", "empiar_10009_64_voxel.xml")])
    # algorithm "sart" wins over its blockSize 5
    assert c["block_size"] == 1 and c["resolution"] == [256, 256, 256] and c["dimension"] == [64, 64, 64]
    assert c["output_voxel_type"] == "unsigned8" and c["output_orientation"] == "xyz_order"


@needs_reference
def test_reference_relative_path_example():
    # ParameterSourceTest XML_Relative_Path: data/setup/config.xml names
    # "..\jason_64_SART\jason64.ali"; the other required keys come from the command line
    cfg = os.path.join(REF_DATA, "setup", "config.xml")
    c = EI.resolve_config(["--config", cfg, "--algorithm", "sart", "--algorithm.numberOfIterations", "1",
                           "--algorithm.lambda", "0.5", "--volume.dimension", "8 8 8"])
    assert norm(c["projections"]) == norm(os.path.join(REF_DATA, "jason_64_SART", "jason64.ali"))
    assert norm(c["output_filename"]) == norm(os.path.join(REF_DATA, "..", "work", "jason64", "nonexistent.mrc"))


BASE = """<?xml version="1.0"?>
<!-- reconstruction parameters -->
<hardware>
    <subVolumeCount value="2" />
</hardware>
<algorithm value="{alg}">
    {block}
    <numberOfIterations value="3" />
    <lambda value="0.25" />
    <useLongObjectCompensation value="1"/>
</algorithm>
<input>
    <projections value="data\\stack.mrc"/>
    <tiltAngles value="data/stack.tlt"/>
    <skipProjections value=" 2, 5-7 ,3"/>
</input>
<output>
    <filename value="out/volume.mrc"/>
    <options><voxelType value="unsigned16"/><invert value="true"/></options>
</output>
<volume>
    <dimension value="64 32 16"/>
    <origin value="1 -2 0.5"/>
    <resolution value="32 16 8"/>
    <initValue value="0.125"/>
</volume>
<optimization>
    <internalVoxelRepresentationType value="half"/>
    <projectionIteration value="maxangle"/>
</optimization>
"""


def write_cfg(tmp_path, alg="blockIterative", block='<blockSize value="6"/>', extra=""):
    p = tmp_path / "cfg" / "r.xml"
    p.parent.mkdir(exist_ok=True)
    p.write_text(BASE.format(alg=alg, block=block) + extra)
    return str(p)


def test_own_parameter_file(tmp_path):
    p = write_cfg(tmp_path)
    c = EI.resolve_config(["--config", p])
    d = os.path.dirname(p)
    assert c["algorithm"] == "blockIterative" and c["block_size"] == 6 and c["iterations"] == 3
    assert c["lambda"] == 0.25 and c["use_long_object_compensation"] and c["subvolume_count"] == 2
    assert c["projections"] == norm(os.path.join(d, "data", "stack.mrc"))
    assert c["tilt_angles"] == norm(os.path.join(d, "data", "stack.tlt"))
    assert c["skip_projections"] == [1, 2, 4, 5, 6]          # 1-based list and ranges
    assert c["output_filename"] == norm(os.path.join(d, "out", "volume.mrc"))
    assert c["output_voxel_type"] == "unsigned16" and c["output_invert"]
    assert c["dimension"] == [64, 32, 16] and c["origin"] == [1, -2, 0.5] and c["resolution"] == [32, 16, 8]
    assert c["init_value"] == 0.125 and c["voxel_representation"] == "half"
    assert c["projection_iteration"] == "maxangle"
    assert c["fp_samples"] == 1 and c["bp_samples"] == 1 and c["profiling_level"] == "overall"


def test_command_line_wins_over_the_file(tmp_path):
    p = write_cfg(tmp_path)
    c = EI.resolve_config(["-c", p, "--algorithm.lambda", "0.75", "--hardware.subVolumeCount=5",
                           "--input.projections", "/abs/dir"])
    assert c["lambda"] == 0.75 and c["subvolume_count"] == 5 and c["projections"] == "/abs/dir"


@pytest.mark.parametrize("alg,block,expect", [
    ("sart", '<blockSize value="6"/>', 1),            # exact "sart": one projection per block
    ("sirt", "", 0xFFFFFFFF),
    ("blockIterative", "", 0xFFFFFFFF),               # no blockSize: the whole stack
    ("SART", "", 0xFFFFFFFF),                         # chosen case-insensitively, sized case-sensitively
])
def test_block_size_rules(tmp_path, alg, block, expect):
    assert EI.resolve_config(["--config", write_cfg(tmp_path, alg, block)])["block_size"] == expect


@pytest.mark.parametrize("override,err", [
    (["--algorithm", "dart"], Unsupported),
    (["--algorithm.blockSize", "0"], InvalidArgument),
    (["--algorithm.lambda", "fast"], InvalidArgument),
    (["--algorithm.useLongObjectCompensation", "yes"], InvalidArgument),
    (["--algorithm.numberOfIterations", "-1"], InvalidArgument),
    (["--volume.dimension", "1 2"], InvalidArgument),
    (["--optimization.internalVoxelRepresentationType", "uchar"], InvalidArgument),
    (["--output.options.voxelType", "float64"], InvalidArgument),
    (["--output.options.orientation", "xxy_order"], InvalidArgument),
    (["--input.skipProjections", "0"], InvalidArgument),
    (["--algorithm.basisFunctions", "blobs"], Unsupported),
    (["--hardware.openclDeviceId", "0", "--hardware.openclVendor", "amd"], InvalidArgument),
])
def test_invalid_parameters(tmp_path, override, err):
    with pytest.raises(err):
        EI.resolve_config(["--config", write_cfg(tmp_path)] + override)


def test_missing_required_parameter(tmp_path):
    p = tmp_path / "m.xml"
    p.write_text('<algorithm value="sart"><lambda value="0.3"/></algorithm>')
    with pytest.raises(InvalidArgument, match="algorithm.numberOfIterations"):
        EI.resolve_config(["--config", str(p)])


def test_malformed_parameter_file_names_the_line(tmp_path):
    p = tmp_path / "bad.xml"
    p.write_text('<algorithm value="sart">\n  <lambda value="0.3">\n</algorithm>\n')
    with pytest.raises(InvalidArgument, match=r"bad.xml:3"):
        EI.resolve_config(["--config", str(p)])


def test_command_line_tool_exit_codes(tmp_path, capfd):
    # main.cpp:71-140: help and the missing-parameter report without --config
    # end with 0, errors with 1
    assert EI.main(["--help"]) == 0
    assert "--config" in capfd.readouterr().out
    assert EI.main(["-H"]) == 0
    assert "algorithm.numberOfIterations" in capfd.readouterr().out
    assert EI.main(["--algorithm", "sart"]) == 0
    assert "Missing required parameter" in capfd.readouterr().out
    p = tmp_path / "m.xml"
    p.write_text('<algorithm value="sart"/>')
    assert EI.main(["--config", str(p)]) == 1
    assert "Error: Missing required parameter algorithm.numberOfIterations" in capfd.readouterr().out
    assert EI.main(["stray"]) == 1
    # an output directory that does not exist is reported before any GPU work
    out = write_cfg(tmp_path)
    assert EI.main(["--config", out, "--output.filename", str(tmp_path / "no" / "such" / "dir" / "v.mrc")]) == 1
    assert "Cannot write to output path" in capfd.readouterr().out
    assert EI.main(["--config", out, "--output.filename", str(tmp_path / "v.raw")]) == 1
    assert "No suitable serializer" in capfd.readouterr().out


def test_native_tool_runs(tmp_path):
    import subprocess
    exe = os.path.join(os.path.dirname(EI.__file__), "etn")
    r = subprocess.run([exe, "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "--config" in r.stdout
