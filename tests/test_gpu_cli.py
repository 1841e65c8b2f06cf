"""The command-line reconstruction on the GPU (etn_main, the `etn` tool): a
parameter file with the reference's keys, an image-stack directory or MRC
stack in, an MRC volume out.

Each run is compared bit for bit with the same reconstruction set up directly
through the C ABI (etn_reconstruction_*, whose parity with the oracle the
reconstruction tests pin), and the reference's own end-to-end bounds of
SyntheticReconstructionTest (SyntheticReconstruction.cpp:132-171: RMS to the
ground truth <= 0.7 / 0.7 / 0.2 after the sart / sirt / blockIterative
parameter files) are asserted on the written files."""
import os
import subprocess
import sys

import numpy as np
import pytest

import ettention_amd as E
from ettention_amd import api as A
from ettention_amd import io as EI
from tests._util import load_golden

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

TEST_BASE = E.ScannerGeometry.make((-64, -64, -128), (-64, -64, 128), (0.5, 0, 0), (0, 0.5, 0), (256, 256))


@pytest.fixture(scope="module")
def synthetic():
    s = load_golden("synthetic_3_parallel")
    return np.ascontiguousarray(s["ground_truth"]), s["angles"], np.ascontiguousarray(s["projections"])


@pytest.fixture(scope="module")
def stack_dir(tmp_path_factory, synthetic):
    # the synthetic_3_parallel directory: metadata.xml + one TIFF per projection
    d = tmp_path_factory.mktemp("synthetic_3_parallel")
    _, tilts, projs = synthetic
    EI.write_image_stack_directory(str(d), projs, tilts, TEST_BASE, (64, 64, 64))
    return str(d)


def parameter_file(directory, name, algorithm, body, output="generated.mrc", projections="."):
    text = ('<?xml version="1.0"?>\n<hardware>\n  <openclVendor value="amd"/>\n  <subVolumeCount value="4"/>\n'
            '</hardware>\n<algorithm value="%s">\n%s\n</algorithm>\n<input>\n  <projections value="%s"/>\n</input>\n'
            '<output>\n  <filename value="%s"/>\n  <options><voxelType value="float32"/></options>\n</output>\n'
            '<volume>\n  <dimension value="64 64 64"/>\n  <initValue value="0.0"/>\n</volume>\n'
            '<debug>\n  <writeIntermediateVolumes value="false"/>\n  <profilingLevel value="normal"/>\n</debug>\n'
            % (algorithm, body, projections, output))
    p = os.path.join(directory, name)
    with open(p, "w") as f:
        f.write(text)
    return p


def api_volume(tilts, projs, iterations, **kw):
    vg = E.VolumeGeometry.make((64, 64, 64), subvolumes=kw.pop("subvolumes", 4))
    r = A.BlockIterativeReconstructionOperator(volume=vg, base_geometry=kw.pop("base", TEST_BASE),
                                               tilt_angles=tilts, projections=torch.from_numpy(projs).cuda(),
                                               iterations=iterations, **kw)
    r.run()
    return r.get_reconstructed_volume(dtype=torch.float32).cpu().numpy()


def max_slice_rms_z(vol, gt):
    d = vol.astype(np.float64) - gt
    return float(np.sqrt((d ** 2).mean(axis=(1, 2))).max())


CASES = [
    # the reference's synthetic_3_parallel parameter files, key for key
    ("sart.xml", "sart", '<numberOfIterations value="2"/><useLongObjectCompensation value="false"/>'
     '<lambda value="0.1"/>', '<backProjection><samples value="9"/></backProjection>'
     '<forwardProjection><samples value="4"/></forwardProjection>',
     dict(algorithm="sart", lam=0.1, fp_samples=4, bp_samples=9), 2, 0.7),
    ("sirt.xml", "sirt", '<numberOfIterations value="25"/><useLongObjectCompensation value="false"/>'
     '<lambda value="1.0"/>', "", dict(algorithm="sirt", lam=1.0), 25, 0.7),
    ("blockIterative.xml", "blockIterative", '<blockSize value="10"/><numberOfIterations value="10"/>'
     '<lambda value="1.0"/>', "", dict(algorithm="blockiterative", block_size=10, lam=1.0), 10, 0.2),
]


@pytest.mark.parametrize("name,alg,body,extra,api,iterations,bound", CASES, ids=[c[0] for c in CASES])
def test_reference_parameter_files(tmp_path, synthetic, stack_dir, name, alg, body, extra, api, iterations, bound):
    gt, tilts, projs = synthetic
    out = str(tmp_path / "generated.mrc")
    p = parameter_file(stack_dir, name, alg, body, output=out)
    with open(p, "a") as f:
        f.write(extra)
    print("etn --config", p, flush=True)
    assert EI.main(["--config", p]) == 0
    vol = EI.read_mrc_volume(out, "xzy_order")
    ref = api_volume(tilts, projs, iterations, projection_iteration="random", **api)
    np.testing.assert_array_equal(vol, ref)
    assert max_slice_rms_z(vol, gt) <= bound


def test_native_tool_and_python_module(tmp_path, synthetic, stack_dir):
    gt, tilts, projs = synthetic
    body = '<blockSize value="30"/><numberOfIterations value="2"/><lambda value="0.8"/>'
    outs = []
    for i, launcher in enumerate(([os.path.join(os.path.dirname(EI.__file__), "etn")],
                                  [sys.executable, "-m", "ettention_amd.cli"])):
        out = str(tmp_path / ("v%d.mrc" % i))
        p = parameter_file(stack_dir, "bi%d.xml" % i, "blockIterative", body, output=out)
        print("launch", launcher, p, flush=True)
        r = subprocess.run(launcher + ["--config", p, "--optimization.projectionIteration", "identity"],
                           capture_output=True, text=True, timeout=300,
                           cwd=os.path.dirname(os.path.dirname(EI.__file__)))
        assert r.returncode == 0, r.stdout + r.stderr
        assert "Reconstruction finished." in r.stdout
        outs.append(EI.read_mrc_volume(out))
    ref = api_volume(tilts, projs, 2, algorithm="blockiterative", block_size=30, lam=0.8,
                     projection_iteration="identity")
    np.testing.assert_array_equal(outs[0], ref)
    np.testing.assert_array_equal(outs[1], ref)


def write_mrc_stack(path, projs):
    h = np.zeros(256, np.int32)
    n, nv, nu = projs.shape
    h[:4] = (nu, nv, n, 2)
    with open(path, "wb") as f:
        f.write(h.tobytes())
        f.write(np.ascontiguousarray(projs, np.float32).tobytes())


def test_mrc_stack_with_every_option(tmp_path, synthetic):
    """MRC stack + tilt file, long-object compensation, skipped projections, a
    prior-knowledge mask, an initial volume, half storage and an inverted
    unsigned8 x-y-z output."""
    gt, tilts, projs = synthetic
    sub = slice(0, 180, 4)
    t, pr = tilts[sub], projs[sub]
    d = str(tmp_path)
    write_mrc_stack(os.path.join(d, "stack.mrc"), pr)
    with open(os.path.join(d, "stack.tlt"), "w") as f:
        f.write("\n".join("%.2f" % a for a in t) + "\n")
    rng = np.random.default_rng(21)
    mask = rng.integers(0, 256, (64, 64, 64)).astype(np.float32)
    mask[:, :, :8] = 0
    EI.write_mrc_volume(os.path.join(d, "mask.mrc"), mask, orientation="xyz_order")
    init = (rng.integers(0, 64, (64, 64, 64)) / 8.0).astype(np.float32)     # exact in half precision
    EI.write_mrc_volume(os.path.join(d, "init.mrc"), init)                  # x-z-y, the default
    out = os.path.join(d, "out.mrc")
    text = ('<algorithm value="blockIterative"><blockSize value="5"/><numberOfIterations value="2"/>'
            '<lambda value="0.6"/><useLongObjectCompensation value="true"/></algorithm>'
            '<input><projections value="stack.mrc"/><tiltAngles value="stack.tlt"/>'
            '<skipProjections value="3, 10-12"/></input>'
            '<output><filename value="out.mrc"/><options><voxelType value="unsigned8"/>'
            '<orientation value="xyz_order"/><invert value="true"/></options></output>'
            '<volume><dimension value="64 64 64"/><initFile value="init.mrc"/></volume>'
            '<priorknowledge><maskVolumeFilename value="mask.mrc"/></priorknowledge>'
            '<optimization><internalVoxelRepresentationType value="half"/>'
            '<projectionIteration value="buckets"/></optimization>'
            '<hardware><subVolumeCount value="2"/></hardware>')
    p = os.path.join(d, "all.xml")
    with open(p, "w") as f:
        f.write(text)
    print("etn --config", p, flush=True)
    assert EI.main(["--config", p]) == 0

    # the same reconstruction through the C ABI
    mrc_base = E.ScannerGeometry.make((-128, -128, 256), (-128, -128, -256), (1, 0, 0), (0, 1, 0), (256, 256))
    tl = np.array(["%.2f" % a for a in t], dtype=np.float32)
    vg = E.VolumeGeometry.make((64, 64, 64), subvolumes=2)
    r = A.BlockIterativeReconstructionOperator(
        volume=vg, base_geometry=mrc_base, tilt_angles=tl, projections=torch.from_numpy(np.ascontiguousarray(pr)).cuda(),
        algorithm="blockiterative", block_size=5, iterations=2, lam=0.6, use_long_object_compensation=True,
        projection_iteration="buckets", skip_projections=(2, 9, 10, 11), voxel_type="half",
        mask=torch.from_numpy(mask.astype(np.uint8)).cuda())
    r.set_volume(torch.from_numpy(init).cuda().half())
    r.run()
    ref = r.get_reconstructed_volume(dtype=torch.float32).cpu().numpy()
    EI.write_mrc_volume(os.path.join(d, "ref.mrc"), ref, voxel_type="unsigned8", orientation="xyz_order", invert=True)
    assert open(out, "rb").read() == open(os.path.join(d, "ref.mrc"), "rb").read()


def test_intermediate_volumes(tmp_path, synthetic, stack_dir):
    gt, tilts, projs = synthetic
    out = str(tmp_path / "final.mrc")
    dbg = str(tmp_path / "dbg")
    p = parameter_file(stack_dir, "dump.xml", "blockIterative",
                       '<blockSize value="60"/><numberOfIterations value="2"/><lambda value="0.5"/>', output=out)
    print("etn --config", p, flush=True)
    assert EI.main(["--config", p, "--debug.writeIntermediateVolumesAfterIterations", "true",
                    "--debug.writeIntermediateVolumes", "true", "--debug.infoPath", dbg,
                    "--optimization.projectionIteration", "identity"]) == 0
    names = sorted(os.listdir(dbg))
    assert names == ["intermediate_i000.mrc", "intermediate_i000_b000.mrc", "intermediate_i000_b001.mrc",
                     "intermediate_i000_b002.mrc", "intermediate_i001.mrc", "intermediate_i001_b000.mrc",
                     "intermediate_i001_b001.mrc", "intermediate_i001_b002.mrc"]
    final = EI.read_mrc_volume(out)
    # intermediate volumes are float32 in x-y-z order
    np.testing.assert_array_equal(EI.read_mrc_volume(os.path.join(dbg, "intermediate_i001.mrc"), "xyz_order"), final)
    np.testing.assert_array_equal(EI.read_mrc_volume(os.path.join(dbg, "intermediate_i001_b002.mrc"), "xyz_order"),
                                  final)
    first = EI.read_mrc_volume(os.path.join(dbg, "intermediate_i000.mrc"), "xyz_order")
    np.testing.assert_array_equal(EI.read_mrc_volume(os.path.join(dbg, "intermediate_i000_b002.mrc"), "xyz_order"),
                                  first)
    # per-block dumps run the generic kernels eagerly
    ref = api_volume(tilts, projs, 2, algorithm="blockiterative", block_size=60, lam=0.5,
                     projection_iteration="identity", flags=E.ETN_RECON_FLAG_GENERIC_KERNELS | E.ETN_RECON_FLAG_NO_GRAPH)
    np.testing.assert_array_equal(final, ref)
    one = api_volume(tilts, projs, 1, algorithm="blockiterative", block_size=60, lam=0.5,
                     projection_iteration="identity", flags=E.ETN_RECON_FLAG_GENERIC_KERNELS | E.ETN_RECON_FLAG_NO_GRAPH)
    np.testing.assert_array_equal(first, one)
