"""HIP kernels through the C ABI against the oracle and the reference fixtures.

FP32 tolerance: the kernels follow the oracle's float operations term by term
(no contraction), so forward and back projections are expected to be
bit-identical; the asserted bound is relative L2 <= 1e-6 (the FP32 parity
bound for whole reconstructions is 1e-5)."""
import numpy as np
import pytest

import ettention_amd as E
from ettention_amd import api as A
from oracle import oracle as O
from tests._util import bp_fixture, image_comparator_ok, load_golden, max_slice_rms, projection_test_phantom, rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-6

torch = pytest.importorskip("torch")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


TEST_BASE = E.ScannerGeometry.make((-64, -64, -128), (-64, -64, 128), (0.5, 0, 0), (0, 0.5, 0), (256, 256))
V64 = E.VolumeGeometry.make((64, 64, 64))


def og(g):
    return O.Scanner.make(g.source, g.detector, g.horizontal_pitch, g.vertical_pitch, g.resolution)


def ov(v):
    return O.Volume.make(tuple(v.resolution), tuple(v.bbox_min), tuple(v.bbox_max), v.subvolume_count)


def gpu_fp(vg, vol_t, g, samples=1, subvolume=0, mask=None, proj=None, length=None):
    nu, nv = g.resolution
    proj = proj if proj is not None else torch.zeros((nv, nu), dtype=torch.float32, device="cuda")
    length = length if length is not None else torch.zeros((nv, nu), dtype=torch.float32, device="cuda")
    A.forward_project(vg, vol_t, g, proj, length, samples=samples, subvolume=subvolume, mask=mask)
    return proj, length


@pytest.fixture(scope="module")
def phantom():
    return projection_test_phantom(1.0)


@pytest.mark.parametrize("angle", [0.0, 45.0, 225.0, -60.0, 17.0, 90.0])
def test_fp_matches_oracle_and_fixture(phantom, angle):
    g = A.rotate(TEST_BASE, angle)
    p, l = gpu_fp(V64, dev(phantom), g)
    op, ol = O.forward_project(ov(V64), phantom, og(g))
    p, l = host(p), host(l)
    assert rel_l2(p, op) <= TOL and rel_l2(l, ol) <= TOL
    if angle in (45.0, 225.0):
        ref = load_golden("forward_projection")["projection_45"]
        assert image_comparator_ok(p, ref)


def test_fp_synthetic_goldens():
    s = load_golden("synthetic_3_parallel")
    gt = dev(s["ground_truth"])
    worst = 0.0
    for a, ref in zip(s["angles"][::9], s["projections"][::9]):
        p, _ = gpu_fp(V64, gt, A.rotate(TEST_BASE, float(a)))
        worst = max(worst, rel_l2(host(p), ref))
    assert worst < 2e-6


@pytest.mark.parametrize("samples", [2, 4])
def test_fp_subsamples_and_subvolumes(phantom, samples):
    v4 = E.VolumeGeometry.make((64, 64, 64), subvolumes=4)
    g = A.rotate(TEST_BASE, 33.0)
    p = torch.zeros((256, 256), dtype=torch.float32, device="cuda")
    l = torch.zeros_like(p)
    vt = dev(phantom)
    for s in range(4):
        gpu_fp(v4, vt, g, samples=samples, subvolume=s, proj=p, length=l)
    op = np.zeros((256, 256), np.float32)
    ol = np.zeros((256, 256), np.float32)
    for s in range(4):
        O.forward_project(ov(v4), phantom, og(g), samples=samples, slab=s, proj=op, length=ol)
    assert rel_l2(host(p), op) <= TOL and rel_l2(host(l), ol) <= TOL


def test_fp_mask(phantom):
    rng = np.random.default_rng(2)
    mask = rng.integers(0, 256, (64, 64, 64), dtype=np.uint8)
    g = A.rotate(TEST_BASE, 21.0)
    p, l = gpu_fp(V64, dev(phantom), g, mask=dev(mask))
    op, ol = O.forward_project(ov(V64), phantom, og(g), mask=mask)
    assert rel_l2(host(p), op) <= TOL and rel_l2(host(l), ol) <= TOL


@pytest.mark.parametrize("dtype,vt", [("float16", O.VOXEL_F16), ("bfloat16", O.VOXEL_BF16)])
def test_fp_reduced_precision_volume(phantom, dtype, vt):
    tv = dev(phantom).to(getattr(torch, dtype))
    raw = host(tv.view(torch.int16)).view(np.uint16)
    g = A.rotate(TEST_BASE, -28.0)
    p, l = gpu_fp(V64, tv, g)
    op, ol = O.forward_project(ov(V64), raw, og(g), voxel_type=vt)
    assert rel_l2(host(p), op) <= TOL


@pytest.mark.parametrize("dtype,fn", [("bfloat16", "orc_f32_to_bf16_rne"), ("float16", "orc_f32_to_f16_rte")])
def test_storage_rounding_bit_exact(dtype, fn):
    """The kernels' rounding to 16-bit storage (etn_convert_volume uses the
    storage types' store) against the oracle's, bit for bit, on ties,
    denormals, overflow and random bit patterns; NaN stays NaN."""
    from tests.test_oracle_rounding import rounding_cases
    f = rounding_cases(n_random=50000, seed=11)
    out = torch.empty(f.size, dtype=getattr(torch, dtype), device="cuda")
    A.convert_volume(dev(f), out)
    got = host(out.view(torch.int16)).view(np.uint16)
    want = np.array([getattr(O.lib(), fn)(float(x)) for x in f], dtype=np.uint16)
    np.testing.assert_array_equal(got, want)
    nan = torch.full((4,), float("nan"), dtype=torch.float32, device="cuda")
    out = torch.empty(4, dtype=getattr(torch, dtype), device="cuda")
    A.convert_volume(nan, out)
    assert bool(torch.isnan(out.float()).all())


@pytest.mark.parametrize("tag,angle", [("z", 0.0), ("y", 90.0), ("45", 45.0)])
def test_bp_matches_oracle_and_fixture(tag, angle):
    proj, ref = bp_fixture(tag)
    g = A.rotate(TEST_BASE, angle)
    rl = torch.zeros((256, 256), dtype=torch.float32, device="cuda")
    A.compute_ray_length(V64, g, rl)
    orl = O.ray_length(ov(V64), og(g))
    assert rel_l2(host(rl), orl) == 0.0
    vol = torch.zeros((64, 64, 64), dtype=torch.float32, device="cuda")
    A.back_project(V64, vol, g, dev(proj), rl)
    ovol = O.back_project(ov(V64), np.zeros((64, 64, 64), np.float32), og(g), proj, orl)
    out = host(vol)
    assert rel_l2(out, ovol) <= TOL
    assert max_slice_rms(out, ref) <= 0.011


@pytest.mark.parametrize("samples,lam,bs", [(1, 0.3, 1), (4, 1.0, 8), (9, 0.1, 3)])
def test_bp_samples_lambda_subvolumes(samples, lam, bs):
    rng = np.random.default_rng(samples)
    g = A.rotate(TEST_BASE, 27.0)
    res = rng.normal(size=(256, 256)).astype(np.float32)
    v4 = E.VolumeGeometry.make((64, 64, 64), subvolumes=4)
    orl = O.ray_length(ov(v4), og(g))
    init = rng.uniform(0, 1, (64, 64, 64)).astype(np.float32)
    vol = dev(init)
    for s in (3, 2, 1, 0):
        A.back_project(v4, vol, g, dev(res), dev(orl), lam=lam, block_size=bs, samples=samples, subvolume=s)
    ovol = init.copy()
    for s in (3, 2, 1, 0):
        O.back_project(ov(v4), ovol, og(g), res, orl, lam=lam, block_size=bs, samples=samples, slab=s)
    assert rel_l2(host(vol), ovol) <= TOL


def test_bp_mask_and_half_storage():
    rng = np.random.default_rng(9)
    g = A.rotate(TEST_BASE, -50.0)
    res = rng.normal(size=(256, 256)).astype(np.float32)
    orl = O.ray_length(ov(V64), og(g))
    mask = rng.integers(0, 256, (64, 64, 64), dtype=np.uint8)
    vol = torch.zeros((64, 64, 64), dtype=torch.float32, device="cuda")
    A.back_project(V64, vol, g, dev(res), dev(orl), mask=dev(mask))
    ovol = O.back_project(ov(V64), np.zeros((64, 64, 64), np.float32), og(g), res, orl, mask=mask)
    assert rel_l2(host(vol), ovol) <= TOL
    for dtype, vt in (("float16", O.VOXEL_F16), ("bfloat16", O.VOXEL_BF16)):
        init = rng.uniform(0, 1, (64, 64, 64)).astype(np.float32)
        tv = dev(init).to(getattr(torch, dtype))
        raw = host(tv.view(torch.int16)).view(np.uint16).copy()
        A.back_project(V64, tv, g, dev(res), dev(orl))
        O.back_project(ov(V64), raw, og(g), res, orl, voxel_type=vt)
        got = host(tv.view(torch.int16)).view(np.uint16)
        # rounding to storage follows the oracle; allow one ulp where the FP32 sum differs in the last bit
        assert np.abs(got.astype(np.int32) - raw.astype(np.int32)).max() <= 1


def test_compare_loc_and_norm():
    rng = np.random.default_rng(4)
    b = rng.normal(size=(128, 96)).astype(np.float32)
    p = rng.normal(size=(128, 96)).astype(np.float32)
    r = torch.zeros((128, 96), dtype=torch.float32, device="cuda")
    A.compare(dev(b), dev(p), r)
    np.testing.assert_array_equal(host(r), b - p)
    f = load_golden("long_object_compensation")
    proj = dev(f["input_virtual_projection"])
    length = dev(f["input_traversal_length"])
    A.long_object_compensate(proj, length, 184.28073)
    op = f["input_virtual_projection"].copy()
    ol = f["input_traversal_length"].copy()
    O.long_object_compensate(op, ol, 184.28073)
    np.testing.assert_array_equal(host(proj), op)
    np.testing.assert_array_equal(host(length), ol)
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    big = rng.normal(size=3_000_001).astype(np.float32)
    big2 = rng.normal(size=3_000_001).astype(np.float32)
    A.sum_squared_difference(dev(big), dev(big2), out)
    expect = float(((big.astype(np.float64) - big2) ** 2).sum())
    assert abs(host(out)[0] - expect) <= 1e-12 * expect


def test_full_size_forward_projection_properties():
    # 512^3 volume, 512^2 detector (the benchmark shape): a constant volume projects to the
    # traversal length, and projections are linear in the volume.
    vg = E.VolumeGeometry.make((512, 512, 512))
    base = E.ScannerGeometry.mrc_base(512, 512)
    ones = torch.ones((512, 512, 512), dtype=torch.float32, device="cuda")
    for angle in (-60.0, 0.0, 37.0):
        g = A.rotate(base, angle)
        p, l = gpu_fp(vg, ones, g)
        pp, ll = host(p), host(l)
        np.testing.assert_allclose(pp, ll, rtol=1e-6, atol=1e-3)
        rl = torch.zeros((512, 512), dtype=torch.float32, device="cuda")
        A.compute_ray_length(vg, g, rl)
        assert rel_l2(ll, host(rl)) < 1e-5
    del ones
    gen = torch.Generator(device="cuda").manual_seed(5)
    a = torch.rand((512, 512, 512), generator=gen, device="cuda")
    b = torch.rand((512, 512, 512), generator=gen, device="cuda")
    g = A.rotate(base, 23.0)
    pa, _ = gpu_fp(vg, a, g)
    pb, _ = gpu_fp(vg, b, g)
    a += b
    pab, _ = gpu_fp(vg, a, g)
    assert rel_l2(host(pab), host(pa) + host(pb)) < 1e-5
