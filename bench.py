"""SART throughput on one MI355X: GUPS on the 512^3 x 121-tilt workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--voxel-type float|half|bfloat16]
                    [--workload metric|c2|c3|c4|c5|wbp] [--generic] [--lds-back-projection] [--mask]
                    [--fp-samples N] [--bp-samples M] [--no-cpu-baseline] [--no-time-to-target]

A step is one full iteration of the block-iterative loop over the synthetic
tilt series (for SART: every projection's forward projection, residual and
back projection).  GUPS = voxels x projections x steps / seconds / 1e9, with
the volume, the measured projections and every buffer resident in HBM when
the timed region starts.  Prints ONE JSON line on stdout; progress on stderr.
--mask runs the metric and c2 workloads with a prior-knowledge mask (weight 1
inside a y-axis cylinder holding the specimen, 0 outside); --fp-samples N
(N x N sub-rays per pixel) and --bp-samples M (an M-point pattern per voxel)
run them oversampled.  The default metric line also carries
`time_to_target`: bf16 SART with the learned jump (network parameters
committed in ettention_amd/data) against plain FP32 SART on the c5 workload's
held-out phantom (--no-time-to-target skips it).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, GB/s

WORKLOADS = {
    # SART on a 512^3 volume, 121 tilts at +-60 deg in 1 deg steps, 512^2 detector
    "metric": dict(res=(512, 512, 512), det=(512, 512), max_angle=60.0, step=1.0, algorithm="sart",
                   block_size=0, order="random", lam=0.3, beads=0,
                   name="SART 512^3 volume, 121 tilts +-60 deg by 1 deg, 512^2 detector, random order"),
    # block-iterative B=8, 512x512x256 with 200 beads
    "c2": dict(res=(512, 512, 256), det=(512, 512), max_angle=60.0, step=1.0, algorithm="blockiterative",
               block_size=8, order="random", lam=0.3, beads=200,
               name="block-iterative B=8 on 512x512x256 + 200 beads, 121 tilts +-60 deg by 1 deg, 512^2 detector"),
    # 64 independent SART reconstructions advanced together
    "c4": dict(members=64, algorithm="sart", order="random", lam=0.3,
               name="batch of 64 SART reconstructions, cube edges uniform over {64,96,128,192,256}, "
                    "tilt counts uniform over [41,121] at +-60 deg, detector = edge^2"),
    # time to the residual plain FP32 SART reaches after 50 iterations: bf16 SART + learned jump vs plain
    # precision variant: the same SART in f32, bf16 and f16 storage, error per iteration against f32
    "c3": dict(res=(1024, 1024, 256), det=(1024, 1024), max_angle=60.0, step=1.0, iterations=10, lam=0.3,
               order="random",
               name="precision variant: SART on 1024x1024x256, 121 tilts +-60 deg by 1 deg, 1024^2 detector, "
                    "f32 vs bf16 and f16 storage, 10 iterations"),
    # weighted back projection (the WBP plugin): ramp filter + ray length + BP with lambda / N
    "wbp": dict(res=(512, 512, 512), det=(512, 512), max_angle=60.0, step=1.0, lam=1.0,
                name="WBP 512^3 volume, 121 tilts +-60 deg by 1 deg, 512^2 detector (RamLak 128 taps)"),
    "c5": dict(res=512, tilts=121, k0=3, k_max=50, crop=128, lam=0.3, order="random",
               name="time-to-target-residual at 512^3, 121 tilts +-60 deg, Poisson 1e4 counts: bf16 SART with a "
                    "learned jump from iterate 3 vs plain FP32 SART to its lowest residual within 50 iterations"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_flags(args):
    import ettention_amd as E
    return ((E.ETN_RECON_FLAG_GENERIC_KERNELS if args.generic else 0) |
            (E.ETN_RECON_FLAG_LDS_BACK_PROJECTION if args.lds_back_projection else 0))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--voxel-type", default="float", choices=["float", "half", "bfloat16"])
    ap.add_argument("--workload", default="metric", choices=sorted(WORKLOADS))
    ap.add_argument("--generic", action="store_true", help="force the generic (unfused) kernels")
    ap.add_argument("--lds-back-projection", action="store_true",
                    help="fused kernels: back projections on the band in LDS instead of in registers")
    ap.add_argument("--workgroup-threads", type=int, default=0,
                    help="metric/c2: fused kernels' workgroup shape (256 or 512) instead of the autotuner's")
    ap.add_argument("--mask", action="store_true", help="metric/c2: reconstruct with a prior-knowledge mask")
    ap.add_argument("--fp-samples", type=int, default=1, help="metric/c2: forward sub-rays per pixel axis")
    ap.add_argument("--bp-samples", type=int, default=1, help="metric/c2: back projection samples per voxel")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-projections", type=int, default=32,
                    help="projection steps the CPU baseline times (about 0.35 s each on 16 threads)")
    ap.add_argument("--train-phantoms", type=int, default=8, help="c5: training phantoms (512^3 each)")
    ap.add_argument("--train-steps", type=int, default=2000, help="c5: Adam steps on 128^3 crops")
    ap.add_argument("--save-params", default="", help="c5: write the trained network's parameters (numpy .npy)")
    ap.add_argument("--no-time-to-target", action="store_true",
                    help="metric: skip the time-to-target measurement with the committed network parameters")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a 1 GiB device copy (traffic calibration for rocprof counters)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the volume shards along y with "
              "no exchange; see the design notes). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    if (args.mask or args.fp_samples != 1 or args.bp_samples != 1) and args.workload not in ("metric", "c2"):
        print("--mask, --fp-samples and --bp-samples apply to the metric and c2 workloads", file=sys.stderr)
        sys.exit(2)
    if args.workload == "c4":
        run_batch(args)
        return
    if args.workload == "c5":
        run_learned(args)
        return
    if args.workload == "c3":
        run_precision(args)
        return
    if args.workload == "wbp":
        run_wbp(args)
        return

    import torch
    import ettention_amd as E
    from ettention_amd import api as A
    from ettention_amd import synthetic as S

    wl = WORKLOADS[args.workload]
    nx, ny, nz = wl["res"]
    nu, nv = wl["det"]
    V = nx * ny * nz
    P = nu * nv
    tilts = S.tilt_series(wl["max_angle"], wl["step"])
    N = len(tilts)
    vg = E.VolumeGeometry.make((nx, ny, nz))
    base = E.ScannerGeometry.mrc_base(nu, nv)
    dev = torch.device("cuda:0")

    t0 = time.time()
    phantom = S.ellipsoid_phantom((nz, ny, nx), seed=1000 + (1 if args.workload == "metric" else 2),
                                  beads=wl["beads"], device=dev)
    projs = torch.zeros((N, nv, nu), dtype=torch.float32, device=dev)
    scratch = torch.zeros((nv, nu), dtype=torch.float32, device=dev)
    for i, t in enumerate(tilts):
        scratch.zero_()
        A.forward_project(vg, phantom, A.rotate(base, float(t)), projs[i], scratch)
    projs = S.add_gaussian_noise(projs, 0.02, seed=7)
    torch.cuda.synchronize()
    log("synthetic input ready in %.1f s (phantom max %.2f)" % (time.time() - t0, float(phantom.max())))
    del phantom, scratch

    mask = support_mask(nx, ny, nz, dev) if args.mask else None
    flags = run_flags(args)
    # the reconstruction runs on a stream of its own, on which the timed
    # region's HIP events are recorded
    stream = torch.cuda.Stream()
    r = A.BlockIterativeReconstructionOperator(volume=vg, base_geometry=base, tilt_angles=tilts, projections=projs,
                                               algorithm=wl["algorithm"], block_size=wl["block_size"],
                                               lam=wl["lam"], projection_iteration=wl["order"],
                                               voxel_type=args.voxel_type, flags=flags, mask=mask,
                                               fp_samples=args.fp_samples, bp_samples=args.bp_samples,
                                               stream=stream.cuda_stream)
    fused = r.uses_fused_path()
    if args.workgroup_threads:
        r.set_workgroup_threads(args.workgroup_threads)
    for _ in range(args.warmup):
        r.run_single_iteration()
    r.synchronize()
    shape = (r.tile() + (r.workgroup_threads(),)) if fused else None
    log("warmup done (fused path: %s, tile (planes, band, threads): %s)" % (fused, shape))

    # timed region: K iterations, synchronised on both sides
    torch.cuda.synchronize()
    ev_start, ev_end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t_start = time.perf_counter()
    ev_start.record(stream)
    r.run_iterations(args.steps)
    ev_end.record(stream)
    r.synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    gpu_ms_per_step = ev_start.elapsed_time(ev_end) / args.steps
    ms_per_step = 1e3 * elapsed / args.steps
    gups = V * N * args.steps / elapsed / 1e9
    log("timed %d iterations: %.2f ms/iteration (%.2f ms between the stream's events), %.1f GUPS"
        % (args.steps, ms_per_step, gpu_ms_per_step, gups))
    t_norm = time.perf_counter()
    norm = r.residual_norm()
    log("residual norm after %d iterations: %.6e (FP64, %.1f ms)"
        % (args.warmup + args.steps, norm, 1e3 * (time.perf_counter() - t_norm)))

    block = wl["block_size"] or (1 if wl["algorithm"] == "sart" else N)
    # the committed PMC traffic summaries cover the 512^3 metric workload (f32 and bf16; f32 masked)
    plain = args.fp_samples == 1 and args.bp_samples == 1
    roofline = kernel_roofline(r, args.voxel_type, fused, V, P, block, masked=mask is not None,
                               pmc=args.workload == "metric" and plain, gpu_ms_per_step=gpu_ms_per_step, N=N)
    if args.fp_samples != 1 or args.bp_samples != 1:
        roofline["kernel"] += ", FP %dx%d sub-rays, BP %d samples" % (args.fp_samples, args.fp_samples,
                                                                     args.bp_samples)
    del r
    torch.cuda.empty_cache()
    # the achievable bandwidth on this box, and the dominant kernel's measured
    # HBM traffic rate beside it
    roofline["stream_copy_gbs"] = stream_copy_gbs(dev)
    if roofline.get("traffic"):
        # the measured HBM bytes per launch at the timed launch rate, against the peak
        roofline["traffic_gbs"] = round(roofline["traffic"] / (roofline["avg_launch_ms"] * 1e-3) / 1e9, 1)
        roofline["traffic_frac"] = round(roofline["traffic_gbs"] / HBM_PEAK_GBS, 4)

    if args.calibrate_copy:
        src = torch.empty(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("calibration copy: 1 GiB read + 1 GiB write")
        del src, dst

    cpu = None
    if not args.no_cpu_baseline:
        # oversampling multiplies the oracle's work per step by about (fp^2 + bp) / 2
        cost = max(1.0, (args.fp_samples ** 2 + args.bp_samples) / 2.0)
        cpu = cpu_baseline(wl, tilts, projs, max(1, int(round(args.cpu_sample_projections / cost))),
                           None if mask is None else mask.cpu().numpy(), args.fp_samples, args.bp_samples)
    del mask

    out = {
        "metric": "SART GUPS (proj+backproj) 512^3 x 121 tilts",
        "value": round(gups, 3),
        "unit": "GUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"float": "f32", "half": "f16-storage/f32-accumulate",
                  "bfloat16": "bf16-storage/f32-accumulate"}[args.voxel_type],
        "data": "synthetic (seeded ellipsoid phantom, FP projections + 2% Gaussian noise)",
        "config": {"workload": wl["name"], "volume": [nx, ny, nz], "detector": [nu, nv], "projections": N,
                   "algorithm": wl["algorithm"], "block_size": wl["block_size"] or (1 if wl["algorithm"] == "sart" else N),
                   "voxel_type": args.voxel_type, "kernels": "fused-plane" if fused else "generic",
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if args.mask:
        out["config"]["mask"] = MASK_NAME
    if args.fp_samples != 1 or args.bp_samples != 1:
        out["config"]["fp_samples"] = args.fp_samples
        out["config"]["bp_samples"] = args.bp_samples
    if (args.workload == "metric" and args.voxel_type == "float" and plain and not args.mask and not args.generic
            and not args.no_time_to_target):
        del projs
        torch.cuda.empty_cache()
        out["time_to_target"] = committed_time_to_target(log)
    print(json.dumps(out), flush=True)


NET_PARAMS = os.path.join(ROOT, "ettention_amd", "data", "correction_net_512.npy")


def committed_time_to_target(log):
    """The metric's second half: time to the target residual (the lowest
    ||Ax - b|| plain FP32 SART reaches within 50 iterations) for bf16-storage
    SART with the learned jump from iterate 3 against plain FP32 SART, on the
    c5 workload's held-out 512^3 phantom (121 tilts, Poisson noise at 1e4
    counts), with the network parameters committed in ettention_amd/data
    (trained by `bench.py --workload c5 --save-params`).  The jump counts only
    when accepted (its residual below plain SART's at equal wall time);
    otherwise plain SART is the result and the speedup is 1."""
    import torch
    from ettention_amd import learned as L
    if not os.path.exists(NET_PARAMS):
        return {"skipped": "no committed network parameters"}
    try:
        wl = WORKLOADS["c5"]
        n, N = wl["res"], wl["tilts"]
        params = np.load(NET_PARAMS, allow_pickle=False)
        t0 = time.time()
        net = L.CorrectionNet((n, n, n), params)
        case, ph = L.make_case((n, n, n), N, seed=5000)
        del ph
        res = L.time_to_target(case, net, 1.0, k_max=wl["k_max"], k0=wl["k0"], lam=wl["lam"], order=wl["order"],
                               log=log)
        del net, case
        torch.cuda.empty_cache()
        speedup = res["speedup"] if (res["accepted"] and res["speedup"]) else 1.0
        log("time to target: %.2f s of measurement" % (time.time() - t0))
        return {"speedup": round(speedup, 4), "unit": "x", "accepted": res["accepted"],
                "target_norm": round(res["target_norm"], 2), "plain_best_iteration": res["plain_best_iteration"],
                "plain_s": round(res["plain_s"], 4), "jump_s": round(res["jump_s"], 4),
                "jump_norm": round(res["jump_norm"], 2), "plain_norm_at_jump_time":
                round(res["plain_norm_at_jump_time"], 2), "learned_reached_target": res["reached"],
                "learned_s": round(res["learned_s"], 4),
                "workload": wl["name"], "params": os.path.relpath(NET_PARAMS, ROOT)}
    except Exception as e:   # the metric line stands on its own
        return {"error": "%s: %s" % (type(e).__name__, e)}


def stream_copy_gbs(dev, floats=1 << 28, reps=5):
    """Device-to-device copy bandwidth (read + write bytes / time) of two 1 GiB
    buffers, HIP events around `reps` copies after one untimed copy."""
    import torch
    src = torch.ones(floats, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        dst.copy_(src)
    b.record()
    b.synchronize()
    gbs = 2 * 4 * floats * reps / (a.elapsed_time(b) * 1e-3) / 1e9
    del src, dst
    torch.cuda.empty_cache()
    log("stream copy: %.1f GB/s (1 GiB read + 1 GiB write per copy)" % gbs)
    return round(gbs, 1)


MASK_NAME = "prior-knowledge mask: 255 inside the y-axis cylinder of radius 0.45 nx centred in x-z, 0 outside"


def support_mask(nx, ny, nz, dev):
    """USE_PRIOR_KNOWLEDGE_MASK input for the benchmark: uint8 (nz, ny, nx), 255
    inside a y-axis cylinder that holds the phantom's support, 0 outside."""
    import torch
    x = torch.arange(nx, device=dev, dtype=torch.float32) + 0.5 - nx / 2
    z = torch.arange(nz, device=dev, dtype=torch.float32) + 0.5 - nz / 2
    inside = (x[None, :] ** 2 + z[:, None] ** 2) <= (0.45 * nx) ** 2
    return (inside.to(torch.uint8) * 255)[:, None, :].expand(nz, ny, nx).contiguous()


def run_precision(args):
    """Precision variant: SART on 1024x1024x256 x 121 tilts with the volume
    stored in f32, bf16 and f16 (16-bit forms rounded to nearest even, and with
    the opt-in stochastic rounding), FP32 ray accumulation throughout.

    Full volume: GUPS of each form and, after every iteration, the relative L2
    distance of each 16-bit volume from the f32 GPU volume.

    Against the FP64 oracle: in this geometry volume plane y is traced only by
    detector row y and back-projected only from it, so the planes are
    independent up to the float rounding of the projection matrix (a plane's
    detector-row weight is 0 only to rounding, so neighbouring planes couple
    weakly), and a y-slab of the volume with its detector rows is one shard's
    reconstruction.  The 16 planes y = 504..519 run as their own 1024x16x256
    problem (the same coordinates: the slab straddles y = 0) on the GPU in every
    form and in the FP64 oracle; after every iteration each GPU slab's relative
    L2 distance from the FP64 slab is reported, and at the end the f32 slab's
    distance from those planes of the full f32 volume."""
    import torch
    import ettention_amd as E
    from ettention_amd import api as A
    from ettention_amd import synthetic as S
    from oracle import oracle as O

    wl = WORKLOADS["c3"]
    nx, ny, nz = wl["res"]
    nu, nv = wl["det"]
    iters = wl["iterations"]
    tilts = S.tilt_series(wl["max_angle"], wl["step"])
    N = len(tilts)
    V, P = nx * ny * nz, nu * nv
    vg = E.VolumeGeometry.make((nx, ny, nz))
    base = E.ScannerGeometry.mrc_base(nu, nv)
    dev = torch.device("cuda:0")
    phantom = S.ellipsoid_phantom((nz, ny, nx), seed=1003, device=dev)
    projs = torch.zeros((N, nv, nu), dtype=torch.float32, device=dev)
    scratch = torch.zeros((nv, nu), dtype=torch.float32, device=dev)
    for i, t in enumerate(tilts):
        scratch.zero_()
        A.forward_project(vg, phantom, A.rotate(base, float(t)), projs[i], scratch)
    projs = S.add_gaussian_noise(projs, 0.02, seed=7)
    del scratch, phantom
    torch_type = {"float": torch.float32, "bfloat16": torch.bfloat16, "half": torch.float16}
    forms = {"float": ("float", 0), "bfloat16": ("bfloat16", 0), "half": ("half", 0),
             "bfloat16_sr": ("bfloat16", E.ETN_RECON_FLAG_STOCHASTIC_ROUNDING),
             "half_sr": ("half", E.ETN_RECON_FLAG_STOCHASTIC_ROUNDING)}

    def make(volume, base_g, p):
        return {f: A.BlockIterativeReconstructionOperator(volume=volume, base_geometry=base_g, tilt_angles=tilts,
                                                          projections=p, algorithm="sart", lam=wl["lam"],
                                                          projection_iteration=wl["order"], voxel_type=vt, flags=fl)
                for f, (vt, fl) in forms.items()}

    ops = make(vg, base, projs)
    for f, r in ops.items():      # tile autotuning and graph capture outside the timed region
        r.run_iterations(1)
        r.set_volume(torch.zeros(r.shape, dtype=torch_type[forms[f][0]], device=dev))
    elapsed = {f: 0.0 for f in ops}
    per_iter = []
    for it in range(1, iters + 1):
        vols = {}
        for f, r in ops.items():
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r.run_iterations(1)
            r.synchronize()
            elapsed[f] += time.perf_counter() - t0
            vols[f] = r.get_reconstructed_volume(dtype=torch.float32)
        ref = vols["float"]
        rn = float(torch.linalg.vector_norm(ref.double()))
        row = {"iteration": it}
        for f in forms:
            if f != "float":
                row[f] = float(torch.linalg.vector_norm((vols[f] - ref).double())) / rn
        per_iter.append(row)
        log("iteration %d: rel L2 vs f32 " % it + "  ".join("%s %.3e" % (f, row[f]) for f in forms if f != "float"))
        del vols, ref
    full_f32_slab = ops["float"].get_reconstructed_volume(dtype=torch.float32)[:, 504:520, :].cpu().numpy()
    gups = {f: V * N * iters / elapsed[f] / 1e9 for f in ops}
    log("GUPS: " + ", ".join("%s %.1f" % kv for kv in gups.items()))
    ms_bf16 = 1e3 * elapsed["bfloat16"] / iters
    roofline = kernel_roofline(ops["bfloat16"], "bfloat16", True, V, P, 1, pmc=False, gpu_ms_per_step=ms_bf16, N=N)
    roofline["kernel"] += ", bf16 storage"
    roofline["time_source"] = "host clock around each synchronised iteration"
    norms = {f: r.residual_norm() for f, r in ops.items()}
    del ops
    torch.cuda.empty_cache()

    # the y-slab against the FP64 oracle
    y0, sh = 504, 16
    svg = E.VolumeGeometry.make((nx, sh, nz))
    sbase = E.ScannerGeometry.mrc_base(nu, sh)
    sprojs = projs[:, y0:y0 + sh, :].contiguous()
    sops = make(svg, sbase, sprojs)
    ov = O.Volume.make((nx, sh, nz))
    spec = O.ReconSpec(ov, O.mrc_base_scanner(nu, sh), tilts, sprojs.cpu().numpy(), block_size=1, order=wl["order"],
                       lam=wl["lam"])
    d64 = np.zeros((nz, sh, nx), np.float64)
    vs_oracle = []
    t_oracle = 0.0
    for it in range(1, iters + 1):
        t0 = time.perf_counter()
        spec.run_f64(d64, 1)
        t_oracle += time.perf_counter() - t0
        dn = np.linalg.norm(d64)
        row = {"iteration": it}
        for f, r in sops.items():
            r.run_iterations(1)
            got = r.get_reconstructed_volume(dtype=torch.float32).cpu().numpy().astype(np.float64)
            row[f] = float(np.linalg.norm(got - d64) / dn)
        vs_oracle.append(row)
        log("slab iteration %d: rel L2 vs FP64 oracle " % it + "  ".join("%s %.3e" % (f, row[f]) for f in forms))
    slab_f32 = sops["float"].get_reconstructed_volume(dtype=torch.float32).cpu().numpy()
    slab_vs_full = float(np.linalg.norm((slab_f32 - full_f32_slab).astype(np.float64)) /
                         np.linalg.norm(full_f32_slab.astype(np.float64)))
    log("FP64 oracle: %d slab iterations in %.1f s; f32 slab vs the full volume's planes: rel L2 %.3e"
        % (iters, t_oracle, slab_vs_full))
    del sops
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(wl, tilts, projs, 6)
    out = {
        "metric": "SART GUPS, bf16 storage, 1024x1024x256 x 121 tilts (precision variant)",
        "value": round(gups["bfloat16"], 3),
        "unit": "GUPS",
        "n_gpus": 1,
        "steps": iters,
        "warmup": 1,
        "ms_per_step": round(ms_bf16, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16-storage/f32-accumulate",
        "data": "synthetic (seeded ellipsoid phantom, FP projections + 2% Gaussian noise)",
        "config": {"workload": wl["name"], "volume": [nx, ny, nz], "detector": [nu, nv], "projections": N,
                   "parallelism": "single GPU"},
        "gups": {f: round(g, 3) for f, g in gups.items()},
        "residual_norm_after": norms,
        "rel_l2_vs_f32_per_iteration": per_iter,
        "rel_l2_vs_fp64_oracle_per_iteration": vs_oracle,
        "fp64_slab": {"planes": [y0, y0 + sh], "volume": [nx, sh, nz], "detector": [nu, sh],
                      "f32_slab_vs_full_volume_planes_rel_l2": slab_vs_full, "oracle_seconds": round(t_oracle, 1)},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def run_wbp(args):
    """Weighted back projection: a step is one whole WBP reconstruction (every
    projection filtered, its ray length computed and back-projected with
    lambda / N).  GUPS = V x N / seconds per reconstruction.  The dominant
    kernel is the fused back projection (eight projections per volume sweep),
    timed with HIP events around its launches in one more run; its
    algorithmic bytes are those of one sweep for B = 8 projections, 2 s V + 8 P B
    (the volume read and written once, residual and length images read)."""
    import torch
    import ettention_amd as E
    from ettention_amd import api as A
    from ettention_amd import synthetic as S

    wl = WORKLOADS["wbp"]
    nx, ny, nz = wl["res"]
    nu, nv = wl["det"]
    V, P = nx * ny * nz, nu * nv
    tilts = S.tilt_series(wl["max_angle"], wl["step"])
    N = len(tilts)
    vg = E.VolumeGeometry.make((nx, ny, nz))
    base = E.ScannerGeometry.mrc_base(nu, nv)
    dev = torch.device("cuda:0")
    phantom = S.ellipsoid_phantom((nz, ny, nx), seed=1003, device=dev)
    projs = torch.zeros((N, nv, nu), dtype=torch.float32, device=dev)
    scratch = torch.zeros((nv, nu), dtype=torch.float32, device=dev)
    for i, t in enumerate(tilts):
        scratch.zero_()
        A.forward_project(vg, phantom, A.rotate(base, float(t)), projs[i], scratch)
    projs = S.add_gaussian_noise(projs, 0.02, seed=7)
    del phantom, scratch
    torch.cuda.synchronize()
    flags = run_flags(args)
    r = A.WeightedBackProjection(volume=vg, base_geometry=base, tilt_angles=tilts, projections=projs, lam=wl["lam"],
                                 flags=flags)
    zero = torch.zeros((nz, ny, nx), dtype=torch.float32, device=dev)
    for _ in range(args.warmup):
        r.set_volume(zero)
        r.run()
    r.synchronize()
    total = 0.0
    for _ in range(args.steps):
        r.set_volume(zero)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r.run()
        r.synchronize()
        torch.cuda.synchronize()
        total += time.perf_counter() - t0
    ms_per_step = 1e3 * total / args.steps
    gups = V * N / (total / args.steps) / 1e9
    log("timed %d WBP reconstructions: %.2f ms each, %.1f GUPS (fused path: %s)"
        % (args.steps, ms_per_step, gups, r.uses_fused_path()))
    kernel, bp_ms, launches = r.time_kernels()
    del r
    torch.cuda.empty_cache()
    if kernel == 0:
        B = 8
        name = "fused_plane_kernel (back projections only, %d projections per launch)" % B
    else:
        B = 1
        name = "back_project_kernel (generic, one projection with lambda / N)"
    alg = 2 * 4 * V + 8 * P * B
    achieved = alg / (bp_ms * 1e-3) / 1e9
    log("%s: %.4f ms per launch over %d launches -> %.1f GB/s algorithmic" % (name, bp_ms, launches, achieved))
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": name,
                "avg_launch_ms": round(bp_ms, 5), "algorithmic_bytes_per_launch": alg, "timed_launches": launches,
                "stream_copy_gbs": stream_copy_gbs(dev)}
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        from oracle import wbp_oracle as W
        k = max(1, args.cpu_sample_projections // 4)
        t0 = time.perf_counter()
        W.wbp(O.Volume.make((nx, ny, nz)), O.mrc_base_scanner(nu, nv), tilts[:k], projs[:k].cpu().numpy(),
              lam=wl["lam"])
        dt = time.perf_counter() - t0
        cpu = {"value": round(V * k / dt / 1e9, 5), "unit": "GUPS", "cores": O.num_threads(), "kind": "port",
               "sample": "WBP of the first %d projections (filter, ray length, BP) on the full %dx%dx%d volume"
                         % (k, nx, ny, nz)}
        log("cpu baseline: %d WBP projections in %.1f s -> %.4f GUPS" % (k, dt, cpu["value"]))
    out = {
        "metric": "WBP GUPS (filter + backproj) 512^3 x 121 tilts",
        "value": round(gups, 3),
        "unit": "GUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded ellipsoid phantom, FP projections + 2% Gaussian noise)",
        "config": {"workload": wl["name"], "volume": [nx, ny, nz], "detector": [nu, nv], "projections": N,
                   "algorithm": "wbp", "kernels": "fused-plane" if kernel == 0 else "generic",
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def run_batch(args):
    """The batched workload: GUPS = sum over members of voxels x projections,
    per iteration, over the wall time of the iterations."""
    import torch
    import ettention_amd as E
    from ettention_amd import api as A
    from ettention_amd import synthetic as S

    wl = WORKLOADS["c4"]
    dev = torch.device("cuda:0")
    comp = S.batch_composition(wl["members"], 1004)
    t0 = time.time()
    members = []
    for i, (e, n) in enumerate(comp):
        vg = E.VolumeGeometry.make((e, e, e))
        base = E.ScannerGeometry.mrc_base(e, e)
        tilts = np.linspace(-60.0, 60.0, n).astype(np.float32)
        ph = S.ellipsoid_phantom((e, e, e), seed=4000 + i, device=dev)
        projs = torch.zeros((n, e, e), dtype=torch.float32, device=dev)
        scratch = torch.zeros((e, e), dtype=torch.float32, device=dev)
        for k, t in enumerate(tilts):
            scratch.zero_()
            A.forward_project(vg, ph, A.rotate(base, float(t)), projs[k], scratch)
        projs = S.add_gaussian_noise(projs, 0.02, seed=7000 + i)
        members.append(dict(volume=vg, base_geometry=base, tilt_angles=tilts, projections=projs,
                            algorithm=wl["algorithm"], lam=wl["lam"], projection_iteration=wl["order"],
                            voxel_type=args.voxel_type))
        del ph, scratch
    torch.cuda.synchronize()
    log("synthetic batch of %d members ready in %.1f s" % (len(members), time.time() - t0))
    flags = run_flags(args)
    b = A.BatchReconstruction(members, flags=flags)
    fused = all(b.uses_fused_path(i) for i in range(len(b)))
    for _ in range(args.warmup):
        b.run_iterations(1)
    b.synchronize()
    tiles = sorted({(comp[i][0],) + b.tile(i) for i in range(len(b))}) if fused else []
    log("warmup done (fused: %s, tiles (edge, planes, band): %s, workgroup threads %s)"
        % (fused, tiles, b.workgroup_threads(0) if fused else None))
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    b.run_iterations(args.steps)
    b.synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    updates = sum(e ** 3 * n for e, n in comp)
    gups = updates * args.steps / elapsed / 1e9
    ms_per_step = 1e3 * elapsed / args.steps
    log("timed %d iterations: %.2f ms/iteration, %.1f GUPS" % (args.steps, ms_per_step, gups))
    roofline = None
    if fused:
        s = 4 if args.voxel_type == "float" else 2
        kernel_ms, launches = b.time_iteration()
        alg = sum((3 * s * e ** 3 + 20 * e * e) * n for e, n in comp)
        achieved = alg / (kernel_ms * 1e-3) / 1e9
        log("fused launches: %d per iteration, %.2f ms of kernel time -> %.1f GB/s algorithmic"
            % (launches, kernel_ms, achieved))
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                    "kernel": "fused_plane_kernel over all members (one launch per pass index)",
                    "avg_launch_ms": round(kernel_ms / max(launches, 1), 5),
                    "algorithmic_bytes_per_launch": int(alg / max(launches, 1)), "timed_launches": launches}
    out = {
        "metric": "batched SART GUPS (proj+backproj), 64 members",
        "value": round(gups, 3),
        "unit": "GUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"float": "f32", "half": "f16-storage/f32-accumulate",
                  "bfloat16": "bf16-storage/f32-accumulate"}[args.voxel_type],
        "data": "synthetic (per-member seeded ellipsoid phantoms, FP projections + 2% Gaussian noise)",
        "config": {"workload": wl["name"], "members": len(comp), "updates_per_iteration": updates,
                   "voxel_type": args.voxel_type, "kernels": "fused-plane batch" if fused else "generic",
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def run_learned(args):
    """Learned correction: trains the network on samples of seeded training
    phantoms (input: bf16 SART iterate 3 and its last update; target: the FP32
    SART minimum-residual iterate within 50), then measures time to the target
    residual on a held-out phantom (learned.time_to_target)."""
    import torch
    from ettention_amd import learned as L

    wl = WORKLOADS["c5"]
    n, N, k0, K, crop = wl["res"], wl["tilts"], wl["k0"], wl["k_max"], wl["crop"]
    t0 = time.time()
    samples = []
    for i in range(args.train_phantoms):
        case, ph = L.make_case((n, n, n), N, seed=5001 + i)
        del ph
        sample, k_best = L.training_sample(case, k0, K, lam=wl["lam"], order=wl["order"])
        samples.append(sample)
        del case
        log("training sample %d ready: late iterate %d (%.1f s)" % (i, k_best, time.time() - t0))
    scale = 1.0
    net = L.CorrectionNet((crop, crop, crop), seed=5000)
    t1 = time.time()
    losses = L.train_on_crops(net, samples, crop, args.train_steps, scale, lr=1e-3, seed=1, log=log)
    train_s = time.time() - t1
    log("trained %d steps on %d^3 crops in %.1f s: mean loss of the first / last 100 steps %.4e -> %.4e"
        % (args.train_steps, crop, train_s, np.mean(losses[:100]), np.mean(losses[-100:])))
    params = net.parameters()
    if args.save_params:
        np.save(args.save_params, params.astype(np.float32), allow_pickle=False)
        log("parameters written to %s" % args.save_params)
    del net, samples
    torch.cuda.empty_cache()
    full = L.CorrectionNet((n, n, n), params)
    case, ph = L.make_case((n, n, n), N, seed=5000)
    del ph
    x = torch.rand((n, n, n), device="cuda")
    full.time_forward(x, scale)
    fwd_ms, flops = full.time_forward(x, scale)
    del x
    log("network forward at %d^3: %.2f ms, %.1f TFLOP/s" % (n, fwd_ms, flops / fwd_ms / 1e9))
    res = L.time_to_target(case, full, scale, k_max=K, k0=k0, lam=wl["lam"], order=wl["order"], log=log)
    tflops = flops / fwd_ms / 1e9
    out = {
        "metric": "time-to-target-residual speedup, bf16 SART + learned correction vs plain FP32 SART",
        "value": round(res["speedup"], 4) if res["speedup"] else 0.0,
        "unit": "x",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": round(1e3 * res["learned_s"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16-storage/f32-accumulate (SART); bf16 MFMA, f32 accumulate (network)",
        "data": "synthetic (seeded ellipsoid phantoms, FP projections + Poisson noise at 1e4 counts)",
        "config": {"workload": wl["name"], "volume": [n, n, n], "projections": N, "k0": k0, "k_max": K,
                   "train_phantoms": args.train_phantoms, "train_steps": args.train_steps, "train_crop": crop,
                   "train_seconds": round(train_s, 2), "parallelism": "single GPU"},
        "result": {k: (round(v, 6) if isinstance(v, float) else v) for k, v in res.items() if k != "plain_norms"},
        "roofline": {"bound": "mfma", "achieved": round(tflops, 1), "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": round(tflops / 2500.0, 4), "traffic": None,
                     "kernel": "conv3_fwd_kernel (whole network forward at %d^3)" % n,
                     "avg_launch_ms": round(fwd_ms, 4), "flops_per_forward": flops},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def kernel_roofline(r, voxel_type, fused, V, P, block=1, masked=False, pmc=True, gpu_ms_per_step=None, N=1):
    """The dominant kernel's algorithmic bytes against its measured time.

    Fused plane kernel: an iteration's launches are all this kernel.  Its
    algorithmic bytes per iteration are the unfused schedule's (the figure the
    roofline is defined by): per block of B projections 3 s V + 20 P B (the FP
    reads the volume, the BP reads and writes it, 20 bytes per detector pixel
    and projection), summed over the ceil(N / B) blocks.  The time is the
    timed region's HIP-event time per iteration, so achieved = bytes per
    iteration / time per iteration; the average launch is that time over the
    launches one iteration makes (counted by the library; for SART N + 1: the
    first launch only forward-projects, the last only back-projects).  A
    launch that back-projects and forward-projects moves 2 s V + 28 P B' bytes
    itself (B' projections each way: the compulsory bytes); the rocprofv3
    counters give the measured traffic.

    Generic kernels (FP: s V + 8 P; BP: 2 s V + 8 P per launch) share the
    timed region with the residual kernels, so their duration comes from one
    extra eager iteration with HIP events around every launch.  A mask adds
    one byte per voxel to each volume read (the FP and the BP each read it)."""
    s = 4 if voxel_type == "float" else 2
    m = 1 if masked else 0
    kernel, eager_ms, launches = r.time_kernels()
    if fused:
        assert kernel == 0
        blocks = -(-N // block)
        alg_iter = blocks * (3 * s + 2 * m) * V + 20 * P * N
        dur_ms = gpu_ms_per_step / launches
        alg_bytes = alg_iter / launches
        kernel_bytes = (2 * s + m) * V + 28 * P * N / launches   # projections per launch, on average
        source = ("timed region: HIP events on the reconstruction stream, time per iteration / %d launches"
                  % launches)
        name = ("fused_plane_kernel (BP of projection k-1 + FP/residual of projection k)" if block == 1 else
                "fused_plane_kernel (BPs of block b-1 + FPs/residuals of block b, %d projections per block)" % block)
    else:
        dur_ms = eager_ms
        source = "one extra eager iteration, HIP events around every launch"
        if kernel == 1:
            name, alg_bytes = "forward_project_kernel (generic DDA)", (s + m) * V + 8 * P
        else:
            name, alg_bytes = "back_project_kernel (generic)", (2 * s + m) * V + 8 * P
        kernel_bytes = alg_bytes
    if masked:
        name += ", masked"
    achieved = alg_bytes / (dur_ms * 1e-3) / 1e9
    log("dominant kernel %s: %.4f ms average launch (%s) -> %.1f GB/s algorithmic"
        % (name, dur_ms, source, achieved))
    traffic, tsource, l2 = measured_traffic(kernel, voxel_type, masked) if pmc else (None, None, None)
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": tsource,
            "l2_hit_rate": l2, "kernel": name,
            "avg_launch_ms": round(dur_ms, 5), "avg_launch_source": source,
            "algorithmic_bytes_per_launch": int(alg_bytes),
            "kernel_compulsory_bytes_per_launch": int(kernel_bytes),
            "compulsory_frac": round(kernel_bytes / (dur_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "launches_per_iteration": launches}


def measured_traffic(kernel, voxel_type, masked=False):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (copy-calibrated), and its L2 hit rate, when
    the summary exists."""
    name = {0: "fused_plane_kernel", 1: "forward_project_kernel", 2: "back_project_kernel"}[kernel]
    vt = {"float": "f32", "half": "f16", "bfloat16": "bf16"}[voxel_type]
    path = os.path.join(ROOT, "profiles", "%s_sart_512cubed_121_%s%s_traffic.md" % (name, vt, "_masked" if masked else ""))
    if not os.path.exists(path):
        return None, None, None
    traffic = l2 = None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            traffic = int(line.split(":")[1])
        if line.startswith("l2_hit_rate:"):
            l2 = float(line.split(":")[1])
    return traffic, os.path.relpath(path, ROOT), l2


def cpu_baseline(wl, tilts, projs, nsample, mask=None, fp_samples=1, bp_samples=1):
    """The oracle (FP32-faithful C restatement of the reference's OpenCL path,
    OpenMP over the host cores) on a bounded sample of the same workload: the
    first `nsample` projection steps (FP + residual + BP) of a SART pass over
    the full-size volume, with the workload's mask if it has one."""
    from oracle import oracle as O
    nx, ny, nz = wl["res"]
    nu, nv = wl["det"]
    v = O.Volume.make((nx, ny, nz))
    base = O.mrc_base_scanner(nu, nv)
    vol = np.zeros((nz, ny, nx), np.float32)
    meas = projs[:nsample].cpu().numpy()
    t0 = time.perf_counter()
    for i in range(nsample):
        g = O.rotate(base, float(tilts[i]))
        p, l = O.forward_project(v, vol, g, samples=fp_samples, mask=mask)
        res = O.compare(meas[i], p)
        O.back_project(v, vol, g, res, l, lam=wl["lam"], samples=bp_samples, mask=mask)
    dt = time.perf_counter() - t0
    gups = nx * ny * nz * nsample / dt / 1e9
    log("cpu baseline: %d projection steps in %.1f s on %d threads -> %.4f GUPS" % (nsample, dt, O.num_threads(), gups))
    return {"value": round(gups, 5), "unit": "GUPS", "cores": O.num_threads(), "kind": "port",
            "sample": "%d SART projection steps (FP+residual+BP) on the full %dx%dx%d volume, %d^2 detector%s%s"
                      % (nsample, nx, ny, nz, nu, ", masked" if mask is not None else "",
                         ", FP %dx%d sub-rays, BP %d samples" % (fp_samples, fp_samples, bp_samples)
                         if fp_samples != 1 or bp_samples != 1 else ""),
            "cpu_model": host_cpu_model(), "build": "gcc -O2 -ffp-contract=off -fopenmp (FP32 as written)"}


def host_cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or None where it is not readable."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


if __name__ == "__main__":
    main()
