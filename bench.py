#!/usr/bin/env python3
"""Benchmark of the MI355X Beta-splat rasteriser: fwd+bwd frames/s.

One step = one training step of the BetaModel render path on one view:
parameter activations, view-direction query, covariance construction,
conditioning, projection (fused into one launch each way by default,
--path chain runs them as the reference's separate ops), tile intersection +
sort + ranges, blend forward, L1 loss against a target image (the fused
loss kernels; --lambda-dssim adds SSIM as the reference's training loss),
and the backward of all of it down to the leaf parameters.  Inputs are synthetic (seeded) and resident in HBM before the
timed region.  Prints one JSON line: throughput, latency percentiles, the
roofline of the dominant kernel and a CPU baseline from the oracle.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2|c3|c1|c4|c5]
"""
from __future__ import annotations

import argparse
import contextlib
import gc
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "universal-beta-splatting_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

METRIC = "frames/s fwd+bwd & batched-serve MPix/s at 1080p, 1M prims; p50/p99 latency"
HBM_PEAK_GBS = 8000.0

CONFIGS = {
    "c2": dict(N=1_000_000, D=6, W=1920, H=1080,
               workload="1M-primitive static 6-D Beta scene, one 1920x1080 view, FP32 fwd+bwd"),
    "c3": dict(N=1_000_000, D=7, W=1352, H=1014,
               workload="1M-primitive dynamic 7-D Beta scene, one 1352x1014 view, FP32 fwd+bwd"),
    "c1": dict(N=10_000, D=6, W=256, H=256,
               workload="10k-primitive static 6-D Beta scene, one 256x256 view, FP32 fwd+bwd"),
    # batched multi-view training step (not the headline metric): 8 views,
    # resolution cycling 960x540, 1280x720, 1920x1080 by view index
    "c4": dict(N=3_000_000, D=6, views=8, sizes=((960, 540), (1280, 720), (1920, 1080)),
               workload="3M-primitive static 6-D Beta scene, 8 views at 960x540/1280x720/1920x1080, "
                        "one batched FP32 fwd+bwd step"),
    # render serving (not the headline metric): 4 resident scenes, views of
    # mixed resolution arriving as a Poisson process, forward only
    "c5": dict(N=2_000_000, scenes=(("static6_a", 6, 0), ("static6_b", 6, 1), ("dynamic7_a", 7, 2),
                                    ("dynamic7_b", 7, 3)),
               sizes=((640, 480), (1280, 720), (1920, 1080)), rates=(50, 100, 200, 400), duration_s=1.5,
               saturation_views=64,
               workload="4 resident scenes of 2M primitives (two 6-D static, two 7-D dynamic), views at "
                        "640x480/1280x720/1920x1080 on a radius U[3,5] sphere, t U[0,1], FP32 forward"),
}


@contextlib.contextmanager
def no_gc():
    """Python's cyclic garbage collector off inside a timed region: a full
    collection over torch's objects pauses the host for ~40 ms at random
    steps, which is not part of the workload."""
    gc.collect()
    gc.disable()
    try:
        yield
    finally:
        gc.enable()


# The blend backward that reuses the forward's packed records and tile order
# (the autograd path's default) is the same stage as the plain backward.
# the blend entry points, timed with HIP events inside the timed steps
BLEND_ENTRIES = ("ubs_rasterize_to_pixels_fwd", "ubs_rasterize_to_pixels_fwd_prepacked", "ubs_rasterize_to_pixels_bwd",
                 "ubs_rasterize_to_pixels_bwd_reuse")
ENTRY_ALIASES = {"ubs_rasterize_to_pixels_bwd_reuse": "ubs_rasterize_to_pixels_bwd",
                 "ubs_rasterize_to_pixels_fwd_prepacked": "ubs_rasterize_to_pixels_fwd",
                 "ubs_beta_project_fwd_rec": "ubs_beta_project_fwd",
                 "ubs_isect_tiles_emit_ordered_cap": "ubs_isect_tiles_emit_ordered",
                 "ubs_isect_tiles_emit_ordered_rec": "ubs_isect_tiles_emit_ordered",
                 "ubs_radix_sort_pairs_flag": "ubs_radix_sort_pairs",
                 "ubs_isect_tiles_count_info": "ubs_isect_tiles_count_range"}


def launch_durations(timer) -> dict:
    out = {}
    for name, v in timer.durations_ms().items():
        out.setdefault(ENTRY_ALIASES.get(name, name), []).extend(v)
    return out


def algorithmic_bytes(name: str, c: dict, n_isect: int, C: int = 1) -> float:
    """Minimal HBM bytes one launch of `name` must move: per-unit bytes x units.

    Blend: 44 B per intersection (id, xy, conic, opacity, beta, 3 colours),
    24 B per pixel (colour, alpha, last id, T), and in the backward 40 B of
    gradient per intersection.  Sort: 8 B key read for the histograms, then
    per 8-bit pass a read and a write of a 12 B (key, value) pair."""
    N, D, W, H = c["N"], c["D"], c["W"], c["H"]
    tiles = math.ceil(W / 16) * math.ceil(H / 16)
    pix = C * W * H
    CD = D - 3
    if name == "ubs_rasterize_to_pixels_fwd":
        return 44.0 * n_isect + 24.0 * pix + 4.0 * C * tiles
    if name == "ubs_rasterize_to_pixels_bwd":
        return 44.0 * n_isect + 24.0 * pix + 40.0 * n_isect
    if name in ("ubs_radix_sort_pairs", "ubs_radix_sort_pairs_skip"):
        end_bit = 32 + int(tiles).bit_length() + int(C).bit_length()
        tile_passes = math.ceil((end_bit - 32) / 8)
        mode = c.get("sort_mode", "onesweep")
        if name == "ubs_radix_sort_pairs" and mode == "depth_first":
            # two calls per step: the C*N depth keys (4 passes), the entries on the camera/tile bits
            return (C * N * (8.0 + 24.0 * 4) + n_isect * (8.0 + 24.0 * tile_passes)) / 2
        if name == "ubs_radix_sort_pairs" and mode == "segmented":
            return n_isect * (8.0 + 24.0 * tile_passes)
        passes = math.ceil(end_bit / 8)
        if name.endswith("_skip") or mode == "onesweep":  # passes over depth bytes every key shares are not run
            passes -= c.get("sort_skipped", 0)
        return n_isect * (8.0 + 24.0 * passes)
    if name == "ubs_isect_offset_encode":
        return 8.0 * n_isect + 4.0 * C * tiles
    if name in ("ubs_isect_tiles_count", "ubs_isect_tiles_count_range"):
        return 28.0 * C * N
    if name == "ubs_isect_tiles_emit":
        return 24.0 * C * N + 12.0 * n_isect
    if name == "ubs_isect_tiles_emit_ordered":  # + the order and prefix reads
        return 36.0 * C * N + 12.0 * n_isect
    if name == "ubs_isect_depth_keys":
        return 20.0 * C * N
    if name == "ubs_isect_scan_ordered":
        return 24.0 * C * N
    if name == "ubs_sort_tile_segments":
        return 24.0 * n_isect
    if name == "ubs_fully_fused_projection_fwd":
        return 36.0 * N + 28.0 * C * N
    if name == "ubs_fully_fused_projection_bwd":
        return 36.0 * N + 40.0 * C * N + 36.0 * N
    if name == "ubs_cond_mean_convariance_opacity_fwd":
        return N * (4.0 * (D + D * D + 1 + 2 * CD) + 52.0)
    if name == "ubs_cond_mean_convariance_opacity_bwd":
        return N * (4.0 * (D + D * D + 1 + 2 * CD) + 52.0 + 4.0 * (D + D * D + 1 + CD))
    if name == "ubs_beta_activate_fwd":  # leaves + quats in; mean, scale, opacity, beta, query, rot out
        return N * 4.0 * ((3 + (D - 3) + D + 1 + (D - 2) + 4) + (D + D + 1 + (D - 2) + (D - 3) + 9))
    if name == "ubs_beta_activate_bwd":  # raw leaves + output grads in; leaf grads out
        return N * 4.0 * ((D + 1 + (D - 2) + 4) + (D + D + 1 + (D - 2) + 9) + (3 + (D - 3) + D + 1 + (D - 2) + 4))
    leaf_floats = 3 + (D - 3) + D + D * (D - 1) // 2 + 1 + (D - 2) + 4  # with quats
    if name == "ubs_beta_project_fwd":  # leaves once; radii, means2d, depth, conic, opacity, beta per view
        return N * 4.0 * leaf_floats + C * N * 36.0
    if name == "ubs_beta_project_bwd":  # leaves once; radii, conic and 5 gradients per view; leaf grads
        return N * 4.0 * 2 * leaf_floats + C * N * 48.0
    if name == "ubs_adam_step":  # every leaf element (with rgb): p, g, m, v read; p, m, v written
        return N * 28.0 * (leaf_floats + 3)
    if name == "ubs_image_loss_fwd":  # L1 path: image and target read once
        return 8.0 * pix * 3 if c.get("lambda_dssim", 0.0) == 0.0 else float("nan")
    if name == "ubs_image_loss_bwd":  # L1 path: image and target read, gradient written
        return 12.0 * pix * 3 if c.get("lambda_dssim", 0.0) == 0.0 else float("nan")
    if name == "ubs_rot_scale_l_triangle_to_covar_fwd":
        return N * 4.0 * (9 + D + D * (D - 1) // 2 + D * D)
    if name == "ubs_rot_scale_l_triangle_to_covar_bwd":
        return N * 4.0 * (2 * (9 + D + D * (D - 1) // 2) + D * D)
    return float("nan")


FP32_PEAK_TFLOPS = 157.3

# FP32 operations per lane evaluation of the reference's per-pair math
# (rasterize_to_pixels_fwd.cu:140-172, rasterize_to_pixels_bwd.cu:196-247):
# an fma counts 2, a transcendental (exp2, log2, reciprocal) 1, a compare 0.
BLEND_FLOP_MODEL = {
    "ubs_rasterize_to_pixels_fwd": "11 per live lane evaluation (offset and sigma) + 8 per in-support one "
                                   "((1-sigma)^beta, alpha, next T) + 7 per contributing one (vis, 3 colours)",
    "ubs_rasterize_to_pixels_bwd": "11 per live lane evaluation (offset and sigma) + 70 per contributing one "
                                   "(alpha again, T recovery, 3 colour and alpha gradients, sigma, conic, xy, "
                                   "opacity and beta gradients, the running colour sum)",
}


def blend_flops(entry: str, pairs: dict) -> float:
    if entry == "ubs_rasterize_to_pixels_fwd":
        f = pairs["fwd"]
        return 11.0 * f["lane_evals_live"] + 8.0 * f["lane_evals_in_support"] + 7.0 * f["lane_evals_contributing"]
    b = pairs["bwd"]
    return 11.0 * b["lane_evals_live"] + 70.0 * b["lane_evals_contributing"]


def measured_counters(entry: str, workload: str, n_isect: int) -> dict:
    """HBM bytes per call and the VALU / LDS use of the entry's main kernel
    from the committed rocprofv3 --pmc summary of this same workload (empty
    when there is none for it)."""
    try:
        with open(os.path.join(ROOT, "profiles", "pmc_traffic.json")) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return {}
    if d.get("workload") != workload or d.get("n_isects") != n_isect or entry not in d.get("entries", {}):
        return {}
    e = d["entries"][entry]
    util = d.get("valu", {}).get(e["kernels"][0], {}) if e.get("kernels") else {}
    return {"traffic": e.get("traffic_bytes"), "valu_busy": util.get("valu_busy"),
            "valu_busy_formula": d.get("valu_busy_formula"), "lds_busy": util.get("lds_util"),
            "valu_insts_per_wave_step": util.get("valu_insts_per_wave_step"),
            "trans_insts_per_wave_step": util.get("trans_insts_per_wave_step")}


def cpu_baseline(cfg: dict, seed: int = 0):
    """The oracle's full step (float32, OpenMP on the host cores) on the same
    workload, one frame."""
    import numpy as np

    from gsplat import scenes
    from oracle import ubs_oracle as O

    O.build()
    sc = scenes.make_scene(cfg["N"], cfg["D"], seed=seed)
    vm, K, ts = scenes.make_cameras(1, cfg["W"], cfg["H"], seed=seed)
    leaves = dict(xyz=sc.xyz, mean_rest=sc.mean_rest, scale_raw=sc.scale_raw, l_triangle=sc.l_triangle,
                  opacity_raw=sc.opacity_raw, beta_raw=sc.beta_raw, rgb=sc.rgb, quats=sc.quats)
    tgt = np.random.Generator(np.random.PCG64(seed + 5)).uniform(0, 1, (cfg["H"], cfg["W"], 3))
    t0 = time.perf_counter()
    O.beta_step(leaves, vm[0], K[0], cfg["W"], cfg["H"], tgt, timestamp=float(ts[0]), dtype=np.float32)
    dt = time.perf_counter() - t0
    return {"value": 1.0 / dt, "unit": "frames/s", "cores": O.num_threads(), "kind": "port",
            "sample": f"one full fwd+bwd frame of the same workload through the oracle (float32, "
                      f"{O.num_threads()} OpenMP threads): {dt:.2f} s"}


def _probe_lib():
    """libubs_probe.so: the measurement build (VALU calibration kernels, the
    blend kernels with pair counters), built beside libubs_hip.so."""
    import ctypes

    path = os.path.join(ROOT, "universal-beta-splatting_amd", "gsplat", "cuda", "libubs_probe.so")
    lib = ctypes.CDLL(path)
    vp, u32, i64, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int64, ctypes.c_int
    lib.ubs_probe_valu.argtypes = [i32, i32, i32, vp, vp]
    lib.ubs_probe_raster_fwd_stats.argtypes = [u32, u32, i64] + [vp] * 5 + [u32] * 4 + [vp] * 8
    lib.ubs_probe_raster_bwd_stats.argtypes = [u32, u32, i64] + [vp] * 5 + [u32] * 4 + [vp] * 10
    for f in (lib.ubs_probe_valu, lib.ubs_probe_raster_fwd_stats, lib.ubs_probe_raster_bwd_stats):
        f.restype = i32
    return lib


def valu_calibration(blocks: int = 4096, iters: int = 200, reps: int = 5) -> None:
    """The VALU calibration kernels (8 independent chains of v_fma_f32, or of
    v_exp_f32, per thread; 128 x iters instructions each): HIP-event time per
    dispatch, for the utilisation formula's check under rocprofv3 --pmc."""
    import torch

    from gsplat.cuda import _wrapper as W

    lib = _probe_lib()
    out = torch.empty(blocks * 256, dtype=torch.float32, device="cuda:0")
    st = W.stream_of(out)
    res = {}
    for kind, name in ((0, "fma_f32"), (1, "exp_f32")):
        assert lib.ubs_probe_valu(kind, blocks, iters, W.ptr(out), st) == 0
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
        for a, b in ev:
            a.record()
            assert lib.ubs_probe_valu(kind, blocks, iters, W.ptr(out), st) == 0
            b.record()
        torch.cuda.synchronize()
        ms = sorted(a.elapsed_time(b) for a, b in ev)[reps // 2]
        waves = blocks * 4
        insts = waves * iters * 128
        res[name] = {"ms": round(ms, 4), "wave_instructions": insts,
                     "wave_insts_per_ns": round(insts / (ms * 1e6), 2)}
    print(json.dumps({"valu_calibration": res, "blocks": blocks, "threads": 256, "iters": iters}))


def blend_pair_stats(params, meta, img, tgt, W_: int, H_: int) -> dict:
    """Counts of the blend's work on this step's tile lists, from the
    counting builds of the two kernels (libubs_probe.so): (wave, primitive)
    steps, lane evaluations and how many of them fall inside the support or
    contribute.  The L1 loss's image gradient drives the backward."""
    import torch

    from gsplat.cuda import _wrapper as W

    lib = _probe_lib()
    dev = img.device
    with torch.no_grad():
        rgb = params.rgb.detach()[None].contiguous()
        m2, cn, op, be = (meta[k].detach().contiguous() for k in ("means2d", "conics", "opacities", "betas"))
        offs, flat = meta["isect_offsets"], meta["flatten_ids"]
        C, N = op.shape
        n = int(flat.numel())
        th, tw = offs.shape[1:3]
        rc = torch.empty((C, H_, W_, 3), device=dev)
        ra = torch.empty((C, H_, W_, 1), device=dev)
        last = torch.empty((C, H_, W_), dtype=torch.int32, device=dev)
        Tf = torch.empty((C, H_, W_), device=dev)
        ws = torch.empty((int(W.lib().ubs_rasterize_workspace_bytes(C, N, 1)),), dtype=torch.uint8, device=dev)
        sf = torch.zeros(5, dtype=torch.int64, device=dev)
        sb = torch.zeros(7, dtype=torch.int64, device=dev)
        st = W.stream_of(rc)
        assert lib.ubs_probe_raster_fwd_stats(C, N, n, W.ptr(m2), W.ptr(cn), W.ptr(rgb), W.ptr(op), W.ptr(be), W_, H_,
                                              tw, th, W.ptr(offs), W.ptr(flat), W.ptr(rc), W.ptr(ra), W.ptr(last),
                                              W.ptr(Tf), W.ptr(ws), W.ptr(sf), st) == 0
        vrc = (torch.sign(img.detach() - tgt) / float(H_ * W_ * 3)).contiguous()
        vra = torch.zeros_like(ra)
        vcol = torch.empty((C, N, 3), device=dev)
        assert lib.ubs_probe_raster_bwd_stats(C, N, n, W.ptr(m2), W.ptr(cn), W.ptr(rgb), W.ptr(op), W.ptr(be), W_, H_,
                                              tw, th, W.ptr(offs), W.ptr(flat), W.ptr(last), W.ptr(Tf), W.ptr(vrc),
                                              W.ptr(vra), W.ptr(vcol), W.ptr(ws), W.ptr(sb), st) == 0
        torch.cuda.synchronize()
        f, b = [int(x) for x in sf.cpu()], [int(x) for x in sb.cpu()]
    return {
        "fwd": {"wave_primitive_steps": f[0], "lane_evals_live": f[1], "lane_evals_in_support": f[2],
                "lane_evals_contributing": f[3], "entries_staged": f[4],
                "in_support_frac": round(f[2] / max(f[1], 1), 4)},
        "bwd": {"wave_primitive_steps": b[0], "lane_evals_live": b[1], "lane_evals_contributing": b[2],
                "reduce_scatter_steps": b[3], "entries_staged": b[4], "reduce_scatter_steps_le4_lanes": b[5],
                "reduce_scatter_steps_le16_lanes": b[6],
                "contributing_frac": round(b[2] / max(b[1], 1), 4)},
        "source": "counting builds of raster_fwd_kernel<16,3> / raster_bwd_kernel<16,3> (libubs_probe.so) on "
                  "this step's tile lists",
    }


def copy_calibration(mib: int, reps: int = 5) -> None:
    """Plain copies of a known size: each reads and writes mib MiB of HBM."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    assert torch.equal(dst[-16:], src[-16:])
    print(json.dumps({"copy_bytes": n * 4, "reps": reps, "ms_per_copy": round(dt * 1e3, 4),
                      "GB_per_s_read_plus_write": round(2 * n * 4 / dt / 1e9, 1)}))


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS),
                    help="c2 (default): the headline 1080p / 1M workload; c3: 7-D; c1: small; "
                         "c4: batched 8-view step at mixed resolutions (reports views/s); "
                         "c5: render serving of 4 resident scenes (reports MPix/s, p50/p99)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-pair-stats", action="store_true",
                    help="leave out the blend's pair counts (counting builds of the blend kernels)")
    ap.add_argument("--no-chain", action="store_true",
                    help="leave out the drop-in chain's fwd+bwd rate (the reference's unchanged call sequence)")
    ap.add_argument("--no-serve", action="store_true",
                    help="leave out the batched-serve measurement (for counter runs that average per kernel)")
    ap.add_argument("--lambda-dssim", type=float, default=0.0,
                    help="image loss (1 - l) L1 + l (1 - SSIM) in the fused loss kernels; 0 (default) is the L1 "
                         "loss of the benchmark definition, 0.2 the reference's training loss")
    ap.add_argument("--adam", action="store_true",
                    help="end each step with the fused Adam update of every leaf (the reference's training "
                         "iteration: render, loss, backward, optimizer.step); off by default, as the headline "
                         "metric is render fwd+bwd")
    ap.add_argument("--path", default="fused", choices=("fused", "chain"),
                    help="fused: one-launch Beta projection (render_views); chain: the reference's op chain "
                         "(render: activations, rotation, covariance, conditioning, projection as separate ops)")
    ap.add_argument("--serve-streams", type=int, default=2, help="c5: HIP streams of the batched server")
    ap.add_argument("--serve-max-views", type=int, default=16, help="c5: most views per batch")
    ap.add_argument("--sort", default="depth_first", choices=("depth_first", "segmented", "onesweep"),
                    help="isect-id sort (same order every way): depth_first (default) sorts the primitives by depth "
                         "once and emits in that order, then onesweep on the camera/tile bits; segmented: onesweep "
                         "on the camera/tile bits then a per-tile depth sort in LDS; onesweep: every byte")
    ap.add_argument("--transmittance", default="recompute", choices=("recompute", "checkpoint"),
                    help="per-pixel transmittance in the blend backward: recomputed by division from the final T "
                         "(default, as the reference) or stored by the forward after every batch of 256 entries")
    ap.add_argument("--fp64-transmittance", action="store_true",
                    help="recover transmittance in fp64 in the blend backward")
    ap.add_argument("--valu-calibration", action="store_true",
                    help="only run the VALU calibration kernels (fma and exp chains) and exit")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (the known-size reference for "
                         "scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}). "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.valu_calibration:
        valu_calibration()
        return

    import numpy as np
    import torch

    from gsplat import beta, scenes
    from gsplat.cuda import _wrapper as W

    W.set_backward_fp64_transmittance(args.fp64_transmittance)
    W.set_isect_sort_mode(args.sort)
    cfg = CONFIGS[args.config]
    if "views" in cfg:
        return batch_bench(args, cfg)
    if "scenes" in cfg:
        return serve_bench(args, cfg)
    dev = torch.device("cuda:0")
    torch.manual_seed(0)
    sc = scenes.make_scene(cfg["N"], cfg["D"], seed=0)
    vms, Ks, ts = scenes.make_cameras(1, cfg["W"], cfg["H"], seed=0)
    params = beta.BetaParams.from_scene(sc, dev)
    vm = torch.tensor(vms[0], device=dev)
    K = torch.tensor(Ks[0], device=dev)
    t_view = float(ts[0])
    tgt = torch.tensor(np.random.Generator(np.random.PCG64(5)).uniform(0, 1, (cfg["H"], cfg["W"], 3)),
                       dtype=torch.float32, device=dev)
    rest = beta.rest_indices(cfg["D"], dev)
    leaves = list(params.leaves().values())
    W.set_backward_transmittance_checkpoint(args.transmittance == "checkpoint")

    vms_t, Ks_t = vm[None].contiguous(), K[None].contiguous()
    ts_t = torch.tensor([t_view], dtype=torch.float32, device=dev)

    def render():
        if args.path == "fused":
            return beta.render_views(params, vms_t, Ks_t, cfg["W"], cfg["H"], timestamps=ts_t)
        return beta.render(params, vm, K, cfg["W"], cfg["H"], timestamp=t_view, rest=rest)

    from gsplat.loss import image_loss

    adam = None
    if args.adam:
        from gsplat.optim import BetaAdam
        adam = BetaAdam.for_model(params)

    def step():
        for p in leaves:
            p.grad = None
        img, alpha, meta = render()
        if args.path == "fused":
            loss = image_loss(img, tgt[None], args.lambda_dssim)  # [1, H, W, 3]: no select in the graph
        else:  # the reference's torch ops
            loss = (img[0] - tgt).abs().mean()
        loss.backward()
        if adam is not None:
            adam.step()
        return meta

    def fwd_only():
        with torch.no_grad():
            return render()

    for _ in range(args.warmup):
        meta = step()
    torch.cuda.synchronize()
    n_isect = int(meta["isect_ids"].numel())
    n_vis = int((meta["radii"] > 0).sum().item())
    dbits = meta["depths"][meta["radii"] > 0].contiguous().view(torch.int32).to(torch.int64) & 0xFFFFFFFF
    lo, hi = int(dbits.min()), int(dbits.max())
    cfg = dict(cfg, lambda_dssim=args.lambda_dssim, sort_mode=args.sort, sort_skipped=bin(W.constant_depth_passes((hi << 32) | (~lo & 0xFFFFFFFF))).count("1"))

    # ---- timed region: K steps; HIP events on the launch stream around the
    # blend launches (the dominant kernels, whose average is the roofline's
    # duration); the other entry points are timed in a pass of their own
    # after the region, so that their event records add no host time to it
    timer = W.LaunchTimer(BLEND_ENTRIES)
    W.set_launch_timer(timer)
    step_ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(args.steps)]
    with no_gc():
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(args.steps):
            step_ev[k][0].record()
            step()
            step_ev[k][1].record()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
    W.set_launch_timer(None)
    launches_in_run = launch_durations(timer)
    step_ms = sorted(s.elapsed_time(e) for s, e in step_ev)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    # ---- every entry point's launches, timed over a few more steps
    stage_timer = W.LaunchTimer()
    W.set_launch_timer(stage_timer)
    stage_steps = min(args.steps, 5)
    for _ in range(stage_steps):
        step()
    W.set_launch_timer(None)
    launches = launch_durations(stage_timer)

    # ---- forward-only rate (not the headline value)
    for _ in range(3):
        fwd_only()
    with no_gc():
        torch.cuda.synchronize()
        f0 = time.perf_counter()
        for _ in range(args.steps):
            fwd_only()
        torch.cuda.synchronize()
        fwd_ms = (time.perf_counter() - f0) * 1e3 / args.steps

    # ---- the drop-in chain (the reference's unchanged BetaModel.render: the
    # _wrapper ops for rotation, covariance and conditioning, then
    # rasterization() with covars) on the same scene, same loss, fwd+bwd
    chain = None
    if args.path == "fused" and not args.no_chain:
        def chain_step():
            for p in leaves:
                p.grad = None
            img, _, _ = beta.render(params, vm, K, cfg["W"], cfg["H"], timestamp=t_view, rest=rest)
            (img[0] - tgt).abs().mean().backward()

        for _ in range(3):
            chain_step()
        with no_gc():
            torch.cuda.synchronize()
            c0 = time.perf_counter()
            for _ in range(args.steps):
                chain_step()
            torch.cuda.synchronize()
            chain_ms = (time.perf_counter() - c0) * 1e3 / args.steps
        chain = {"frames_per_s": round(1e3 / chain_ms, 3), "ms_per_step": round(chain_ms, 3),
                 "path": "reference call chain: l_triangle_to_rotmat, rot_scale_l_triangle_to_covar, "
                         "cond_mean_convariance_opacity, rasterization(covars=...), torch L1"}

    # ---- batched serving of the same scene: 8 views at the same size in one
    # pass (fused projection, one isect/sort/blend over all views) against
    # the same 8 views rendered one at a time; forward only
    serve = None if args.no_serve else serve_views(params, cfg, args.steps)

    # ---- counts of the blend's work on this step's lists (counting builds)
    pairs = None
    if not args.no_pair_stats and args.path == "fused":
        img_s, _, meta_s = fwd_only()
        pairs = blend_pair_stats(params, meta_s, img_s, tgt, cfg["W"], cfg["H"])
        del img_s, meta_s

    # ---- roofline of the dominant kernel launch
    per_call = {k: (sum(v) / len(v), len(v) / stage_steps) for k, v in launches.items()}
    dominant = max(per_call, key=lambda k: per_call[k][0] * per_call[k][1])
    in_run = launches_in_run.get(dominant)
    avg_ms = sum(in_run) / len(in_run) if in_run else per_call[dominant][0]
    avg_src = "HIP events in the timed region" if in_run else "HIP events in the stage-timing steps"
    alg = algorithmic_bytes(dominant, cfg, n_isect)
    achieved = alg / (avg_ms * 1e-3) / 1e9
    pmc = measured_counters(dominant, cfg["workload"], n_isect)
    hbm = {"achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": pmc.get("traffic"),
           "algorithmic_bytes_per_launch": alg}
    blend = dominant in ("ubs_rasterize_to_pixels_fwd", "ubs_rasterize_to_pixels_bwd")
    if blend and pairs is not None:
        # the blend does no dense contraction: its bound is the FP32 VALU (with
        # the transcendental unit beside it), priced at the published peak
        flop = blend_flops(dominant, pairs)
        tf = flop / (avg_ms * 1e-3) / 1e12
        roofline = {"bound": "valu", "kernel": dominant, "achieved": round(tf, 2), "peak": FP32_PEAK_TFLOPS,
                    "unit": "TFLOP/s", "frac": round(tf / FP32_PEAK_TFLOPS, 4), "traffic": pmc.get("traffic"),
                    "avg_launch_ms": round(avg_ms, 4), "avg_launch_timing": avg_src,
                    "algorithmic_flop_per_launch": flop,
                    "flop_model": BLEND_FLOP_MODEL[dominant], "hbm": hbm}
    else:
        roofline = {"bound": "hbm", "kernel": dominant, **hbm, "avg_launch_ms": round(avg_ms, 4),
                    "avg_launch_timing": avg_src}
    if pmc.get("traffic") is not None:
        roofline["traffic_source"] = ("profiles/pmc_traffic.json: rocprofv3 FETCH_SIZE and WRITE_SIZE of this "
                                      "workload, scaled against a 1 GiB device copy")
    for k in ("valu_busy", "valu_busy_formula", "valu_insts_per_wave_step", "trans_insts_per_wave_step",
              "lds_busy"):
        if pmc.get(k) is not None:
            roofline[k] = pmc[k]
    def gbps(k, ms):
        r = algorithmic_bytes(k, cfg, n_isect) / (ms * 1e-3) / 1e9
        return round(r, 1) if math.isfinite(r) else None  # JSON has no NaN
    stages = {k.replace("ubs_", ""): {"avg_ms": round(v[0], 4), "per_step": v[1], "GB_per_s": gbps(k, v[0])}
              for k, v in sorted(per_call.items(), key=lambda kv: -kv[1][0] * kv[1][1])}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(cfg)

    # ---- the render-serving workload, compact (its own line: --config c5)
    serving = None
    if not args.no_serve and args.config == "c2":
        torch.cuda.empty_cache()
        serving = serve_summary(args)

    line = {
        "metric": METRIC,
        "value": round(1e3 / ms_per_step, 3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded scene and camera generator, gsplat/scenes.py)",
        "config": {"workload": cfg["workload"], "primitives": cfg["N"], "dims": cfg["D"],
                   "width": cfg["W"], "height": cfg["H"], "views": 1, "tile_size": 16,
                   "n_isects": n_isect, "visible_primitives": n_vis, "projection_path": args.path,
                   "loss": "L1" if args.lambda_dssim == 0 else f"{1 - args.lambda_dssim:g} L1 + {args.lambda_dssim:g} (1 - SSIM)",
                   "sort_passes_skipped": cfg["sort_skipped"], "optimizer_step": "adam" if args.adam else None,
                   "transmittance": args.transmittance, "sort": args.sort},
        "fwd_frames_per_s": round(1e3 / fwd_ms, 3),
        "chain_frames_per_s": None if chain is None else chain["frames_per_s"],
        "drop_in_chain": chain,
        "fwd_bwd_mpix_per_s": round(cfg["W"] * cfg["H"] * 1e3 / ms_per_step / 1e6, 2),
        "latency_ms": {"p50": round(step_ms[len(step_ms) // 2], 3),
                       "p99": round(step_ms[min(len(step_ms) - 1, int(math.ceil(0.99 * len(step_ms))) - 1)], 3)},
        "batched_serve": serve,
        "serving_c5": serving,
        "blend_pairs": pairs,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "stages": stages,
    }
    print(json.dumps(line))


def serve_views(params, cfg, steps: int, views: int = 8) -> dict:
    """Forward-only rate of `views` cameras of the scene rendered as one
    batch (beta.render_views over all views) and one view at a time."""
    import torch

    from gsplat import beta, scenes

    dev = params.xyz.device
    W, H = cfg["W"], cfg["H"]
    vms, Ks, ts = scenes.make_cameras(views, W, H, seed=7)
    vms_t, Ks_t = torch.tensor(vms, device=dev), torch.tensor(Ks, device=dev)
    ts_t = torch.tensor(ts, device=dev) if cfg["D"] == 7 else None

    def batched():
        with torch.no_grad():
            return beta.render_views(params, vms_t, Ks_t, W, H, timestamps=ts_t)

    def one_at_a_time():
        with torch.no_grad():
            for c in range(views):
                beta.render_views(params, vms_t[c:c + 1], Ks_t[c:c + 1], W, H,
                                  timestamps=None if ts_t is None else ts_t[c:c + 1])

    out = {}
    for name, fn in (("batched", batched), ("per_view", one_at_a_time)):
        for _ in range(2):
            fn()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        with no_gc():
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for k in range(steps):
                ev[k][0].record()
                fn()
                ev[k][1].record()
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) * 1e3 / steps
        lat = sorted(a.elapsed_time(b) for a, b in ev)
        out[name] = {"ms_per_batch": round(ms, 3), "mpix_per_s": round(views * W * H * 1e3 / ms / 1e6, 1),
                     "frames_per_s": round(views * 1e3 / ms, 1), "p50_ms": round(lat[len(lat) // 2], 3),
                     "p99_ms": round(lat[min(len(lat) - 1, int(math.ceil(0.99 * len(lat))) - 1)], 3)}
    return {"views_per_batch": views, "width": W, "height": H, **out,
            "speedup": round(out["per_view"]["ms_per_batch"] / out["batched"]["ms_per_batch"], 3)}


def serve_measure(args, cfg, rates, n_sat: int, duration_s: float, modes=None, log=sys.stderr):
    """Render serving (gsplat.serve) of cfg's resident scenes: Poisson
    arrivals of view requests of mixed resolution and a saturation run with
    every request pending at once, per server mode:
      graphs   -- batches of one image size over any resident scenes as
                  replayed hipGraphs with device-sized lists (the default);
      batched  -- one scene's views per batch through render_batch (host
                  read of the intersection count per resolution group);
      per_view -- one view at a time in arrival order (the reference's
                  structure, BetaModel.view per request).
    Returns (store, saturation, poisson) with latency_summary dicts."""
    import numpy as np
    import torch

    from gsplat import beta, scenes
    from gsplat.serve import RenderServer, SceneStore, ViewRequest, latency_summary

    dev = torch.device("cuda:0")
    store = SceneStore(dev)
    for name, D, seed in cfg["scenes"]:
        sc = scenes.make_scene(cfg["N"], D, seed=seed)
        store.add(name, beta.BetaParams.from_scene(sc, dev, requires_grad=False, use_quats=(D == 6)))
        del sc
        vm = torch.tensor(scenes.look_at_viewmat(np.array([0.0, 0.4, 4.0])), dtype=torch.float32)
        store.calibrate(name, vm, torch.tensor(scenes.intrinsics(1920, 1080), dtype=torch.float32), t=0.5)
        print(f"scene {name}: {cfg['N']} primitives, {D}-D, {store.isects_per_prim[name]:.1f} isects/primitive",
              file=log, flush=True)
    rng = np.random.Generator(np.random.PCG64(11))
    names = [n for n, _, _ in cfg["scenes"]]

    def make_request(i):
        W, H = cfg["sizes"][int(rng.integers(0, len(cfg["sizes"])))]
        d = rng.normal(0, 1, 3)
        vm = scenes.look_at_viewmat(rng.uniform(3, 5) * d / np.linalg.norm(d))
        return ViewRequest(names[int(rng.integers(0, len(names)))], torch.tensor(vm, dtype=torch.float32),
                           torch.tensor(scenes.intrinsics(W, H), dtype=torch.float32), W, H,
                           t=float(rng.uniform(0, 1)), rid=i)

    all_modes = {"graphs": dict(max_views=args.serve_max_views, streams=args.serve_streams, graphs=True),
                 "batched": dict(max_views=args.serve_max_views, streams=args.serve_streams),
                 "per_view": dict(max_views=1, streams=1)}
    modes = {k: all_modes[k] for k in (modes or all_modes)}
    servers = {m: RenderServer(store, keep_images=False, **kw) for m, kw in modes.items()}
    for m, srv in servers.items():
        t0 = time.perf_counter()
        made = srv.prepare_graphs(cfg["sizes"])
        if made:
            print(f"{m}: captured {made} batch graphs in {time.perf_counter() - t0:.1f} s", file=log, flush=True)

    def run(reqs, arrivals, mode):
        with no_gc():
            return _run(reqs, arrivals, mode)

    def _run(reqs, arrivals, mode):
        srv = servers[mode]  # kept across runs: its captured graphs are reused
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        i, done = 0, []
        while i < len(reqs) or srv.busy():
            now = time.perf_counter() - t0
            while i < len(reqs) and arrivals[i] <= now:
                r = reqs[i]
                srv.submit(ViewRequest(r.scene, r.viewmat, r.K, r.width, r.height, r.t, arrival=t0 + arrivals[i],
                                       rid=r.rid))
                i += 1
            if not srv.poll() and not srv.busy() and i < len(reqs):
                time.sleep(min(1e-4, max(0.0, arrivals[i] - (time.perf_counter() - t0))))
            done += srv.results
            srv.results = []
        wall = max(r.done for r in done) - t0
        out = latency_summary(done, wall)
        out["overflows"] = srv.overflows
        return out

    # warm-up: every (size, slot count) batch shape the runs can form is
    # captured before the timed runs
    for mode in modes:
        warm = [make_request(-1 - k) for k in range(3 * args.serve_max_views)]
        run(warm, [0.0] * len(warm), mode)
        for v in (1, 2, 3, 5):
            warm = [make_request(-100 - k) for k in range(v * len(cfg["sizes"]))]
            run(warm, [0.0] * len(warm), mode)
    saturation = {}
    sat_reqs = [make_request(k) for k in range(n_sat)]
    for mode in modes:
        saturation[mode] = run(sat_reqs, [0.0] * n_sat, mode)
        print(f"saturation {mode}: {saturation[mode]}", file=log, flush=True)
    poisson = {}
    for lam in rates:
        gaps = rng.exponential(1.0 / lam, size=int(lam * duration_s * 2) + 8)
        arr = np.cumsum(gaps)
        arr = arr[arr < duration_s]
        reqs = [make_request(k) for k in range(arr.size)]
        poisson[str(lam)] = {}
        for mode in modes:
            poisson[str(lam)][mode] = run(reqs, list(arr), mode)
            print(f"poisson {lam}/s {mode}: {poisson[str(lam)][mode]}", file=log, flush=True)
    return store, saturation, poisson


def serve_bench(args, cfg):
    """The render-serving workload as its own bench line: value = MPix/s of
    the graph-batched server with every request pending at once
    (saturation); p50/p99 latency per arrival rate, per server mode."""
    import torch

    store, saturation, poisson = serve_measure(args, cfg, cfg["rates"], cfg["saturation_views"], cfg["duration_s"])
    b, pv = saturation["graphs"], saturation["per_view"]
    line = {
        "metric": "batched-serve MPix/s (4 resident 2M-primitive scenes, mixed resolution); p50/p99 latency",
        "value": b["mpix_per_s"], "unit": "MPix/s", "n_gpus": 1, "steps": cfg["saturation_views"], "warmup": 0,
        "ms_per_step": round(1e3 / b["frames_per_s"], 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded scenes, cameras and arrivals, gsplat/scenes.py)",
        "config": {"workload": cfg["workload"], "primitives_per_scene": cfg["N"], "scenes": len(cfg["scenes"]),
                   "resident_gb": round(store.resident_bytes() / 1e9, 3), "max_views_per_batch": args.serve_max_views,
                   "sort": args.sort, "streams": args.serve_streams,
                   "peak_hbm_gb": round(torch.cuda.max_memory_allocated() / 1e9, 2)},
        "saturation": saturation,
        "speedup_vs_per_view": round(b["mpix_per_s"] / pv["mpix_per_s"], 3),
        "speedup_batched_eager_vs_per_view": round(saturation["batched"]["mpix_per_s"] / pv["mpix_per_s"], 3),
        "poisson": poisson,
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line))


def serve_summary(args) -> dict:
    """A compact run of the render-serving workload for the headline line:
    4 resident 2M-primitive scenes, 48 pending requests at once
    (saturation) and 1 s of Poisson arrivals at 400 requests/s, graph-batched
    server against one view at a time."""
    import io

    cfg = CONFIGS["c5"]
    log = io.StringIO()
    store, sat, poi = serve_measure(args, cfg, (400,), 48, 1.0, modes=("graphs", "batched", "per_view"), log=log)
    g, pv = sat["graphs"], sat["per_view"]
    p400 = poi["400"]
    out = {"workload": cfg["workload"],
           "saturated_mpix_per_s": {m: v["mpix_per_s"] for m, v in sat.items()},
           "speedup_vs_per_view": round(g["mpix_per_s"] / pv["mpix_per_s"], 3),
           "speedup_batched_eager_vs_per_view": round(sat["batched"]["mpix_per_s"] / pv["mpix_per_s"], 3),
           "poisson_400_per_s": {m: {"p50_ms": v["p50_ms"], "p99_ms": v["p99_ms"], "views": v["views"],
                                     "mean_batch_views": v["mean_batch_views"]} for m, v in p400.items()},
           "overflows": g.get("overflows", 0)}
    del store
    return out


def batch_bench(args, cfg):
    """The batched multi-view training step: one fused projection for all
    views (per-view conditioning and image size), isect + sort + blend per
    resolution group, L1 loss per view, backward to the leaves."""
    import numpy as np
    import torch

    from gsplat import beta, scenes
    from gsplat.cuda import _wrapper as W

    dev = torch.device("cuda:0")
    C = cfg["views"]
    sizes = [cfg["sizes"][c % len(cfg["sizes"])] for c in range(C)]
    sc = scenes.make_scene(cfg["N"], cfg["D"], seed=0)
    params = beta.BetaParams.from_scene(sc, dev)
    vms, _, _ = scenes.make_cameras(C, sizes[0][0], sizes[0][1], seed=0)
    Ks = np.stack([scenes.intrinsics(w, h) for w, h in sizes]).astype(np.float32)
    vms_t, Ks_t = torch.tensor(vms, device=dev), torch.tensor(Ks, device=dev)
    g = np.random.Generator(np.random.PCG64(5))
    tgts = [torch.tensor(g.uniform(0, 1, (h, w, 3)), dtype=torch.float32, device=dev) for w, h in sizes]
    leaves = list(params.leaves().values())

    def step():
        for p in leaves:
            p.grad = None
        imgs, _ = beta.render_batch(params, vms_t, Ks_t, sizes)
        loss = sum((im - t).abs().mean() for im, t in zip(imgs, tgts))
        loss.backward()

    # the blend backward's per-pixel transmittance, both ways: recomputed by
    # division (the default, as the reference) and stored per batch
    modes = {}
    for mode in ("checkpoint", "recompute"):
        W.set_backward_transmittance_checkpoint(mode == "checkpoint")
        torch.cuda.reset_peak_memory_stats()
        for _ in range(args.warmup):
            step()
        with no_gc():
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                step()
            torch.cuda.synchronize()
            t1 = time.perf_counter()
        # every entry point's launches, timed over steps of their own
        timer = W.LaunchTimer()
        W.set_launch_timer(timer)
        stage_steps = min(args.steps, 3)
        for _ in range(stage_steps):
            step()
        W.set_launch_timer(None)
        launches = launch_durations(timer)
        modes[mode] = {"ms_per_step": round((t1 - t0) * 1e3 / args.steps, 3),
                       "views_per_s": round(C * 1e3 * args.steps / ((t1 - t0) * 1e3), 3),
                       "peak_hbm_gb": round(torch.cuda.max_memory_allocated() / 1e9, 2)}
    W.set_backward_transmittance_checkpoint(False)
    ms = modes["recompute"]["ms_per_step"]
    stages = {k.replace("ubs_", ""): {"ms_per_step": round(sum(v) / stage_steps, 4),
                                      "calls_per_step": len(v) / stage_steps}
              for k, v in sorted(launches.items(), key=lambda kv: -sum(kv[1]))}

    # the reference's structure for comparison: one view at a time
    # (BetaModel.render per camera, train.py:111-128), same kernels
    def step_per_view():
        for p in leaves:
            p.grad = None
        for c, (w, h) in enumerate(sizes):
            img, _, _ = beta.render_views(params, vms_t[c:c + 1], Ks_t[c:c + 1], w, h)
            (img[0] - tgts[c]).abs().mean().backward()

    for _ in range(max(1, args.warmup // 2)):
        step_per_view()
    with no_gc():
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step_per_view()
        torch.cuda.synchronize()
        ms_per_view_loop = (time.perf_counter() - t0) * 1e3 / args.steps
    line = {
        "metric": "views/s fwd+bwd, batched multi-view step (3M prims, 8 views at 540p/720p/1080p)",
        "value": round(C * 1e3 / ms, 3), "unit": "views/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic (seeded scene and camera generator, gsplat/scenes.py)",
        "config": {"workload": cfg["workload"], "primitives": cfg["N"], "dims": cfg["D"], "views": C,
                   "sizes": sizes, "transmittance": "recompute", "sort": args.sort, "peak_hbm_gb": modes["recompute"]["peak_hbm_gb"]},
        "roofline": None, "cpu_baseline": None, "stages": stages, "transmittance_modes": modes,
        "per_view_loop_ms_per_step": round(ms_per_view_loop, 3),
        "speedup_vs_per_view_loop": round(ms_per_view_loop / ms, 3),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
