"""Benchmark of the gfx950 path tracer on the 1M-triangle procedural mesh scene.

Workload (one MI355X): plain GPU PT+NEE over a seeded scene of 200 displaced
icospheres (1,024,000 FP32 triangles, half Lambert / half GGX conductor) on a
ground quad under a native_sky, 1920x1080, max depth 8 (Atrc trace_std:
NEE over every light, MIS-BSDF ray, Russian roulette from depth 6 with
cont_prob 0.9).  One step = one wavefront pass of 1 sample per pixel over the
whole frame; the timed region renders ``--steps`` samples per pixel with the
scene, BVHs and path state already resident in HBM.

value = (closest-hit + any-hit ray queries) / wall time of the timed region.

roofline: the dominant kernel is the closest-hit traversal.  Its algorithmic
bytes = node fetches * node size (128 B BVH4 node with 4 child boxes, or 32 B
Atrc binary node per box test) + triangle tests * 48 B + hits * 96 B, counted
by a stats pass over the first samples (per-ray averages, scaled to the timed
rays), divided by that kernel's device time measured with HIP events on the
library's stream during the timed region; peak = 8 TB/s HBM.  traffic = HBM
bytes per launch of the same kernel from the committed rocprofv3 FETCH_SIZE /
WRITE_SIZE summary of this workload (profiles/).

cpu_baseline: the CPU restatement of Atrc `pt` (FP32, Atrc's 32x32 tile
scheduler, one worker per core) on a band of rows of the same frame at 1 spp,
on this host.

guided: equal-time relMSE of guided vs plain GPU PT on the door-gap room
(10 x 10 x 4 Lambert room lit only through a 0.05-wide gap from a second room,
500k clutter triangles, seed 7).  Plain and guided renders each run for
``--guided-seconds`` of wall time in chunks of ``--guided-chunk`` spp.  The
guided run starts from a freshly initialised network, trains on every pass for
the first ``--guided-train-frac`` of the time (one sample per pixel per pass,
``--guided-steps`` Adam steps on a ring of ``--guided-batch`` samples), then
renders with the network frozen; every sample, warm-up included, goes into its
image.  The reference is plain PT with a different seed for
``--guided-ref-factor`` times as long.  relMSE = mean over pixels and
channels of (x - y)^2 / (y^2 + 1e-2) after scaling both images by 1 / mean(y)
(the room is dark).  MLP utilisation = 61,440 FLOP per evaluated vertex / the
inference kernel's device time, and 3 x 61,440 FLOP per trained sample and
step / the training device time (gather, matrix-core forward and backward,
reduction, Adam), both against the 2.5 PFLOP/s dense BF16 peak.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mrays/s; equal-time relMSE, guided vs plain GPU PT, 1 MI355X"
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
FP32_PEAK_TFLOPS = 157.3
MLP_FORWARD_FLOP = 2 * (256 * 64 + 3 * 64 * 64 + 64 * 32)   # 61,440 per evaluated vertex
MLP_TRAIN_FLOP = 3 * MLP_FORWARD_FLOP                       # forward + backward per sample and step


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def build_scene(args):
    from atrc_amd import scenes

    if args.config == "cornell":
        s, cam, tr = scenes.cornell_box()
        W, H = 256, 256
        workload = "plain GPU PT+NEE, Cornell-like scene (5 quads + 2 boxes + quad light), 256x256, depth 5"
    elif args.config == "volumetric":
        s, cam, tr = scenes.volumetric_room(grid_n=args.grid)
        W, H = 1024, 1024
        workload = (f"plain GPU PT+NEE with media: door-gap room (500k clutter triangles) filled with a "
                    f"{args.grid}^3 fBm heterogeneous medium (sigma_t max 2, albedo 0.9, g 0.3) and a homogeneous "
                    f"slab (sigma_t 0.5), 1024x1024, depth 16")
    else:
        s, cam, tr = scenes.procedural_spheres(n_spheres=200, subdiv=4, seed=1)
        W, H = 1920, 1080
        workload = ("plain GPU PT+NEE, 1M-triangle procedural mesh scene (200 displaced icospheres, "
                    "Lambert + GGX conductor, ground quad, native_sky), 1920x1080, depth 8")
    return s, cam, tr, W, H, workload


def cpu_baseline(desc, cam, tr, W, H, target_s: float):
    from atrc_amd import scenes
    from oracle import py_oracle as po

    threads = max(1, min(16, os.cpu_count() or 1))
    ref = po.OracleScene(desc, 32)
    # size the sample from a 16-row probe: a band of rows at 1 spp, or the whole
    # frame at several spp when one frame takes less than the target
    probe = 16
    r = ref.render(scenes.render_params(W, H, 1, cam, seed=777, **tr), threads=threads,
                   rows=((H - probe) // 2, (H - probe) // 2 + probe), aovs=False)
    per_row = max(r["seconds"], 1e-4) / probe
    rows_needed = target_s / per_row
    if rows_needed <= H:
        rows, spp = int(max(probe, rows_needed)), 1
    else:
        rows, spp = H, int(min(256, max(1, round(rows_needed / H))))
    r0 = (H - rows) // 2
    p = scenes.render_params(W, H, spp, cam, seed=777, **tr)
    r = ref.render(p, threads=threads, rows=(r0, r0 + rows), aovs=False)
    c = r["counters"]
    rays = c["closest"] + c["shadow"]
    return {
        "value": rays / r["seconds"] / 1e6,
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"rows {r0}..{r0 + rows} of the {W}x{H} frame at {spp} spp ({rays} ray queries, "
                  f"{r['seconds']:.1f} s on {threads} threads)",
        "samples_per_s": rows * W * spp / r["seconds"],
    }


def relmse_scaled(x, ref) -> float:
    import numpy as np

    k = 1.0 / max(float(ref.mean()), 1e-30)
    return float(np.mean((x * k - ref * k) ** 2 / ((ref * k) ** 2 + 1e-2)))


def guided_leg(ctx, args) -> dict:
    """Equal-time relMSE, guided vs plain, on the door-gap room."""
    import numpy as np

    from atrc_amd import capi, scenes

    t0 = time.time()
    s, cam, tr = scenes.guided_room(n_clutter_tris=args.guided_tris, seed=7)
    s.accel = capi.ACCEL_BVH4
    desc, keep = s.to_desc()
    ctx.upload_scene(desc)
    W = H = args.guided_res
    log(f"[bench] guided leg: door-gap room, {s.triangle_count()} triangles, {W}x{H} "
        f"(built + uploaded in {time.time() - t0:.1f} s)")

    def run(seconds: float, seed: int, guide: bool, train_s: float = 0.0):
        acc = np.zeros((H, W, 3), np.float64)
        spp = 0
        st_sum = {"guide_infer_ms": 0.0, "guide_train_ms": 0.0, "guide_rows": 0, "guide_samples": 0,
                  "guide_trained": 0, "total_ms": 0.0, "rays": 0}
        t_start = time.perf_counter()
        while time.perf_counter() - t_start < seconds:
            train = guide and (spp == 0 or time.perf_counter() - t_start < train_s)
            p = scenes.render_params(W, H, args.guided_chunk, cam, seed=seed, sample_begin=spp, guide=guide,
                                     guide_reset=guide and spp == 0, guide_train=train,
                                     guide_steps=args.guided_steps, guide_batch=args.guided_batch, **tr)
            img = ctx.render(p, aovs=False)["image"]
            acc += img.astype(np.float64) * args.guided_chunk
            spp += args.guided_chunk
            st = ctx.stats()
            for k in ("guide_infer_ms", "guide_train_ms", "guide_rows", "guide_samples", "guide_trained", "total_ms"):
                st_sum[k] += st[k]
            st_sum["rays"] += st["closest_rays"] + st["shadow_rays"]
        wall = time.perf_counter() - t_start
        return acc / spp, spp, wall, st_sum

    ctx.render(scenes.render_params(W, H, 1, cam, seed=9, **tr), aovs=False)   # warm the kernels
    ref, ref_spp, ref_wall, _ = run(args.guided_seconds * args.guided_ref_factor, 1001, False)
    log(f"[bench] guided leg: reference plain {ref_spp} spp in {ref_wall:.1f} s")
    plain, plain_spp, plain_wall, pst = run(args.guided_seconds, 2002, False)
    guided, guided_spp, guided_wall, gst = run(args.guided_seconds, 3003, True,
                                               train_s=args.guided_train_frac * args.guided_seconds)
    ep, eg = relmse_scaled(plain, ref), relmse_scaled(guided, ref)
    # time-normalised: relMSE ~ 1 / time for an unbiased estimator
    ratio_t = (ep * plain_wall) / (eg * guided_wall) if eg > 0 else None
    log(f"[bench] guided leg: plain {plain_spp} spp / {plain_wall:.2f} s relMSE {ep:.4e}; "
        f"guided {guided_spp} spp / {guided_wall:.2f} s relMSE {eg:.4e}; ratio {ep / eg:.2f}")
    infer_tflops = gst["guide_rows"] * MLP_FORWARD_FLOP / max(gst["guide_infer_ms"], 1e-9) / 1e9
    train_tflops = gst["guide_trained"] * MLP_TRAIN_FLOP / max(gst["guide_train_ms"], 1e-9) / 1e9
    return {
        "workload": (f"door-gap room (10x10x4 Lambert room lit through a 0.05-wide gap, "
                     f"{s.triangle_count()} triangles), {W}x{H}, max depth {tr['max_depth']}"),
        "seconds": args.guided_seconds, "chunk_spp": args.guided_chunk,
        "train_frac": args.guided_train_frac, "adam_steps_per_pass": args.guided_steps,
        "sample_ring": args.guided_batch,
        "plain": {"spp": plain_spp, "wall_s": round(plain_wall, 3), "relmse": ep,
                  "Mrays_per_s": pst["rays"] / plain_wall / 1e6},
        "guided": {"spp": guided_spp, "wall_s": round(guided_wall, 3), "relmse": eg,
                   "Mrays_per_s": gst["rays"] / guided_wall / 1e6},
        "reference": {"spp": ref_spp, "wall_s": round(ref_wall, 3)},
        "ratio": ep / eg if eg > 0 else None,
        "ratio_time_normalised": ratio_t,
        "mlp": {
            "rows": gst["guide_rows"], "infer_ms": gst["guide_infer_ms"],
            "infer_tflops": infer_tflops, "infer_frac_bf16_peak": infer_tflops / BF16_PEAK_TFLOPS,
            "train_samples": gst["guide_samples"], "trained_sample_steps": gst["guide_trained"],
            "train_ms": gst["guide_train_ms"],
            "train_tflops": train_tflops, "train_frac_bf16_peak": train_tflops / BF16_PEAK_TFLOPS,
            "device_share": (gst["guide_infer_ms"] + gst["guide_train_ms"]) / max(gst["total_ms"], 1e-9),
        },
    }


def batched_main(args) -> int:
    """Batched workload: seeded scenes (triangles log-uniform in [1e4, 2e6],
    resolution from {256, 512, 1024} squared, 32 spp each) rendered by one
    agx_render_batch call; value = all ray queries / wall time of the call
    (host BVH builds and uploads included)."""
    import numpy as np
    from atrc_amd import capi, scenes

    t0 = time.time()
    descs, keep, params, tris, pix = [], [], [], 0, 0
    for i in range(args.batch_scenes):
        s, cam, tr, res = scenes.batch_scene(100 + i)
        s.accel = capi.ACCEL_BVH4 if args.accel == "bvh4" else capi.ACCEL_ATRC
        d, k = s.to_desc()
        descs.append(d)
        keep.append((k, s))
        params.append(scenes.render_params(res, res, 32, cam, seed=42 + i, **tr))
        tris += s.triangle_count()
        pix += res * res
    log(f"[bench] batched: {len(descs)} scenes, {tris} triangles, {pix} pixels x 32 spp "
        f"(generated in {time.time() - t0:.1f} s)")
    ctx = capi.Context(0)
    # warm the kernels on the smallest scene
    small = min(range(len(descs)), key=lambda i: params[i].width)
    ctx.upload_scene(descs[small])
    ctx.render(scenes.render_params(64, 64, 1, {"pos": (0, -10, 5), "dst": (0, 0, 0), "fov": 40.0}), aovs=False)
    t0 = time.perf_counter()
    imgs, recs, rep = ctx.render_batch(descs, params, workers=args.batch_workers)
    wall = time.perf_counter() - t0
    ctx.close()
    bad = [i for i, r in enumerate(recs) if r["status"] != capi.AGX_OK]
    if bad:
        raise RuntimeError(f"batched items failed: {bad}")
    log(f"[bench] batched: {rep['rays']} rays in {wall:.2f} s with {rep['workers']} workers "
        f"(autotuner probe rays/s by workers: {rep['tune_rays_per_s']})")
    # the same batch captured into one hipGraph (scenes built first, then one launch)
    graph = None
    if not args.no_batch_graph:
        ctx = capi.Context(0)
        t0 = time.perf_counter()
        gimgs, grecs, grep_ = ctx.render_batch(descs, params, workers=4, graph=True)
        gwall = time.perf_counter() - t0
        ctx.close()
        same = all(np.array_equal(a, b) for a, b in zip(imgs, gimgs))
        graph = {"branches": grep_["workers"], "wall_s": round(gwall, 4), "rays": grep_["rays"],
                 "Mrays_per_s": round(grep_["rays"] / gwall / 1e6, 3),
                 "build_ms_max": max(r["build_ms"] for r in grecs),
                 "graph_run_ms": grecs[0]["render_ms"] if grecs else 0.0,
                 "Mrays_per_s_graph_run": round(grep_["rays"] / max(grecs[0]["render_ms"], 1e-9) / 1e3, 3)
                 if grecs else 0.0,
                 "images_identical_to_worker_batch": bool(same)}
        log(f"[bench] batched graph: {grep_['rays']} rays in {gwall:.2f} s "
            f"(graph run {graph['graph_run_ms']:.0f} ms = {graph['Mrays_per_s_graph_run']:.0f} Mrays/s; "
            f"images identical to the worker batch: {same})")
    out = {
        "metric": METRIC, "value": round(rep["rays"] / wall / 1e6, 3), "unit": "Mrays/s", "n_gpus": 1,
        "steps": 1, "warmup": 0, "ms_per_step": round(wall * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": (f"batched plain GPU PT+NEE: {len(descs)} seeded procedural mesh scenes "
                                f"(triangles log-uniform in [1e4, 2e6]; {tris} in all), resolutions from "
                                f"256^2 / 512^2 / 1024^2, 32 spp each, one agx_render_batch call "
                                f"(host BVH builds included)"),
                   "scenes": len(descs), "triangles": tris, "pixels": pix, "spp": 32, "accel": args.accel},
        "batch": {"workers": rep["workers"], "tuned": rep["tuned"], "tune_rays_per_s": rep["tune_rays_per_s"],
                  "build_ms_sum": sum(r["build_ms"] for r in recs),
                  "render_ms_sum": sum(r["render_ms"] for r in recs)},
        "batch_graph": graph,
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    return 0


def pmc_traffic(kernel: str, workload_key: str):
    """HBM bytes per launch of ``kernel`` from the committed rocprofv3 --pmc summary of
    this workload (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by the copy kernel)."""
    f = ROOT / "profiles" / f"{kernel}_{workload_key}_pmc.txt"
    if not f.exists():
        return None, None
    for line in f.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1].strip()), str(f.relative_to(ROOT))
    return None, None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="mesh1m", choices=["mesh1m", "cornell", "volumetric", "batched"])
    ap.add_argument("--grid", type=int, default=256, help="density grid resolution of the volumetric config")
    ap.add_argument("--batch-scenes", type=int, default=64, help="scenes of the batched config (seeds 100..)")
    ap.add_argument("--batch-workers", type=int, default=0, help="scenes in flight (0 = autotuner)")
    ap.add_argument("--accel", default="bvh4", choices=["bvh4", "atrc"],
                    help="acceleration structure: bvh4 (SAH BVH4, production) or atrc (Atrc's own BVHs)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--stats-steps", type=int, default=2)
    ap.add_argument("--no-replay", action="store_true", help="skip the traversal-only ray replay")
    ap.add_argument("--no-guided", action="store_true", help="skip the guided vs plain equal-time leg")
    ap.add_argument("--guided-res", type=int, default=512)
    ap.add_argument("--guided-tris", type=int, default=500_000)
    ap.add_argument("--guided-seconds", type=float, default=4.0)
    ap.add_argument("--guided-chunk", type=int, default=4)
    ap.add_argument("--guided-ref-factor", type=float, default=16.0)
    ap.add_argument("--guided-train-frac", type=float, default=0.3)
    ap.add_argument("--guided-steps", type=int, default=8)
    ap.add_argument("--guided-batch", type=int, default=65536)
    ap.add_argument("--no-batch-graph", action="store_true",
                    help="batched config: skip the hipGraph-captured run of the same batch")
    ap.add_argument("--copy-calibration-mb", type=int, default=256,
                    help="run one plain copy kernel of this size first (traffic-counter reference), 0 = off")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "runs on one GPU only (--gpus 1); sharding the render over several "
                          "GPUs is not implemented", "n_gpus": args.gpus}), flush=True)
        return 2

    import numpy as np
    from atrc_amd import capi, scenes

    if args.config == "batched":
        return batched_main(args)
    t0 = time.time()
    s, cam, tr, W, H, workload = build_scene(args)
    s.accel = capi.ACCEL_BVH4 if args.accel == "bvh4" else capi.ACCEL_ATRC
    desc, keep = s.to_desc()
    log(f"[bench] scene built: {s.triangle_count()} triangles, {len(s.entities)} entities ({time.time() - t0:.1f} s)")
    ctx = capi.Context(0)
    t0 = time.time()
    ctx.upload_scene(desc)
    log(f"[bench] scene uploaded (BVH build + copy) in {time.time() - t0:.1f} s")

    copy = None
    if args.copy_calibration_mb > 0:
        nbytes = args.copy_calibration_mb << 20
        ms = ctx.copy_kernel(nbytes)
        copy = {"bytes_read": nbytes, "bytes_written": nbytes, "ms": ms, "GB/s": 2 * nbytes / (ms * 1e-3) / 1e9}
        log(f"[bench] copy calibration kernel: {nbytes >> 20} MiB in {ms:.3f} ms ({copy['GB/s']:.0f} GB/s r+w)")
    # one host image buffer for every render, as a render loop keeps it
    img = np.empty((H, W, 3), np.float32)
    # warmup: renders of the timed configuration (same pass size, so the wavefront
    # buffers the timed render uses are allocated before it), on other samples
    for i in range(args.warmup):
        ctx.render(scenes.render_params(W, H, args.steps, cam, seed=42, sample_begin=i * args.steps, **tr),
                   aovs=False, out=img)
    if args.warmup > 0:
        log(f"[bench] warmup: {args.warmup} renders of {args.steps} spp done")
    warm = args.warmup * args.steps
    # counting pass: per-ray algorithmic BVH bytes
    st_steps = max(1, min(args.stats_steps, args.steps))
    ctx.render(scenes.render_params(W, H, st_steps, cam, seed=42, collect_stats=True,
                                    sample_begin=warm, **tr), aovs=False, out=img)
    cs = ctx.stats()
    bytes_per_closest = cs["closest_bvh_bytes"] / max(1, cs["closest_rays"])
    bytes_per_shadow = cs["shadow_bvh_bytes"] / max(1, cs["shadow_rays"])
    log(f"[bench] stats pass: {bytes_per_closest:.0f} B/closest ray, {bytes_per_shadow:.0f} B/shadow ray, "
        f"{cs['closest_node_visits'] / max(1, cs['closest_rays']):.1f} node tests/closest ray")

    # timed region: the library synchronises its stream before returning
    p = scenes.render_params(W, H, args.steps, cam, seed=42, sample_begin=warm + st_steps, **tr)
    t0 = time.perf_counter()
    ctx.render(p, aovs=False, out=img)   # the image comes back to the host; the AOV planes are not asked for
    t1 = time.perf_counter()
    st = ctx.stats()
    wall = t1 - t0
    rays = st["closest_rays"] + st["shadow_rays"]
    value = rays / wall / 1e6
    log(f"[bench] timed {args.steps} spp: {wall:.3f} s wall, device {st['total_ms']:.1f} ms, "
        f"closest {st['trace_closest_ms']:.1f} ms, shadow {st['trace_shadow_ms']:.1f} ms, shade {st['shade_ms']:.1f} ms; "
        f"host call {st['host_ms']:.1f} ms of which image readback {st['readback_ms']:.1f} ms")

    closest_bytes = bytes_per_closest * st["closest_rays"]
    achieved = closest_bytes / (st["trace_closest_ms"] * 1e-3) / 1e9
    traffic, traffic_src = pmc_traffic("k_trace_closest", f"{args.config}_{args.accel}")
    roofline = {
        "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic, "traffic_source": traffic_src,
        "kernel": "k_trace_closest",
        "avg_launch_ms": st["trace_closest_ms"] / max(1, st["trace_closest_launches"]),
        "bytes_per_launch": closest_bytes / max(1, st["trace_closest_launches"]),
    }
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": workload, "width": W, "height": H, "spp_per_step": 1,
                   "triangles": s.triangle_count(), "max_depth": tr["max_depth"], "min_depth": tr["min_depth"],
                   "cont_prob": tr["cont_prob"], "accel": args.accel},
        "rays": {"closest": st["closest_rays"], "shadow": st["shadow_rays"], "paths": st["paths"]},
        "samples_per_s": st["paths"] / wall,
        "device_ms": {k: st[k] for k in ("total_ms", "trace_closest_ms", "trace_shadow_ms", "shade_ms")},
        "host_ms": {"call": round(st["host_ms"], 3), "image_readback": round(st["readback_ms"], 3)},
        "roofline": roofline,
        "guided_vs_plain_relmse_ratio": None,
        "guided": None,
        "copy_calibration": copy,
        "traversal_stack": "lds" if os.environ.get("AGX_TRAVERSAL_STACK") == "lds" else "private",
    }
    if not args.no_replay:
        # BVH-bound throughput of the same ray distribution: one frame's closest-hit and any-hit
        # queries captured from the wavefront and traced again by the traversal kernel alone
        rr = ctx.replay_rays(scenes.render_params(W, H, 1, cam, seed=4242, **tr))
        n_r = rr["closest_rays"] + rr["shadow_rays"]
        replay_mrays = n_r / ((rr["replay_closest_ms"] + rr["replay_shadow_ms"]) * 1e-3) / 1e6
        out["replay"] = {
            "rays": n_r, "closest_rays": rr["closest_rays"], "shadow_rays": rr["shadow_rays"],
            "closest_ms": rr["replay_closest_ms"], "shadow_ms": rr["replay_shadow_ms"],
            "render_closest_ms": rr["render_closest_ms"], "render_shadow_ms": rr["render_shadow_ms"],
            "Mrays_per_s": replay_mrays, "pt_fraction": value / replay_mrays,
        }
        log(f"[bench] ray replay: {n_r} rays, traversal alone {replay_mrays:.1f} Mrays/s; "
            f"PT reaches {value / replay_mrays:.1%} of it")
    if not args.no_guided:
        g = guided_leg(ctx, args)
        out["guided"] = g
        out["guided_vs_plain_relmse_ratio"] = g["ratio"]
    ctx.close()
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (oracle, FP32) ...")
        out["cpu_baseline"] = cpu_baseline(desc, cam, tr, W, H, args.cpu_seconds)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
