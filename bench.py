"""Benchmark of the gfx950 path tracer (one MI355X).

Headline workload: plain GPU PT+NEE
on the seeded 1M-triangle procedural mesh scene -- 200 displaced icospheres
(1,024,000 FP32 triangles, half Lambert / half GGX conductor) on a ground
quad under a native_sky and the sun, 1920x1080, 64 spp, max depth 8 (Atrc
trace_std: NEE over every light, MIS-BSDF ray, Russian roulette after depth 5
with cont_prob 0.9).  Atrc has no directional light, so the sun is a distant
emissive disk subtending 0.53 degrees (scenes.procedural_spheres(sun=True)).
One step = one full frame (1920x1080 at 64 spp, four wavefront passes of 16
sample layers) with the scene, BVH and path state resident in HBM; the image
comes back to the host at the end of every step.

value = (closest-hit + any-hit ray queries) / wall time of the timed steps.

roofline: the dominant kernel is the closest-hit traversal (k_trace_closest).
Algorithmic bytes per closest-hit ray = node fetches x 64 B (BVH4 node with
four quantized child boxes) + triangle tests x 48 B + hits x 96 B (hit
record), counted by a stats pass over the first samples; x the closest-hit
rays of the timed steps = algorithmic bytes of the timed launches, / that
kernel's device time (HIP events on the library's stream around every launch)
against the 8 TB/s HBM peak.  traffic = HBM bytes per launch of the same
kernel in the same command, from the committed rocprofv3 FETCH_SIZE /
WRITE_SIZE summary (profiles/k_trace_closest_mesh1m_bvh4_pmc.txt); both
per-launch figures average over every launch of the frame's bounces.

cpu_baseline: the CPU restatement of Atrc `pt` (the oracle; FP32, Atrc's
32x32 tile scheduler, one worker per core) on a band of rows of the same
frame at 1 spp, on this host's cores: the generic build, a -march=native
build made on this host, and a one-thread run; nproc and the CPU model are
reported.

Further legs (each can be switched off):
* watertight: the headline frame again in the watertight BVH4 mode
  (Woop-Benthin-Wald triangle and quad tests over the same tree).
* cornell: Cornell-like scene, 256x256, 16 spp, depth 5, on the
  GPU and on the CPU oracle, timed, each with its relMSE against a 4096-spp
  GPU reference.
* volumetric (plain PT): the door-gap room filled with a 256^3
  fBm heterogeneous medium and a homogeneous slab, 1024x1024, depth 16.
* batched: 64 seeded scenes (1e4..2e6 triangles, 256^2..1024^2,
  32 spp) in one agx_render_batch_graph call: host BVH builds on every core
  overlapped with one hipGraph launch per group of 8 scenes, configuration
  (branches, pass size, traversal block size) picked by the library's
  autotuner.
* guided: equal-time relMSE of guided vs plain GPU PT on the
  door-gap room (500k clutter triangles) at 1024x1024; guided_volumetric:
  the same comparison in the fog-filled room (FP16 density
  grid, depth 16) with the learned density at probability --vol-guided-alpha;
  --vol-guided-medium-depth > 0 also guides scattering points up to that
  depth, each with its own network evaluation.  Plain and guided runs
  each take --guided-seconds of wall time in chunks of --guided-chunk spp; the
  guided run starts from a fresh network, trains on every pass for the first
  --guided-train-frac of its time, then renders with the network frozen, and
  every sample goes into its image.  The reference is plain PT with another
  seed for --guided-ref-factor times as long, rendered as two halves so that
  its own noise can be subtracted.  relMSE = mean over pixels and channels of
  (x - y)^2 / (y^2 + 1e-2) after scaling both images by 1 / mean(y).  MLP
  utilisation = 61,440 FLOP per evaluated vertex / the inference kernel's
  device time, and 3 x 61,440 FLOP per trained sample and step / the
  training device time, against the 2.5 PFLOP/s dense BF16 peak.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import tempfile
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mrays/s; equal-time relMSE, guided vs plain GPU PT, 1 MI355X"
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
MLP_FORWARD_FLOP = 2 * (256 * 64 + 3 * 64 * 64 + 64 * 32)   # 61,440 per evaluated vertex
MLP_TRAIN_FLOP = 3 * MLP_FORWARD_FLOP                       # forward + backward per sample and step


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def host_info() -> dict:
    model = platform.processor() or ""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    return {"nproc": os.cpu_count() or 1, "affinity": affinity, "cpu_model": model}


def cpu_threads() -> int:
    """Worker threads for the CPU runs: every core this process may run on, or
    the job's thread share when the environment states one (OMP_NUM_THREADS)."""
    info = host_info()
    n = info["affinity"]
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def relmse(x, y) -> float:
    import numpy as np

    return float(np.mean((x - y) ** 2 / (y ** 2 + 1e-2)))


def relmse_scaled(x, ref) -> float:
    import numpy as np

    k = 1.0 / max(float(ref.mean()), 1e-30)
    return float(np.mean((x * k - ref * k) ** 2 / ((ref * k) ** 2 + 1e-2)))


# ----------------------------------------------------------------------------- headline

def build_headline(args):
    from atrc_amd import capi, scenes

    if args.config == "cornell":
        s, cam, tr = scenes.cornell_box()
        W, H = 256, 256
        workload = "plain GPU PT+NEE, Cornell-like scene (5 quads + 2 boxes + quad light), 256x256, depth 5"
    elif args.config == "volumetric":
        s, cam, tr = scenes.volumetric_room(grid_n=args.grid)
        W, H = 1024, 1024
        workload = (f"plain GPU PT+NEE with media: door-gap room (500k clutter triangles) filled with a "
                    f"{args.grid}^3 fBm heterogeneous medium (sigma_t max 2, albedo 0.9, g 0.3) and a "
                    f"homogeneous slab (sigma_t 0.5), 1024x1024, depth 16")
    else:
        s, cam, tr = scenes.procedural_spheres(n_spheres=200, subdiv=4, seed=1, sun=True)
        W, H = 1920, 1080
        workload = ("plain GPU PT+NEE, 1M-triangle procedural mesh scene (200 displaced icospheres, "
                    "Lambert + GGX conductor, ground quad, native_sky + sun as a 0.53-degree emissive disk), "
                    f"1920x1080, {args.spp} spp per frame, depth 8")
    s.accel = {"bvh4": capi.ACCEL_BVH4, "atrc": capi.ACCEL_ATRC,
               "bvh4_watertight": capi.ACCEL_BVH4_WATERTIGHT}[args.accel]
    return s, cam, tr, W, H, workload


def pmc_traffic(kernel: str, workload_key: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc summary of
    this workload (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by the copy kernel)."""
    f = ROOT / "profiles" / f"{kernel}_{workload_key}_pmc.txt"
    if not f.exists():
        return None, None
    for line in f.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1].strip()), str(f.relative_to(ROOT))
    return None, None


def headline(ctx, args):
    import numpy as np
    from atrc_amd import scenes

    t0 = time.time()
    s, cam, tr, W, H, workload = build_headline(args)
    desc, keep = s.to_desc()
    log(f"[bench] scene built: {s.triangle_count()} triangles, {len(s.entities)} entities ({time.time() - t0:.1f} s)")
    t0 = time.time()
    ctx.upload_scene(desc)
    log(f"[bench] scene uploaded (BVH build + copy) in {time.time() - t0:.1f} s")
    copy = None
    if args.copy_calibration_mb > 0:
        nbytes = args.copy_calibration_mb << 20
        ms = ctx.copy_kernel(nbytes)
        copy = {"bytes_read": nbytes, "bytes_written": nbytes, "ms": ms, "GB/s": 2 * nbytes / (ms * 1e-3) / 1e9}
        log(f"[bench] copy calibration kernel: {nbytes >> 20} MiB in {ms:.3f} ms ({copy['GB/s']:.0f} GB/s r+w)")
    img = np.empty((H, W, 3), np.float32)      # one host image buffer, as a render loop keeps it
    spp = args.spp
    for i in range(args.warmup):               # the timed configuration: its buffers exist before timing
        ctx.render(scenes.render_params(W, H, spp, cam, seed=42, sample_begin=i * spp, **tr), aovs=False, out=img)
    if args.warmup:
        log(f"[bench] warmup: {args.warmup} frames of {spp} spp done")
    base = args.warmup * spp
    # counting pass: algorithmic BVH bytes per ray
    ctx.render(scenes.render_params(W, H, args.stats_spp, cam, seed=42, collect_stats=True, sample_begin=base,
                                    **tr), aovs=False, out=img)
    cs = ctx.stats()
    bytes_per_closest = cs["closest_bvh_bytes"] / max(1, cs["closest_rays"])
    bytes_per_shadow = cs["shadow_bvh_bytes"] / max(1, cs["shadow_rays"])
    log(f"[bench] stats pass: {bytes_per_closest:.0f} B/closest ray, {bytes_per_shadow:.0f} B/shadow ray, "
        f"{cs['closest_node_visits'] / max(1, cs['closest_rays']):.2f} node fetches/closest ray")
    base += args.stats_spp
    acc = {k: 0.0 for k in ("closest_rays", "shadow_rays", "paths", "total_ms", "trace_closest_ms",
                            "trace_shadow_ms", "shade_ms", "trace_closest_launches", "trace_shadow_launches",
                            "host_ms", "readback_ms")}
    t0 = time.perf_counter()
    for k in range(args.steps):                # the library synchronises its stream before returning
        ctx.render(scenes.render_params(W, H, spp, cam, seed=42, sample_begin=base + k * spp, **tr),
                   aovs=False, out=img)
        st = ctx.stats()
        for key in acc:
            acc[key] += st[key]
    wall = time.perf_counter() - t0
    rays = acc["closest_rays"] + acc["shadow_rays"]
    value = rays / wall / 1e6
    log(f"[bench] timed {args.steps} frames of {spp} spp: {wall:.3f} s wall ({wall * 1e3 / args.steps:.1f} ms/frame), "
        f"device {acc['total_ms']:.1f} ms: closest {acc['trace_closest_ms']:.1f}, shadow {acc['trace_shadow_ms']:.1f}, "
        f"shade {acc['shade_ms']:.1f}; image readback {acc['readback_ms']:.1f} ms; {value:.1f} Mrays/s")
    launches = max(1, int(acc["trace_closest_launches"]))
    closest_bytes = bytes_per_closest * acc["closest_rays"]
    achieved = closest_bytes / (acc["trace_closest_ms"] * 1e-3) / 1e9
    traffic, traffic_src = pmc_traffic("k_trace_closest", f"{args.config}_{args.accel}")
    bytes_per_launch = closest_bytes / launches
    roofline = {
        "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
        "kernel": "k_trace_closest",
        "algorithmic_bytes_per_ray": round(bytes_per_closest, 2),
        "bytes_per_launch": bytes_per_launch,
        "avg_launch_ms": acc["trace_closest_ms"] / launches,
        "launches": launches,
        "traffic_source": traffic_src,
        "traffic_over_algorithmic": (traffic / bytes_per_launch) if traffic else None,
        "shadow": {"kernel": "k_trace_shadow", "algorithmic_bytes_per_ray": round(bytes_per_shadow, 2),
                   "achieved": round(bytes_per_shadow * acc["shadow_rays"] / max(acc["trace_shadow_ms"], 1e-9)
                                     / 1e6, 2),
                   "frac": round(bytes_per_shadow * acc["shadow_rays"] / max(acc["trace_shadow_ms"], 1e-9) / 1e6
                                 / HBM_PEAK_GBS, 5)},
    }
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": workload, "width": W, "height": H, "spp_per_step": spp,
                   "triangles": s.triangle_count(), "max_depth": tr["max_depth"], "min_depth": tr["min_depth"],
                   "cont_prob": tr["cont_prob"], "accel": args.accel, "lights": 2 if args.config == "mesh1m" else None},
        "rays": {"closest": int(acc["closest_rays"]), "shadow": int(acc["shadow_rays"]), "paths": int(acc["paths"])},
        "samples_per_s": acc["paths"] / wall,
        "device_ms": {k: acc[k] for k in ("total_ms", "trace_closest_ms", "trace_shadow_ms", "shade_ms")},
        "host_ms": {"call": round(acc["host_ms"], 3), "image_readback": round(acc["readback_ms"], 3)},
        "roofline": roofline,
        "copy_calibration": copy,
    }
    if not args.no_replay:
        # BVH-bound throughput of the same ray distribution: one 1-spp frame's queries
        # captured from the wavefront and traced again by the traversal kernel alone
        rr = ctx.replay_rays(scenes.render_params(W, H, 1, cam, seed=4242, **tr))
        n_r = rr["closest_rays"] + rr["shadow_rays"]
        replay_mrays = n_r / ((rr["replay_closest_ms"] + rr["replay_shadow_ms"]) * 1e-3) / 1e6
        out["replay"] = {
            "rays": n_r, "closest_ms": rr["replay_closest_ms"], "shadow_ms": rr["replay_shadow_ms"],
            "render_closest_ms": rr["render_closest_ms"], "render_shadow_ms": rr["render_shadow_ms"],
            "Mrays_per_s": replay_mrays, "pt_fraction": value / replay_mrays,
        }
        log(f"[bench] ray replay: {n_r} rays, traversal alone {replay_mrays:.1f} Mrays/s; "
            f"PT reaches {value / replay_mrays:.1%} of it")
    return out, (desc, keep, cam, tr, W, H)


def leg_watertight(ctx, args, head: dict, desc, cam, tr, W: int, H: int) -> dict:
    """The headline frame again with the watertight BVH4 mode (Woop-Benthin-Wald
    triangle and quad tests over the same tree): its cost against the default
    Moller-Trumbore mode."""
    import numpy as np
    from atrc_amd import capi, scenes

    desc.accel = capi.ACCEL_BVH4_WATERTIGHT
    t0 = time.time()
    ctx.upload_scene(desc)
    build_s = time.time() - t0
    img = np.empty((H, W, 3), np.float32)
    ctx.render(scenes.render_params(W, H, args.spp, cam, seed=42, **tr), aovs=False, out=img)
    rays = closest_ms = 0.0
    t0 = time.perf_counter()
    for k in range(args.watertight_frames):
        ctx.render(scenes.render_params(W, H, args.spp, cam, seed=42, sample_begin=(k + 1) * args.spp, **tr),
                   aovs=False, out=img)
        st = ctx.stats()
        rays += st["closest_rays"] + st["shadow_rays"]
        closest_ms += st["trace_closest_ms"]
    wall = time.perf_counter() - t0
    desc.accel = capi.ACCEL_BVH4
    value = rays / wall / 1e6
    log(f"[bench] watertight BVH4: {args.watertight_frames} frames of {args.spp} spp, {value:.1f} Mrays/s "
        f"({value / head['value']:.3f} of the default mode), closest hit {closest_ms / args.watertight_frames:.1f} "
        f"ms/frame")
    return {"workload": "the headline scene and frame, accel bvh4_watertight", "frames": args.watertight_frames,
            "Mrays_per_s": value, "ms_per_frame": wall * 1e3 / args.watertight_frames,
            "closest_ms_per_frame": closest_ms / args.watertight_frames, "upload_s": build_s,
            "vs_default_mode": value / head["value"]}


# ----------------------------------------------------------------------------- CPU baseline

def build_native_oracle():
    """The oracle built with -march=native on this host, or None."""
    out = Path(tempfile.mkdtemp(prefix="oracle_native_"))
    try:
        subprocess.run(["make", "-s", "-B", "-C", str(ROOT / "oracle"), "native", f"OUT={out}"], check=True,
                       timeout=300, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE)
    except (subprocess.SubprocessError, OSError) as e:
        log(f"[bench] native oracle build failed ({e}); generic build only")
        return None
    return out / "liboracle_native.so"


def cpu_run(po, desc, cam, tr, W, H, target_s, threads, seed=777):
    from atrc_amd import scenes

    ref = po.OracleScene(desc, 32)
    probe = 8
    r = ref.render(scenes.render_params(W, H, 1, cam, seed=seed, **tr), threads=threads,
                   rows=((H - probe) // 2, (H - probe) // 2 + probe), aovs=False)
    per_row = max(r["seconds"], 1e-4) / probe
    rows_needed = target_s / per_row
    # a band of rows at 1 spp, or the whole frame at several spp when one frame
    # takes less than the target time
    if rows_needed <= H:
        rows, spp = int(max(probe, rows_needed)), 1
    else:
        rows, spp = H, int(min(64, max(1, round(rows_needed / H))))
    r0 = (H - rows) // 2
    r = ref.render(scenes.render_params(W, H, spp, cam, seed=seed, **tr), threads=threads, rows=(r0, r0 + rows),
                   aovs=False)
    c = r["counters"]
    rays = c["closest"] + c["shadow"]
    return {"Mrays_per_s": rays / r["seconds"] / 1e6, "seconds": r["seconds"], "rows": (r0, r0 + rows),
            "spp": spp, "rays": rays, "samples_per_s": rows * W * spp / r["seconds"], "threads": threads}


def cpu_baseline(desc, cam, tr, W, H, args) -> dict:
    from oracle import py_oracle as po

    info = host_info()
    threads = cpu_threads()
    gen = cpu_run(po, desc, cam, tr, W, H, args.cpu_seconds, threads)
    log(f"[bench] CPU oracle (generic build), {threads} threads: {gen['Mrays_per_s']:.2f} Mrays/s "
        f"({gen['rays']} rays in {gen['seconds']:.1f} s)")
    native = None
    if not args.no_native_cpu:
        lib = build_native_oracle()
        if lib is not None:
            po.load(lib)
            native = cpu_run(po, desc, cam, tr, W, H, args.cpu_seconds, threads)
            log(f"[bench] CPU oracle (-march=native), {threads} threads: {native['Mrays_per_s']:.2f} Mrays/s")
    one = cpu_run(po, desc, cam, tr, W, H, args.cpu_seconds / 2, 1)
    log(f"[bench] CPU oracle, 1 thread: {one['Mrays_per_s']:.3f} Mrays/s")
    best = native if native and native["Mrays_per_s"] > gen["Mrays_per_s"] else gen
    return {
        "value": round(best["Mrays_per_s"], 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
        "sample": (f"rows {best['rows'][0]}..{best['rows'][1]} of the {W}x{H} frame at {best['spp']} spp "
                   f"({best['rays']} ray queries, {best['seconds']:.1f} s on {threads} threads, "
                   f"{'-march=native' if best is native else 'generic -O3'} build)"),
        "build": "-march=native" if best is native else "generic -O3",
        "generic": gen, "native": native, "one_thread": one,
        "host": dict(info, threads_used=threads),
    }


# ----------------------------------------------------------------------------- further legs

def leg_cornell(ctx, args) -> dict:
    """The Cornell-like scene at 256x256, 16 spp, on the GPU and on the CPU oracle."""
    import numpy as np
    from atrc_amd import capi, scenes
    from oracle import py_oracle as po

    s, cam, tr = scenes.cornell_box()
    s.accel = capi.ACCEL_BVH4
    desc, keep = s.to_desc()
    ctx.upload_scene(desc)
    W = H = 256
    ref = ctx.render(scenes.render_params(W, H, 4096, cam, seed=9001, **tr), aovs=False)["image"].astype(np.float64)
    p = scenes.render_params(W, H, 16, cam, seed=42, **tr)
    ctx.render(p, aovs=False)                           # warm
    t0 = time.perf_counter()
    g = ctx.render(p, aovs=False)["image"].astype(np.float64)
    gw = time.perf_counter() - t0
    gst = ctx.stats()
    threads = cpu_threads()
    c = po.OracleScene(desc, 32).render(scenes.render_params(W, H, 16, cam, seed=4242, **tr), threads=threads,
                                        aovs=False)
    cc = c["counters"]
    g_rays = gst["closest_rays"] + gst["shadow_rays"]
    c_rays = cc["closest"] + cc["shadow"]
    out = {
        "workload": "Cornell-like scene, 256x256, 16 spp, depth 5 (min_depth 2, cont_prob 0.9), BVH4",
        "gpu": {"wall_s": gw, "Mrays_per_s": g_rays / gw / 1e6, "relmse_vs_ref": relmse(g, ref)},
        "cpu": {"wall_s": c["seconds"], "threads": threads, "Mrays_per_s": c_rays / c["seconds"] / 1e6,
                "relmse_vs_ref": relmse(c["image"], ref), "kind": "oracle FP32, generic build"},
        "reference": "GPU 4096 spp, another seed",
    }
    out["relmse_gpu_over_cpu"] = out["gpu"]["relmse_vs_ref"] / max(out["cpu"]["relmse_vs_ref"], 1e-30)
    out["speedup"] = c["seconds"] / gw
    log(f"[bench] cornell 256^2 16 spp: GPU {gw * 1e3:.1f} ms ({out['gpu']['Mrays_per_s']:.0f} Mrays/s, relMSE "
        f"{out['gpu']['relmse_vs_ref']:.3e}) vs CPU {c['seconds']:.2f} s on {threads} threads "
        f"({out['cpu']['Mrays_per_s']:.1f} Mrays/s, relMSE {out['cpu']['relmse_vs_ref']:.3e})")
    return out


def leg_volumetric(ctx, args) -> dict:
    """Plain PT in the door-gap room filled with media, 1024x1024, depth 16."""
    import numpy as np
    from atrc_amd import capi, scenes

    t0 = time.time()
    s, cam, tr = scenes.volumetric_room(grid_n=args.grid)
    s.accel = capi.ACCEL_BVH4
    desc, keep = s.to_desc()
    ctx.upload_scene(desc)
    log(f"[bench] volumetric room built and uploaded in {time.time() - t0:.1f} s")
    W = H = 1024
    img = np.empty((H, W, 3), np.float32)
    ctx.render(scenes.render_params(W, H, 1, cam, seed=5, **tr), aovs=False, out=img)
    p = scenes.render_params(W, H, args.vol_spp, cam, seed=42, **tr)
    t0 = time.perf_counter()
    ctx.render(p, aovs=False, out=img)
    wall = time.perf_counter() - t0
    st = ctx.stats()
    rays = st["closest_rays"] + st["shadow_rays"]
    log(f"[bench] volumetric 1024^2 {args.vol_spp} spp: {wall:.3f} s, {rays / wall / 1e6:.0f} Mrays/s "
        f"(shade {st['shade_ms']:.0f} ms of {st['total_ms']:.0f} ms device)")
    return {"workload": f"door-gap room + {args.grid}^3 fBm medium + homogeneous slab, 1024x1024, "
                        f"{args.vol_spp} spp, depth 16, BVH4",
            "wall_s": wall, "Mrays_per_s": rays / wall / 1e6, "rays": rays,
            "device_ms": {k: st[k] for k in ("total_ms", "trace_closest_ms", "trace_shadow_ms", "shade_ms")},
            "finite": bool(np.isfinite(img).all())}


def leg_batched(args) -> dict:
    """Seeded scenes of uneven size rendered in one hipGraph launch."""
    from atrc_amd import capi, scenes

    t0 = time.time()
    descs, keep, params, tris, pix = [], [], [], 0, 0
    for i in range(args.batch_scenes):
        s, cam, tr, res = scenes.batch_scene(100 + i)
        s.accel = capi.ACCEL_BVH4
        d, k = s.to_desc()
        descs.append(d)
        keep.append((k, s))
        params.append(scenes.render_params(res, res, 32, cam, seed=42 + i, **tr))
        tris += s.triangle_count()
        pix += res * res
    log(f"[bench] batched: {len(descs)} scenes, {tris} triangles, {pix} pixels x 32 spp "
        f"(generated in {time.time() - t0:.1f} s)")
    ctx = capi.Context(0)
    try:
        t0 = time.perf_counter()
        imgs, recs, rep = ctx.render_batch(descs, params, workers=args.batch_branches, graph=True)
        wall = time.perf_counter() - t0
    finally:
        ctx.close()
    bad = [i for i, r in enumerate(recs) if r["status"] != capi.AGX_OK]
    if bad:
        raise RuntimeError(f"batched items failed: {bad}")
    # device time of the graphs: each group's graph time, summed once per group
    group_ms = {}
    for r in recs:
        group_ms.setdefault(round(r["render_ms"], 6), r["render_ms"])
    graph_ms = sum(group_ms.values())
    out = {"workload": (f"{len(descs)} seeded procedural mesh scenes (triangles log-uniform in [1e4, 2e6]; {tris} in "
                        f"all), 256^2..1024^2, 32 spp each, one agx_render_batch_graph call (BVH4; host builds "
                        f"overlapped with the group graphs)"),
           "scenes": len(descs), "triangles": tris, "pixels": pix, "rays": rep["rays"],
           "wall_s": wall, "Mrays_per_s": rep["rays"] / wall / 1e6,
           "tuned": rep["tuned"], "branches": rep["workers"], "pass_paths_log2": rep["pass_paths_log2"],
           "trace_block": rep["trace_block"], "tune_configs": rep["tune_configs"],
           "tune_rays_per_s": rep["tune_rays_per_s"], "groups": rep["groups"], "build_threads": rep["build_threads"],
           "build_wall_s": rep["build_wall_ms"] / 1e3, "render_s": rep["render_ms"] / 1e3,
           "graph_device_ms": graph_ms, "Mrays_per_s_graph_device": rep["rays"] / max(graph_ms, 1e-9) / 1e3,
           "build_ms_max": max(r["build_ms"] for r in recs)}
    log(f"[bench] batched graph: {rep['rays']} rays in {wall:.2f} s = {out['Mrays_per_s']:.0f} Mrays/s whole call "
        f"(builds on {rep['build_threads']} threads done at {rep['build_wall_ms'] / 1e3:.2f} s; {rep['groups']} group "
        f"graphs, {graph_ms:.0f} ms device = {out['Mrays_per_s_graph_device']:.0f} Mrays/s); tuned: "
        f"{rep['workers']} branches, passes of 2^{rep['pass_paths_log2']}, {rep['trace_block']}-thread traversal "
        f"blocks ({rep['tune_configs']} configurations timed)")
    return out


def leg_guided(ctx, args, volumetric: bool = False) -> dict:
    """The door-gap room, or the same room filled with media:
    equal-time relMSE, guided vs plain."""
    import numpy as np
    from atrc_amd import capi, scenes

    t0 = time.time()
    if volumetric:
        s, cam, tr = scenes.volumetric_room(n_clutter_tris=args.guided_tris, grid_n=args.grid)
        seconds, ref_factor = args.vol_guided_seconds, args.vol_guided_ref_factor
        gopts = {"guide_alpha": args.vol_guided_alpha, "guide_medium_depth": args.vol_guided_medium_depth}
    else:
        gopts = {}
        s, cam, tr = scenes.guided_room(n_clutter_tris=args.guided_tris, seed=7)
        seconds, ref_factor = args.guided_seconds, args.guided_ref_factor
    s.accel = capi.ACCEL_BVH4
    desc, keep = s.to_desc()
    ctx.upload_scene(desc)
    W = H = args.guided_res
    label = "guided volumetric leg" if volumetric else "guided leg"
    log(f"[bench] {label}: {'fog-filled ' if volumetric else ''}door-gap room, {s.triangle_count()} triangles, "
        f"{W}x{H} (built + uploaded in {time.time() - t0:.1f} s)")
    img = np.empty((H, W, 3), np.float32)

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
                                     guide_steps=args.guided_steps, guide_batch=args.guided_batch,
                                     **(gopts if guide else {}), **tr)
            ctx.render(p, aovs=False, out=img)
            acc += img.astype(np.float64) * args.guided_chunk
            spp += args.guided_chunk
            st = ctx.stats()
            for k in ("guide_infer_ms", "guide_train_ms", "guide_rows", "guide_samples", "guide_trained", "total_ms"):
                st_sum[k] += st[k]
            st_sum["rays"] += st["closest_rays"] + st["shadow_rays"]
        return acc / max(spp, 1), spp, time.perf_counter() - t_start, st_sum

    ctx.render(scenes.render_params(W, H, 1, cam, seed=9, **tr), aovs=False, out=img)   # warm the kernels
    half = seconds * ref_factor / 2
    ref_a, spp_a, wall_a, _ = run(half, 1001, False)
    ref_b, spp_b, wall_b, _ = run(half, 1002, False)
    ref = (ref_a * spp_a + ref_b * spp_b) / (spp_a + spp_b)
    # the reference's own noise: var(ref) = E|a - b|^2 / 4 for equal halves
    ref_noise = relmse_scaled(ref_a, ref_b) / 4.0 * (4.0 * spp_a * spp_b / (spp_a + spp_b) ** 2)
    log(f"[bench] {label}: reference plain {spp_a + spp_b} spp in {wall_a + wall_b:.1f} s "
        f"(own relMSE about {ref_noise:.3e})")
    plain, plain_spp, plain_wall, pst = run(seconds, 2002, False)
    guided, guided_spp, guided_wall, gst = run(seconds, 3003, True, train_s=args.guided_train_frac * seconds)
    ep_raw, eg_raw = relmse_scaled(plain, ref), relmse_scaled(guided, ref)
    ep, eg = max(ep_raw - ref_noise, 1e-30), max(eg_raw - ref_noise, 1e-30)
    log(f"[bench] {label}: plain {plain_spp} spp / {plain_wall:.2f} s relMSE {ep:.4e}; guided {guided_spp} spp / "
        f"{guided_wall:.2f} s relMSE {eg:.4e}; ratio {ep / eg:.2f} (uncorrected {ep_raw / eg_raw:.2f})")
    infer_tflops = gst["guide_rows"] * MLP_FORWARD_FLOP / max(gst["guide_infer_ms"], 1e-9) / 1e9
    train_tflops = gst["guide_trained"] * MLP_TRAIN_FLOP / max(gst["guide_train_ms"], 1e-9) / 1e9
    return {
        "workload": (f"door-gap room (10x10x4 Lambert room lit through a 0.05-wide gap, "
                     f"{s.triangle_count()} triangles)"
                     + (f" filled with a {args.grid}^3 fBm medium (FP16 grid, sigma_t max 2, albedo 0.9, g 0.3) "
                        f"and a homogeneous slab (sigma_t 0.5); mixture alpha {args.vol_guided_alpha}, medium points "
                        f"guided to depth {args.vol_guided_medium_depth}" if volumetric else "")
                     + f", {W}x{H}, max depth {tr['max_depth']}, BVH4"),
        "seconds": seconds, "chunk_spp": args.guided_chunk,
        "train_frac": args.guided_train_frac, "adam_steps_per_pass": args.guided_steps,
        "sample_ring": args.guided_batch, "guide_alpha": gopts.get("guide_alpha", 0.5),
        "guide_medium_depth": gopts.get("guide_medium_depth", 0),
        "plain": {"spp": plain_spp, "wall_s": round(plain_wall, 3), "relmse": ep, "relmse_uncorrected": ep_raw,
                  "Mrays_per_s": pst["rays"] / plain_wall / 1e6},
        "guided": {"spp": guided_spp, "wall_s": round(guided_wall, 3), "relmse": eg, "relmse_uncorrected": eg_raw,
                   "Mrays_per_s": gst["rays"] / guided_wall / 1e6},
        "reference": {"spp": spp_a + spp_b, "wall_s": round(wall_a + wall_b, 3), "own_relmse": ref_noise},
        "ratio": ep / eg,
        "ratio_uncorrected": ep_raw / eg_raw,
        "ratio_time_normalised": (ep * plain_wall) / (eg * guided_wall),
        "mlp": {
            "rows": gst["guide_rows"], "infer_ms": gst["guide_infer_ms"],
            "infer_tflops": infer_tflops, "infer_frac_bf16_peak": infer_tflops / BF16_PEAK_TFLOPS,
            "train_samples": gst["guide_samples"], "trained_sample_steps": gst["guide_trained"],
            "train_ms": gst["guide_train_ms"],
            "train_tflops": train_tflops, "train_frac_bf16_peak": train_tflops / BF16_PEAK_TFLOPS,
            "device_share": (gst["guide_infer_ms"] + gst["guide_train_ms"]) / max(gst["total_ms"], 1e-9),
        },
    }


# ----------------------------------------------------------------------------- main

def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed frames")
    ap.add_argument("--warmup", type=int, default=2, help="untimed frames before timing")
    ap.add_argument("--spp", type=int, default=64, help="samples per pixel of one frame (one step)")
    ap.add_argument("--config", default="mesh1m", choices=["mesh1m", "cornell", "volumetric"],
                    help="headline workload (default: the 1M-triangle scene)")
    ap.add_argument("--accel", default="bvh4", choices=["bvh4", "atrc", "bvh4_watertight"],
                    help="acceleration structure: bvh4 (SAH BVH4, production), atrc (Atrc's own BVHs) or "
                         "bvh4_watertight (BVH4 with watertight triangle and quad tests)")
    ap.add_argument("--no-watertight", action="store_true",
                    help="skip the leg that renders the headline scene in the watertight BVH4 mode")
    ap.add_argument("--watertight-frames", type=int, default=3)
    ap.add_argument("--grid", type=int, default=256, help="density grid resolution of the volumetric scene")
    ap.add_argument("--stats-spp", type=int, default=2)
    ap.add_argument("--copy-calibration-mb", type=int, default=256,
                    help="run one plain copy kernel of this size first (traffic-counter reference), 0 = off")
    ap.add_argument("--no-replay", action="store_true", help="skip the traversal-only ray replay")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-native-cpu", action="store_true", help="CPU baseline: skip the -march=native build")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cornell", action="store_true")
    ap.add_argument("--no-volumetric", action="store_true")
    ap.add_argument("--vol-spp", type=int, default=16)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batch-scenes", type=int, default=64, help="scenes of the batched leg (seeds 100..)")
    ap.add_argument("--batch-branches", type=int, default=0,
                    help="graph branches of the batched leg (0: chosen by the library's autotuner)")
    ap.add_argument("--no-guided", action="store_true", help="skip the guided vs plain equal-time leg")
    ap.add_argument("--guided-res", type=int, default=1024)
    ap.add_argument("--guided-tris", type=int, default=500_000)
    ap.add_argument("--guided-seconds", type=float, default=10.0)
    ap.add_argument("--guided-chunk", type=int, default=4)
    ap.add_argument("--guided-ref-factor", type=float, default=8.0)
    ap.add_argument("--guided-train-frac", type=float, default=0.3)
    ap.add_argument("--guided-steps", type=int, default=8)
    ap.add_argument("--guided-batch", type=int, default=65536)
    ap.add_argument("--no-guided-volumetric", action="store_true", help="skip the guided vs plain leg in media")
    ap.add_argument("--vol-guided-seconds", type=float, default=6.0)
    ap.add_argument("--vol-guided-ref-factor", type=float, default=6.0)
    ap.add_argument("--vol-guided-alpha", type=float, default=0.25,
                    help="probability of the learned density in the fog room's mixture")
    ap.add_argument("--vol-guided-medium-depth", type=int, default=0,
                    help="medium scattering points guided up to this path depth (0: surfaces only)")
    ap.add_argument("--skip-headline", action="store_true",
                    help="run only the further legs (profiling); prints no headline line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU (--gpus 1); sharding a render over several GPUs "
                                   "is not implemented", "n_gpus": args.gpus}), flush=True)
        return 2

    from atrc_amd import capi

    ctx = capi.Context(0)
    if args.skip_headline:
        out, desc = {"metric": METRIC, "headline": None}, None
        args.no_cpu_baseline = True
    else:
        out, (desc, keep, cam, tr, W, H) = headline(ctx, args)
    legs = {}
    if desc is not None and not args.no_watertight and args.accel == "bvh4":
        legs["watertight"] = leg_watertight(ctx, args, out, desc, cam, tr, W, H)
    if not args.no_cornell:
        legs["cornell"] = leg_cornell(ctx, args)
    if not args.no_volumetric:
        legs["volumetric"] = leg_volumetric(ctx, args)
    if not args.no_guided:
        legs["guided"] = leg_guided(ctx, args)
    if not args.no_guided_volumetric:
        legs["guided_volumetric"] = leg_guided(ctx, args, volumetric=True)
    ctx.close()
    if not args.no_batched:
        legs["batched"] = leg_batched(args)
    out["legs"] = legs
    out["guided_vs_plain_relmse_ratio"] = legs["guided"]["ratio"] if "guided" in legs else None
    out["cpu_baseline"] = None
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (oracle, FP32) ...")
        out["cpu_baseline"] = cpu_baseline(desc, cam, tr, W, H, args)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
