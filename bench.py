"""Benchmark of the render path on one MI355X.

Workload: the 5M-triangle environment-lit terrain (1582^2 fBm heightfield,
8 octaves, amplitude 20, + 1000 instanced GGX spheres, 2048x1024 log-normal
HDR environment with three 2-degree hot spots), 1920x1080, 32 spp, 8 bounces.
A step renders the full 32-spp frame (66.4M pixel samples) from frame 1.

Reported `value`: wavefront Mrays/s (extension + shadow rays) over the timed
steps; the megakernel is timed the same way and reported beside it.  Ray
counts come from the device counters of an untimed counting pass (they are
deterministic: both renderers trace the same rays).

`roofline`: the dominant kernel of the wavefront renderer, with its
algorithmic bytes per launch (112 B per wide BVH node visit: four child boxes
and references, 48 B per watertight triangle test, 48 B per instance entry,
plus its per-ray queue and record traffic) divided by its average launch duration, measured with HIP events on
the renderer's own stream during an extra, event-timed step.

`cpu_baseline`: the CPU restatement of the same path (the oracle, FP32) on
all host cores over a bounded sample of the same frame (1 spp over a band of
rows), in Mrays/s.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s + spp/s @1080p (wavefront vs megakernel); bwd/fwd time ratio"
HBM_PEAK_GBS = 8000.0
WIDE_NODE_BYTES = 112  # four child boxes (96 B) + four child references (16 B) per wide-node visit


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def traversal_bytes(st, prefix, tri_test):
    nodes = st[prefix + "node_visits"]
    tris = st[prefix + "tri_tests"]
    inst = st[prefix + "instance_entries"]
    second = st[prefix + "tri_second_fetch"]
    tri_bytes = 48 * tris if tri_test == 0 else 32 * tris + 16 * second
    return WIDE_NODE_BYTES * nodes + tri_bytes + 48 * inst


def cpu_baseline(wl, target_s):
    from oracle.binding import OracleScene
    o = OracleScene.from_workload(wl)
    cores = os.cpu_count() or 1
    W, H = wl.camera.w, wl.camera.h
    rows = max(1, H // 32)
    t = time.time()
    o.render(1, 1, threads=cores, rect=(0, 0, W, rows))
    dt = max(time.time() - t, 1e-6)
    # the sample: a band of rows at 1 spp, or the whole frame at several spp
    want = target_s / dt * rows  # rows-at-1-spp that fit the time budget
    if want <= H:
        want_rows, spp = max(rows, int(want)), 1
    else:
        want_rows, spp = H, max(1, min(wl.spp, int(want / H)))
    y0 = (H - want_rows) // 2
    t = time.time()
    _, cnt = o.render(1, spp, threads=cores, rect=(0, y0, W, y0 + want_rows))
    dt = time.time() - t
    rays = cnt.extension_rays + cnt.shadow_rays
    log(f"cpu baseline: {want_rows} rows x {W} px x {spp} spp, {rays} rays in {dt:.1f}s on {cores} threads")
    return {
        "value": round(rays / dt / 1e6, 3),
        "unit": "Mrays/s",
        "cores": cores,
        "kind": "port",
        "sample": f"rows {y0}-{y0 + want_rows} of the {W}x{H} frame at {spp} spp (frames 1-{spp}), "
                  f"FP32 oracle, {rays} rays in {dt:.1f} s",
        "spp_per_s": round(want_rows * W * spp / dt, 1),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="terrain", choices=["terrain", "procedural", "cornell"])
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--skip-mega", action="store_true")
    ap.add_argument("--queue", type=int, default=0)
    ap.add_argument("--block", type=int, default=64)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(pixels shard trivially; see the sharding design); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from noor_amd import scenes
    from noor_amd.renderer import Renderer

    t0 = time.time()
    fn = getattr(scenes, args.workload)
    kw = dict(width=args.width, height=args.height)
    if args.spp:
        kw["spp"] = args.spp
    wl = fn(**kw)
    spp = wl.spp
    tri_test = wl.options.tri_test
    log(f"scene {args.workload}: {wl.payload.view.num_triangles} triangles, BVH {wl.payload.stats.num_nodes} nodes, "
        f"build {wl.payload.stats.build_seconds:.1f}s, total setup {time.time() - t0:.1f}s")

    r = Renderer.from_workload(wl)
    r.set_wavefront_params(args.queue, args.block)
    pixels = wl.camera.w * wl.camera.h

    # counting pass (untimed): rays and traversal events of one step
    wl.options.stats = 1
    r.update_options(wl.options)
    r.reset_stats()
    r.render_wavefront(spp=spp, frame=1)
    st = r.stats()
    wl.options.stats = 0
    r.update_options(wl.options)
    rays_per_step = st["extension_rays"] + st["shadow_rays"]
    log(f"counting pass: {st['extension_rays']} extension + {st['shadow_rays']} shadow rays per step, "
        f"{st['node_visits'] / max(1, st['extension_rays']):.1f} nodes/ext ray")

    import numpy as np

    def timed(render, steps):
        r.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            render(spp=spp, frame=1)
        r.synchronize()
        return time.perf_counter() - t

    for _ in range(args.warmup):
        r.render_wavefront(spp=spp, frame=1)
    log("wavefront warm-up done")
    t_wave = timed(r.render_wavefront, args.steps)
    wave_film = r.film()
    log(f"wavefront: {args.steps} steps in {t_wave:.3f}s")

    t_mega = None
    mega_film = None
    if not args.skip_mega:
        for _ in range(args.warmup):
            r.render(spp=spp, frame=1)
        t_mega = timed(r.render, args.steps)
        mega_film = r.film()
        log(f"megakernel: {args.steps} steps in {t_mega:.3f}s")

    # event-timed step for per-kernel durations (the roofline)
    r.set_timing(True)
    r.reset_kernel_times()
    r.render_wavefront(spp=spp, frame=1)
    if not args.skip_mega:
        r.render(spp=spp, frame=1)
    kt = r.kernel_times()
    r.set_timing(False)
    log("kernel times: " + ", ".join(f"{k} {v[0]}x {v[1]:.2f}ms" for k, v in kt.items()))

    ext_rays, shd_rays = st["extension_rays"], st["shadow_rays"]
    bytes_kernel = {
        "wave_extend": traversal_bytes(st, "", tri_test) + ext_rays * (4 + 32 + 24 + 4 + 16),
        "wave_shadow": traversal_bytes(st, "shadow_", tri_test) + shd_rays * (4 + 32 + 32 + 32),
    }
    dom = max((k for k in kt if k in bytes_kernel), key=lambda k: kt[k][1])
    launches, ms = kt[dom]
    avg_s = ms / 1e3 / launches
    per_launch = bytes_kernel[dom] / launches
    achieved = per_launch / avg_s / 1e9
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "bytes_per_launch": int(per_launch), "avg_launch_ms": round(avg_s * 1e3, 4), "launches": launches}
    mega_roof = None
    if "megakernel" in kt:
        mb = traversal_bytes(st, "", tri_test) + traversal_bytes(st, "shadow_", tri_test)
        ms_m = kt["megakernel"][1] / kt["megakernel"][0]
        mega_roof = {"achieved": round(mb / (ms_m / 1e3) / 1e9, 1), "avg_launch_ms": round(ms_m, 3),
                     "bytes_per_launch": int(mb)}

    wave_mrays = rays_per_step * args.steps / t_wave / 1e6
    out = {
        "metric": METRIC,
        "value": round(wave_mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_wave / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene and environment map)",
        "config": {"workload": f"{args.workload} {wl.payload.view.num_triangles} tris, "
                               f"{wl.camera.w}x{wl.camera.h}, {spp} spp, {wl.spec.bounces} bounces",
                   "renderer": "wavefront (material-sorted queues, persistent traversal, hipGraph)",
                   "triangles": wl.payload.view.num_triangles, "width": wl.camera.w, "height": wl.camera.h,
                   "spp": spp, "bounces": wl.spec.bounces},
        "spp_per_s": round(pixels * spp * args.steps / t_wave, 1),
        "rays_per_step": {"extension": ext_rays, "shadow": shd_rays},
        "wavefront": {"mrays_per_s": round(wave_mrays, 2), "ms_per_step": round(t_wave / args.steps * 1e3, 2)},
        "roofline": roofline,
        "kernel_ms": {k: round(v[1] / v[0], 4) for k, v in kt.items()},
    }
    if t_mega is not None:
        mega_mrays = rays_per_step * args.steps / t_mega / 1e6
        out["megakernel"] = {"mrays_per_s": round(mega_mrays, 2), "ms_per_step": round(t_mega / args.steps * 1e3, 2),
                             "spp_per_s": round(pixels * spp * args.steps / t_mega, 1), "roofline": mega_roof}
        out["wavefront_vs_megakernel"] = round(wave_mrays / mega_mrays, 3)
        diff = np.abs(wave_film[..., :3].astype(np.float64) - mega_film[..., :3])
        out["image_mean_rel_err_wave_vs_mega"] = float(np.mean(diff / (np.abs(mega_film[..., :3]) + 1e-3)))
    if not args.no_cpu:
        wl.options.stats = 0
        out["cpu_baseline"] = cpu_baseline(wl, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    r.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
