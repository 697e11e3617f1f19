"""Benchmark of the render path on one MI355X.

Headline workload (C2): the 5M-triangle environment-lit terrain (1582^2 fBm
heightfield, 8 octaves, amplitude 20, + 1000 instanced GGX spheres, 2048x1024
log-normal HDR environment with three 2-degree hot spots), 1920x1080, 32 spp,
8 bounces.  A step renders the full 32-spp frame (66.4M pixel samples) from
frame 1.

Reported `value`: wavefront Mrays/s (extension + shadow rays) over the timed
steps; the megakernel is timed the same way and reported beside it.  Ray
counts come from the device counters of an untimed counting pass (they are
deterministic: both renderers trace the same rays).

`roofline`: the dominant kernel of the wavefront renderer (the closest-hit
traversal) against the HBM peak.  `achieved` = the reference-layout
algorithmic bytes of one launch divided by its average duration (HIP events
on the renderer's own stream, an extra event-timed step).  Reference-layout
bytes count the events of the reference's binary-BVH traversal of the same
rays, from the oracle's counters over the CPU sample: 32 B per node visit,
32 B per triangle test plus 16 B when the second Wald record is read, 48 B per
instance entry (per extension ray, scaled to the step's extension rays).
`achieved_build_layout` counts the renderer's own fetches instead (56 B per
compressed wide-node visit, 48 B per watertight triangle test, 48 B per
instance entry, 72 B of queue and record traffic per ray).  `traffic` and
`limiter` are the HBM bytes per dispatch (FETCH_SIZE + WRITE_SIZE, scaled by a
copy of known size) and the VALU rate, lane activity and wave waiting of the
same kernel, read from
profiles/pmc_traffic_terrain_1080p.md when that profile was measured on the
current kernel sources (matched by a hash of noor_amd/csrc, include and the
build flags), else null.

`cpu_baseline`: the CPU restatement of the same path (the oracle, FP32) on
all host cores over a bounded sample of the same frame (a band of rows, or
the whole frame at several spp, about 15 s of CPU work), in Mrays/s; the
sample's pixels are compared with the GPU's film of the same rows and frames.

`configs`: the other configurations at their stated sizes: C0 (Cornell +
81,920-triangle displaced sphere, 512^2, 16 spp; the CPU baseline over the
whole frame, and the FP64 oracle at 256 spp on a crop with the RMSE), C1 (1M
triangles, Hosek sky, 1920x1080, 64 spp, with a CPU sample), C3 (`batch`: 48
views of the C1 scene as one autotuned wavefront batch against the megakernel
view by view), C4 (`adjoint`: forward and path-replay backward times of the
512^2 one-sample differentiable render, bwd/fwd ratio).
"""
import argparse
import hashlib
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s + spp/s @1080p (wavefront vs megakernel); bwd/fwd time ratio"
HBM_PEAK_GBS = 8000.0
WIDE_NODE_BYTES = 56  # per wide-node visit: origin + steps (16 B), 8-bit child planes (24 B), references (16 B)
PMC_PROFILE = os.path.join(ROOT, "profiles", "pmc_traffic_terrain_1080p.md")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def source_hash():
    """sha256 of the kernel and library sources and the build script: a
    measured profile applies to the binary these produce."""
    h = hashlib.sha256()
    files = []
    for d in ("noor_amd/csrc", "include"):
        for f in sorted(os.listdir(os.path.join(ROOT, d))):
            files.append(os.path.join(d, f))
    files.append("noor_amd/build.py")
    for f in files:
        h.update(f.encode())
        with open(os.path.join(ROOT, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


def measured_profile():
    """The JSON block of the PMC profile if it was measured on these sources."""
    try:
        text = open(PMC_PROFILE).read()
    except OSError:
        return None
    m = re.search(r"```json\n(\{.*?\})\n```", text, re.S)
    if not m:
        return None
    d = json.loads(m.group(1))
    return d if d.get("source_sha256") == source_hash() else None


def traversal_bytes(st, prefix, tri_test):
    """The renderer's own fetches (compressed wide nodes)."""
    nodes = st[prefix + "node_visits"]
    tris = st[prefix + "tri_tests"]
    inst = st[prefix + "instance_entries"]
    second = st[prefix + "tri_second_fetch"]
    tri_bytes = 48 * tris if tri_test == 0 else 32 * tris + 16 * second
    return WIDE_NODE_BYTES * nodes + tri_bytes + 48 * inst


def reference_bytes(cnt, shadow=False):
    """Reference-layout traversal bytes of the oracle's counted events (the
    reference's binary BVH and Wald records)."""
    p = "shd_" if shadow else ""
    return (32 * getattr(cnt, p + "node_visits") + 32 * getattr(cnt, p + "tri_tests") +
            16 * getattr(cnt, p + "tri_second_fetch") + 48 * getattr(cnt, p + "instance_entries"))


def shade_roofline(st, paths, kt, prof):
    """k_shade's algorithmic bytes per launch and its achieved rate.  Per
    shaded path: path state read (class-queue entry 4, L 16, beta and bounce
    16, RNG 16, hit record 32, ray 32; a path's first ray reads no L and
    beta) and written (L, beta, RNG: 40); per continuing ray the next ray and
    its queue entry (36); per shadow task the NEE candidate, shadow ray and
    queue entry (68); per ended path its free-queue entry (4); per hit the
    shading fetch (attributes 16, three vertices, normals and uvs 120,
    object-to-world and normal rows 96: 232) and the material record (128);
    per miss the environment lookup (64)."""
    if "wave_shade" not in kt:
        return None
    # (the totals below are over every path; the last factor keeps k_shade's share)
    ext, shd, hits = st["extension_rays"], st["shadow_rays"], st["extension_hits"]
    cont = ext - paths  # every extension ray after a path's first
    total = (ext * (116 + 40) - paths * 32 + cont * 36 + shd * 68 + paths * 4 + hits * (232 + 128) +
             (ext - hits) * 64)
    # k_shade shades the hits of the rays k_extend traced (the drain the rest)
    total *= (ext - st["drain_extension_rays"]) / max(1, ext)
    launches, ms = kt["wave_shade"]
    per_launch = total / launches
    avg_s = ms / 1e3 / launches
    kp = (prof or {}).get("kernels", {}).get("wave_shade")
    return {"bytes_per_launch": int(per_launch), "avg_launch_ms": round(avg_s * 1e3, 4), "launches": launches,
            "achieved": round(per_launch / avg_s / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(per_launch / avg_s / 1e9 / HBM_PEAK_GBS, 4),
            "bytes_per_path": round(total / max(1, ext), 1),
            "traffic": kp["hbm_bytes_per_dispatch"] if kp else None,
            "traffic_over_algorithmic": round(kp["hbm_bytes_per_dispatch"] / per_launch, 3) if kp else None}


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
    film, cnt = o.render(1, spp, threads=cores, rect=(0, y0, W, y0 + want_rows))
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
    }, film, (y0, y0 + want_rows), spp, cnt


def _count_rays(r, wl, go, spp):
    """Rays of one render (device counters, an untimed counting pass)."""
    wl.options.stats = 1
    r.update_options(wl.options)
    r.reset_stats()
    go(spp=spp, frame=1)
    st = r.stats()
    wl.options.stats = 0
    r.update_options(wl.options)
    return st["extension_rays"], st["shadow_rays"]


def _time_render(r, go, spp, reps):
    go(spp=spp, frame=1)  # warm-up
    r.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        go(spp=spp, frame=1)
    r.synchronize()
    return (time.perf_counter() - t) / reps


def bench_config(wl, name, cpu_seconds, reps, cpu_full=False):
    """One configuration at its stated size: megakernel and wavefront Mrays/s
    and spp/s, their films compared, and the CPU baseline (the whole frame when
    `cpu_full`, else a bounded sample) compared pixel by pixel with the GPU."""
    import numpy as np
    from noor_amd.renderer import Renderer
    from oracle.binding import OracleScene
    r = Renderer.from_workload(wl)
    spp = wl.spp
    ext, shd = _count_rays(r, wl, r.render_wavefront, spp)
    rays = ext + shd
    px = wl.camera.w * wl.camera.h
    row = {"workload": f"{name}: {wl.payload.view.num_triangles} tris, {wl.camera.w}x{wl.camera.h}, {spp} spp, "
                       f"{wl.spec.bounces} bounces", "rays_per_render": {"extension": ext, "shadow": shd}}
    films = {}
    for mode, go in (("megakernel", r.render), ("wavefront", r.render_wavefront)):
        dt = _time_render(r, go, spp, reps)
        films[mode] = r.film()
        row[mode] = {"ms": round(dt * 1e3, 2), "mrays_per_s": round(rays / dt / 1e6, 1),
                     "spp_per_s": round(px * spp / dt, 1)}
    row["wavefront_vs_megakernel"] = round(row["wavefront"]["mrays_per_s"] / row["megakernel"]["mrays_per_s"], 3)
    row["wavefront_equals_megakernel"] = bool(np.array_equal(films["megakernel"], films["wavefront"]))
    if cpu_seconds > 0 or cpu_full:
        cores = os.cpu_count() or 1
        if cpu_full:
            t = time.time()
            cf, cnt = OracleScene.from_workload(wl).render(1, spp, threads=cores)
            dt = time.time() - t
            cr = cnt.extension_rays + cnt.shadow_rays
            row["cpu_baseline"] = {"value": round(cr / dt / 1e6, 3), "unit": "Mrays/s", "cores": cores,
                                   "kind": "port", "sample": f"the whole frame at {spp} spp, FP32 oracle, "
                                                             f"{cr} rays in {dt:.1f} s",
                                   "spp_per_s": round(px * spp / dt, 1)}
            g = films["megakernel"]
        else:
            cb, cf, rows, cspp, _ = cpu_baseline(wl, cpu_seconds)
            row["cpu_baseline"] = cb
            r.render(spp=cspp, frame=1)
            g = r.film()[rows[0]:rows[1]]
        row["cpu_vs_gpu_identical_pixels"] = float(np.mean(np.all(g == cf, axis=-1)))
        row["cpu_vs_gpu_mean_rel_err"] = float(np.mean(np.abs(g[..., :3].astype(np.float64) - cf[..., :3]) /
                                                       (np.abs(cf[..., :3]) + 1e-3)))
    log(f"{name}: megakernel {row['megakernel']['mrays_per_s']} Mrays/s, wavefront {row['wavefront']['mrays_per_s']} "
        f"Mrays/s, films equal {row['wavefront_equals_megakernel']}"
        + (f", cpu {row['cpu_baseline']['value']} Mrays/s" if "cpu_baseline" in row else ""))
    r.close()
    return row


def bench_fp64_rmse(wl, spp=256, rect=(208, 288, 304, 384)):
    """The GPU's FP32 film against the FP64 oracle on the same RNG streams, at
    high spp on a crop: RMSE, relative RMSE and mean relative error."""
    import numpy as np
    from noor_amd.renderer import Renderer
    from oracle.binding import OracleScene
    r = Renderer.from_workload(wl)
    r.render(spp=spp, frame=1)
    g = r.film()[..., :3].astype(np.float64)
    r.close()
    t = time.time()
    f64, _ = OracleScene.from_workload(wl).render(1, spp, rect=rect, fp64=True)
    dt = time.time() - t
    x0, y0, x1, y1 = rect
    crop, ref = g[y0:y1, x0:x1], f64[..., :3]
    rmse = float(np.sqrt(np.mean((crop - ref) ** 2)))
    out = {"spp": spp, "crop": list(rect), "rmse": rmse, "relative_rmse": rmse / float(np.sqrt(np.mean(ref ** 2))),
           "mean_rel_err": float(np.mean(np.abs(crop - ref) / (np.abs(ref) + 1e-3))), "fp64_seconds": round(dt, 1)}
    log(f"FP64 oracle: RMSE {rmse:.3e} (relative {out['relative_rmse']:.3e}) at {spp} spp on {rect}")
    return out


def bench_batch(wl=None, log_prefix="batch"):
    """48 views of the 1M-triangle procedural scene (Hosek sky), resolution in
    {256^2, 512^2, 1280x720, 1920x1080} and spp in {1, 4, 16, 64}: one
    wavefront batch (after autotuning on a one-sample copy) against rendering
    the views one by one with the megakernel."""
    from noor_amd import scenes
    from noor_amd.renderer import Renderer
    if wl is None:
        wl = scenes.procedural(width=256, height=256)
    views = scenes.batch_views(48)
    samples = sum(c.w * c.h * spp for c, _, spp, _ in views)
    r = Renderer.from_workload(wl)
    t = time.perf_counter()
    tune = r.autotune([(c, 1, 1, sd) for c, _, _, sd in views])
    t_tune = time.perf_counter() - t
    best = min(tune, key=lambda x: x[3])
    log(f"{log_prefix}: autotune {len(tune)} settings in {t_tune:.1f}s, best capacity {best[0]} block {best[1]} "
        f"lds {best[2]} ({best[3]:.1f} ms for the one-sample batch)")
    # rays of the batch (counting pass, untimed)
    wl.options.stats = 1
    r.update_options(wl.options)
    r.reset_stats()
    r.render_batch(views)
    st = r.stats()
    wl.options.stats = 0
    r.update_options(wl.options)
    rays = st["extension_rays"] + st["shadow_rays"]
    r.render_batch(views)  # warm-up
    r.synchronize()
    t = time.perf_counter()
    r.render_batch(views)
    r.synchronize()
    t_batch = time.perf_counter() - t
    log(f"{log_prefix}: {samples} samples, {rays} rays in {t_batch:.3f}s (wavefront batch)")
    t_mega = 0.0
    for c, f, spp, _ in views:
        r.update_camera(c)
        r.render(spp=1, frame=1)  # allocation and warm-up of this size
        r.synchronize()
        t = time.perf_counter()
        r.render(spp=spp, frame=1)
        r.synchronize()
        t_mega += time.perf_counter() - t
    log(f"{log_prefix}: megakernel view by view {t_mega:.3f}s")
    r.close()
    return {"views": len(views), "samples": samples, "rays": rays,
            "wavefront_batch_s": round(t_batch, 4), "megakernel_per_view_s": round(t_mega, 4),
            "wavefront_mrays_per_s": round(rays / t_batch / 1e6, 2),
            "megakernel_mrays_per_s": round(rays / t_mega / 1e6, 2),
            "batch_vs_per_view": round(t_mega / t_batch, 3),
            "autotune": {"settings": len(tune), "seconds": round(t_tune, 1),
                         "best": {"queue_capacity": best[0], "block": best[1], "stack_lds": best[2]}}}


def bench_textured(width=1920, height=1080, spp=8, tex=512):
    """The image-texture pipeline at 1080p: the textured scene (mip-mapped
    colour, roughness and emittance textures at the ray-differential
    footprint, bump and alpha maps) against its untextured twin (every image
    replaced by its mean), megakernel and wavefront, with the films of the two
    renderers compared."""
    import numpy as np
    from noor_amd import scenes
    from noor_amd.renderer import Renderer
    res = {}
    for name, flat in (("textured", False), ("flat", True)):
        wl = scenes.textured(width=width, height=height, tex=tex, flat=flat)
        r = Renderer.from_workload(wl)
        wl.options.stats = 1
        r.update_options(wl.options)
        r.reset_stats()
        r.render(spp=spp, frame=1)
        st = r.stats()
        rays = st["extension_rays"] + st["shadow_rays"]
        wl.options.stats = 0
        r.update_options(wl.options)
        row = {"rays": rays}
        films = {}
        for mode in ("megakernel", "wavefront"):
            go = r.render if mode == "megakernel" else r.render_wavefront
            go(spp=spp, frame=1)  # warm-up
            r.synchronize()
            t = time.perf_counter()
            go(spp=spp, frame=1)
            r.synchronize()
            dt = time.perf_counter() - t
            films[mode] = r.film()
            row[mode] = {"ms": round(dt * 1e3, 2), "mrays_per_s": round(rays / dt / 1e6, 1)}
        row["wavefront_equals_megakernel"] = bool(np.array_equal(films["megakernel"], films["wavefront"]))
        r.set_timing(True)
        r.reset_kernel_times()
        r.render_wavefront(spp=spp, frame=1)
        r.synchronize()
        kt = r.kernel_times()
        r.set_timing(False)
        row["wave_kernel_ms"] = {k: round(v[1], 3) for k, v in kt.items() if k.startswith("wave_")}
        res[name] = row
        log(f"textured leg ({name}): {rays} rays, megakernel {row['megakernel']['mrays_per_s']} Mrays/s, "
            f"wavefront {row['wavefront']['mrays_per_s']} Mrays/s, films equal {row['wavefront_equals_megakernel']}")
        r.close()
    res["workload"] = (f"textured open Cornell scene, {width}x{height}, {spp} spp, 6 bounces, {tex}^2 textures; "
                       f"flat: the same scene with every image replaced by its mean")
    res["wavefront_cost_of_textures"] = round(res["flat"]["wavefront"]["mrays_per_s"] /
                                              res["textured"]["wavefront"]["mrays_per_s"], 3)
    return res


def bench_adjoint(passes=10):
    """Forward (one-sample render) and path-replay backward on the 512^2,
    depth-4 Cornell scene with 64 glossy materials and a 512x256 environment."""
    import ctypes as C
    import torch
    from noor_amd import scenes
    from noor_amd._lib import lib
    from noor_amd.renderer import Renderer
    wl = scenes.adjoint_scene()
    M = wl.info["n_materials"]
    eh, ew = wl.info["env"].shape[:2]
    dev = torch.device("cuda:0")
    r = Renderer.from_workload(wl)
    alb = torch.full((M, 3), 0.5, device=dev)
    rough = torch.full((M,), 0.5, device=dev)
    env = torch.full((eh, ew, 3), 0.5, device=dev)
    img = torch.empty((wl.camera.h, wl.camera.w, 4), device=dev)
    grad = torch.ones_like(img) * 1e-3
    ga, gr, ge = torch.zeros_like(alb), torch.zeros_like(rough), torch.zeros_like(env)
    torch.cuda.synchronize()
    P = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
    L = lib()
    assert L.noor_set_params_device(r._h, P(alb), P(rough), M, P(env), ew, eh) == 0
    for s in range(1, 3):
        assert L.noor_render_sample(r._h, s, P(img)) == 0
        assert L.noor_adjoint(r._h, s, P(img), P(grad), P(ga), P(gr), M, P(ge)) == 0
    r.set_timing(True)
    r.reset_kernel_times()
    for s in range(3, 3 + passes):
        assert L.noor_render_sample(r._h, s, P(img)) == 0
        assert L.noor_adjoint(r._h, s, P(img), P(grad), P(ga), P(gr), M, P(ge)) == 0
    kt = r.kernel_times()
    r.set_timing(False)
    fwd = kt["megakernel"][1] / kt["megakernel"][0]
    bwd = kt["adjoint"][1] / kt["adjoint"][0]
    log(f"adjoint: forward {fwd:.3f} ms, backward {bwd:.3f} ms per one-sample pass ({passes} passes)")
    r.close()
    return {"workload": f"adjoint cornell {wl.payload.view.num_triangles} tris, {wl.camera.w}x{wl.camera.h}, "
                        f"1 spp, {wl.spec.bounces} bounces, {M} materials, env {ew}x{eh}",
            "fwd_ms": round(fwd, 4), "bwd_ms": round(bwd, 4), "bwd_over_fwd": round(bwd / fwd, 3)}


def pmc_probe():
    """Short run for hardware counter passes (rocprofv3 --pmc): one wavefront
    step of the headline workload, then a 1 GiB device-to-device copy whose
    counter readings calibrate the byte counters (bytes moved are known)."""
    import torch  # device memory for the calibration copy
    from noor_amd import scenes
    from noor_amd.renderer import Renderer
    wl = scenes.terrain()
    r = Renderer.from_workload(wl)
    r.render_wavefront(spp=wl.spp, frame=1)
    r.synchronize()
    r.close()
    n = 1 << 28  # floats: 1 GiB
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    log(f"pmc probe: one wavefront step, then a copy of {n * 4} bytes")


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
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-textured", action="store_true")
    ap.add_argument("--no-configs", action="store_true", help="skip the C0, C1 and FP64 legs")
    ap.add_argument("--tri", default="watertight", choices=["watertight", "wald"],
                    help="triangle test (wald: the reference's projection test)")
    ap.add_argument("--textured-only", action="store_true", help="run only the texture-pipeline leg (profiling)")
    ap.add_argument("--no-build-compare", action="store_true")
    ap.add_argument("--pmc-probe", action="store_true", help="short run for rocprofv3 --pmc counter passes")
    args = ap.parse_args()
    if args.pmc_probe:
        pmc_probe()
        return
    if args.textured_only:
        print(json.dumps({"textured": bench_textured()}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: the renderer has no multi-GPU path "
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
    if args.tri == "wald":
        wl.options.tri_test = 1
    tri_test = wl.options.tri_test
    log(f"scene {args.workload}: {wl.payload.view.num_triangles} triangles, BVH {wl.payload.stats.num_nodes} nodes, "
        f"build {wl.payload.stats.build_seconds:.1f}s, total setup {time.time() - t0:.1f}s")
    build = None
    if args.workload in ("terrain", "procedural") and not args.no_build_compare:
        # the same payload built with the bottom-level trees on the GPU, twice
        # (the second reuses the builder's device memory); untimed by the steps
        import numpy as np
        times, same = [], True
        for _ in range(2):
            wd = fn(build_device=0, **kw)
            times.append(wd.payload.stats.build_seconds)
            a, b = wl.payload.arrays(), wd.payload.arrays()
            same = same and all(np.array_equal(a[k], b[k]) for k in a)
            del wd
        build = {"triangles": wl.payload.view.num_triangles, "host_ms": round(wl.payload.stats.build_seconds * 1e3, 1),
                 "device_ms_first": round(times[0] * 1e3, 1), "device_ms_repeat": round(times[1] * 1e3, 1),
                 "identical_payload": bool(same)}
        log(f"scene build: host {build['host_ms']} ms, with device trees {build['device_ms_first']} ms "
            f"(repeat {build['device_ms_repeat']} ms), identical payload {same}")

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
    # the CPU baseline (untimed by the steps): also the oracle's count of the
    # reference's traversal events over the same rays
    cpu = None
    if not args.no_cpu:
        wl.options.stats = 0
        cb, cpu_film, rows, cpu_spp, cpu_cnt = cpu_baseline(wl, args.cpu_seconds)
        cpu = (cb, cpu_film, rows, cpu_spp, cpu_cnt)
    # per ray beside the traversal: extension: queue entry 4, ray 32, hit record
    # 32, material class 4; shadow: queue entry 4, ray 32, NEE candidate 32,
    # radiance read and write 32
    # the drain kernel traces the last paths of a group: the persistent
    # kernels' share of the per-ray events is the rays they traced themselves
    ext_k = ext_rays - st["drain_extension_rays"]
    shd_k = shd_rays - st["drain_shadow_rays"]
    build_bytes = {
        "wave_extend": (traversal_bytes(st, "", tri_test) / max(1, ext_rays) + (4 + 32 + 32 + 4)) * ext_k,
        "wave_shadow": (traversal_bytes(st, "shadow_", tri_test) / max(1, shd_rays) + (4 + 32 + 32 + 32)) * shd_k,
    }
    dom = max((k for k in kt if k in build_bytes), key=lambda k: kt[k][1])
    launches, ms = kt[dom]
    avg_s = ms / 1e3 / launches
    build_per_launch = build_bytes[dom] / launches
    ref_per_launch = None
    if cpu is not None:
        c = cpu[4]
        if dom == "wave_extend":
            ref_per_launch = reference_bytes(c) / max(1, c.extension_rays) * ext_k / launches
        else:
            ref_per_launch = reference_bytes(c, shadow=True) / max(1, c.shadow_rays) * shd_k / launches
    per_launch = ref_per_launch if ref_per_launch is not None else build_per_launch
    achieved = per_launch / avg_s / 1e9
    prof = measured_profile() if (args.queue == 0 and args.block == 64 and args.workload == "terrain") else None
    kprof = prof.get("kernels", {}).get(dom) if prof else None
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": kprof["hbm_bytes_per_dispatch"] if kprof else None,
                "bytes_per_launch": int(per_launch),
                "bytes_layout": "reference (binary BVH, Wald records; oracle-counted events)"
                if ref_per_launch is not None else "the renderer's wide BVH (no CPU sample to count the reference's)",
                "achieved_build_layout": round(build_per_launch / avg_s / 1e9, 1),
                "bytes_per_launch_build_layout": int(build_per_launch),
                "avg_launch_ms": round(avg_s * 1e3, 4), "launches": launches,
                "rays_per_launch": int((ext_k if dom == "wave_extend" else shd_k) / launches),
                "limiter": ({"valu_rate": kprof["valu_rate"], "lanes_active": kprof["lanes_active"],
                             "wave_wait_fraction": kprof["wave_wait_fraction"],
                             "what": f"latency of the traversal's dependent fetches with "
                                     f"{100 * kprof['lanes_active']:.0f}% of the lanes active "
                                     "(hardware counters, see traffic source)"}
                            if kprof else None),
                "traffic_source": (os.path.relpath(PMC_PROFILE, ROOT) if kprof else None)}
    if ref_per_launch is not None:
        roofline["note"] = (f"reference-layout bytes: the events the reference's binary BVH traversal of the same rays "
                            f"would fetch; the renderer's four-wide BVH fetches {build_per_launch / ref_per_launch:.2f} "
                            f"of them, so the reference-layout rate can exceed the HBM peak")
    shade = shade_roofline(st, pixels * spp, kt, prof)
    mega_roof = None
    if "megakernel" in kt:
        mb = traversal_bytes(st, "", tri_test) + traversal_bytes(st, "shadow_", tri_test)
        ms_m = kt["megakernel"][1] / kt["megakernel"][0]
        mega_roof = {"achieved_build_layout": round(mb / (ms_m / 1e3) / 1e9, 1), "avg_launch_ms": round(ms_m, 3),
                     "bytes_per_launch_build_layout": int(mb)}
        if cpu is not None:
            c = cpu[4]
            rb = (reference_bytes(c) / max(1, c.extension_rays) * ext_rays +
                  reference_bytes(c, shadow=True) / max(1, c.shadow_rays) * shd_rays)
            mega_roof["achieved"] = round(rb / (ms_m / 1e3) / 1e9, 1)
            mega_roof["bytes_per_launch"] = int(rb)

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
        "shade_roofline": shade,
        "events_per_ray": {
            "extension": {k: round(st[k] / max(1, ext_rays), 3)
                          for k in ("node_visits", "tri_tests", "tri_second_fetch", "instance_entries")},
            "shadow": {k: round(st["shadow_" + k] / max(1, shd_rays), 3)
                       for k in ("node_visits", "tri_tests", "tri_second_fetch", "instance_entries")},
            "extension_hit_fraction": round(st["extension_hits"] / max(1, ext_rays), 4),
            "drained_fraction": {"extension": round(st["drain_extension_rays"] / max(1, ext_rays), 4),
                                 "shadow": round(st["drain_shadow_rays"] / max(1, shd_rays), 4)},
            "layout": "the renderer's four-wide BVH (device counters of the counting pass)"},
    }
    if t_mega is not None:
        mega_mrays = rays_per_step * args.steps / t_mega / 1e6
        out["megakernel"] = {"mrays_per_s": round(mega_mrays, 2), "ms_per_step": round(t_mega / args.steps * 1e3, 2),
                             "spp_per_s": round(pixels * spp * args.steps / t_mega, 1), "roofline": mega_roof}
        out["wavefront_vs_megakernel"] = round(wave_mrays / mega_mrays, 3)
        diff = np.abs(wave_film[..., :3].astype(np.float64) - mega_film[..., :3])
        out["image_mean_rel_err_wave_vs_mega"] = float(np.mean(diff / (np.abs(mega_film[..., :3]) + 1e-3)))
    if cpu is not None:
        cb, cpu_film, rows, cpu_spp, cpu_cnt = cpu
        out["cpu_baseline"] = cb
        out["reference_events_per_ray"] = {
            "extension": {k: round(getattr(cpu_cnt, k) / max(1, cpu_cnt.extension_rays), 3)
                          for k in ("node_visits", "tri_tests", "tri_second_fetch", "instance_entries")},
            "shadow": {k: round(getattr(cpu_cnt, "shd_" + k) / max(1, cpu_cnt.shadow_rays), 3)
                       for k in ("node_visits", "tri_tests", "tri_second_fetch", "instance_entries")}}
        # the CPU sample against the GPU film of the same rows and samples
        if cpu_spp != spp:
            r.render(spp=cpu_spp, frame=1)
            ref = r.film()
        else:
            ref = wave_film
        g = ref[rows[0]:rows[1]]
        out["cpu_vs_gpu_identical_pixels"] = float(np.mean(np.all(g == cpu_film, axis=-1)))
        out["cpu_vs_gpu_mean_rel_err"] = float(np.mean(np.abs(g[..., :3].astype(np.float64) - cpu_film[..., :3]) /
                                                       (np.abs(cpu_film[..., :3]) + 1e-3)))
    else:
        out["cpu_baseline"] = None
    r.close()
    del r
    out["scene_build"] = build
    c1 = None
    if not args.no_configs:
        from noor_amd import scenes as sc
        cfg = {}
        c0 = sc.cornell()
        cfg["C0"] = bench_config(c0, "C0 Cornell + displaced sphere", 0, 5, cpu_full=not args.no_cpu)
        cfg["C0"]["fp64_oracle"] = bench_fp64_rmse(c0)
        del c0
        c1 = sc.procedural()
        cfg["C1"] = bench_config(c1, "C1 procedural, Hosek sky", 0 if args.no_cpu else 10.0, 2)
        out["configs"] = cfg
    if not args.no_batch:
        out["batch"] = bench_batch(c1)
    del c1
    if not args.no_adjoint:
        out["adjoint"] = bench_adjoint()
    if not args.no_textured:
        out["textured"] = bench_textured()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
