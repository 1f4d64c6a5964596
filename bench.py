"""Benchmark of the plain HIP path on one MI355X.

Workload (default): one 1920x1080 frame of a 1,000,024-triangle scene (400
instanced icospheres / subdivided cubes, 4 point lights, 2048^2 atlas with its
mip chain), rendered prearrange -> depth -> id -> shade.  A step is one frame
with the scene, atlas and lights already resident in HBM.

Prints one JSON line: frames/s (value), Mpix/s, per-pass device times from
HIP events on the render stream, the shading kernel's roofline against
MI355X's 8 TB/s HBM, and the oracle timed on this host's cores.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
UINT_MAX = 0xFFFFFFFF


def hip_sync():
    lib = C.CDLL("libamdhip64.so")
    rc = lib.hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed ({rc})")


def traffic_from_profiles(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE, WRITE_SIZE scaled by the copy-kernel factor)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        rec = json.load(open(path))
        ent = rec[workload][kernel]
        return float(ent["bytes_per_launch"])
    except Exception:
        return None


def shade_algorithmic_bytes(W, H, V, O, x_tex, x_shadow, L):
    """Minimal field bytes of one shading launch: 20 B per pixel (depth 4, id
    4, next-depth clear 4, RGBA8 4, normal 4), 132 B per distinct visible
    triangle (112 B of vertex fields + 20 B fragment record), 52 B per
    distinct visible object, 4 B per distinct atlas / cube-map texel and 56 B
    per light."""
    return 20 * W * H + 132 * V + 52 * O + 4 * x_tex + 4 * x_shadow + 56 * L


def cpu_baseline(sc, view, gpu_out, seconds):
    from oracle.binding import Oracle
    cores = min(os.cpu_count() or 1, 16)
    o = Oracle()
    atlas = o.atlas(sc.textures)
    maps = o.shadow_maps(sc, cores) if np.any(sc.lights["shadow"]) else None
    times = []
    ref = None
    t_end = time.perf_counter() + seconds
    while True:
        t0 = time.perf_counter()
        r = o.render(sc, view, atlas=atlas, maps=maps, nthreads=cores)
        times.append(time.perf_counter() - t0)
        ref = ref or r
        if time.perf_counter() > t_end or len(times) >= 8:
            break
    parity = None
    if gpu_out is not None:
        parity = bool(np.array_equal(gpu_out["depth"], ref["depth"]) and
                      np.array_equal(gpu_out["tri_id"], ref["tri_id"]) and
                      int(np.abs(gpu_out["rgba"].astype(int) - ref["rgba"].astype(int)).max()) <= 1)
    med = float(np.median(times))
    return {"value": 1.0 / med, "unit": "frames/s", "cores": cores, "kind": "port",
            "sample": f"{len(times)} full {view['width'][0]}x{view['height'][0]} frames of the same scene, "
                      f"oracle on {cores} OpenMP threads, median {med * 1e3:.0f} ms/frame",
            "ms_per_frame": med * 1e3}, parity


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", type=int, default=2, choices=[1, 2])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--calib", action="store_true", help="also run the copy kernel used to scale PMC counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is design-only)", "n_gpus": args.gpus}))
        sys.exit(2)

    from openclrenderer_amd import runtime, scenes
    if args.config == 1:
        sc, view = scenes.config1()
        workload = "800x600 2480 tris (icosphere + 10x10 cubes), 1 light, 256^2 texture"
    else:
        sc, view = scenes.config2()
        workload = "1920x1080 1M tris (400 instances), 4 point lights, 2048^2 atlas"
    W, H = int(view["width"][0]), int(view["height"][0])

    ctx = runtime.Context(0)
    g = runtime.GpuScene(ctx, sc)
    fr = runtime.Frame(ctx, W, H)
    out = fr.render_read(g, view)          # sizes the work lists, reads the frame once
    for _ in range(args.warmup):
        fr.render(g, view)
    hip_sync()
    fr.set_profiling(True)
    hip_sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        fr.render(g, view)
    hip_sync()
    t1 = time.perf_counter()
    st = fr.stats()
    if st["overflow"]:
        raise RuntimeError("work list overflow inside the timed region")
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    fps = 1e3 / ms_per_step

    # algorithmic bytes of the shading launch
    tri = out["tri_id"]
    vis = tri[tri != UINT_MAX]
    V = int(np.unique(vis).size)
    O = int(np.unique(sc.tris["v"]["object_id"][vis, 0]).size) if V else 0
    x_tex, x_shadow = fr.count_texels(g, view)
    B = shade_algorithmic_bytes(W, H, V, O, x_tex, x_shadow, int(sc.lights.shape[0]))
    achieved = B / (st["ms_shade"] * 1e-3) / 1e9
    traffic = traffic_from_profiles("k_shade", workload)

    if args.calib:
        cal = C.CDLL(os.path.join(ROOT, "openclrenderer_amd", "libocr_calib.so"))
        cal.ocr_calib_copy.argtypes = [C.c_size_t, C.c_int]
        if cal.ocr_calib_copy(256 << 20, 3) != 0:
            raise RuntimeError("calibration copy failed")

    cpu, parity = (None, None) if args.no_cpu else cpu_baseline(sc, view, out, args.cpu_seconds)
    line = {
        "metric": "frames/s + Mpix/s at 1080p/1M tris; p50/p99 latency on mixed-size request stream",
        "value": fps,
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": workload, "width": W, "height": H, "triangles": sc.ntri,
                   "lights": int(sc.lights.shape[0]), "path": "plain"},
        "mpix_per_s": W * H / (ms_per_step * 1e-3) / 1e6,
        "passes_ms": {k: st[k] for k in ("ms_clear", "ms_prearrange", "ms_depth", "ms_id", "ms_shade", "ms_total")},
        "work": {"cut_tris": st["n_cut"], "fragments": st["n_frag"], "visible_tris": V, "visible_objs": O,
                 "atlas_texels": x_tex, "shadow_texels": x_shadow},
        "roofline": {"bound": "hbm", "kernel": "k_shade", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes": B},
        "cpu_baseline": cpu,
        "parity_vs_oracle": parity,
    }
    if cpu:
        line["speedup_vs_cpu"] = fps / cpu["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
