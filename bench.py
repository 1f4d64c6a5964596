"""Benchmark of the HIP path on one MI355X.

Workload (default, --config 2): one 1920x1080 frame of a 1,000,024-triangle
scene (400 instanced icospheres / subdivided cubes, 4 point lights, 2048^2
atlas with its mip chain), rendered prearrange -> depth -> id -> shade.  A
step is one frame with the scene, atlas and lights already resident in HBM.

Prints one JSON line: frames/s (value), Mpix/s, per-pass device times from
HIP events on the render stream, the shading kernel's roofline against
MI355X's 8 TB/s HBM and against the VALU issue rate measured on the device,
the raster passes' depth tests per second against the measured atomicMin
rate, the oracle timed on this host's cores, and two bounded legs of the
metric's second half: the batched form (64 views over 8 scenes, batched vs
sequential) and the served form (Poisson streams at 1000 and 2000 req/s,
p50/p99 latency, batched vs one request at a time).

Other workloads (one JSON line each):
  --config 1  800x600, 2,480 triangles, 1 light
  --config 3  3840x2160, 4M triangles, 16 shadow-casting lights: a step is
              the shadow pass (16 cube maps of 6 x 1024^2) plus the frame
  --config 4  64 views of mixed size over 8 resident scenes (50k-2M
              triangles): one batched render per step vs 64 plain renders
  --config 5  served requests (Poisson arrivals at several rates) against 12
              resident scenes: throughput and p50/p99 latency, batched vs one
              request at a time

python bench.py [--gpus 1] [--steps K] [--warmup W] [--config C] [--no-cpu] [--plain-only]
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
UINT_MAX = 0xFFFFFFFF


_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        _hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        _hip.hipEventSynchronize.argtypes = [C.c_void_p]
    return _hip


def hip_sync():
    rc = hip().hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed ({rc})")


class EventPair:
    """Two HIP events on the null stream (the stream the passes run on)."""

    def __init__(self):
        self.a, self.b = C.c_void_p(), C.c_void_p()
        for e in (self.a, self.b):
            if hip().hipEventCreate(C.byref(e)) != 0:
                raise RuntimeError("hipEventCreate failed")

    def start(self):
        hip().hipEventRecord(self.a, None)

    def stop(self):
        hip().hipEventRecord(self.b, None)

    def ms(self):
        hip().hipEventSynchronize(self.b)
        t = C.c_float()
        hip().hipEventElapsedTime(C.byref(t), self.a, self.b)
        return t.value


def profile_entry(kernel, workload):
    """The committed rocprofv3 --pmc summary of `kernel` on `workload`: HBM
    bytes per launch (FETCH_SIZE, WRITE_SIZE scaled by the copy-kernel
    factor) and VALU wave-instructions per launch (SQ_INSTS_VALU)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        return json.load(open(path))[workload][kernel]
    except Exception:
        return {}


def calibration():
    """Measured ceilings on this GPU (calib.hip): the VALU issue rate of
    independent binary32 and binary64 FMA chains at 8 wavefronts per SIMD,
    and the u32 atomicMin rate on a 1080p-sized buffer, uniform random and
    64-address clusters (the depth pass's neighbourhoods)."""
    cal = C.CDLL(os.path.join(ROOT, "openclrenderer_amd", "libocr_calib.so"))
    cal.ocr_calib_valu.argtypes = [C.c_int, C.c_uint32]
    cal.ocr_calib_valu.restype = C.c_double
    cal.ocr_calib_atomics.argtypes = [C.c_int, C.c_uint32, C.c_size_t, C.c_uint32, C.c_int]
    cal.ocr_calib_atomics.restype = C.c_double
    f32 = cal.ocr_calib_valu(0, 4096)
    f64 = cal.ocr_calib_valu(1, 4096)
    n = 1 << 26
    words = 1920 * 1080
    us_rand = cal.ocr_calib_atomics(0, words, n, 1, 5)
    us_clus = cal.ocr_calib_atomics(0, words, n, 64, 5)
    if min(f32, f64, us_rand, us_clus) <= 0:
        raise RuntimeError("calibration kernels failed")
    return {"valu_f32_wave_insts_per_s": f32, "valu_f64_wave_insts_per_s": f64,
            "atomic_min_per_s_uniform": n / (us_rand * 1e-6), "atomic_min_per_s_clustered": n / (us_clus * 1e-6)}


def shadow_group_size(ntri):
    """Lights per shadow prearrange launch (the runtime renders up to 16
    lights' maps in one launch, whatever the triangle count)."""
    return 16


def shade_algorithmic_bytes(W, H, V, O, x_tex, x_shadow, L):
    """Minimal field bytes of one shading launch: 16 B per pixel (depth 4,
    triangle id 4, RGBA8 4, normal 4; the next render's clear is k_clear's),
    132 B per distinct visible triangle (112 B of vertex fields + 20 B
    fragment record), 52 B per distinct visible object, 4 B per distinct
    atlas / cube-map texel and 56 B per light."""
    return 16 * W * H + 132 * V + 52 * O + 4 * x_tex + 4 * x_shadow + 56 * L


def host_cpu():
    """CPU model, logical CPUs the process may use, and physical cores."""
    model, cores = "unknown", set()
    phys = core = None
    try:
        for ln in open("/proc/cpuinfo"):
            k, _, v = ln.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name" and model == "unknown":
                model = v
            elif k == "physical id":
                phys = v
            elif k == "core id":
                core = v
            elif not k and phys is not None and core is not None:
                cores.add((phys, core))
                phys = core = None
        if phys is not None and core is not None:
            cores.add((phys, core))
    except OSError:
        pass
    try:
        logical = len(os.sched_getaffinity(0))
    except AttributeError:
        logical = os.cpu_count() or 1
    return model, logical, len(cores) or logical


def timing_oracle_path():
    """The oracle's fastest build this host can run: the same source and
    arithmetic contract at -O3 for x86-64-v4 (AVX-512) or -v3 (AVX2 + FMA),
    else the test build."""
    flags = set()
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("flags"):
                flags = set(ln.split(":", 1)[1].split())
                break
    except OSError:
        pass
    base = os.path.join(ROOT, "oracle", "_build")
    v3 = {"avx2", "fma", "bmi1", "bmi2", "movbe", "f16c", "lzcnt" if "lzcnt" in flags else "abm"}
    v4 = v3 | {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"}
    for name, need in (("libocr_oracle_v4.so", v4), ("libocr_oracle_v3.so", v3)):
        path = os.path.join(base, name)
        if need <= flags and os.path.exists(path):
            return path, name
    return os.path.join(base, "libocr_oracle.so"), "libocr_oracle.so"


def cpu_baseline(sc, view, gpu_out, seconds, with_shadows=False):
    """The oracle on this host over the same workload: whole frames (with the
    shadow pass when a step has one), on the host cores this GPU's share
    allows (OMP_NUM_THREADS, at most the physical cores) after 2 warm-ups,
    median of 5; then the same on one thread (1 warm-up, median of 3).
    Stops early when `seconds` of CPU time are used up."""
    from oracle.binding import Oracle
    model, logical, physical = host_cpu()
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or physical
    cores = max(1, min(share, physical, logical))
    path, build = timing_oracle_path()
    o = Oracle(path)
    atlas = o.atlas(sc.textures)
    has_shadow = bool(np.any(sc.lights["shadow"]))
    maps = None if with_shadows or not has_shadow else o.shadow_maps(sc, cores)
    t_end = time.perf_counter() + seconds

    shadow_samples = covered = None

    def frames(nthreads, warm, timed):
        nonlocal maps, shadow_samples, covered
        times, first = [], None
        for k in range(warm + timed):
            t0 = time.perf_counter()
            if with_shadows:
                dyn, stat, shadow_samples = o.shadow_maps(sc, nthreads, with_samples=True)
                maps = (dyn, stat)
                covered = int(np.count_nonzero(dyn != 0xFFFFFFFF))
            r = o.render(sc, view, atlas=atlas, maps=maps, nthreads=nthreads)
            if k >= warm:
                times.append(time.perf_counter() - t0)
            first = first or r
            if times and time.perf_counter() > t_end:
                break
        return times, first

    times, ref = frames(cores, 2, 5)
    times1, _ = frames(1, 1, 3) if time.perf_counter() < t_end else ([], None)
    parity = None
    if gpu_out is not None:
        parity = bool(np.array_equal(gpu_out["depth"], ref["depth"]) and
                      np.array_equal(gpu_out["tri_id"], ref["tri_id"]) and
                      int(np.abs(gpu_out["rgba"].astype(int) - ref["rgba"].astype(int)).max()) <= 1)
    med = float(np.median(times))
    med1 = float(np.median(times1)) if times1 else None
    what = "shadow pass + frame" if with_shadows else "frames"
    return {"value": 1.0 / med, "unit": "frames/s", "cores": cores, "kind": "port",
            "sample": f"{view['width'][0]}x{view['height'][0]} {what} of the same scene: {len(times)} timed after "
                      f"2 warm-ups on {cores} OpenMP threads (median {med * 1e3:.0f} ms), {len(times1)} on 1 "
                      f"thread" + (f" (median {med1 * 1e3:.0f} ms)" if med1 else ""),
            "ms_per_frame": med * 1e3, "single_thread_ms_per_frame": med1 * 1e3 if med1 else None,
            "single_thread_value": 1.0 / med1 if med1 else None,
            "host": {"cpu_model": model, "nproc": logical, "physical_cores": physical,
                     "oracle_build": build}}, parity, dict(ref["stats"], shadow_samples=shadow_samples,
                                                        shadow_texels_covered=covered)


def animated_descriptors(objs, n, seed=5):
    """n descriptor arrays in which every object turns about its own axis and
    drifts a little more each frame (object::g_flush of every object)."""
    rng = np.random.default_rng(seed)
    m = objs.shape[0]
    axis = rng.normal(size=(m, 3))
    axis /= np.linalg.norm(axis, axis=1, keepdims=True)
    drift = rng.normal(size=(m, 3)) * 0.5
    q0 = objs["world_rot_quat"].astype(np.float64)        # (x, y, z, w)
    out = []
    for k in range(n):
        a = 0.02 * (k + 1)
        r = np.concatenate([axis * np.sin(a / 2), np.full((m, 1), np.cos(a / 2))], axis=1)
        x1, y1, z1, w1 = r.T
        x2, y2, z2, w2 = q0.T
        q = np.stack([w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2, w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                      w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2, w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2], axis=1)
        o = objs.copy()
        o["world_rot_quat"] = q.astype(np.float32)
        o["world_pos"][:, :3] = objs["world_pos"][:, :3] + (drift * (k + 1)).astype(np.float32)
        out.append(o)
    return out


def run_animated(g, fr, sc, view, steps, warmup):
    """Frames per second when every frame first re-uploads all object
    descriptors on the render stream (ocr_scene_update_objs_async), then
    renders; the scene is restored afterwards."""
    sets = animated_descriptors(sc.objs, steps + warmup)
    for k in range(warmup):
        g.update_objs_async(0, sets[k])
        fr.render(g, view)
    hip_sync()
    t0 = time.perf_counter()
    for k in range(steps):
        g.update_objs_async(0, sets[warmup + k])
        fr.render(g, view)
    hip_sync()
    t1 = time.perf_counter()
    g.update_objs(0, sc.objs)
    ms = (t1 - t0) * 1e3 / steps
    return {"value": 1e3 / ms, "unit": "frames/s", "ms_per_step": ms, "objects_updated_per_frame": int(sc.objs.shape[0]),
            "note": "every frame: all descriptors re-uploaded (a new geometry version: descriptors and the "
                    "world-space copy, built while the previous render shades), then rendered; same scene and "
                    "view otherwise"}


def run_blend(ctx, g, fr, sc, view, W, H, reps=100):
    """blend_screens_with_depth (ocr_frame_blend_with_depth) of a second
    view of the scene onto the frame: average device time per launch and the
    HBM rate of its 20 algorithmic bytes per pixel (two depths and two
    colours read, one colour written); do_pseudo_aa (ocr_frame_pseudo_aa,
    colour snapshot + kernel) and do_motion_blur (ocr_frame_motion_blur,
    snapshot + kernel; the previous camera 40 units to the side and turned
    by 0.005 rad, frame ids alternating) on the frame."""
    from openclrenderer_amd import runtime
    other = view.copy()
    other["c_pos"][0, 0] += 40.0
    f2 = runtime.Frame(ctx, W, H)
    f2.render_read(g, other)
    fr.render_read(g, view)
    ev = EventPair()
    f2.blend_from(fr)
    hip_sync()
    ev.start()
    for _ in range(reps):
        f2.blend_from(fr)
    ev.stop()
    us = ev.ms() * 1e3 / reps
    fr.pseudo_aa()
    hip_sync()
    ev.start()
    for _ in range(reps):
        fr.pseudo_aa()
    ev.stop()
    aa_us = ev.ms() * 1e3 / reps
    fr.render_read(g, view)
    c_pos_old = view["c_pos"][0] + np.array([40.0, 0.0, 0.0, 0.0], np.float32)
    c_rot_old = view["c_rot"][0] + np.array([0.0, 0.005, 0.0, 0.0], np.float32)
    fr.motion_blur(g, view, c_pos_old, c_rot_old, 0)
    hip_sync()
    ev.start()
    for i in range(reps):
        fr.motion_blur(g, view, c_pos_old, c_rot_old, i)
    ev.stop()
    mb_us = ev.ms() * 1e3 / reps
    fr.render_read(g, view)
    return {"kernel": "k_blend_depth", "us_per_launch": us, "bytes_per_pixel": 20,
            "pseudo_aa_us_per_call": aa_us, "motion_blur_us_per_call": mb_us,
            "achieved_GBs": 20.0 * W * H / (us * 1e-6) / 1e9, "peak_GBs": HBM_PEAK_GBS,
            "note": "back-to-back launches over the same two frames, whose 20 B/pixel stay in the "
                    "256 MB Infinity Cache, so the rate can exceed the HBM peak"}


def base_line(args, metric, value, unit, ms_per_step, config, dtype="f32"):
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config}


METRIC = "frames/s + Mpix/s at 1080p/1M tris; p50/p99 latency on mixed-size request stream"


def run_plain(args):
    from openclrenderer_amd import runtime, scenes
    if args.config == 1:
        sc, view = scenes.config1()
        workload = "800x600 2480 tris (icosphere + 10x10 cubes), 1 light, 256^2 texture"
    elif args.config == 2:
        sc, view = scenes.config2()
        workload = "1920x1080 1M tris (400 instances), 4 point lights, 2048^2 atlas"
    else:
        sc, view = scenes.config3()
        workload = "3840x2160 4M tris (1600 instances), 16 shadow-casting point lights (6x1024^2 maps), 8 1024^2 textures"
    W, H = int(view["width"][0]), int(view["height"][0])
    shadows = bool(np.any((sc.lights["shadow"] == 1) & (sc.lights["is_static"] == 0)))

    ctx = runtime.Context(0)
    g = runtime.GpuScene(ctx, sc)
    fr = runtime.Frame(ctx, W, H)
    if shadows:
        g.shadow_pass()
    out = fr.render_read(g, view)          # sizes the work lists, reads the frame once
    for _ in range(args.warmup):
        if shadows:
            g.shadow_pass()
        fr.render(g, view)
    hip_sync()
    # the timed region times the shading pass of every render (two events
    # per render, on the stream it runs on)
    fr.set_profiling(True, shade_only=True)
    ev = [EventPair() for _ in range(args.steps)] if shadows else []
    hip_sync()
    t0 = time.perf_counter()
    for i in range(args.steps):
        if shadows:
            ev[i].start()
            g.shadow_pass()
            ev[i].stop()
        fr.render(g, view)
    hip_sync()
    t1 = time.perf_counter()
    st = fr.stats()
    if st["overflow"]:
        raise RuntimeError("work list overflow inside the timed region")
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    fps = 1e3 / ms_per_step
    ms_shadow = float(np.mean([e.ms() for e in ev])) if ev else 0.0
    # every pass of the same back-to-back stream, in a run of its own (the
    # events around every pass add time, so not in the timed region)
    fr.set_profiling(True)
    for i in range(min(args.steps, 100)):
        if shadows:
            g.shadow_pass()
        fr.render(g, view)
    hip_sync()
    st_all = fr.stats()
    # Renders of one frame overlap (render N+1's raster passes run beside
    # render N's shading), so the timed region's per-pass durations include
    # co-running work.  The same passes with the host waiting after every
    # render give each kernel's duration alone.
    fr.set_profiling(True)
    for _ in range(max(5, min(args.steps, 50))):
        fr.render(g, view)
        hip_sync()
    st_iso = fr.stats()
    ms_shadow_iso = 0.0
    if shadows:
        evi = [EventPair() for _ in range(max(5, min(args.steps, 20)))]
        for e in evi:
            hip_sync()
            e.start()
            g.shadow_pass()
            e.stop()
        ms_shadow_iso = float(np.mean([e.ms() for e in evi]))

    # algorithmic bytes of the shading launch
    tri = out["tri_id"]
    vis = tri[tri != UINT_MAX]
    V = int(np.unique(vis).size)
    O = int(np.unique(sc.tris["v"]["object_id"][vis, 0]).size) if V else 0
    x_tex, x_shadow = fr.count_texels(g, view)
    B = shade_algorithmic_bytes(W, H, V, O, x_tex, x_shadow, int(sc.lights.shape[0]))
    achieved = B / (st["ms_shade"] * 1e-3) / 1e9
    # the shading kernel the path launches: without shadow-casting lights the
    # one with the occlusion code compiled out
    shade_kernel = "k_shade<false>" if np.any(sc.lights["shadow"] != 0) else "k_shade_unshadowed"
    prof = profile_entry(shade_kernel, workload)
    traffic = prof.get("bytes_per_launch")
    achieved_iso = B / (st_iso["ms_shade"] * 1e-3) / 1e9
    cal = calibration()
    valu = None
    if "valu_insts_per_launch" in prof:
        n = float(prof["valu_insts_per_launch"])
        peak = cal["valu_f32_wave_insts_per_s"]
        valu = {"insts_per_launch": n, "peak_insts_per_s": peak,
                "peak_source": "measured: independent binary32 FMA chains, 8 wavefronts per SIMD (calib.hip)",
                "frac": n / (st["ms_shade"] * 1e-3) / peak,
                "frac_isolated": n / (st_iso["ms_shade"] * 1e-3) / peak}

    # the legs below time whole calls; per-pass timing events stay off
    fr.set_profiling(False)
    # the first update of a scene allocates its second geometry version and
    # the first waits across streams create their events: a longer warm-up
    extras = args.config == 2 and not args.plain_only
    animated = run_animated(g, fr, sc, view, max(args.steps, 200), max(args.warmup, 20)) if extras else None
    post = run_blend(ctx, g, fr, sc, view, W, H) if extras else None

    if args.calib:
        copy_lib = C.CDLL(os.path.join(ROOT, "openclrenderer_amd", "libocr_calib.so"))
        copy_lib.ocr_calib_copy.argtypes = [C.c_size_t, C.c_int]
        if copy_lib.ocr_calib_copy(256 << 20, 3) != 0:
            raise RuntimeError("calibration copy failed")

    cpu, parity, ref_stats = (None, None, None) if args.no_cpu else \
        cpu_baseline(sc, view, out, args.cpu_seconds, shadows)
    # depth tests per second of the raster passes (one per covered sample:
    # the reference's atomic_min count; the oracle counts them) against the
    # measured atomicMin rate
    atomics = None
    if ref_stats is not None:
        S = int(ref_stats["n_samples"])
        peak = cal["atomic_min_per_s_clustered"]
        atomics = {"covered_samples": S, "peak_atomic_min_per_s": peak,
                   "peak_source": "measured: u32 atomicMin over a 1080p-sized buffer in 64-address clusters "
                                  f"(uniform random: {cal['atomic_min_per_s_uniform']:.3g}/s)"}
        for name, key in (("depth", "ms_depth"), ("id", "ms_id")):
            rate = S / (st_iso[key] * 1e-3)
            atomics[name] = {"ms_isolated": st_iso[key], "tests_per_s": rate, "frac": rate / peak}
        if shadows and ms_shadow > 0 and ref_stats.get("shadow_samples"):
            rate = ref_stats["shadow_samples"] / (ms_shadow * 1e-3)
            atomics["shadow"] = {"covered_samples": int(ref_stats["shadow_samples"]), "ms": ms_shadow,
                                 "tests_per_s": rate, "frac": rate / peak}
    served = batched = None
    if extras:
        batched = batched_leg(ctx, steps=10, warmup=2, autotune_reps=2)
        served = served_leg(ctx, n_req=400, rates=(1000, 2000), modes=("8x4", "1x1"))
    passes = {k: st_all[k] for k in ("ms_clear", "ms_prearrange", "ms_depth", "ms_id", "ms_shade", "ms_total")}
    passes["ms_shade_timed_region"] = st["ms_shade"]
    shadow = None
    if shadows:
        passes["ms_shadow"] = ms_shadow
        Ls = int(sc.light_buffer_dim)
        nl = int(np.sum(sc.lights["shadow"] == 1))
        # algorithmic bytes of one pass: the triangle positions and object
        # ids read once per group of lights, every map texel cleared once,
        # one 4-byte write per texel that some triangle covers
        groups = -(-nl // shadow_group_size(sc.ntri))
        covered = ref_stats.get("shadow_texels_covered") if ref_stats else None
        Bs = groups * sc.ntri * 40 + nl * 6 * Ls * Ls * 4 + (covered or 0) * 4
        shadow = {"ms_isolated": ms_shadow_iso, "ms_in_stream": ms_shadow, "lights": nl, "groups": groups,
                  "map_dim": Ls, "covered_texels": covered, "algorithmic_bytes": Bs,
                  "achieved_gbs": Bs / (ms_shadow_iso * 1e-3) / 1e9,
                  "frac_hbm": Bs / (ms_shadow_iso * 1e-3) / 1e9 / HBM_PEAK_GBS}
    passes_iso = {k: st_iso[k] for k in ("ms_clear", "ms_prearrange", "ms_depth", "ms_id", "ms_shade", "ms_total")}
    line = base_line(args, METRIC, fps, "frames/s", ms_per_step,
                     {"workload": workload, "width": W, "height": H, "triangles": sc.ntri,
                      "lights": int(sc.lights.shape[0]), "shadow_lights": int(np.sum(sc.lights["shadow"] == 1)),
                      "path": "plain"})
    line.update({
        "mpix_per_s": W * H / (ms_per_step * 1e-3) / 1e6,
        "passes_ms": passes,
        "passes_ms_isolated": passes_iso,
        "work": {"cut_tris": st["n_cut"], "fragments": st["n_frag"], "visible_tris": V, "visible_objs": O,
                 "atlas_texels": x_tex, "shadow_texels": x_shadow},
        "roofline": {"bound": "hbm", "kernel": shade_kernel, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes": B, "valu_issue": valu,
                     "limiter": "valu" if valu and valu["frac"] > achieved / HBM_PEAK_GBS else "hbm",
                     "isolated": {"ms": st_iso["ms_shade"], "achieved": achieved_iso,
                                  "frac": achieved_iso / HBM_PEAK_GBS}},
        "atomics": atomics,
        "shadow_pass": shadow,
        "calibration": cal,
        "post_blend_with_depth": post,
        "animated": animated,
        "batched": batched,
        "served": served,
        "cpu_baseline": cpu,
        "parity_vs_oracle": parity,
    })
    if cpu:
        line["speedup_vs_cpu"] = fps / cpu["value"]
    print(json.dumps(line), flush=True)


def batched_leg(ctx, steps, warmup, autotune_reps=3):
    """Config 4: the 64 views rendered as one autotuned, graph-replayed batch
    per step against the same views as 64 plain render sequences."""
    from openclrenderer_amd import runtime, scenes
    scs, views = scenes.config4_workload()
    gs = [runtime.GpuScene(ctx, s) for s in scs]
    frames = [runtime.Frame(ctx, int(v["width"][0]), int(v["height"][0])) for _, v in views]
    owners = [gs[i] for i, _ in views]
    vlist = [v for _, v in views]
    b = runtime.Batch(ctx, len(views))
    b.set_views(owners, vlist, frames)
    b.render_checked()
    tuning = b.autotune(reps=autotune_reps)
    b.render_checked(graph=True)
    for _ in range(warmup):
        b.launch_graph()
    hip_sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.launch_graph()
    hip_sync()
    t_batch = (time.perf_counter() - t0) / steps
    b.stats()
    # the same 64 views rendered one plain launch sequence at a time
    plain = [runtime.Frame(ctx, f.width, f.height) for f in frames]
    for f, g, v in zip(plain, owners, vlist):
        f.render_read(g, v)
    hip_sync()
    seq_steps = max(1, steps // 2)
    t0 = time.perf_counter()
    for _ in range(seq_steps):
        for f, g, v in zip(plain, owners, vlist):
            f.render(g, v)
    hip_sync()
    t_seq = (time.perf_counter() - t0) / seq_steps
    same = all(np.array_equal(a.read()["rgba"], p.read()["rgba"]) for a, p in zip(frames, plain))
    npix = sum(f.width * f.height for f in frames)
    out = {"views_per_s": len(views) / t_batch, "ms_per_batch": t_batch * 1e3, "mpix_per_s": npix / t_batch / 1e6,
           "sequential_views_per_s": len(views) / t_seq, "sequential_ms": t_seq * 1e3,
           "speedup_vs_sequential": t_seq / t_batch,
           "tuning": {"raster_blocks_per_cu": tuning[0], "shade_tile_w_log2": tuning[1], "groups": tuning[2]},
           "batched_equals_sequential": bool(same), "views": len(views), "scenes": len(scs),
           "triangles": [s.ntri for s in scs],
           "workload": "64 views (320x240 .. 2560x1440) over 8 scenes of 50k-2M tris, 4 lights each"}
    b.close()
    for f in frames + plain:
        f.close()
    for g in gs:
        g.close()
    return out


def run_batched(args):
    from openclrenderer_amd import runtime
    ctx = runtime.Context(0)
    r = batched_leg(ctx, args.steps, args.warmup)
    line = base_line(args, "views/s, 64 mixed-size views over 8 resident scenes (batched vs sequential plain)",
                     r["views_per_s"], "views/s", r["ms_per_batch"],
                     {"workload": r["workload"], "views": r["views"], "scenes": r["scenes"],
                      "triangles": r["triangles"], "path": "batched (hipGraph, autotuned)"})
    line.update({k: r[k] for k in ("mpix_per_s", "sequential_views_per_s", "speedup_vs_sequential", "tuning",
                                   "batched_equals_sequential")})
    line["sequential_ms_per_step"] = r["sequential_ms"]
    print(json.dumps(line), flush=True)


def served_leg(ctx, n_req, rates, modes):
    """Config 5: Poisson arrivals of n_req requests per rate against 12
    resident scenes through a server in each mode ("VxF": at most V views per
    batch, F batches in flight; "1x1" renders one request at a time)."""
    from openclrenderer_amd import runtime, scenes
    scs, rng = scenes.config5_workload()
    gs = [runtime.GpuScene(ctx, s) for s in scs]

    def schedule(rate):
        return scenes.config5_requests(rng, n_req, rate, len(scs))

    def drive(srv, slots, sched):
        lat, qwait, batch = [], [], []
        t_base = srv.now_ns()
        refused = failed = i = done = 0
        while done < len(sched):
            now = (srv.now_ns() - t_base) * 1e-9
            while i < len(sched) and sched[i][0] <= now:
                rid = srv.submit(slots[sched[i][1]], sched[i][2])
                if rid is None:
                    refused += 1
                    break                      # retried after completions are released
                i += 1
            for c in srv.poll(max_n=512, wait_ms=0):
                lat.append((c.t_done_ns - c.t_submit_ns) * 1e-6)
                qwait.append((c.t_start_ns - c.t_submit_ns) * 1e-6)
                batch.append(c.batch_views)
                failed += c.status != 0
                srv.release(c.request_id)
                done += 1
            if i < len(sched) and done < len(sched):
                gap = sched[i][0] - (srv.now_ns() - t_base) * 1e-9
                if gap > 2e-4:
                    time.sleep(min(gap, 1e-3) * 0.5)
        wall = (srv.now_ns() - t_base) * 1e-9
        lat = np.asarray(lat)
        return {"completed_per_s": len(sched) / wall, "p50_ms": float(np.percentile(lat, 50)),
                "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()),
                "queue_p50_ms": float(np.percentile(qwait, 50)), "mean_batch": float(np.mean(batch)),
                "refused_submits": refused, "failed": failed}

    results = {}
    for m in modes:
        v, f = (int(x) for x in m.split("x"))
        name = "one_at_a_time" if (v, f) == (1, 1) else f"batched_{v}x{f}"
        with runtime.Server(ctx, max_batch_views=v, max_inflight=f, max_batch_tris=24_000_000, hbm_fraction=0.25,
                            read_mask=runtime.READ_RGBA) as srv:
            slots = [srv.add_scene(g) for g in gs]
            warm = schedule(2000)[:64]
            drive(srv, slots, [(0.0, k, vw) for _, k, vw in warm])
            results[name] = {}
            for rate in rates:
                r = drive(srv, slots, schedule(rate))
                r["splits_so_far"] = srv.info()["n_splits"]
                results[name][str(rate)] = r
                print(json.dumps({"progress": name, "rate": rate, **r}), file=sys.stderr, flush=True)
    for g in gs:
        g.close()
    top = str(max(rates))
    batched = [k for k in results if k != "one_at_a_time"]
    best_mode = min(batched, key=lambda k: results[k][top]["p99_ms"]) if batched else "one_at_a_time"
    out = {"workload": "12 resident scenes (50k-4M tris, 1-16 shadowed lights), resolutions "
                       "640x480/1280x720/1920x1080/3840x2160 at .4/.3/.2/.1, Poisson arrivals",
           "requests_per_rate": n_req, "rates": list(rates), "scenes_tris": [s.ntri for s in scs],
           "modes": results, "best_mode": best_mode, "best": results[best_mode][top],
           "p50_ms_at_top_rate": results[best_mode][top]["p50_ms"],
           "p99_ms_at_top_rate": results[best_mode][top]["p99_ms"]}
    if "one_at_a_time" in results:
        alone = results["one_at_a_time"][top]
        out["one_at_a_time_p99_ms_at_top_rate"] = alone["p99_ms"]
        out["throughput_vs_one_at_a_time"] = results[best_mode][top]["completed_per_s"] / alone["completed_per_s"]
    return out


def run_served(args):
    from openclrenderer_amd import runtime
    ctx = runtime.Context(0)
    rates = [int(r) for r in args.rates.split(",")]
    r = served_leg(ctx, args.requests, rates, args.modes.split(","))
    best = r["best"]
    line = base_line(args, "served requests/s with p50/p99 latency (Poisson arrivals, mixed resolutions)",
                     best["completed_per_s"], "requests/s", 1e3 / best["completed_per_s"],
                     {"workload": r["workload"], "requests_per_rate": r["requests_per_rate"], "rates": rates,
                      "scenes_tris": r["scenes_tris"], "path": "served (continuous batching)"})
    line.update({"served": r["modes"], "best_mode": r["best_mode"], "p99_ms_at_top_rate": r["p99_ms_at_top_rate"],
                 "one_at_a_time_p99_ms_at_top_rate": r.get("one_at_a_time_p99_ms_at_top_rate")})
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 200 frames after 20 warm-up ones: the two-deep render pipeline's fill and
    # drain are amortised (the timed region is still well under a second)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", type=int, default=2, choices=[1, 2, 3, 4, 5])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--calib", action="store_true", help="also run the copy kernel used to scale PMC counters")
    ap.add_argument("--plain-only", action="store_true",
                    help="config 2: skip the animated and post-pass legs (profiling runs)")
    ap.add_argument("--requests", type=int, default=2000, help="served: requests per arrival rate")
    ap.add_argument("--rates", default="50,100,200,500,1000,2000", help="served: arrival rates (req/s)")
    ap.add_argument("--modes", default="8x4,16x3,1x1",
                    help="served: server modes, max views per batch x batches in flight (1x1 = one at a time)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is design-only)", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.config in (1, 2, 3):
        run_plain(args)
    elif args.config == 4:
        run_batched(args)
    else:
        run_served(args)


if __name__ == "__main__":
    main()
