#!/usr/bin/env python3
"""Benchmark: Mrays/s (primary + AO + shadow) of the gfx950 per-ray path.

Workload (default): the molecular surface -- the Gaussian-density isosurface
(sigma 1 A, iso 0.5, 0.5 A grid) of a 40,000-atom synthetic protein as 4.03 M
indexed triangles at alpha 0.6 over its 40,000 opaque CPK spheres, with the
view's object-space silhouette edges (sphere rings clipped to the union
boundary, mesh silhouettes, creases and borders) as chained open curves;
hard shadows from one distant light, 16 cosine AO rays per shaded hit (8 A),
1920x1080 at 1 spp. One step = one full frame: every transparency layer's
primary rays, hit setup, AO, shadows, POV shading and compositing, then the
resolve, with the scene and its BVH already resident in HBM and the frame
left in HBM (no host copy). `--workload ballstick` runs the ball-and-stick
frame (20,000 atoms + half-bond cylinders, 1920x1080 ss2, AO16, fog).

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mrays/s primary+AO+shadow; time to 40 dB PSNR, denoised vs plain, 1 MI355X"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpus() -> dict:
    """The host's processor as lscpu reports it, and the CPUs this process
    may run on."""
    import subprocess
    info = {"affinity_cpus": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
    except (OSError, subprocess.SubprocessError):
        return info
    keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
            "Thread(s) per core": "threads_per_core", "CPU(s)": "cpus"}
    for line in out.splitlines():
        k, _, v = line.partition(":")
        if k.strip() in keys and keys[k.strip()] not in info:
            v = v.strip()
            info[keys[k.strip()]] = int(v) if v.isdigit() else v
    return info


def cpu_baseline(scene, opt, target_s: float):
    """Oracle (CPU restatement, FP32) on the host: the full frame, median of
    5 runs after 1 warm-up, when that fits the time target; otherwise a row
    sample sized to about target_s seconds. Threads: every CPU this process
    may use, capped by the host share the job is given (OMP_NUM_THREADS when
    set: one GPU's share of a shared host)."""
    from oracle import binding as ob
    cpus = host_cpus()
    threads = cpus["affinity_cpus"]
    share = os.environ.get("OMP_NUM_THREADS", "")
    if share.isdigit() and int(share) > 0:
        threads = min(threads, int(share))
    ob.set_scene_cache(True)  # the oracle's BVH is built once, outside the timed window
    H = opt.height * opt.supersample
    W = opt.width * opt.supersample
    step = 64
    f = ob.render(scene, opt, hires_only=True, threads=threads, rows=(0, H, step))
    cal = max(1e-3, f.render_seconds)
    log(f"cpu calibration: {f.total_rays} rays in {cal:.2f}s ({f.total_rays / cal / 1e6:.2f} Mrays/s)")
    est_full = cal * step
    if 6 * est_full <= 2 * target_s:
        times = []
        for i in range(6):
            f = ob.render(scene, opt, hires_only=True, threads=threads)
            if i > 0:
                times.append(f.render_seconds)
        sec = float(np.median(times))
        rays = f.total_rays
        sample = (f"full frame ({H} x {W} hi-res px, {rays} rays), median of 5 runs after 1 warm-up "
                  f"({', '.join(f'{t:.3f}' for t in times)} s)")
    else:
        step2 = max(1, min(step, int(step * cal / target_s)))
        f = ob.render(scene, opt, hires_only=True, threads=threads, rows=(0, H, step2))
        rays, sec = f.total_rays, f.render_seconds
        sample = f"hi-res rows y = 0 mod {step2} ({len(range(0, H, step2))} of {H} rows x {W} px), {rays} rays"
    ob.set_scene_cache(False)
    log(f"cpu baseline: {rays} rays in {sec:.3f}s with {threads} threads ({cpus})")
    return {
        "value": rays / sec / 1e6,
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "host": cpus,
        "sample": sample + "; CPU restatement (not Embree), FP32, own binned-SAH BVH, std::thread over rows; "
                           "time window = trace + shade + fog (BVH build excluded)",
    }


TRAFFIC_TABLE = ROOT / "profiles" / "hbm_traffic_ballstick_surface.md"


def measured_traffic(workload: str, layout: str, kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE scaled by a 256 MiB copy kernel's
    readings); None when the table has no row for this configuration."""
    if not TRAFFIC_TABLE.exists():
        return None
    for line in TRAFFIC_TABLE.read_text().splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 6 and cells[0] == workload and cells[1] == layout and cells[2] == kernel:
            try:
                return float(cells[5]) * 1e6
            except ValueError:
                return None
    return None


COMPUTE_TABLE = ROOT / "profiles" / "pmc_valu_lds_l2_surface_gi_unet.md"
COMPUTE_SECTIONS = {"surface-edges": "## Molecular surface + object-space edges", "gi": "## pt1 GI on the molecular surface"}


def measured_compute(workload: str, kernel: str):
    """VALU issue share, lane utilization and L2 hit rate of `kernel` from the
    committed rocprofv3 counter summary; None without a row for it."""
    head = COMPUTE_SECTIONS.get(workload)
    if head is None or not COMPUTE_TABLE.exists():
        return None
    inside = False
    for line in COMPUTE_TABLE.read_text().splitlines():
        if line.startswith("## "):
            inside = line.startswith(head)
            continue
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if inside and len(cells) >= 10 and cells[0].strip("`").startswith(kernel + "<false"):
            try:
                return {"valu_issue_share": round(float(cells[4]) / 100, 3),
                        "valu_lane_utilization": round(float(cells[5]) / 100, 3),
                        "l2_hit": round(float(cells[8]) / 100, 3),
                        "source": COMPUTE_TABLE.name}
            except ValueError:
                return None
    return None


def time_to_40db(dev) -> dict:
    """The metric's second half on the path-traced GI workload: the 40,000-
    atom molecular surface as a GGX dielectric over GGX conductor / dielectric
    CPK spheres, 64 spherical area lights (RIS + ReSTIR, MIS with BSDF rays),
    4 bounces, 1920x1080. Time for a frame to reach 40 dB PSNR (sRGB, the
    whole frame) against a 4096-sample render: the plain path tracer at
    1, 2, 4, ... samples per pixel and the learned denoiser
    (umbreon_amd/weights/pt_unet.npz) on the same frames, each until it
    reaches 40 dB."""
    from umbreon_amd.denoise_train import load
    from umbreon_amd.pt_train import WEIGHTS, evaluate_full
    net = load(dev, str(WEIGHTS))
    try:
        return evaluate_full(dev, net, log=log)
    finally:
        net.close()


def serving(dev, n_requests: int = 256, rate: float = 40.0) -> dict:
    """Batched serving: the seeded stream of n_requests renders (Poisson
    arrivals at `rate` per second; atom counts log-uniform 500..200,000 over
    a pool of 8 molecule sizes, surfaces up to 36,106 atoms; CPK, ball-and-
    stick or molecular-surface representation; 512x512, 1024x768 or
    1920x1080; AO 16, hard shadows) served in real time by umbreon_amd.serve:
    scenes built on the GPU on first use and cached, requests admitted by an
    HBM budget into batches rendered by ubg_render_batch, tuning picked by
    the autotuner on calibration requests first. Latency = completion (the
    image on the host) - arrival."""
    from umbreon_amd import serve
    stream = serve.request_stream(n_requests, rate=rate, seed=0)
    pool = serve.MoleculePool()
    t0 = time.perf_counter()
    for key in sorted({(r.atoms, r.rep) for r in stream}):
        pool.get(*key)
    gen_s = time.perf_counter() - t0
    log(f"serving: {len(pool.scenes)} molecules generated on the host in {gen_s:.1f}s")
    budget = 96 << 30
    calib = serve.request_stream(32, rate=rate, seed=1)
    t0 = time.perf_counter()
    tun = serve.autotune(dev, pool, calib, budget, log=log)
    tune_s = time.perf_counter() - t0
    log(f"autotuned in {tune_s:.1f}s: {tun}")
    # batching's gain on the calibration requests: one request per batch vs the tuned batches
    single = serve.Server(dev, pool, budget, serve.Tuning(ao_waves_per_simd=tun.ao_waves_per_simd, max_batch=1,
                                                          leaf_max=dict(tun.leaf_max)))
    try:
        serve.offline_throughput(single, calib)
        one = serve.offline_throughput(single, calib)
    finally:
        single.close()
    srv = serve.Server(dev, pool, budget, tun)
    try:
        res = srv.run(stream, log=None)
    finally:
        srv.close()
    sizes = serve.pool_sizes()
    res.update({"workload": f"{n_requests} seeded requests, Poisson {rate:g}/s, atoms log-uniform 500-200000 "
                            f"(pool {sizes}, surfaces <= {max(v for v in sizes if v <= 40000)}), "
                            f"CPK / ball-and-stick / surface, "
                            f"512x512 / 1024x768 / 1920x1080, AO16, hard shadows",
                "tuning": {"ao_waves_per_simd": tun.ao_waves_per_simd, "max_batch": tun.max_batch,
                           "leaf_max": tun.leaf_max},
                "calibration_images_per_s": {"one_request_per_launch": round(one, 1)},
                "host_generation_s": round(gen_s, 1), "autotune_s": round(tune_s, 1)})
    for k in ("images_per_s", "p50_latency_ms", "p99_latency_ms", "mean_latency_ms", "makespan_s", "mrays_per_s",
              "scene_build_s"):
        res[k] = round(res[k], 3)
    log(f"serving: {res}")
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["ballstick", "spheres", "surface", "gi"], default="surface")
    ap.add_argument("--no-edges", action="store_true",
                    help="surface workload without the object-space silhouette edges")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--layout", choices=["bvh8", "bvh8gpu", "bvh2"], default="bvh8",
                    help="acceleration structure: compressed 8-wide built on the host (default) or on the GPU, "
                         "or full-precision binary")
    ap.add_argument("--no-denoise-metric", action="store_true",
                    help="skip the time-to-40-dB measurement of the learned GI denoiser")
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the batched-serving measurement (256-request Poisson stream)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (a copy kernel of known "
                         "traffic, for scaling FETCH_SIZE / WRITE_SIZE counter readings)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X: this build runs on a single GPU; "
              "multi-GPU runs (--gpus > 1) are not supported", file=sys.stderr)
        return 2

    from umbreon_amd import capi
    from umbreon_amd.render import BuiltScene, Device
    from umbreon_amd.scenes import ball_and_stick_scene, molecular_surface_scene, spheres_scene

    t0 = time.perf_counter()
    if args.workload == "ballstick":
        scene, opt = ball_and_stick_scene()
        workload = (f"ball-and-stick: {len(scene.sphere_center)} atom spheres r0.4 + "
                    f"{len(scene.cyl_p0)} capped half-bond cylinders r0.2, 1920x1080 ss2, "
                    f"AO16 (8 A), hard shadows, linear fog 60 A")
    elif args.workload == "surface":
        scene, opt = molecular_surface_scene()
        workload = (f"molecular surface: Gaussian isosurface (sigma 1 A, iso 0.5, 0.5 A grid) of a "
                    f"{len(scene.sphere_center)}-atom protein, {len(scene.index)} indexed triangles at alpha 0.6 "
                    f"over {len(scene.sphere_center)} CPK spheres, 1920x1080 ss1, AO16 (8 A), hard shadows")
    elif args.workload == "gi":
        scene, opt = molecular_surface_scene()
        opt.ao_samples = 0
        opt.gi, opt.gi_integrator, opt.pt1_spp, opt.gi_bounces, opt.pt1_denoise = True, 1, 8, 2, True
        workload = (f"pt1 GI on the molecular surface: {len(scene.index)} indexed triangles at alpha 0.6 over "
                    f"{len(scene.sphere_center)} CPK spheres, 1920x1080 ss1, full-res gather 8 spp, 2 bounces, "
                    f"NEE + hard shadows, a-trous E denoise")
    else:
        scene, opt = spheres_scene()
        workload = "2,000 CPK spheres, 800x600 ss1, hard shadows, no AO"
    log(f"scene generated in {time.perf_counter() - t0:.1f}s: {workload}")

    dev = Device(0)
    edges_info = None
    if args.workload == "surface" and not args.no_edges:
        from umbreon_amd.edges import ObjectSpaceEdgeOptions, append_edges, object_space_edges
        t0 = time.perf_counter()
        e = object_space_edges(scene, ObjectSpaceEdgeOptions(enable=True), dev)
        edge_s = time.perf_counter() - t0
        scene = append_edges(scene, e)
        edges_info = {"cylinders": len(e["group"]), "generate_ms": round(edge_s * 1e3, 1)}
        workload += (f", object-space edges ({len(e['group'])} open curve segments r{e['radius']:g}: "
                     f"clipped sphere rings, mesh silhouettes, creases, borders)")
        log(f"object-space edges: {len(e['group'])} in {edge_s * 1e3:.0f} ms")
    bs = BuiltScene(dev, scene, {"bvh8": capi.BVH8, "bvh8gpu": capi.BVH8_GPU_BUILD, "bvh2": capi.BVH2}[args.layout])
    info = bs.info
    log(f"BVH ({args.layout}): {info['n_nodes']} nodes of {info['node_bytes']} B, depth {info['bvh_depth']}, "
        f"build {info['build_seconds'] * 1e3:.1f} ms, SAH {info['sah_cost']:.1f}")

    # visit counts on the same BVH and rays (instrumented kernels, untimed)
    vc = bs.render(scene, opt, outputs=False, count_visits=True)
    st = vc.stats
    rays_per_frame = vc.total_rays
    log(f"rays/frame {rays_per_frame}: primary {st['primary_rays']}, AO traced {st['ao_rays_traced']} "
        f"sampled {st['ao_rays_sampled']}, shadow {st['shadow_rays']}, GI gather {st['gi_rays']} "
        f"NEE {st['gi_nee_rays']} G-buffer {st['gi_gbuffer_rays']}")

    for _ in range(args.warmup):
        bs.render(scene, opt, outputs=False)
    ks = {"primary": 0.0, "ao": 0.0, "resolve": 0.0, "post": 0.0}
    kn = {k: 0 for k in ks}
    t1 = time.perf_counter()
    for _ in range(args.steps):
        fr = bs.render(scene, opt, outputs=False, kernel_times=True)
        for k in ks:
            ks[k] += fr.kernel_ms[k]
            kn[k] += fr.kernel_launches[k]
    elapsed = time.perf_counter() - t1
    assert fr.total_rays == rays_per_frame
    ms_per_step = elapsed / args.steps * 1e3
    value = rays_per_frame * args.steps / elapsed / 1e6
    log(f"{args.steps} frames in {elapsed:.3f}s: {ms_per_step:.2f} ms/frame, {value:.1f} Mrays/s; "
        f"kernel ms/launch: " + ", ".join(f"{k} {ks[k] / max(1, kn[k]):.3f}" for k in ks))

    # roofline of the dominant kernel
    dom = max(("primary", "ao", "resolve"), key=lambda k: ks[k])
    cls = {"primary": 0, "ao": 1, "resolve": 2}[dom]
    # (the GI gather runs in the AO slot and counts its visits there)
    rays_in = {"primary": st["primary_rays"],
               "ao": st["ao_rays_traced"] + st["ao_rays_sampled"] + st["gi_rays"] + st["gi_nee_rays"],
               "resolve": st["shadow_rays"]}[dom]
    nodes, tests = st["node_visits"][cls], st["prim_tests"][cls]
    node_b, prim_b = info["node_bytes"], info["prim_bytes"]
    # per frame: a frame with transparency layers launches each kernel once per layer
    bytes_per_frame = 48 * rays_in + node_b * nodes + prim_b * tests
    dur_s = ks[dom] / args.steps / 1e3
    achieved = bytes_per_frame / dur_s / 1e9
    kname = {"primary": "k_primary", "ao": "k_ao" if args.layout == "bvh2" else "k_ao_persistent",
             "resolve": "k_resolve"}[dom]
    if args.workload == "gi" and dom == "ao":
        kname = "k_gi_gather"
    traffic_key = "surface-edges" if (args.workload == "surface" and not args.no_edges) else args.workload
    per_launch = measured_traffic(traffic_key, args.layout, kname)
    launches = kn[dom] // args.steps
    roofline = {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        # HBM bytes per launch (mean over the frame's launches, one per
        # transparency layer) beside the algorithmic bytes per launch
        "traffic": None if per_launch is None else int(per_launch),
        "algorithmic_bytes_per_launch": int(bytes_per_frame / max(1, launches)),
        "traffic_per_frame": None if per_launch is None else int(per_launch * launches),
        "kernel": kname,
        "compute": measured_compute(traffic_key, kname),
        "kernel_ms_per_frame": round(dur_s * 1e3, 4),
        "launches_per_frame": kn[dom] // args.steps,
        "algorithmic_bytes_per_frame": int(bytes_per_frame),
        "per_ray": {"node_visits": round(nodes / max(1, rays_in - st["ao_rays_sampled"] if dom == "ao" else rays_in), 2),
                    "prim_tests": round(tests / max(1, rays_in - st["ao_rays_sampled"] if dom == "ao" else rays_in), 2),
                    "node_bytes": node_b, "prim_bytes": prim_b, "ray_bytes": 48},
    }

    # the same frame through the BVH8 built on the GPU (the C++ facade's
    # builder): build time and frame rate
    gpu_build = None
    if args.layout == "bvh8":
        bg = BuiltScene(dev, scene, capi.BVH8_GPU_BUILD)
        try:
            for _ in range(max(1, args.warmup)):
                bg.render(scene, opt, outputs=False)
            tg = time.perf_counter()
            for _ in range(args.steps):
                fg = bg.render(scene, opt, outputs=False)
            eg = time.perf_counter() - tg
            assert fg.total_rays == rays_per_frame
            gpu_build = {"build_ms": round(bg.info["build_seconds"] * 1e3, 1), "bvh_nodes": bg.info["n_nodes"],
                         "bvh_depth": bg.info["bvh_depth"], "ms_per_frame": round(eg / args.steps * 1e3, 3),
                         "mrays_per_s": round(rays_per_frame * args.steps / eg / 1e6, 1)}
            log(f"GPU-built BVH8: build {gpu_build['build_ms']} ms, {gpu_build['ms_per_frame']} ms/frame, "
                f"{gpu_build['mrays_per_s']} Mrays/s")
        finally:
            bg.close()

    if args.calib_copy_mib > 0:
        from umbreon_amd.counters import calib_copy
        calib_copy(args.calib_copy_mib, log=log)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(scene, opt, args.cpu_seconds)

    t40 = None
    if not args.no_denoise_metric:
        t40 = time_to_40db(dev)

    serve_res = None
    if not args.no_serve:
        serve_res = serving(dev)

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded protein generator, seed 0)",
        "config": {
            "workload": workload,
            "hi_res": [opt.width * opt.supersample, opt.height * opt.supersample],
            "rays_per_frame": rays_per_frame,
            "layers": fr.layers,
            "ray_split": {"primary": st["primary_rays"], "ao_traced": st["ao_rays_traced"],
                          "ao_sampled": st["ao_rays_sampled"], "shadow": st["shadow_rays"],
                          **({"gi_gather": st["gi_rays"], "gi_nee": st["gi_nee_rays"],
                              "gi_gbuffer": st["gi_gbuffer_rays"]} if args.workload == "gi" else {})},
            "bvh_layout": args.layout,
            "bvh_build_ms": round(info["build_seconds"] * 1e3, 2),
            "gpu_built_bvh8": gpu_build,
            "bvh_nodes": info["n_nodes"],
            "bvh_depth": info["bvh_depth"],
            "object_space_edges": edges_info,
            "parallelism": "single GPU",
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "time_to_40db": t40,
        "serving": serve_res,
    }
    print(json.dumps(out), flush=True)
    bs.close()
    dev.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
