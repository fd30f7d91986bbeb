#!/usr/bin/env python3
"""Benchmark of the path-tracing hot path on one MI355X.

Workload (default): unidirectional MIS path tracing (light + BSDF sampling, power heuristic) of
the room with 200 displaced GGX icospheres (about 1.3M triangles), 1024 x 1024 pixels, 64 samples
per pixel, at most 8 bounces, FP32. One step = one render of the whole film; the scene (BVH,
triangles, materials, lights) is resident in HBM before the timed region and the film stays in
HBM. Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--workload spheres|box|indirect] [--form plain]
                    [--spp N] [--no-cpu-baseline]
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

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector (dense)
# Algorithmic work per ray query (layout independent): an interior visit reads two child boxes
# (2 x 24 B) and their links (2 x 4 B) and does 2 x (6 sub + 6 mul); a triangle test reads three
# vertices (36 B) and an id (4 B) and does the 53 flops of ray::intersect (tensor.h:688-725).
BYTES_PER_NODE, BYTES_PER_TEST = 56, 40
FLOPS_PER_NODE, FLOPS_PER_TEST = 24, 53


def workload(name, spp):
    from x3dcore2_amd import scenes

    if name == "spheres":
        d = scenes.spheres_scene()
    elif name == "box":
        d = scenes.box_scene()
    elif name == "indirect":
        d = scenes.indirect_scene()
    else:
        raise SystemExit(f"unknown workload {name}")
    if spp:
        d.spp = spp
    return d


def cpu_baseline(desc, budget_s=12.0):
    """Times the reference renderer itself (oracle/_ref, built from the reference's sources) on
    the host: unidirect_lt1 (its NEE tracer) on the same scene and camera at a reduced resolution,
    threads = host cores as pt_image_renderer does. Returns None when the build is absent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import bindings

    if not bindings.reference_available():
        return None
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    ref = bindings.Reference(desc)
    side = 32
    t0 = time.perf_counter()
    _, taken = ref.render(bindings.TR_UNIDIRECT_LT1, side, side, threads, threads=threads)
    dt = time.perf_counter() - t0
    # scale the pixel count to spend about budget_s seconds
    scale = max(1.0, budget_s / max(dt, 1e-3))
    side = int(min(desc.width, max(32, side * scale ** 0.5)))
    t0 = time.perf_counter()
    _, taken = ref.render(bindings.TR_UNIDIRECT_LT1, side, side, threads, threads=threads)
    dt = time.perf_counter() - t0
    samples = side * side * taken
    return {
        "value": samples / dt / 1e6,
        "unit": "Msamples/s",
        "cores": threads,
        "kind": "reference",
        "sample": (f"reference unidirect_lt1 (Russian roulette, NEE, no MIS) on the same scene and camera at "
                   f"{side}x{side}, {taken} samples per pixel ({threads} tasks x 1), {samples} samples in {dt:.1f} s "
                   "incl. its serial first-hit pass; reference BVH build not timed"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="spheres")
    ap.add_argument("--form", default="plain", choices=["plain", "wavefront"])
    ap.add_argument("--spp", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run one device-to-device copy of BYTES (to calibrate rocprof byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        from x3dcore2_amd.device import DeviceArray, copy_device, synchronize

        a, b = DeviceArray(args.copy_calibration), DeviceArray(args.copy_calibration)
        copy_device(b, a, args.copy_calibration)
        synchronize()
        print(json.dumps({"copy_calibration_bytes": args.copy_calibration}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the renderer runs on a single MI355X "
              "(sample-split sharding across GPUs is not implemented)", file=sys.stderr)
        sys.exit(2)

    import x3dcore2_amd as x3d
    from x3dcore2_amd.device import DeviceArray, synchronize

    desc = workload(args.workload, args.spp)
    w, h, spp, depth = desc.width, desc.height, desc.spp, desc.max_depth
    print(f"[bench] workload {desc.name}: {desc.num_tris} tris, {w}x{h}, {spp} spp, depth {depth}", file=sys.stderr,
          flush=True)
    scene = x3d.Scene(desc)
    info = scene.info()
    print(f"[bench] BVH build {info.build_seconds:.2f} s, upload {info.upload_seconds:.2f} s, "
          f"{info.device_bytes / 2**20:.1f} MiB in HBM, max leaf depth {info.max_leaf_depth}", file=sys.stderr,
          flush=True)
    form = x3d.FORM_PLAIN if args.form == "plain" else x3d.FORM_WAVEFRONT
    r = x3d.Renderer(x3d.PT_UNIDIRECT_MIS, form=form, max_depth=depth, seed=1)
    film = DeviceArray(w * h * 16)
    for k in range(args.warmup):
        st = r.render_device(scene, desc.camera, w, h, spp, film.ptr)
        print(f"[bench] warmup {k}: {st.kernel_ms:.1f} ms", file=sys.stderr, flush=True)
    synchronize()
    stats = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        st = r.render_device(scene, desc.camera, w, h, spp, film.ptr)
        stats.append(st)
        print(f"[bench] step {k}: {st.kernel_ms:.1f} ms (trace {st.trace_ms:.1f} ms)", file=sys.stderr, flush=True)
    synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    samples = w * h * spp * args.steps
    rays = sum(s.closest_queries + s.anyhit_queries for s in stats)
    img = film.to_host(np.float32, (h * w, 4))
    if not np.isfinite(img).all():
        raise SystemExit("non-finite film")

    # Dominant kernel: the sample kernel (bounce loop). Its algorithmic bytes / flops per launch
    # come from the traversal counters of the same launch; its duration from HIP events on its
    # stream.
    s0 = stats[-1]
    nodes = (s0.closest_nodes - s0.first_hit_nodes) + s0.anyhit_nodes
    tests = (s0.closest_tests - s0.first_hit_tests) + s0.anyhit_tests
    trace_ms = float(np.mean([s.trace_ms for s in stats]))
    alg_bytes = BYTES_PER_NODE * nodes + BYTES_PER_TEST * tests
    alg_flops = FLOPS_PER_NODE * nodes + FLOPS_PER_TEST * tests
    achieved_gbs = alg_bytes / (trace_ms * 1e-3) / 1e9
    achieved_tflops = alg_flops / (trace_ms * 1e-3) / 1e12
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{desc.name}_{args.form}.md")
    if os.path.exists(tpath):
        for line in open(tpath):
            if line.startswith("hbm_bytes_per_launch:"):
                traffic = float(line.split(":", 1)[1])

    out = {
        "metric": "Msamples/s & Mrays/s at 1024^2 vs plain GPU path at equal RMSE; grad evals/s",
        "value": samples / elapsed / 1e6,
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene)",
        "config": {"workload": f"{desc.name}: room + 200 displaced GGX icospheres" if desc.name == "spheres"
                   else desc.name, "triangles": desc.num_tris, "width": w, "height": h, "spp": spp,
                   "max_depth": depth, "tracer": "unidirectional MIS (NEE + BSDF, power heuristic)",
                   "form": args.form},
        "mrays_per_s": rays / elapsed / 1e6,
        "rays_per_sample": rays / samples,
        "kernel_ms_per_step": float(np.mean([s.kernel_ms for s in stats])),
        "roofline": {
            "bound": "hbm",
            "kernel": "k_render_mis" if args.form == "plain" else "wavefront",
            "achieved": achieved_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved_gbs / HBM_PEAK_GBS,
            "traffic": traffic,
            "alg_bytes_per_launch": alg_bytes,
            "launch_ms": trace_ms,
            "valu_tflops": achieved_tflops,
            "valu_frac": achieved_tflops / FP32_PEAK_TFLOPS,
            "node_visits_per_launch": nodes,
            "tri_tests_per_launch": tests,
        },
        "bvh_build_s": info.build_seconds,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(desc)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
