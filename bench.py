#!/usr/bin/env python3
"""Benchmark of the path-tracing hot path on one MI355X.

Workload (default): unidirectional MIS path tracing (light + BSDF sampling, power heuristic) of
the room with 200 displaced GGX icospheres (about 1.3M triangles), 1024 x 1024 pixels, 64 samples
per pixel, at most 8 bounces, FP32. One step = one render of the whole film; the scene (BVH,
triangles, materials, lights) is resident in HBM before the timed region and the film stays in
HBM. Both execution forms are timed on the same workload — the plain one-thread-per-path kernel
(the drop-in) and the autotuned wavefront form — and their films are checked to be bit-identical
(so the RMSE against any reference is equal). `value` is the wavefront form's samples/s. The
reference's own lt1 tracer is also timed in both forms at 16 spp (`lt1_tracer`), and the line
carries two more configurations' rates, a few seconds each: the bidirectional tracer on the
indirectly lit room at 512 x 512 x 32 spp with the reference's own BDPT timed on the host beside
it (`bdpt_512x512x32`), gradient evaluations of the 256 x 256 x 32 spp inverse-rendering room
(`adjoint_256x256x32`), and the 64-scene batch in one call against its jobs one after another
(`batch_64scenes`, about 20 s of it building the scenes). Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--workload spheres|box|indirect|batch|adjoint] [--spp N]
                    [--no-cpu-baseline] [--no-autotune | --wf-config WG,DEPTH,SORT,CAP]
                    [--probe-spp N] [--form both|plain|wavefront] [--no-ablation] [--no-legs]
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
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector (dense; counts an FMA as 2 and packed dual issue)
# VALU issue ceiling: 256 CUs x 4 SIMDs, each issuing one 64-lane VALU instruction per 4 cycles
# (16 lanes per cycle) at 2.4 GHz: 39.3 T lane-operations/s (= 157.3 TFLOP/s / 2 (FMA) / 2 (packed)).
VALU_LANE_OPS_PEAK_T = 1024 * 16 * 2.4e9 / 1e12
# Algorithmic work per ray query (layout independent): an interior visit reads two child boxes
# (2 x 24 B) and their links (2 x 4 B) and does 2 x (6 sub + 6 mul); a triangle test reads three
# vertices (36 B) and an id (4 B) and does the 53 flops of ray::intersect (tensor.h:688-725).
BYTES_PER_NODE, BYTES_PER_TEST = 56, 40
FLOPS_PER_NODE, FLOPS_PER_TEST = 24, 53


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


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


def cpu_model() -> str:
    """The host CPU as /proc/cpuinfo names it, with its x86 family / model numbers (a virtual
    machine's model name may be generic) and the logical CPUs the OS shows."""
    fields = {}
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            fields.setdefault(k.strip(), v.strip())
    except OSError:
        return "unknown"
    name = fields.get("model name", "unknown")
    extra = [f"family {fields['cpu family']} model {fields['model']}"] if "cpu family" in fields and "model" in fields else []
    extra.append(f"{os.cpu_count()} logical CPUs visible")
    return f"{name} ({', '.join(extra)})"


PROFILE_KEYS = ("kernel", "config", "hbm_bytes_per_launch", "valu_insts_per_launch", "valu_lane_ops_per_launch",
                "cycles_per_launch")


def profile_fields(name, kernel, config):
    """The machine-readable `key: value` lines at the end of profiles/<name>, or None unless the
    summary was taken on the same kernel instance and launch configuration."""
    tpath = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(tpath):
        return None
    fields = {}
    for line in open(tpath):
        k, sep, v = line.partition(":")
        if sep and k in PROFILE_KEYS:
            fields[k] = v.strip()
    if fields.get("kernel") != kernel or fields.get("config") != config:
        return None
    return fields


def counters_from_profile(name, kernel, config):
    """Per-launch VALU counters recorded in profiles/<name> (rocprofv3 --pmc SQ_INSTS_VALU,
    SQ_THREAD_CYCLES_VALU, GRBM_GUI_ACTIVE), when taken on the same kernel and configuration."""
    fields = profile_fields(name, kernel, config)
    try:
        return {k: float(fields[k]) for k in ("valu_insts_per_launch", "valu_lane_ops_per_launch", "cycles_per_launch")}
    except (KeyError, ValueError, TypeError):
        return None


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
    scale = max(1.0, budget_s / max(dt, 1e-3))  # scale the pixel count to ~budget_s seconds
    side = int(min(desc.width, max(32, side * scale ** 0.5)))
    t0 = time.perf_counter()
    _, taken = ref.render(bindings.TR_UNIDIRECT_LT1, side, side, threads, threads=threads)
    dt = time.perf_counter() - t0
    samples = side * side * taken
    return {
        "value": samples / dt / 1e6,
        "unit": "Msamples/s",
        "cores": threads,
        "cpu_model": cpu_model(),
        "kind": "reference",
        "sample": (f"reference unidirect_lt1 (Russian roulette, NEE, no MIS) on the same scene and camera at "
                   f"{side}x{side}, {taken} samples per pixel ({threads} tasks x 1), {samples} samples in {dt:.1f} s "
                   "incl. its serial first-hit pass; reference BVH build not timed"),
    }


def timed_renders(x3d, r, scene, desc, film, steps, warmup, name):
    from x3dcore2_amd.device import synchronize

    w, h, spp = desc.width, desc.height, desc.spp
    for k in range(warmup):
        st = r.render_device(scene, desc.camera, w, h, spp, film.ptr)
        log(f"{name} warmup {k}: {st.kernel_ms:.1f} ms")
    synchronize()
    stats = []
    t0 = time.perf_counter()
    for k in range(steps):
        st = r.render_device(scene, desc.camera, w, h, spp, film.ptr)
        stats.append(st)
        log(f"{name} step {k}: {st.kernel_ms:.1f} ms (sample kernels {st.trace_ms:.1f} ms)")
    synchronize()
    return time.perf_counter() - t0, stats


def traffic_from_profile(name, kernel, config):
    """HBM bytes per launch recorded in profiles/<name> (copy-calibrated FETCH_SIZE + WRITE_SIZE),
    used only when the summary was taken on the same kernel instance and launch configuration."""
    fields = profile_fields(name, kernel, config)
    try:
        return float(fields["hbm_bytes_per_launch"])
    except (KeyError, ValueError, TypeError):
        return None


def leg_bdpt(x3d, steps, warmup, spp=0, cpu=True, budget_s=8.0, plain=True):
    """The bidirectional tracer with the reference's uniform strategy weights on the indirectly lit
    room with four 20k-triangle GGX spheres, 512 x 512, 32 spp, subpaths of up to 6 vertices (film in
    HBM), in the wavefront form (the default) and the plain form, whose films must be identical;
    beside it the reference's own bidirect_mis_path_tracer (oracle/_ref, its fixed 8-vertex
    subpaths) on the host cores at a reduced resolution sized to ~budget_s. plain=False skips the
    plain form."""
    from x3dcore2_amd import scenes
    from x3dcore2_amd.device import DeviceArray, synchronize

    desc = scenes.indirect_scene()
    scene = x3d.Scene(desc)
    w, h, spp = desc.width, desc.height, spp or desc.spp

    def timed(form, n):
        r = x3d.Renderer(x3d.PT_BIDIRECT_MIS, form=form, max_path_len=6, seed=1)
        film = DeviceArray(w * h * 16)
        for _ in range(warmup):
            r.render_device(scene, desc.camera, w, h, spp, film.ptr)
        synchronize()
        t0 = time.perf_counter()
        stats = [r.render_device(scene, desc.camera, w, h, spp, film.ptr) for _ in range(n)]
        synchronize()
        return time.perf_counter() - t0, stats, film.to_host(np.float32, (w * h, 4))

    dt, stats, img = timed(x3d.FORM_WAVEFRONT, steps)
    rays = sum(st.closest_queries + st.anyhit_queries for st in stats)
    res = {"workload": f"indirect room + 4 GGX spheres ({desc.num_tris} tris)", "width": w, "height": h, "spp": spp,
           "max_path_len": 6, "form": "wavefront", "msamples_per_s": w * h * spp * steps / dt / 1e6,
           "ms_per_step": dt * 1e3 / steps, "mrays_per_s": rays / dt / 1e6,
           "rays_per_sample": rays / (w * h * spp * steps)}
    if plain:
        dp, _, img_p = timed(x3d.FORM_PLAIN, max(1, steps // 2))
        res["plain_ms_per_step"] = dp * 1e3 / max(1, steps // 2)
        res["films_bit_identical"] = bool(np.array_equal(img.view(np.uint32), img_p.view(np.uint32)))
        res["speedup_vs_plain"] = res["plain_ms_per_step"] / res["ms_per_step"]
        if not res["films_bit_identical"]:
            raise SystemExit("plain and wavefront bidirectional films differ")
    log(f"bdpt leg: {res['ms_per_step']:.1f} ms per render ({res.get('plain_ms_per_step', 0):.1f} plain), "
        f"{res['msamples_per_s']:.1f} Msamples/s")
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import bindings

        if bindings.reference_available():
            threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
            ref = bindings.Reference(desc)
            side = 16
            t0 = time.perf_counter()
            ref.render(bindings.TR_BIDIRECT_MIS, side, side, threads, threads=threads)
            d0 = time.perf_counter() - t0
            side = int(min(w, max(16, side * (budget_s / max(d0, 1e-3)) ** 0.5)))
            t0 = time.perf_counter()
            _, taken = ref.render(bindings.TR_BIDIRECT_MIS, side, side, threads, threads=threads)
            d1 = time.perf_counter() - t0
            res["cpu_reference"] = {"msamples_per_s": side * side * taken / d1 / 1e6, "cores": threads,
                                    "cpu_model": cpu_model(), "kind": "reference",
                                    "sample": f"reference bidirect_mis_path_tracer (8-vertex subpaths) at {side}x{side}, "
                                              f"{taken} samples per pixel, {d1:.1f} s"}
    return res


def leg_batch(x3d, steps, warmup):
    """The 64-scene batch (1k-2M triangles per scene, 14.2M in all, 128^2-1024^2, 4-64 spp,
    depth 6) in one x3d_render_batch call in the wavefront form, against the same jobs rendered
    one after another in the wavefront form (one untimed pass first)."""
    from x3dcore2_amd import scenes
    from x3dcore2_amd.device import DeviceArray, synchronize

    descs = scenes.batch_scenes()
    built = [x3d.Scene(d) for d in descs]
    films = [DeviceArray(d.width * d.height * 16) for d in descs]
    items = [(built[k], d.camera, d.width, d.height, d.spp) for k, d in enumerate(descs)]
    ptrs = [f.ptr for f in films]
    samples = sum(d.width * d.height * d.spp for d in descs)
    r = x3d.Renderer(x3d.PT_UNIDIRECT_MIS, form=x3d.FORM_WAVEFRONT, max_depth=6, seed=1)
    for timed in (False, True):
        synchronize()
        t0 = time.perf_counter()
        for k, d in enumerate(descs):
            r.render_device(built[k], d.camera, d.width, d.height, d.spp, ptrs[k])
        synchronize()
    seq = time.perf_counter() - t0
    for _ in range(warmup):
        r.render_batch_device(items, ptrs)
    synchronize()
    t0 = time.perf_counter()
    res = [r.render_batch_device(items, ptrs) for _ in range(steps)]
    synchronize()
    dt = (time.perf_counter() - t0) / steps
    st, lat = res[-1]
    out = {"workload": f"64 seeded scenes, {sum(d.num_tris for d in descs)} tris", "samples": samples,
           "max_depth": 6, "form": "wavefront (every job's paths in shared generations)",
           "msamples_per_s": samples / dt / 1e6, "ms_per_step": dt * 1e3, "kernel_ms": st.kernel_ms,
           "generations": st.generations, "sequential_wavefront_ms": seq * 1e3, "speedup_vs_sequential": seq / dt,
           "latency_ms": {"p50": float(np.percentile(lat, 50)), "p90": float(np.percentile(lat, 90))}}
    log(f"batch leg: {dt * 1e3:.1f} ms per batch ({seq * 1e3:.1f} ms one after another)")
    return out


def leg_adjoint(x3d, steps, warmup, spp=0):
    """Gradient evaluations of the inverse-rendering room (16 surfaces x (albedo rgb, GGX
    roughness), 256 x 256, 32 spp, depth 4): forward render, L2 loss on the host, path-replay
    backward, per step."""
    from x3dcore2_amd import scenes

    ta, tr = scenes.adjoint_truth()
    desc = scenes.adjoint_scene()
    scene = x3d.Scene(desc)
    w, h, spp, depth = desc.width, desc.height, spp or desc.spp, desc.max_depth
    r = x3d.Renderer(x3d.PT_UNIDIRECT_MIS, max_depth=depth, seed=1)
    scenes.set_adjoint_params(scene, ta, tr)
    target, _ = r.forward(scene, desc.camera, w, h, spp)
    scenes.set_adjoint_params(scene, np.full((16, 3), 0.5), np.full(16, 0.4))

    def step():
        img, sf = r.forward(scene, desc.camera, w, h, spp)
        dl = np.zeros_like(img)
        dl[:, :3] = 2 * (img[:, :3] - target[:, :3])
        g, sb = r.backward(scene, desc.camera, w, h, spp, dl)
        return sf, sb

    for _ in range(warmup):
        step()
    t0 = time.perf_counter()
    st = [step() for _ in range(steps)]
    dt = time.perf_counter() - t0
    bwd = float(np.mean([b.kernel_ms for _, b in st]))
    res = {"workload": "box room, 16 surfaces x (albedo rgb, GGX roughness)", "width": w, "height": h, "spp": spp,
           "max_depth": depth, "parameters": 64, "grad_evals_per_s": steps / dt, "ms_per_step": dt * 1e3 / steps,
           "forward_kernel_ms": float(np.mean([a.kernel_ms for a, _ in st])), "backward_kernel_ms": bwd,
           "backward_msamples_per_s": w * h * spp / (bwd * 1e-3) / 1e6}
    log(f"adjoint leg: {res['grad_evals_per_s']:.1f} grad evals/s")
    return res


def bench_secondary(args):
    """Secondary workloads (one JSON line each; not the headline metric):
    indirect: the bidirectional tracer on the indirectly lit room with four 20k-triangle GGX
              spheres, 512 x 512, 32 spp, subpaths of up to 6 vertices;
    batch:    64 seeded scenes (1k-2M triangles each, 128^2-1024^2, 4-64 spp, depth 6) in one
              x3d_render_batch call (wavefront form: every job's paths in shared generations),
              against rendering the same jobs one after another and the plain form's batch;
    adjoint:  gradient evaluations (forward render, L2 loss on the host, path-replay backward)
              of the 16-surface inverse-rendering room, 256 x 256, 32 spp, depth 4."""
    import x3dcore2_amd as x3d
    from x3dcore2_amd import scenes
    from x3dcore2_amd.device import DeviceArray, synchronize

    out = {"n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded procedural scenes)"}
    if args.workload == "indirect":
        res = leg_bdpt(x3d, args.steps, args.warmup, args.spp, cpu=not args.no_cpu_baseline,
                       plain=args.form != "wavefront")
        out.update(metric="bidirectional path tracing Msamples/s", unit="Msamples/s", value=res["msamples_per_s"],
                   ms_per_step=res["ms_per_step"], mrays_per_s=res["mrays_per_s"],
                   rays_per_sample=res["rays_per_sample"], plain_ms_per_step=res.get("plain_ms_per_step"),
                   speedup_vs_plain=res.get("speedup_vs_plain"), films_bit_identical=res.get("films_bit_identical"),
                   cpu_reference=res.get("cpu_reference"),
                   config={"workload": res["workload"], "width": res["width"], "height": res["height"],
                           "spp": res["spp"], "max_path_len": 6, "form": "wavefront",
                           "tracer": "bidirectional, reference uniform strategy weights"})
        print(json.dumps(out), flush=True)
        return
    if args.workload == "adjoint":
        res = leg_adjoint(x3d, args.steps, args.warmup, args.spp)
        out.update(metric="grad evals/s (forward + path-replay backward, host loss)", unit="grad evals/s",
                   value=res["grad_evals_per_s"], ms_per_step=res["ms_per_step"],
                   forward_kernel_ms=res["forward_kernel_ms"], backward_kernel_ms=res["backward_kernel_ms"],
                   backward_msamples_per_s=res["backward_msamples_per_s"],
                   config={k: res[k] for k in ("workload", "width", "height", "spp", "max_depth", "parameters")})
        print(json.dumps(out), flush=True)
        return
    descs = scenes.batch_scenes()
    t0 = time.perf_counter()
    built = [x3d.Scene(d) for d in descs]
    scene_bytes = sum(b.info().device_bytes for b in built)
    log(f"batch: {len(descs)} scenes, {sum(d.num_tris for d in descs)} tris, built in {time.perf_counter() - t0:.1f} s, "
        f"{scene_bytes / 2**20:.0f} MiB in HBM")
    films = [DeviceArray(d.width * d.height * 16) for d in descs]
    items = [(built[k], d.camera, d.width, d.height, d.spp) for k, d in enumerate(descs)]
    ptrs = [f.ptr for f in films]
    samples = sum(d.width * d.height * d.spp for d in descs)
    rp = x3d.Renderer(x3d.PT_UNIDIRECT_MIS, max_depth=6, seed=1)  # plain form
    rw = x3d.Renderer(x3d.PT_UNIDIRECT_MIS, form=x3d.FORM_WAVEFRONT, max_depth=6, seed=1)

    def sequential(rr):
        """One job after another through the single-job entry point (one untimed pass first)."""
        for timed in (False, True):
            synchronize()
            t0 = time.perf_counter()
            for k, d in enumerate(descs):
                rr.render_device(built[k], d.camera, d.width, d.height, d.spp, ptrs[k])
            synchronize()
        return time.perf_counter() - t0

    seq = sequential(rp)
    seq_wf = sequential(rw)
    log(f"sequential single-job renders: plain {seq * 1e3:.1f} ms, wavefront {seq_wf * 1e3:.1f} ms")

    def batches(rr, steps, warmup):
        for _ in range(warmup):
            rr.render_batch_device(items, ptrs)
        synchronize()
        t0 = time.perf_counter()
        res = [rr.render_batch_device(items, ptrs) for _ in range(steps)]
        synchronize()
        return time.perf_counter() - t0, res[-1]

    # the plain form's batch (shared launch of the small jobs, large jobs alone), for comparison
    dt_plain, (st_plain, _) = batches(rp, 1, 1)
    log(f"plain-form batch: {dt_plain * 1e3:.1f} ms ({st_plain.alone_jobs} jobs alone)")
    tune = {}
    if not args.no_autotune:
        # probe with an eighth of each job's samples: same scenes and packing problem
        rep = rw.autotune_batch([(sc, cam, w, h, max(1, spp // 8)) for (sc, cam, w, h, spp) in items])
        tune = {"workgroup_size": rep.workgroup_size, "stack_depth": rep.stack_depth,
                "candidates": rep.candidates, "best_ms": rep.best_ms, "default_ms": rep.default_ms}
        log(f"batch autotune: {tune}")
    # the wavefront form's batch (every job's paths in shared generations): the measured line
    dt, (st, lat) = batches(rw, args.steps, args.warmup)
    lat = np.array(lat)
    out.update(metric="batched rendering Msamples/s (64 scenes)", unit="Msamples/s",
               value=samples * args.steps / dt / 1e6, ms_per_step=dt * 1e3 / args.steps,
               kernel_ms=st.kernel_ms, sequential_ms=seq * 1e3, sequential_wavefront_ms=seq_wf * 1e3,
               plain_form_batch_ms=dt_plain * 1e3,
               speedup_vs_sequential=min(seq, seq_wf) / (dt / args.steps),
               mrays_per_s=(st.closest_queries + st.anyhit_queries) / (st.kernel_ms * 1e-3) / 1e6,
               latency_ms={"p50": float(np.percentile(lat, 50)), "p90": float(np.percentile(lat, 90)),
                           "max": float(lat.max())},
               autotune=tune,
               config={"workload": f"64 seeded scenes, {sum(d.num_tris for d in descs)} tris", "samples": samples,
                       "max_depth": 6, "form": "wavefront (every job's paths in shared generations)",
                       "generations": st.generations, "tracer": "unidirectional MIS",
                       "scene_bytes": scene_bytes})
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="spheres")
    ap.add_argument("--form", default="both", choices=["both", "plain", "wavefront"])
    ap.add_argument("--spp", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-ablation", action="store_true", help="skip the sorted-ray-queue ablation renders")
    ap.add_argument("--no-legs", action="store_true",
                    help="only the headline workload: skip the compositor, lt1, bidirectional and adjoint legs")
    ap.add_argument("--wf-config", default="", metavar="WG,DEPTH,SORT,CAP",
                    help="fixed wavefront configuration instead of autotuning (SORT: material|none)")
    ap.add_argument("--probe-spp", type=int, default=16,
                    help="samples per pixel of each autotuner probe render (16 at 1024^2 = 16M paths)")
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

    if args.workload in ("indirect", "batch", "adjoint"):
        bench_secondary(args)
        return
    import x3dcore2_amd as x3d
    from x3dcore2_amd.device import DeviceArray

    desc = workload(args.workload, args.spp)
    w, h, spp, depth = desc.width, desc.height, desc.spp, desc.max_depth
    log(f"workload {desc.name}: {desc.num_tris} tris, {w}x{h}, {spp} spp, depth {depth}")
    scene = x3d.Scene(desc)
    info = scene.info()
    log(f"BVH build {info.build_seconds:.2f} s, upload {info.upload_seconds:.2f} s, "
        f"{info.device_bytes / 2**20:.1f} MiB in HBM, max leaf depth {info.max_leaf_depth}")
    samples = w * h * spp * args.steps
    out = {
        "metric": "Msamples/s & Mrays/s at 1024^2 vs plain GPU path at equal RMSE; grad evals/s",
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene)",
        "config": {"workload": (f"{desc.name}: room + 200 displaced GGX icospheres" if desc.name == "spheres"
                                else desc.name), "triangles": desc.num_tris, "width": w, "height": h, "spp": spp,
                   "max_depth": depth, "tracer": "unidirectional MIS (NEE + BSDF, power heuristic)"},
        "bvh_build_s": info.build_seconds,
    }
    films = {}
    results = {}
    forms = ["plain", "wavefront"] if args.form == "both" else [args.form]
    for form in forms:
        film = DeviceArray(w * h * 16)
        if form == "plain":
            r = x3d.Renderer(x3d.PT_UNIDIRECT_MIS, max_depth=depth, seed=1)
        else:
            r = x3d.Renderer(x3d.PT_UNIDIRECT_MIS, form=x3d.FORM_WAVEFRONT, max_depth=depth, seed=1)
            if args.wf_config:
                wg, sd, so, cap = args.wf_config.split(",")
                out["autotune"] = {"workgroup_size": int(wg), "stack_depth": int(sd), "sort_key": so,
                                   "queue_capacity": int(cap), "candidates": 0}
                r = x3d.Renderer(x3d.PT_UNIDIRECT_MIS, form=x3d.FORM_WAVEFRONT, max_depth=depth, seed=1,
                                 workgroup_size=int(wg), stack_depth=int(sd), queue_capacity=int(cap),
                                 sort_key={"material": 1, "none": 2}[so])
            elif not args.no_autotune:
                rep = r.autotune(scene, desc.camera, w, h, args.probe_spp)
                out["autotune"] = {"workgroup_size": rep.workgroup_size, "stack_depth": rep.stack_depth,
                                   "sort_key": ["default", "material", "none"][rep.sort_key],
                                   "queue_capacity": rep.queue_capacity, "candidates": rep.candidates,
                                   "probe_ms_best": rep.best_ms, "probe_ms_default": rep.default_ms}
                log(f"autotune: {out['autotune']}")
        elapsed, stats = timed_renders(x3d, r, scene, desc, film, args.steps, args.warmup, form)
        films[form] = film.to_host(np.float32, (h * w, 4))
        if not np.isfinite(films[form]).all():
            raise SystemExit(f"non-finite film ({form})")
        rays = sum(s.closest_queries + s.anyhit_queries for s in stats)
        results[form] = dict(elapsed=elapsed, stats=stats, rays=rays, renderer=r, film=film)
        out[f"{form}_msamples_per_s"] = samples / elapsed / 1e6
        out[f"{form}_mrays_per_s"] = rays / elapsed / 1e6
        out[f"{form}_ms_per_step"] = elapsed * 1e3 / args.steps

    best = "wavefront" if "wavefront" in results else "plain"
    R = results[best]
    out["value"] = samples / R["elapsed"] / 1e6
    out["ms_per_step"] = R["elapsed"] * 1e3 / args.steps
    out["mrays_per_s"] = R["rays"] / R["elapsed"] / 1e6
    out["rays_per_sample"] = R["rays"] / samples
    out["config"]["form"] = best
    if len(results) == 2:
        out["speedup_vs_plain"] = results["plain"]["elapsed"] / results["wavefront"]["elapsed"]
        out["films_bit_identical"] = bool(np.array_equal(films["plain"].view(np.uint32),
                                                         films["wavefront"].view(np.uint32)))
        if not out["films_bit_identical"]:
            raise SystemExit("plain and wavefront films differ")

    # Roofline of the dominant kernel. Algorithmic bytes / flops per launch come from the
    # traversal counters of the same launches (56 B + 24 flop per node visit, 40 B + 53 flop per
    # triangle test); the duration from HIP events on the kernel's stream.
    s0 = R["stats"][-1]
    if best == "wavefront":
        # closest-hit traversal kernel: one profiled render with the same configuration (events
        # around each launch, no graph)
        tuned = out.get("autotune", {})
        wg, sd = tuned.get("workgroup_size", 0), tuned.get("stack_depth", 0)
        sort = {"default": 0, "material": 1, "none": 2}[tuned.get("sort_key", "default")]
        rp = x3d.Renderer(x3d.PT_UNIDIRECT_MIS, form=x3d.FORM_WAVEFRONT, max_depth=depth, seed=1, workgroup_size=wg,
                          queue_capacity=tuned.get("queue_capacity", 0), sort_key=sort, stack_depth=sd, profile=1)
        sp = rp.render_device(scene, desc.camera, w, h, spp, R["film"].ptr)
        launches = max(sp.trace_launches, 1)
        nodes = (sp.closest_nodes - sp.first_hit_nodes) / launches
        tests = (sp.closest_tests - sp.first_hit_tests) / launches
        launch_ms = sp.extend_ms / launches
        # the closest-hit kernel's own LDS stack depth: 12 beside the any-hit kernels' 16
        ed = 12 if (sd or 16) == 16 else sd
        kernel = f"k_wf_extend<{wg or 64},{ed}>"
        launch_cfg = (f"{wg or 64},{sd or 16},{'material' if sort == 1 else 'none'},"
                      f"{tuned.get('queue_capacity', 0) or 1 << 26}")
        traffic = traffic_from_profile(f"traffic_{desc.name}_wavefront.md", kernel, launch_cfg)
        counters_profile = f"counters_{desc.name}_wavefront.md"
        out["wavefront_kernel_ms"] = {"extend": sp.extend_ms, "shadow": sp.shadow_ms, "launches": sp.trace_launches,
                                      "all_sample_kernels": sp.trace_ms}
    else:
        nodes = (s0.closest_nodes - s0.first_hit_nodes) + s0.anyhit_nodes
        tests = (s0.closest_tests - s0.first_hit_tests) + s0.anyhit_tests
        launch_ms = float(np.mean([s.trace_ms for s in R["stats"]]))
        # one "launch" here is one render (its generations' k_render_mis_items launches summed)
        kernel = "k_render_mis_items"
        launch_cfg = "plain"
        counters_profile = f"k_render_mis_items_{desc.name}_{w}x{h}x{spp}.md"
        traffic = traffic_from_profile(counters_profile, kernel, launch_cfg)
    if best == "wavefront" and not args.no_ablation:
        # ablation: the same configuration with both ray queues sorted by key before traversal
        ra = x3d.Renderer(x3d.PT_UNIDIRECT_MIS, form=x3d.FORM_WAVEFRONT, max_depth=depth, seed=1, workgroup_size=wg,
                          queue_capacity=tuned.get("queue_capacity", 0), sort_key=sort, stack_depth=sd,
                          ray_order=x3d.RAY_ORDER_SORTED)
        ea, _ = timed_renders(x3d, ra, scene, desc, R["film"], 1, 1, "wavefront, sorted rays")
        spa = x3d.Renderer(x3d.PT_UNIDIRECT_MIS, form=x3d.FORM_WAVEFRONT, max_depth=depth, seed=1, workgroup_size=wg,
                           queue_capacity=tuned.get("queue_capacity", 0), sort_key=sort, stack_depth=sd,
                           ray_order=x3d.RAY_ORDER_SORTED, profile=1).render_device(scene, desc.camera, w, h, spp,
                                                                                    R["film"].ptr)
        out["ray_order_ablation"] = {"append_order_ms_per_step": out["ms_per_step"], "sorted_ms_per_step": ea * 1e3,
                                     "sorted_extend_ms": spa.extend_ms, "sorted_shadow_ms": spa.shadow_ms}
        del ra
    alg_bytes = BYTES_PER_NODE * nodes + BYTES_PER_TEST * tests
    alg_flops = FLOPS_PER_NODE * nodes + FLOPS_PER_TEST * tests
    achieved_gbs = alg_bytes / (launch_ms * 1e-3) / 1e9
    achieved_tflops = alg_flops / (launch_ms * 1e-3) / 1e12
    # The ceiling that binds the traversal kernel is VALU issue (the counters show the vector ALUs
    # busy most cycles with under half of each instruction's lanes active), not HBM: most of the
    # node and triangle bytes hit in L2 / the Infinity Cache. `frac` is the kernel's VALU lane
    # operations per second (per-launch count from the committed counter profile of this kernel
    # and configuration, over this run's launch time) against the issue ceiling; the algorithmic
    # HBM and FP32 fractions stay beside it.
    counters = counters_from_profile(counters_profile, kernel, launch_cfg)
    roof = {"kernel": kernel, "launch_ms": launch_ms}
    if counters:
        lane_ops_t = counters["valu_lane_ops_per_launch"] / (launch_ms * 1e-3) / 1e12
        roof.update(bound="valu", achieved=lane_ops_t, peak=VALU_LANE_OPS_PEAK_T, unit="T VALU lane-ops/s",
                    frac=lane_ops_t / VALU_LANE_OPS_PEAK_T,
                    valu_issue_frac=counters["valu_insts_per_launch"] * 4 / (1024 * counters["cycles_per_launch"]),
                    active_lane_frac=counters["valu_lane_ops_per_launch"] / (64 * counters["valu_insts_per_launch"]))
    else:
        roof.update(bound="hbm", achieved=achieved_gbs, peak=HBM_PEAK_GBS, unit="GB/s", frac=achieved_gbs / HBM_PEAK_GBS)
    roof.update({
        "traffic": traffic,
        # algorithmic node / triangle bytes per launch over the launch time, against the HBM peak
        # (above 1: the bytes are served from L2 and the Infinity Cache)
        "hbm_alg_frac": achieved_gbs / HBM_PEAK_GBS,
        # real HBM bytes (copy-calibrated PMC) per launch over the launch time, against the peak
        "hbm_traffic_frac": (traffic / (launch_ms * 1e-3) / 1e9 / HBM_PEAK_GBS) if traffic else None,
        "alg_bytes_per_launch": alg_bytes,
        "alg_tflops": achieved_tflops,
        "alg_flop_frac": achieved_tflops / FP32_PEAK_TFLOPS,
        "node_visits_per_launch": nodes,
        "tri_tests_per_launch": tests,
    })
    out["roofline"] = roof
    if not args.no_legs:
        out["compose"] = bench_compose(x3d, DeviceArray, R["film"], films[best], w, h, args.no_cpu_baseline)
        out["lt1_tracer"] = bench_lt1(x3d, scene, desc, R["film"])
        # the other configurations' headline rates, a few seconds each
        out["bdpt_512x512x32"] = leg_bdpt(x3d, 3, 1, cpu=not args.no_cpu_baseline)
        out["adjoint_256x256x32"] = leg_adjoint(x3d, 20, 3)
        out["batch_64scenes"] = leg_batch(x3d, 2, 1)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(desc)
    print(json.dumps(out), flush=True)


def bench_lt1(x3d, scene, desc, film_d, spp=16):
    """The reference's own unidirect_lt1 tracer (NEE without MIS, Russian roulette) on the same
    scene and camera at 16 spp: the plain kernel and the wavefront form side by side (device
    kernel time, best of two renders after a warm-up), and whether their films are identical."""
    w, h = desc.width, desc.height
    res = {"tracer": "unidirect_lt1", "spp": spp}
    films = {}
    for form, f in (("plain", x3d.FORM_PLAIN), ("wavefront", x3d.FORM_WAVEFRONT)):
        r = x3d.Renderer(x3d.PT_UNIDIRECT_LT1, form=f, seed=1)
        films[form], _ = r.render(scene, desc.camera, w, h, spp)
        ms = min(r.render_device(scene, desc.camera, w, h, spp, film_d.ptr).kernel_ms for _ in range(2))
        res[form + "_ms"] = ms
        res[form + "_msamples_per_s"] = w * h * spp / ms / 1e3
        log(f"lt1 {form}: {ms:.2f} ms, {w * h * spp / ms / 1e3:.1f} Msamples/s")
    res["speedup"] = res["plain_ms"] / res["wavefront_ms"]
    res["films_bit_identical"] = bool(np.array_equal(films["plain"].view(np.uint32),
                                                     films["wavefront"].view(np.uint32)))
    return res


def bench_compose(x3d, DeviceArray, film_d, film_host, w, h, no_cpu, calls=20):
    """The step after the film: ACES with auto exposure, gamma 2.2 -> 8-bit RGBA, on the film in
    HBM (two HBM-bound passes: the exposure's log-luminance sum reads 16 B per pixel; the
    tonemap reads 16 B and writes 4 B per pixel). Device times from the compositor's HIP events."""
    frame_d = DeviceArray(w * h * 4)
    c = x3d.Compositor(x3d.COMPOSITOR_ACES)
    c.commit_device(film_d.ptr, w, h, frame_d.ptr)  # warm-up
    st = [c.commit_device(film_d.ptr, w, h, frame_d.ptr) for _ in range(calls)]
    e_ms = float(np.mean([s.exposure_ms for s in st]))
    c_ms = float(np.mean([s.compose_ms for s in st]))
    npix = w * h
    res = {"compositor": "aces, auto exposure", "film": f"{w}x{h}", "exposure": st[-1].exposure,
           "exposure_pass_ms": e_ms, "compose_pass_ms": c_ms, "call_ms": float(np.mean([s.total_ms for s in st])),
           "exposure_pass_gbs": 16 * npix / (e_ms * 1e-3) / 1e9, "compose_pass_gbs": 20 * npix / (c_ms * 1e-3) / 1e9}
    res["frac"] = 36 * npix / ((e_ms + c_ms) * 1e-3) / 1e9 / HBM_PEAK_GBS
    if not no_cpu:
        # the reference's compositor arithmetic (float std::log / exp / pow), one core
        from oracle import bindings as ob

        t0 = time.perf_counter()
        ob.compose(film_host, w, h, ob.COMPOSE_ACES, libm=True)
        res["cpu_ms"] = (time.perf_counter() - t0) * 1e3
    log(f"compose: {res}")
    return res


if __name__ == "__main__":
    main()
