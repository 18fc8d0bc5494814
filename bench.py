"""Benchmark: Mpaths/s of the wavefront path tracer on the 720x720 Cornell box
with a mirror and a Cauchy-dispersive glass sphere, depth 8, 64 samples per
pixel, 4 shadow rays, spectral (4 hero wavelengths), FP32.

One step = one full frame (720*720*64 = 33,177,600 paths) rendered on the GPU
with the scene already resident in HBM; the frame's FP64 accumulator stays on
the device during timing.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOADS = {
    # the headline: 720x720 Cornell box with a mirror and a dispersive glass sphere
    "c1": dict(scene="cornell-c1", width=720, height=720, depth=8, spp=64, shadow=4, hero=4,
               metric="Mpaths/s 720^2 Cornell d8 (mirror + Cauchy glass spheres, 64 spp, 4 shadow rays, spectral)",
               desc="cornell-c1 720x720 depth 8, 64 spp, 4 shadow rays, 4 hero wavelengths",
               data="synthetic scene (cornell-c1: 555-unit box, two r=90 spheres), per-step seeds"),
    # traversal stress: the box with an 81,932-triangle displaced icosphere
    "c2": dict(scene="cornell-c2", width=1024, height=1024, depth=8, spp=256, shadow=4, hero=4,
               metric="Mpaths/s 1024^2 Cornell d8 + 81,932-triangle displaced icosphere (256 spp, 4 shadow rays, "
                      "spectral)",
               desc="cornell-c2 1024x1024 depth 8, 256 spp, 4 shadow rays, 4 hero wavelengths",
               data="synthetic scene (cornell-c2: seeded 3-octave value-noise icosphere, subdivision 6), "
                    "per-step seeds"),
}
WORKLOAD = WORKLOADS["c1"]
# Algorithmic bytes per unit of work (minimal SoA records, each written once
# and read once across a kernel boundary): per path 64 B (state init + pixel
# accumulate), per extension ray 176 B (ray 56 + hit 32 + state 72 + queue
# indices 16), per shadow ray 120 B, per emissive hit 24 B.
BYTES_PATH, BYTES_EXT, BYTES_SHADOW, BYTES_EMIT = 64, 176, 120, 24
HBM_PEAK_GBS = 8000.0
# HBM bytes per k_bounce launch, from rocprofv3 FETCH_SIZE and WRITE_SIZE
# passes scaled by a 1 GiB copy (factors 2.0 and 1.0), for the pass size they
# were measured with (the queue records a launch reads and writes depend on the
# paths per pass, not on the block shape):
# profiles/k_bounce_cornell-c1_720x720_d8_64spp_hbm_traffic.md.
TRAFFIC_BYTES_PER_LAUNCH = 9541160182
TRAFFIC_CONFIG = dict(target_pass_paths=64 << 20)
SCHEDULES = {0: "graph", 1: "persistent", 2: "staged"}
# The reference CLI, compiled from the reference's own sources by oracle/ref.mk.
REF_CLI = os.path.join(ROOT, "oracle", "_ref", "pcr-ref-cli")


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU-baseline sample duration")
    p.add_argument("--block", type=int, default=0)
    p.add_argument("--blocks-per-cu", type=int, default=0)
    p.add_argument("--pass-paths", type=int, default=0)
    p.add_argument("--no-autotune", action="store_true", help="keep the library's default launch configuration")
    p.add_argument("--workload", choices=sorted(WORKLOADS), default="c1",
                   help="c1 (default, the headline metric) or c2 (the 81,932-triangle traversal workload)")
    p.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                   help="only run 3 device-to-device copies of MIB MiB each (the reference workload for "
                        "scaling the FETCH_SIZE / WRITE_SIZE counters) and exit")
    return p.parse_args()


def copy_calibration(mib):
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_per_copy": 4 * n, "copies": 3}), flush=True)


def copy_gbs(mib=1024, reps=5):
    """Attainable HBM bandwidth: (read + write) bytes of a device-to-device
    copy over its event-timed duration."""
    import torch
    a = torch.ones(mib * (1 << 18), device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    e1.synchronize()
    gbs = 2.0 * a.numel() * 4 * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del a, b
    torch.cuda.empty_cache()
    return gbs


def cpu_baseline(scene_path, target_s):  # the oracle port
    """The CPU oracle's non-branching estimator on a bounded sample of the same
    frame: the full frame at the first k samples per pixel, k chosen
    from a 1-sample calibration run so the sample takes about target_s; all
    host threads (at most 16, the box's CPU share)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import scene_io
    osc = scene_io.build_oracle_scene(scene_io.parse_scene(scene_path))
    threads = max(1, min(16, os.cpu_count() or 1))
    W, H, d, S, spp = WORKLOAD["width"], WORKLOAD["height"], WORKLOAD["depth"], WORKLOAD["shadow"], WORKLOAD["spp"]

    def run(k):
        t0 = time.perf_counter()
        _, cnt = osc.render_paths(W, H, d, spp, S, 7, s_range=(0, k), threads=threads)
        return time.perf_counter() - t0, int(cnt[0])

    t1, _ = run(1)
    k = max(1, min(spp, int(round(target_s / max(t1, 1e-3)))))
    dt, paths = run(k)
    return {"value": paths / dt / 1e6, "unit": "Mpaths/s", "cores": threads, "kind": "port",
            "sample": "oracle non-branching estimator: the %dx%d frame at samples 0..%d of %d, depth %d, "
                      "%d shadow rays, spectral; %d paths in %.1f s on %d threads"
                      % (W, H, k - 1, spp, d, S, paths, dt, threads)}


def cpu_baseline_reference(target_s):
    """The reference CLI's CPU renderer on the same scene and settings (depth,
    shadow rays, spectral with 4 hero wavelengths), one indirect sample per hit
    (code/Renderer.cpp:468-889) and k jittered primary rays per pixel
    (--aa --aa-samples k), k chosen from a 1-ray calibration run so the sample
    takes about target_s.  The CLI runs one thread per hardware thread
    (code/Renderer.cpp:158-161); its CPU affinity is limited to 16 cores, the
    host's share.  Time is the CLI's own "Render took N ms" clock (scene load
    and PNG write excluded)."""
    W, H, d, S = WORKLOAD["width"], WORKLOAD["height"], WORKLOAD["depth"], WORKLOAD["shadow"]
    scenes = os.path.join(ROOT, "pcr_amd", "scenes")
    cores = sorted(os.sched_getaffinity(0))[:16]

    def run(k):
        with tempfile.TemporaryDirectory() as od:
            args = [REF_CLI, "--scene", WORKLOAD["scene"], "--scenes-dir", scenes, "-d", str(d), "-s", "1",
                    "-S", str(S), "-w", str(W), "--height", str(H), "--spectral", "--hero-samples",
                    str(WORKLOAD["hero"]), "-o", od]
            if k > 1:
                args += ["--aa", "--aa-samples", str(k)]
            keep = os.sched_getaffinity(0)
            os.sched_setaffinity(0, cores)  # inherited by the child
            try:
                p = subprocess.run(args, cwd=scenes, capture_output=True, text=True, timeout=600)
            finally:
                os.sched_setaffinity(0, keep)
        m = re.search(r"Render took (\d+) ms", p.stdout + p.stderr)
        if p.returncode != 0 or not m:
            raise RuntimeError("reference CLI failed (%d): %s" % (p.returncode, (p.stdout + p.stderr)[-400:]))
        return int(m.group(1)) * 1e-3

    t1 = run(1)
    k = max(2, min(256, int(round(target_s / max(t1, 1e-3)))))
    dt = run(k)
    paths = W * H * k
    return {"value": paths / dt / 1e6, "unit": "Mpaths/s", "cores": len(cores), "kind": "reference",
            "sample": "reference CLI (oracle/_ref/pcr-ref-cli) CPU renderer: the %dx%d %s frame, %d jittered "
                      "primary rays per pixel, 1 indirect sample per hit, depth %d, %d shadow rays, spectral; "
                      "%d paths in %.1f s (its own render clock), affinity %d cores"
                      % (W, H, WORKLOAD["scene"], k, d, S, paths, dt, len(cores))}


def main():
    global WORKLOAD
    a = parse()
    WORKLOAD = WORKLOADS[a.workload]
    if a.gpus != 1:
        print("bench.py measures one GPU: every box in this pool has a single MI355X; "
              "multi-GPU runs are not supported by this benchmark.", file=sys.stderr)
        sys.exit(2)
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
        return
    import pcr_amd
    scene_path = os.path.join(ROOT, "pcr_amd", "scenes", WORKLOAD["scene"] + ".json")
    ctx = pcr_amd.Context(0)
    cfg = {}
    if a.block:
        cfg["block"] = a.block
    if a.blocks_per_cu:
        cfg["blocks_per_cu"] = a.blocks_per_cu
    if a.pass_paths:
        cfg["target_pass_paths"] = a.pass_paths
    ds = ctx.upload(pcr_amd.HostScene(scene_path))
    kw = dict(width=WORKLOAD["width"], height=WORKLOAD["height"], depth=WORKLOAD["depth"], spp=WORKLOAD["spp"],
              shadow=WORKLOAD["shadow"], hero=WORKLOAD["hero"], spectral=True)
    if not a.no_autotune and not cfg:
        # untimed: the library's autotuner picks block shape, grid, schedule and pass size
        ds.autotune(trial_paths=16 << 20, seed=999, **kw)
    if cfg:
        ctx.set_config(**cfg)
    tuned = ctx.config()
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:
        sync = lambda: None  # noqa: E731
    for i in range(a.warmup):
        ds.render(seed=1000 + i, **kw)
    ctx.profile(enable=True, reset=True)
    sync()
    t0 = time.perf_counter()
    stats = []
    for i in range(a.steps):
        _, st = ds.render(seed=2000 + i, **kw)
        stats.append(st)
    sync()
    dt = time.perf_counter() - t0
    prof = ctx.profile(enable=False, reset=True)
    paths = sum(s.paths for s in stats)
    ext = sum(s.ext_rays for s in stats)
    shadow = sum(s.shadow_rays for s in stats)
    emit = sum(s.emit_hits for s in stats)
    alg_bytes = BYTES_PATH * paths + BYTES_EXT * ext + BYTES_SHADOW * shadow + BYTES_EMIT * emit
    achieved = alg_bytes / (prof.bounce_ms * 1e-3) / 1e9 if prof.bounce_ms > 0 else None
    value = paths / dt / 1e6
    out = {
        "metric": WORKLOAD["metric"],
        "value": round(value, 3),
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(dt / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": WORKLOAD["data"],
        "config": {"workload": WORKLOAD["desc"],
                   "paths_per_step": paths // a.steps, "ext_rays_per_step": ext // a.steps,
                   "shadow_rays_per_step": shadow // a.steps, "passes_per_step": stats[0].passes,
                   "launch": {"block": tuned.block, "blocks_per_cu": tuned.blocks_per_cu,
                              "target_pass_paths": tuned.target_pass_paths,
                              "schedule": SCHEDULES.get(tuned.schedule),
                              "autotuned": not a.no_autotune and not cfg,
                              "note": "profiled passes issue the graph schedule's kernels as direct launches"},
                   "parallelism": "1 GPU"},
        "roofline": {
            "bound": "hbm",
            "kernel": "k_bounce (fused intersect + shade + NEE + compaction)",
            "achieved": round(achieved, 1) if achieved else None,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
            "traffic": TRAFFIC_BYTES_PER_LAUNCH if a.workload == "c1" and all(
                getattr(tuned, k) == v for k, v in TRAFFIC_CONFIG.items()) else None,
            "traffic_source": "rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0 per k_bounce dispatch",
            "copy_gbs": round(copy_gbs(), 1),
            "launches": prof.bounce_launches,
            "avg_launch_ms": round(prof.bounce_ms / max(1, prof.bounce_launches), 4),
            "algorithmic_bytes_per_launch": alg_bytes // max(1, prof.bounce_launches),
            "queue_bytes_per_launch": prof.queue_bytes // max(1, prof.bounce_launches),
        },
    }
    ds.release()
    ctx.close()
    if not a.no_cpu_baseline:
        port = cpu_baseline(scene_path, a.cpu_seconds)
        ref = None
        if os.path.exists(REF_CLI):
            try:
                ref = cpu_baseline_reference(a.cpu_seconds)
            except RuntimeError as e:  # e.g. a scene primitive the reference's loader does not know
                port["note"] = "reference CLI not used: %s" % str(e).splitlines()[0][:200]
        if ref:
            out["cpu_baseline"] = ref
            out["cpu_port"] = port
        else:
            out["cpu_baseline"] = port
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
