"""Benchmark of the RadFoam cell walk on one MI355X.

Headline (``value``): forward + adjoint throughput, Mrays/s, of the training
workload at 1,048,576 cells: the mixture-on-shell scene (SH degree 3), 64 ring
cameras at 1024x768, 256 batches of 4,096 random (camera, pixel) rays.  One
step = one pass over the 256 batches; per batch the fast forward walk, the L2
loss gradient against target colours rendered from a perturbed copy of the
scene, and the fast adjoint that adds density / SH / position gradients into
the gradient tables (zeroed once per step); the fast path runs the three as one
fused kernel per batch (bv_train_step_l2).  Inputs are resident in HBM before
the timed region.  The same step with the plain one-ray-per-thread port of the
reference WGSL walk (forward and adjoint) is timed beside it; their ratio is
``speedup_vs_plain``.  The 256 batches of a step are independent (gradients
accumulate; no parameter update inside a step), so both kernels issue them
round-robin on 4 HIP streams, each batch still one launch; the one-stream
figures of both are reported beside.

Also reported: forward-only Mrays/s over all 64 cameras (50,331,648 rays) for
both kernels; the roofline of the dominant kernel (the fused training kernel)
from its bytes per launch and the HIP-event durations of its calls inside the
timed steps, with the committed rocprofv3 traffic and L2 hit rate; the
trainer's own GPU path over the same rays (batched path record + volumetric
graph forward and reverse); the reference's CPU-runnable case (16,384 cells,
one 128x128 view) on the GPU and on the host; the oracle (CPU restatement of
trace.rs) on all host threads for a bounded sample; the fused step on the
sigma = 0.05 cluster reading of the scene; the Gaussian frame and adjoint
(524,288 particles, 1024x1024); the PowerFoam splat frame (1M sites, 1080p) and
path recorder (one 1024x768 camera, 64-interval rows, with and without the
Jacobian streams); and the serving loop
(``serving``): four resident PowerFoam scenes of 65,536 to 2,097,152 sites,
relit splat views at 64x64 to 1920x1080 with 1-8 point lights, Poisson
arrivals at 60% of the measured burst capacity, p50 / p99 latency and Mrays/s.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--cells N]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "blade-volume_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0
METRIC = "Mrays/s fwd+adjoint @1M cells vs plain port; p50/p99 view latency (serving)"


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "train_l2_fused_1M_cells_traffic.txt")


def measured_counters() -> dict:
    """Per-launch counter figures of the fused kernel from the committed
    rocprofv3 summary: traffic_bytes_per_launch (FETCH_SIZE + WRITE_SIZE,
    copy-calibrated) and l2_hit_rate (TCC_HIT / (TCC_HIT + TCC_MISS)); empty
    when the file is absent."""
    out = {}
    try:
        kv = dict(line.split(":", 1) for line in open(TRAFFIC_FILE) if ":" in line)
    except OSError:
        return out
    for key, conv in (("traffic_bytes_per_launch", int), ("l2_hit_rate", float), ("fetch_kib_per_launch", float),
                      ("write_kib_per_launch", float)):
        try:
            out[key] = conv(kv[key].strip())
        except (KeyError, ValueError):
            pass
    return out


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=1 << 20)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--batches", type=int, default=256)
    ap.add_argument("--cameras", type=int, default=64)
    ap.add_argument("--width", type=int, default=1024)
    ap.add_argument("--height", type=int, default=768)
    ap.add_argument("--plain-steps", type=int, default=1, help="timed steps of the plain port")
    ap.add_argument("--streams", type=int, default=32, help="HIP streams the batches of a step are spread over")
    ap.add_argument("--merge", type=int, default=256,
                    help="consecutive batches of a step passed to one train_step_l2 call (one launch); the "
                         "default carries the whole step in one launch")
    ap.add_argument("--forward-cameras", type=int, default=None, help="cameras in the forward-only pass")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="budget of the oracle CPU sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-survey-reading", action="store_true",
                    help="skip the fused step on the sigma = 0.05 cluster reading of the shell scene")
    ap.add_argument("--no-graph", action="store_true", help="skip the trainer-graph phase (batched record + graph)")
    ap.add_argument("--graph-steps", type=int, default=256, help="L of the trainer-graph phase's path rows")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop phase")
    ap.add_argument("--no-gauss", action="store_true", help="skip the Gaussian ray-tracing phase")
    ap.add_argument("--gaussians", type=int, default=524_288)
    ap.add_argument("--gauss-res", type=int, default=1024)
    ap.add_argument("--serve-scenes", default="65536,262144,1048576,2097152",
                    help="PowerFoam scene sizes resident in the server")
    ap.add_argument("--serve-requests", type=int, default=120, help="Poisson-arrival requests timed")
    ap.add_argument("--serve-load", type=float, default=0.6, help="arrival rate as a fraction of burst capacity")
    ap.add_argument("--serve-rate", type=float, default=0.0,
                    help="arrival rate in requests per second (overrides --serve-load when set)")
    ap.add_argument("--gauss-probe", action="store_true",
                    help="for rocprofv3 runs of the Gaussian kernels: only the Gaussian phase (forward, then "
                         "forward + adjoint), printed to stderr, no JSON line")
    ap.add_argument("--splat-probe", action="store_true",
                    help="only the PowerFoam splat phase (1080p frame, then the path recorder over one camera), "
                         "printed to stderr, no JSON line")
    ap.add_argument("--compact-probe", action="store_true",
                    help="for rocprofv3 runs of the path compaction: record one camera's 128-step rows at the "
                         "training scene and time bv_compact_paths over them, printed to stderr, no JSON line")
    ap.add_argument("--serve-probe", action="store_true",
                    help="only the serving phase (its scenes, burst and Poisson mix), printed to stderr, no JSON line")
    ap.add_argument("--objects-probe", action="store_true",
                    help="for rocprofv3 runs of the multi-object scene trace: one camera frame over four instances "
                         "of the training scene, printed to stderr, no JSON line")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="for rocprofv3 runs: a 256 MiB device copy (the byte-count calibration) and the step's fused "
                         "training calls (--probe-reps times), then exit")
    ap.add_argument("--probe-reps", type=int, default=1, help="passes over the step's calls in --pmc-probe")
    return ap.parse_args()


SERVE_RESOLUTIONS = ((64, 64), (256, 256), (512, 512), (1280, 720), (1920, 1080))


def serve_phase(args, bv, scenes, np, torch) -> dict:
    """Serving loop on resident PowerFoam scenes: relit splat views (1-8 point
    lights, each casting a visibility ray with probability 1/2) at resolutions from 64x64 to
    1920x1080, submitted with Poisson arrivals at a fraction of the measured
    burst capacity; latency is submit -> image in host memory."""
    sizes = [int(x) for x in args.serve_scenes.split(",") if x]
    t0 = time.time()
    clouds = []
    for k, n in enumerate(sizes):
        h = scenes.powerfoam_scene(n, seed=4 + k, cech=bv.compute_cech)
        clouds.append((n, bv.RadFoamGpuCloud(h)))
        del h
    log(f"serving scenes {sizes} built in {time.time() - t0:.1f}s")
    rng = np.random.default_rng(5)

    def request(srv, ids, r):
        k = int(r.integers(len(clouds)))
        n = clouds[k][0]
        half = (n / 65536.0) ** (1.0 / 3.0)
        v = r.normal(size=3)
        v /= np.linalg.norm(v)
        w, h = SERVE_RESOLUTIONS[int(r.integers(len(SERVE_RESOLUTIONS)))]
        cam = scenes.looking_at(2.4 * half * v, (0.0, 0.0, 0.0), math.radians(55.0), w / h, 100.0)
        lights = scenes.random_lights(int(r.integers(1, 9)), r, radius=3.0 * half)
        for light in lights:  # each light with an optional visibility ray
            light["visibility"] = bool(r.integers(2))
        t, _ = srv.submit(ids[k], cam, (w, h), bv.EVAL_RELIGHT, lights, False, renderer=bv.RENDER_SPLAT)
        return t, w * h

    def new_server():
        srv = bv.Server(hbm_budget_bytes=1 << 30, max_active=64)
        ids = [srv.add_scene(c) for _, c in clouds]
        return srv, ids

    # warm-up (BVH builds at first submission) and burst capacity
    srv, ids = new_server()
    for k in range(len(clouds)):
        t, _ = srv.submit(ids[k], scenes.looking_at((3.0, 0.0, 0.0), (0, 0, 0), 1.0, 1.0, 100.0), (64, 64),
                          bv.EVAL_RELIGHT, scenes.random_lights(1, rng), False, renderer=bv.RENDER_SPLAT)
        srv.wait(t)
    burst = [request(srv, ids, rng) for _ in range(max(8, args.serve_requests // 3))]
    tb = time.perf_counter()
    for t, _ in burst:
        srv.wait(t)
    burst_s = time.perf_counter() - tb
    srv.close()
    capacity = len(burst) / burst_s
    rate = args.serve_rate if args.serve_rate > 0 else args.serve_load * capacity
    log(f"serving burst: {len(burst)} requests in {burst_s:.2f}s ({capacity:.1f} req/s); Poisson rate {rate:.1f}/s")

    srv, ids = new_server()
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=args.serve_requests))
    tickets = []
    start = time.perf_counter()
    for a in arrivals:
        now = time.perf_counter() - start
        if a > now:
            time.sleep(a - now)
        tickets.append(request(srv, ids, rng))
    lat = [srv.wait(t) for t, _ in tickets]
    wall = time.perf_counter() - start
    st = srv.stats()
    srv.close()
    rays = sum(px for _, px in tickets)
    for _, c in clouds:
        c.deinit()
    out = {
        "workload": "relit PowerFoam splat views, Poisson arrivals, continuous batching",
        "scenes": sizes, "resolutions": [list(r) for r in SERVE_RESOLUTIONS], "lights": "1-8 uniform",
        "visibility": "per light, on for half the lights", "requests": len(tickets),
        "arrival_rate_hz": round(rate, 2), "load_fraction_of_burst_capacity": round(rate / capacity, 3),
        "burst_capacity_req_s": round(capacity, 2),
        "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "mean_ms": round(float(np.mean(lat)), 3), "max_ms": round(float(np.max(lat)), 3),
        "mrays_per_s": round(rays / wall / 1e6, 3), "server_p50_ms": round(st["p50_ms"], 3),
        "server_p99_ms": round(st["p99_ms"], 3), "peak_arena_bytes": int(st["peak_bytes"]),
    }
    log(f"serving: p50 {out['p50_ms']} ms, p99 {out['p99_ms']} ms, {out['mrays_per_s']} Mrays/s")
    return out


def gaussian_phase(args, bv, scenes, np, torch) -> dict:
    """Gaussian ray tracing fwd and fwd+bwd: mixture-on-shell means, log-uniform
    scales in [1e-3, 5e-2], random unit quaternions, opacity sigmoid(N(0,1)),
    SH degree 3; one 1024x1024 camera at distance 3 looking at the origin.  The
    forward renders the camera frame; forward + adjoint run on the camera's
    rays as a batch ordered in 8x8 pixel tiles."""
    t0 = time.time()
    g = scenes.gaussian_scene(args.gaussians, seed=3, sh_degree=3)
    cloud = bv.GaussianGpuCloud(g.mean, g.opacity, g.rotation, g.scale, g.sh, 3, 0.01)
    torch.cuda.synchronize()
    log(f"gaussians: {g.n} particles, BVH built in {time.time() - t0:.1f}s")
    R = args.gauss_res
    cam = scenes.looking_at((0.0, 0.4, 3.0), (0.0, 0.0, 0.0), math.radians(60.0), 1.0, 100.0)
    st = bv.GaussianTraceSettings(0.01, 0.01, 0.0, 100.0)
    # forward: the viewer's frame, one ray per pixel (bv_gaussians_trace_camera)
    tracer = bv.GaussianTracer(cloud, st)
    out = tracer.dispatch(cam, (R, R), stats=True)
    torch.cuda.synchronize()
    stats = out["stats"].to(torch.int64)
    bytes_fwd = int(stats[:, 0].sum().item())
    hits = float(out["hits"].float().mean().item())
    reps = 3
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        tracer.dispatch(cam, (R, R), out=out)
    e1.record()
    torch.cuda.synchronize()
    fwd_ms = e0.elapsed_time(e1) / reps
    # forward + adjoint: the same camera's rays as a ray batch, in 8x8 pixel tiles
    o, d = bv.camera_rays(cam, R, R)
    tiles = torch.arange(R * R, device=o.device).view(R // 8, 8, R // 8, 8).permute(0, 2, 1, 3).reshape(-1)
    o, d = o[tiles].contiguous(), d[tiles].contiguous()
    rout = bv.gaussian_trace_rays(cloud, o, d, st)
    grads = bv.gaussian_zero_grads(cloud)
    d_rgba = torch.randn((R * R, 4), dtype=torch.float32, device=o.device)
    e0.record()
    for _ in range(reps):
        bv.gaussian_trace_rays(cloud, o, d, st, out=rout)
        bv.gaussian_adjoint(cloud, o, d, rout["rgba"], d_rgba, grads, st)
    e1.record()
    torch.cuda.synchronize()
    fb_ms = e0.elapsed_time(e1) / reps
    res = {
        "workload": "Gaussian ray tracing, 3DGRT max-response blend (k-buffer 16)",
        "particles": g.n, "rays": R * R, "sh_degree": 3,
        "forward": "camera frame, one ray per pixel", "fwd_bwd": "the camera's rays as a batch in 8x8 pixel tiles",
        "forward_mrays_s": round(R * R / fwd_ms / 1e3, 2), "forward_ms": round(fwd_ms, 3),
        "fwd_bwd_mrays_s": round(R * R / fb_ms / 1e3, 2), "fwd_bwd_ms": round(fb_ms, 3),
        "mean_hits_per_ray": round(hits, 2),
        # node, particle and SH bytes the traversal reads, over the frame time:
        # not HBM traffic (the 176 MiB cloud stays in L2 and the infinity cache,
        # 99.8% L2 hits), so no HBM roofline fraction; the kernel is bound by
        # VALU issue (profiles/gauss_trace_adjoint_524k_1024sq.md)
        "forward_algorithmic_gb_s": round(bytes_fwd / (fwd_ms * 1e-3) / 1e9, 1),
        "forward_bound": "valu",
        "hbm_mib": round(cloud.device_bytes / 2**20, 1),
    }
    log(f"gaussians: fwd {res['forward_mrays_s']} Mrays/s, fwd+bwd {res['fwd_bwd_mrays_s']} Mrays/s, "
        f"{hits:.1f} hits/ray")
    cloud.deinit()
    return res


def splat_phase(bv, scenes, np, torch, n: int = 1 << 20) -> dict:
    """PowerFoam splat forward (SH) over a 1920x1080 camera at a 1M-site scene."""
    h = scenes.powerfoam_scene(n, seed=6, cech=bv.compute_cech)
    cloud = bv.RadFoamGpuCloud(h)
    half = (n / 65536.0) ** (1.0 / 3.0)
    cam = scenes.looking_at((2.4 * half, 0.3 * half, 0.5 * half), (0, 0, 0), math.radians(55.0), 16 / 9, 100.0)
    tr = bv.PowerFoamGpuSplatTracer(cloud)
    img = tr.dispatch(cam, (1920, 1080), outputs=("rgba", "stats"))
    torch.cuda.synchronize()
    nbytes = int(img["stats"].to(torch.int64)[:, 0].sum().item())
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(3):
        tr.dispatch(cam, (1920, 1080), out=img)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 3
    res = {"workload": "PowerFoam splats, SH forward, 1920x1080", "sites": n,
           "edges": int(h.neighbors.size), "mrays_s": round(1920 * 1080 / ms / 1e3, 2), "ms": round(ms, 3),
           "algorithmic_gb_s": round(nbytes / (ms * 1e-3) / 1e9, 1)}
    log(f"splats: {res['mrays_s']} Mrays/s at 1080p, {res['algorithmic_gb_s']} GB/s algorithmic")
    res["recorder"] = splat_record_phase(bv, torch, cloud, cam)
    cloud.deinit()
    return res


def splat_record_phase(bv, torch, cloud, cam, W: int = 1024, H: int = 768, L: int = 64) -> dict:
    """The PowerFoam path recorder over one camera's rays: [P, L] rows of
    (cell, near, dt), then the same rows with the Jacobian and reference-tangent
    streams (72 B per row slot)."""
    o, d = bv.camera_rays(cam, W, H)
    rec = bv.splat_record(cloud, o, d, L)
    jac = bv.splat_record_jacobians(cloud, o, d, L)
    torch.cuda.synchronize()
    counts = rec["count"].to(torch.int64)
    records = int(counts.sum().item())
    status = jac["status"].to(torch.int64)
    res = {"workload": f"PowerFoam path recorder, one {W}x{H} camera, {L}-interval rows", "rows": W * H,
           "intervals_per_row": round(records / (W * H), 2),
           "full_rows_fraction": round(float((counts >= L).to(torch.float64).mean().item()), 4)}
    for name, call in (("record", lambda: bv.splat_record(cloud, o, d, L, out=rec)),
                       ("record_jacobians", lambda: bv.splat_record_jacobians(cloud, o, d, L, out=jac))):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(3):
            call()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 3
        res[name] = {"ms": round(ms, 3), "mrays_s": round(W * H / ms / 1e3, 2),
                     "mrecords_s": round(records / ms / 1e3, 1)}
    res["jacobian_rows_match"] = bool(torch.equal(status & 0x7FFFFFFF, counts))
    log(f"splat recorder: {res['record']['ms']} ms rows, {res['record_jacobians']['ms']} ms with Jacobians, "
        f"{res['intervals_per_row']} intervals per row")
    return res


def objects_phase(bv, scenes, np, torch, cloud, W: int, H: int) -> dict:
    """Multi-object scene trace (bv_trace_objects) of one W x H camera frame.
    Four instances of the training scene on a row at x = -3, -1, 1, 3 (the
    second rotated 0.8 rad about y, the third scaled by 0.5) seen from
    (0, 0.5, 7); beside it one identity instance from (0, 0, 3) and the
    single-scene walk (bv_trace_rays) over the same rays.  Times are HIP events
    on the launching stream, averaged over 5 frames after 2 warmups."""
    st = bv.TraceSettings(weight_threshold=0.001, max_steps=1024, depth=100.0)
    q = (0.0, math.sin(0.4), 0.0, math.cos(0.4))

    def timed(fn) -> float:
        for _ in range(2):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(5):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 5

    row = bv.Scene()
    for k, x in enumerate((-3.0, -1.0, 1.0, 3.0)):
        row.add_radfoam(cloud, bv.Transform(position=(x, 0.0, 0.0), rotation=q if k == 1 else (0.0, 0.0, 0.0, 1.0),
                                            scale=(0.5, 0.5, 0.5) if k == 2 else (1.0, 1.0, 1.0)))
    cam = scenes.looking_at(np.array([0.0, 0.5, 7.0], np.float32), np.zeros(3, np.float32), math.radians(60.0),
                            W / H, 100.0)
    row.update_camera(cam.position)
    o, d = bv.camera_rays(cam, W, H)
    out = torch.empty((W * H, 4), dtype=torch.float32, device=o.device)
    ms_row = timed(lambda: row.trace_rays(o, d, st, out=out))
    covered = float((out[:, 3] > 0).float().mean())
    one = bv.Scene()
    one.add_radfoam(cloud)
    cam1 = scenes.looking_at(np.array([0.0, 0.0, 3.0], np.float32), np.zeros(3, np.float32), math.radians(60.0),
                             W / H, 100.0)
    one.update_camera(cam1.position)
    o1, d1 = bv.camera_rays(cam1, W, H)
    ms_one = timed(lambda: one.trace_rays(o1, d1, st, out=out))
    start = cloud.containing_point(cam1.position)
    ms_walk = timed(lambda: bv.trace_rays(cloud, o1, d1, start_cell=start, settings=st))
    n = W * H
    return {"objects_probe": {"rays": n, "four_objects_ms": round(ms_row, 3),
                              "four_objects_Mrays_s": round(n / ms_row / 1e3, 1), "four_objects_covered": round(covered, 3),
                              "one_object_ms": round(ms_one, 3), "one_object_Mrays_s": round(n / ms_one / 1e3, 1),
                              "single_scene_walk_ms": round(ms_walk, 3),
                              "single_scene_walk_Mrays_s": round(n / ms_walk / 1e3, 1)}}


def compact_phase(bv, torch, cloud, cam, start, W: int, H: int, L: int = 128) -> dict:
    """bv_compact_paths over one camera's recorded 128-step rows at the training
    scene (the recorder's rows, as the trainer consumes them).  Algorithmic
    bytes: 4 B of status per path, then per record the 4 B cell read and 16 B
    written (slot, cell, pixel, active)."""
    o, d = bv.camera_rays(cam, W, H)
    rec = bv.PathRecorder(cloud).dispatch(o, d, start_cell=start, max_steps=L, depth=100.0, max_path_dt=50.0)
    out = bv.PathCompactBuffers(W * H, L)
    out.clear()
    comp = bv.PathCompactor()
    comp.dispatch(rec, out)
    records = out.active_count()
    reps = 10
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        out.count.zero_()
        comp.dispatch(rec, out)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    nbytes = 4 * W * H + 20 * records
    res = {"workload": f"path-record compaction, one {W}x{H} camera, {L}-step rows", "paths": W * H,
           "records": records, "ms": round(ms, 4), "algorithmic_gb_s": round(nbytes / (ms * 1e-3) / 1e9, 1),
           "frac_of_hbm_peak": round(nbytes / (ms * 1e-3) / 8e12, 3),
           "timing": f"{reps} calls between two events, each after a 4-byte count fill"}
    log(f"compaction: {records} records of {W * H} paths in {res['ms']} ms, {res['algorithmic_gb_s']} GB/s algorithmic")
    return res


def make_rays(bv, np, torch, cams, starts, cam_idx, px, py, W: int, H: int):
    """Device origins, directions and start cells of (camera, pixel) rays:
    each camera's pixel-centre rays (bv_camera_rays), gathered."""
    dev = torch.device("cuda:0")
    n = len(cam_idx)
    O = torch.empty((n, 3), dtype=torch.float32, device=dev)
    D = torch.empty((n, 3), dtype=torch.float32, device=dev)
    cam_idx_t = torch.from_numpy(cam_idx.astype(np.int64)).to(dev)
    pix_t = torch.from_numpy((py.astype(np.int64) * W + px.astype(np.int64))).to(dev)
    for k, c in enumerate(cams):
        sel = torch.nonzero(cam_idx_t == k).flatten()
        if sel.numel() == 0:
            continue
        co, cd = bv.camera_rays(c, W, H)
        O[sel] = co[pix_t[sel]]
        D[sel] = cd[pix_t[sel]]
        del co, cd
    S = torch.tensor(np.asarray(starts, np.uint32).view(np.int32), device=dev)[cam_idx_t].contiguous()
    return O, D, S


def target_colours(bv, torch, target_scene, O, D, S, st):
    """Target colours: the rays traced through the perturbed copy of the scene."""
    tcloud = bv.RadFoamGpuCloud(target_scene)
    target = bv.trace_rays(tcloud, O, D, S, settings=st)["rgba"][:, :3].contiguous()
    tcloud.deinit()
    return target


def survey_reading_phase(args, bv, scenes, np, torch, cams, cam_idx, px, py, W: int, H: int, st) -> dict:
    """The fused training step on the other reading of the scene's cluster
    spread: each mixture sample centre + N(0, 0.05^2 I) before re-projection
    (the headline scene uses 0.5), same cameras, rays and schedule."""
    t0 = time.time()
    scene = scenes.shell_scene(args.cells, seed=1, sh_degree=3, angular_sigma=0.05)
    build_s = time.time() - t0
    cloud = bv.RadFoamGpuCloud(scene)
    starts = [cloud.containing_point(c.position) for c in cams]
    O, D, S = make_rays(bv, np, torch, cams, starts, cam_idx, px, py, W, H)
    target = target_colours(bv, torch, scenes.perturbed(scene, seed=2), O, D, S, st)
    grads = bv.zero_grads(cloud, O.device)
    n = O.shape[0]

    def step():
        for g in grads.values():
            g.zero_()
        bv.train_step_l2(cloud, O, D, S, target, grads, settings=st, kernel=bv.KERNEL_FAST)

    step()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t) / args.steps
    r = bv.trace_rays(cloud, O, D, S, settings=st, kernel=bv.KERNEL_FAST, outputs=("stats", "steps"))
    steps = float(r["steps"].to(torch.float64).mean().item())
    segs = float(r["stats"][:, 1].to(torch.float64).mean().item())
    cloud.deinit()
    out = dict(angular_sigma=0.05, radial_sigma=0.05, mrays_s=round(n / sec / 1e6, 4), ms_per_step=round(sec * 1e3, 4),
               steps_per_ray=round(steps, 2), integrated_segments_per_ray=round(segs, 2),
               edges=int(scene.neighbors.size), scene_build_s=round(build_s, 1))
    log(f"sigma 0.05 reading: {out}")
    return out


def graph_phase(args, bv, np, torch, cloud, scene, cams, starts, cam_idx, px, py, W: int, H: int, target) -> dict:
    """The reference trainer's GPU path for the same step's rays: one batched
    record over the 64 camera slices (bv_record_paths_batch, rays generated
    in the recorder from pixel indices) into [P, L] rows, then the volumetric
    graph's forward + reverse over the rows (bv_volume_graph_step, positions
    dt, ReLU density, L2 loss), gradients into density / position / SH
    tables.  Timed per stage with HIP events on the current stream."""
    L = args.graph_steps
    order = np.argsort(cam_idx, kind="stable")
    counts = np.bincount(cam_idx, minlength=len(cams))
    offs = np.concatenate([[0], np.cumsum(counts)[:-1]])
    pix = (py.astype(np.int64) * W + px.astype(np.int64))[order]
    n = len(order)
    rec = bv.PathRecorder(cloud)
    buf = rec.buffers(n, L, rays=True)
    buf.write_pixel_indices(pix)
    views = [bv.RecordPathsArgs(c, W, H, int(counts[k]), L, start_point=int(starts[k]), pixel_offset=int(offs[k]),
                                max_path_dt=50.0, depth=100.0) for k, c in enumerate(cams)]
    dev = buf.cells.device
    labels = target[torch.from_numpy(order).to(dev)].contiguous()
    K = 16
    params = dict(log_density=torch.from_numpy(scene.density).to(dev), positions=torch.from_numpy(scene.xyz).to(dev),
                  sh=torch.from_numpy(scene.sh).to(dev), sh_degree=3)
    grads = dict(d_log_density=torch.zeros(scene.n, device=dev), d_positions=torch.zeros((scene.n, 3), device=dev),
                 d_sh=torch.zeros((scene.n, 3 * K), device=dev))
    batch = dict(cells=buf.cells, next_cells=buf.next_cells, dts=buf.dts, mask=buf.mask, status=buf.status,
                 ray_origin=buf.ray_origins, ray_dir=buf.ray_dirs, labels=labels)
    opts = bv.VolumeGraphOptions(loss=bv.LOSS_L2)
    stream = torch.cuda.current_stream()
    ev = []

    def step(timed: bool):
        e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        e[0].record(stream)
        rec.dispatch_batch(buf, views)
        e[1].record(stream)
        for g in grads.values():
            g.zero_()
        volume_out = bv.volume_graph_step(params, batch, opts, grads, outputs=("loss",))
        e[2].record(stream)
        if timed:
            ev.append(e)
        return volume_out

    step(False)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        out = step(True)
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t) / args.steps
    rec_ms = sum(e[0].elapsed_time(e[1]) for e in ev) / len(ev)
    graph_ms = sum(e[1].elapsed_time(e[2]) for e in ev) / len(ev)
    status = buf.status.to(torch.int64) & 0x7FFFFFFF
    trunc = float(((buf.status.to(torch.int64) & 0x80000000) != 0).to(torch.float64).mean().item())
    res = dict(mrays_s=round(n / sec / 1e6, 4), ms_per_step=round(sec * 1e3, 4), record_ms=round(rec_ms, 4),
               graph_ms=round(graph_ms, 4), max_steps=L, rows_per_step=n, slices=len(views),
               recorded_steps_per_ray=round(float(status.to(torch.float64).mean().item()), 2),
               truncated_fraction=round(trunc, 6), loss=float(out["loss"].item()),
               note="bv_record_paths_batch over 64 camera slices + bv_volume_graph_step (positions dt, ReLU "
                    "density, L2), forward and reverse, gradients added; one call each per step")
    log(f"trainer graph path: {res}")
    del buf, batch, grads
    return res


def c1_phase(bv, scenes, np, torch, oracle_mod) -> dict:
    """The reference's CPU-runnable case: 16,384 uniform sites, one 128x128
    camera (60 degree fov) at distance 3 looking at the origin; the fast walk
    on the GPU and, when the oracle is given, on all host threads."""
    scene = scenes.uniform_scene(16_384, seed=0)
    cam = scenes.looking_at((0.0, 0.0, -3.0), (0.0, 0.0, 0.0), math.radians(60.0), 1.0, 100.0)
    cloud = bv.RadFoamGpuCloud(scene)
    start = cloud.containing_point(cam.position)
    tr = bv.RadFoamGpuTracer(cloud, bv.RadFoamTraceSettings(1024, 0.001), bv.KERNEL_FAST)
    img = bv._alloc_outputs(128 * 128, ("rgba",), torch.device("cuda:0"))
    tr.dispatch(cam, (128, 128), out=img, start_cell=start)
    torch.cuda.synchronize()
    reps = 200
    t = time.perf_counter()
    for _ in range(reps):
        tr.dispatch(cam, (128, 128), out=img, start_cell=start)
    torch.cuda.synchronize()
    out = dict(gpu_mrays_s=round(reps * 128 * 128 / (time.perf_counter() - t) / 1e6, 3),
               gpu_note="one 128x128 camera per call, 200 calls back to back")
    cloud.deinit()
    if oracle_mod is not None:
        osc = oracle_mod.of_host(scene)
        o, d = oracle_mod.camera_rays(cam.position, cam.orientation, cam.tan_half(), cam.principal, 128, 128)
        s = np.full(128 * 128, start, np.uint32)
        t = time.perf_counter()
        done = 0
        while done == 0 or time.perf_counter() - t < 2.0:
            oracle_mod.trace_rays(osc, o, d, s, weight_threshold=0.001, max_steps=1024, depth=100.0, threads=0)
            done += 1
        out["cpu_mrays_s"] = round(done * 128 * 128 / (time.perf_counter() - t) / 1e6, 4)
        out["cpu_threads"] = oracle_mod.max_threads()
    log(f"C1 forward: {out}")
    return out


def main() -> int:
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the walk needs no collective, and the kernels run on a single "
              "MI355X; sharding rays or cells over several GPUs is not implemented", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import bv
    from bv import scenes

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()

    if args.serve_probe:
        log(json.dumps(serve_phase(args, bv, scenes, np, torch)))
        return 0
    if args.gauss_probe:
        log(json.dumps(gaussian_phase(args, bv, scenes, np, torch)))
        return 0
    if args.splat_probe:
        log(json.dumps(splat_phase(bv, scenes, np, torch)))
        return 0

    # ------------------------------------------------------------ scene + rays (untimed)
    t0 = time.time()
    scene = scenes.shell_scene(args.cells, seed=1, sh_degree=3)
    scene_build_s = time.time() - t0
    log(f"scene: {scene.n} sites, {scene.neighbors.size} directed edges, built (sites + Delaunay) in "
        f"{scene_build_s:.1f}s")
    cloud = bv.RadFoamGpuCloud(scene)
    target_scene = scenes.perturbed(scene, seed=2)
    cams = scenes.ring_cameras(args.cameras, 3.0, math.radians(60.0), args.width / args.height, 100.0)
    starts = [cloud.containing_point(c.position) for c in cams]
    W, H = args.width, args.height
    if args.compact_probe:
        log(json.dumps(compact_phase(bv, torch, cloud, cams[0], starts[0], W, H)))
        return 0
    if args.objects_probe:
        log(json.dumps(objects_phase(bv, scenes, np, torch, cloud, W, H)))
        return 0
    rng = np.random.default_rng(2)
    n_rays = args.batch * args.batches
    cam_idx, px, py = scenes.random_rays(cams, W, H, n_rays, rng)
    O, D, S = make_rays(bv, np, torch, cams, starts, cam_idx, px, py, W, H)
    st = bv.TraceSettings(weight_threshold=0.001, max_steps=1024, depth=100.0)
    target = target_colours(bv, torch, target_scene, O, D, S, st)
    torch.cuda.synchronize()
    log(f"setup done in {time.time() - t0:.1f}s; scene holds {cloud.device_bytes / 2**20:.0f} MiB of HBM")

    B = args.batch
    batches = [(O[i * B:(i + 1) * B], D[i * B:(i + 1) * B], S[i * B:(i + 1) * B], target[i * B:(i + 1) * B])
               for i in range(args.batches)]
    # calls of a step: `merge` consecutive batches each (their rays are contiguous)
    def calls_of(merge: int) -> list:
        m = max(1, min(merge, args.batches))
        return [(i * B, min((i + m) * B, n_rays)) for i in range(0, args.batches, m)]

    calls = calls_of(args.merge)
    per_batch = calls_of(1)
    grads = bv.zero_grads(cloud, dev)
    rgba_out = torch.empty((n_rays, 4), dtype=torch.float32, device=dev)
    loss = torch.zeros(1, dtype=torch.float32, device=dev)

    # The batches of one step are independent (gradients accumulate, no
    # parameter update inside a step).  The headline schedule passes `merge`
    # batches to one call (by default the whole step: one launch, whose
    # persistent waves claim rays until the step's rays run out); calls go
    # round-robin over HIP streams.  The per-batch schedule (one call per
    # 4,096-ray batch, as the reference dispatches) is timed beside it on 32
    # streams and on one.  Both kernels (fast: one fused kernel per call;
    # plain: trace, loss, adjoint) run every schedule.
    side = [torch.cuda.Stream() for _ in range(max(args.streams, 32))]

    def step(kernel: int, streams: int = 1, sched=None, evs=None) -> None:
        sched = calls if sched is None else sched
        # no more streams than calls: a step of one call runs on the main
        # stream, without the fan-out / fan-in events
        streams = max(1, min(streams, len(sched)))
        for g in grads.values():
            g.zero_()
        loss.zero_()
        ready = torch.cuda.Event()
        ready.record(stream)
        for ss in side[:streams]:
            ss.wait_event(ready)
        for i, (a, b) in enumerate(sched):
            ss = stream if streams == 1 else side[i % streams]
            if evs is not None:
                ea, eb = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                ea.record(ss)
            bv.train_step_l2(cloud, O[a:b], D[a:b], S[a:b], target[a:b], grads, settings=st, rgba_out=rgba_out[a:b],
                             loss=loss, kernel=kernel, stream=ss)
            if evs is not None:
                eb.record(ss)
                evs.append((ea, eb))
        if streams > 1:
            for ss in side[:streams]:
                done = torch.cuda.Event()
                done.record(ss)
                stream.wait_event(done)

    def timed(kernel: int, steps: int, warmup: int, streams: int, sched=None, evs=None) -> float:
        for _ in range(warmup):
            step(kernel, streams=streams, sched=sched)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            step(kernel, streams=streams, sched=sched, evs=evs)
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps

    if args.pmc_probe:
        # Calibration copy: 256 MiB read + 256 MiB written by one copy kernel.
        src = torch.empty(64 << 20, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        # one step's calls (all of them: `--merge 256` makes it one launch), --probe-reps times
        for _ in range(max(1, args.probe_reps)):
            for a, b in calls:
                bv.train_step_l2(cloud, O[a:b], D[a:b], S[a:b], target[a:b], grads, settings=st,
                                 kernel=bv.KERNEL_FAST)
        torch.cuda.synchronize()
        log("pmc probe done")
        return 0

    # ------------------------------------------------------------ headline: fast fwd+adjoint
    n_streams = max(1, min(args.streams, len(calls)))
    log(f"timing fast fwd+adjoint: {args.warmup} warmup + {args.steps} steps of {n_rays} rays, "
        f"{len(calls)} call(s) per step on {n_streams} streams")
    launch_events = []
    fast_s = timed(bv.KERNEL_FAST, args.steps, args.warmup, n_streams, evs=launch_events)
    value = n_rays / fast_s / 1e6
    # the fused calls' own durations, from HIP events recorded around each call
    # on its stream inside the timed steps (each includes the call's small
    # clear / add-back kernels)
    adj_ms = sum(a.elapsed_time(b) for a, b in launch_events) / len(launch_events)
    log(f"fast: {fast_s * 1e3:.2f} ms/step, {value:.2f} Mrays/s")
    log(f"timing plain fwd+adjoint: {args.plain_steps} step(s), same schedule")
    plain_s = timed(bv.KERNEL_PLAIN, args.plain_steps, 1, n_streams)
    plain_value = n_rays / plain_s / 1e6
    log(f"plain: {plain_s * 1e3:.2f} ms/step, {plain_value:.2f} Mrays/s")
    # per-batch schedule: one call per batch on 32 streams, and on one stream
    fast_pb_s = timed(bv.KERNEL_FAST, args.steps, 1, 32, per_batch)
    plain_pb_s = timed(bv.KERNEL_PLAIN, args.plain_steps, 1, 32, per_batch)
    log(f"per batch, 32 streams: fast {n_rays / fast_pb_s / 1e6:.2f}, plain {n_rays / plain_pb_s / 1e6:.2f} Mrays/s")
    fast_serial_s = timed(bv.KERNEL_FAST, 1, 0, 1, per_batch)
    plain_serial_s = timed(bv.KERNEL_PLAIN, 1, 0, 1, per_batch)
    log(f"per batch, one stream: fast {n_rays / fast_serial_s / 1e6:.2f}, plain {n_rays / plain_serial_s / 1e6:.2f} "
        f"Mrays/s")

    # ------------------------------------------------------------ dominant kernel: the fused training kernel
    # algorithmic bytes per adjoint launch from the instrumented walk (untimed pass)
    K = 16
    cache_steps = int(os.environ.get("BV_PATH_CACHE_STEPS", "256"))
    fwd_bytes = 0
    segs = 0
    steps_total = 0
    kernel_walk_bytes = 0  # what the fused kernel itself moves besides gradients (see below)
    for o, d, s, _ in batches:
        r = bv.trace_rays(cloud, o, d, s, settings=st, kernel=bv.KERNEL_FAST, outputs=("stats", "steps"))
        stt = r["stats"].to(torch.int64)
        nst = r["steps"].to(torch.int64)
        fwd_bytes += int(stt[:, 0].sum().item())
        segs += int(stt[:, 1].sum().item())
        steps_total += int(nst.sum().item())
        cached = nst <= cache_steps
        # forward walk + 12 B target; a cached ray writes and reads one 64 B
        # path-cache record per step, a longer ray walks again (no rgba store)
        kernel_walk_bytes += int(stt[:, 0].sum().item()) + 12 * int(nst.numel()) \
            + int((128 * nst[cached]).sum().item()) + int((stt[~cached, 0] - 16).sum().item())
    # fused batch = forward walk (incl. the 16 B rgba store) + 12 B target read + adjoint replay of the walk
    # + per integrated segment 4 (d sigma) + 12K (d SH) + 24 (d xyz) logical gradient bytes
    # (fwd_bytes holds 16 B per ray for the rgba store; the replay does not store it)
    adj_bytes = 2 * fwd_bytes - 16 * n_rays + 12 * n_rays + segs * (4 + 12 * K + 24)
    adj_bytes_per_launch = adj_bytes / len(calls)
    achieved = adj_bytes_per_launch / (adj_ms * 1e-3) / 1e9
    seg_per_ray = segs / n_rays
    kernel_bytes = kernel_walk_bytes + segs * (4 + 12 * K + 24)
    kernel_bytes_per_launch = kernel_bytes / len(calls)
    counters = measured_counters()
    log(f"fused launch {adj_ms:.3f} ms: {achieved:.1f} GB/s algorithmic, "
        f"{kernel_bytes_per_launch / (adj_ms * 1e-3) / 1e9:.1f} GB/s of kernel-needed bytes")

    # ------------------------------------------------------------ the autograd op's path
    # bv_trace_rays + L2 gradient + bv_adjoint (what RadFoamRender runs for a
    # trainer with its own loss), one call each over the step's rays
    ag_rgba = torch.empty((n_rays, 4), dtype=torch.float32, device=dev)
    ag_d = torch.empty((n_rays, 4), dtype=torch.float32, device=dev)
    ag_out = {"rgba": ag_rgba}

    def autograd_step() -> None:
        for g in grads.values():
            g.zero_()
        bv.trace_rays(cloud, O, D, S, settings=st, out=ag_out, stream=stream)
        bv.l2_loss_grad(ag_rgba, target, d_rgba=ag_d, stream=stream)
        bv.adjoint(cloud, O, D, S, ag_rgba, ag_d, grads, settings=st, stream=stream)

    autograd_step()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        autograd_step()
    torch.cuda.synchronize()
    ag_s = (time.perf_counter() - t) / args.steps
    log(f"autograd path (trace + L2 gradient + adjoint): {n_rays / ag_s / 1e6:.2f} Mrays/s")

    # ------------------------------------------------------------ one camera's rays in one call
    # every ray starts in the camera's cell and most share their next cells:
    # the fused call and the fast adjoint over one whole 1024x768 image
    cO, cD = bv.camera_rays(cams[0], W, H)
    cS = torch.full((W * H,), int(starts[0]), dtype=torch.int32, device=dev)
    cT = target_colours(bv, torch, target_scene, cO, cD, cS, st)
    c_rgba = torch.empty((W * H, 4), dtype=torch.float32, device=dev)
    c_d = torch.empty((W * H, 4), dtype=torch.float32, device=dev)

    def cam_train() -> None:
        bv.train_step_l2(cloud, cO, cD, cS, cT, grads, settings=st, rgba_out=c_rgba, stream=stream)

    def cam_adjoint() -> None:
        bv.adjoint(cloud, cO, cD, cS, c_rgba, c_d, grads, settings=st, stream=stream)

    cam_res = {"workload": f"one {W}x{H} camera's rays in one call", "rays": W * H}
    for name, fn in (("train_step_l2", cam_train), ("adjoint", cam_adjoint)):
        fn()
        if name == "train_step_l2":
            bv.l2_loss_grad(c_rgba, cT, d_rgba=c_d, stream=stream)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        cam_res[f"{name}_mrays_s"] = round(3 * W * H / (time.perf_counter() - t) / 1e6, 2)
    log(f"one camera in one call: {cam_res}")

    # ------------------------------------------------------------ forward-only over the camera ring
    fcams = cams[: (args.forward_cameras or len(cams))]
    fwd_img = bv._alloc_outputs(W * H, ("rgba",), dev)
    fwd_res = {}
    for name, kern, ncam in (("fast", bv.KERNEL_FAST, len(fcams)), ("plain", bv.KERNEL_PLAIN, min(8, len(fcams)))):
        tracer = bv.RadFoamGpuTracer(cloud, bv.RadFoamTraceSettings(1024, 0.001), kern)
        tracer.dispatch(fcams[0], (W, H), out=fwd_img, start_cell=starts[0])
        torch.cuda.synchronize()
        t = time.perf_counter()
        for k in range(ncam):
            tracer.dispatch(fcams[k], (W, H), out=fwd_img, start_cell=starts[k])
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        fwd_res[name] = dict(value=ncam * W * H / dt / 1e6, cameras=ncam, seconds=dt)
        log(f"forward {name}: {fwd_res[name]['value']:.1f} Mrays/s over {ncam} cameras")

    # ------------------------------------------------------------ path recording + compaction of one camera
    compaction = compact_phase(bv, torch, cloud, cams[0], starts[0], W, H)

    # ------------------------------------------------------------ the trainer's own GPU path: record + graph
    graph = None
    if not args.no_graph:
        graph = graph_phase(args, bv, np, torch, cloud, scene, cams, starts, cam_idx, px, py, W, H, target)

    # ------------------------------------------------------------ CPU baseline: oracle on the host
    cpu = None
    oracle_mod = None
    if not args.no_cpu:
        import oracle as oracle_mod

        osc = oracle_mod.of_host(scene)
        tgt_h = target.cpu().numpy()
        o_h, d_h, s_h = O.cpu().numpy(), D.cpu().numpy(), S.cpu().numpy().view(np.uint32)
        threads = oracle_mod.max_threads()
        t = time.perf_counter()
        done = 0
        while done < args.batches and (time.perf_counter() - t) < args.cpu_seconds:
            sl = slice(done * B, (done + 1) * B)
            r = oracle_mod.trace_rays(osc, o_h[sl], d_h[sl], s_h[sl], weight_threshold=0.001, max_steps=1024,
                                      depth=100.0, threads=0)
            g = np.concatenate([2 * (r["rgba"][:, :3] - tgt_h[sl]), np.zeros((B, 1), np.float32)], 1)
            oracle_mod.adjoint(osc, o_h[sl], d_h[sl], s_h[sl], g, weight_threshold=0.001, max_steps=1024,
                               depth=100.0, threads=0)
            done += 1
        el = time.perf_counter() - t
        cpu = dict(value=done * B / el / 1e6, unit="Mrays/s", cores=threads, kind="port",
                   sample=f"{done} batches x {B} rays of the same workload (fwd + adjoint) on {threads} OpenMP "
                          f"threads (host reports {os.cpu_count()} CPUs), {el:.1f} s",
                   host_cpus=os.cpu_count())
        log(f"cpu oracle: {cpu['value']:.4f} Mrays/s on {done} batches, {threads} threads")
    c1 = c1_phase(bv, scenes, np, torch, oracle_mod)
    if cpu is not None:
        cpu["c1_forward_mrays_s"] = c1.get("cpu_mrays_s")

    survey = None
    if not args.no_survey_reading:
        survey = survey_reading_phase(args, bv, scenes, np, torch, cams, cam_idx, px, py, W, H, st)

    gauss = None
    splats = None
    if not args.no_gauss:
        gauss = gaussian_phase(args, bv, scenes, np, torch)
        splats = splat_phase(bv, scenes, np, torch)

    serving = None
    if not args.no_serve:
        cloud_bytes = cloud.device_bytes
        serving = serve_phase(args, bv, scenes, np, torch)
        serving["training_scene_resident_mib"] = round(cloud_bytes / 2**20, 1)

    line = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(fast_s * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded mixture-on-shell scene, Delaunay adjacency from bv_compute_delaunay, "
                "random-init parameters)",
        "config": {
            "workload": "RadFoam training pass: forward + L2 loss gradient + adjoint",
            "cells": scene.n, "edges": int(scene.neighbors.size), "sh_degree": 3,
            "batch": B, "batches": args.batches, "rays_per_step": n_rays,
            "cameras": args.cameras, "image": [W, H], "max_steps": 1024, "weight_threshold": 0.001,
            "scene": {"components": 8, "angular_sigma": 0.5, "radial_sigma": 0.05, "background_fraction": 0.1,
                      "density": "LogNormal(0, 1.5)", "sh": "N(0, 0.3)", "seed": 1,
                      "note": "each mixture sample is centre + N(0, angular_sigma^2 I) re-projected to the unit "
                              "sphere, radius N(1, radial_sigma^2); survey_cluster_reading times angular_sigma "
                              "0.05"},
            "scene_build_s": round(scene_build_s, 1),
        },
        "roofline": {
            "kernel": "k_train_l2_coop<16,2,3> (forward + L2 gradient + adjoint; persistent, one launch per call)",
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": counters.get("traffic_bytes_per_launch"),
            "l2_hit_rate": counters.get("l2_hit_rate"),
            "counter_source": "profiles/train_l2_fused_1M_cells_traffic.txt (rocprofv3 --pmc over bench.py "
                              "--pmc-probe: FETCH_SIZE, WRITE_SIZE copy-calibrated; TCC_HIT, TCC_MISS)",
            "algorithmic_bytes_per_launch": int(adj_bytes_per_launch),
            "algorithmic_bytes_per_ray": round(adj_bytes / n_rays, 1),
            "kernel_bytes_per_launch": int(kernel_bytes_per_launch),
            "kernel_bytes_per_ray": round(kernel_bytes / n_rays, 1),
            "kernel_bytes_frac": round(kernel_bytes_per_launch / (adj_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "bytes_note": "algorithmic = the walk model's bytes for forward + replay walk + target + gradients "
                          "(what achieved / frac use); kernel = what this kernel must move: the forward walk, "
                          "target, a 64 B path-cache record written and read per step (rays over "
                          f"{cache_steps} steps walk again) and the gradients",
            "avg_launch_ms": round(adj_ms, 4),
            "launch_timing": f"HIP events around each of the {len(launch_events)} fused calls of the {args.steps} "
                             f"timed steps, on the call's stream ({args.batches // len(calls)} batch(es) of {B} "
                             f"rays per call; includes the call's clear / add-back kernels)",
            "step_algorithmic_gb_s": round(adj_bytes / fast_s / 1e9, 2),
            "step_frac": round(adj_bytes / fast_s / 1e9 / HBM_PEAK_GBS, 4),
        },
        "cpu_baseline": cpu,
        "plain_port": {"value": round(plain_value, 4), "ms_per_step": round(plain_s * 1e3, 4)},
        "speedup_vs_plain": round(value / plain_value, 3),
        "schedule": {"batches_per_call": args.batches // len(calls), "calls_per_step": len(calls),
                     "streams": n_streams,
                     "note": "batches of a step are independent (gradients accumulate); calls round-robin over HIP "
                             "streams; same schedule for both kernels"},
        "per_batch_32_streams": {"fast_mrays_s": round(n_rays / fast_pb_s / 1e6, 4),
                                 "plain_mrays_s": round(n_rays / plain_pb_s / 1e6, 4),
                                 "speedup_vs_plain": round(plain_pb_s / fast_pb_s, 3)},
        "one_stream": {"fast_mrays_s": round(n_rays / fast_serial_s / 1e6, 4),
                       "plain_mrays_s": round(n_rays / plain_serial_s / 1e6, 4),
                       "speedup_vs_plain": round(plain_serial_s / fast_serial_s, 3),
                       "note": "one call per batch, in series"},
        "autograd_path": {"mrays_s": round(n_rays / ag_s / 1e6, 4), "ms_per_step": round(ag_s * 1e3, 4),
                          "note": "bv_trace_rays + L2 gradient + bv_adjoint, one call each over the step's rays"},
        "forward_only": {k: {"value": round(v["value"], 2), "unit": "Mrays/s", "cameras": v["cameras"]}
                         for k, v in fwd_res.items()},
        "one_camera_call": cam_res,
        "serving": serving,
        "gaussians": gauss,
        "splats": splats,
        "path_compaction": compaction,
        "trainer_graph": graph,
        "c1_forward": c1,
        "survey_cluster_reading": survey,
        "walk": {"steps_per_ray": round(steps_total / n_rays, 2),
                 "integrated_segments_per_ray": round(seg_per_ray, 2),
                 "mean_degree": round(scene.neighbors.size / scene.n, 3)},
    }
    print(json.dumps(line), flush=True)
    cloud.deinit()
    return 0


if __name__ == "__main__":
    sys.exit(main())
