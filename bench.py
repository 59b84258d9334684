#!/usr/bin/env python3
"""Benchmark: occluder quads/s + AABB queries/s at 1920x1080 on one MI355X.

Workload: the procedural city on terrain -- 64x64 box buildings on a 512x512
fBm heightfield, 285,697 quads in 878 SAH-batched occluders -- rendered with
the reference's frame sequence (Main.cpp:180-206: clear,
setModelViewProjection, query-gated rasterize of every occluder in
front-to-back order) for K frames of a 256-frame camera path, then 65,536
AABB queries against each frame's buffer.  A step is one frame.  The draw
order of each frame is computed before timing; inputs are resident in HBM.

  value         = real submitted quads per second over the K frames, with
                  the frames in flight together: F frames per persistent
                  launch, each frame query-gated on its own buffer and then
                  queried (bit-exact with the single-view path from the same
                  buffer)
  queries_per_s = AABB queries per second of the same timed launches
  single_frame  = the same frames and queries one at a time through
                  rz_draw_frame and rz_query_batch_device

Also reports the roofline of the dominant kernel (the multi-view raster
kernel of the frames-in-flight launch), timed with HIP events on its own
stream, the CPU oracle (a scalar port of the reference) on this host's cores
as the CPU baseline, and under "batched" the 64-view mixed-resolution
workload through the batched persistent launch (occluder order and learned
order) and through the plain single-view path view by view.
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

METRIC = "occluder quads/s + AABB queries/s at 1080p, 1 MI355X: plain vs batched+learned"
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes_per_frame(quads: int, width: int, height: int) -> int:
    """16 B per submitted quad + 2 B per pixel of depth + 2 B per block of hiZ
    + the 64-byte matrix, per view-frame."""
    return 16 * quads + 2 * width * height + 2 * (width // 8) * (height // 8) + 64


def measured_traffic(kernel_profile: str, frames: int):
    """HBM bytes per launch of the multi-view raster kernel for `frames`
    frames in flight, from the calibrated FETCH_SIZE/WRITE_SIZE runs recorded
    under profiles/ (rocprofv3 cannot run inside the timed process): the
    entry measured at this frame count, else the nearest measured count
    scaled per frame.  Returns (bytes, source) or (None, reason)."""
    path = os.path.join(ROOT, "profiles", kernel_profile)
    try:
        vals = dict(line.strip().split("=", 1) for line in open(path) if "=" in line and not line.startswith("#"))
    except OSError:
        return None, f"{kernel_profile} missing"
    measured = {int(k[len("traffic_bytes_f"):]): float(v) for k, v in vals.items() if k.startswith("traffic_bytes_f")}
    if not measured:
        return None, f"{kernel_profile} has no traffic_bytes_f<frames> entry"
    if frames in measured:
        return measured[frames], f"measured at {frames} frames per launch ({kernel_profile})"
    near = min(measured, key=lambda f: abs(f - frames))
    return measured[near] * frames / near, f"measured at {near} frames per launch, scaled per frame ({kernel_profile})"


def cpu_baseline(scene, orders, threads: int, frames: int, query_reps: int, path_frames: int):
    """The oracle on the host cores over the frames the timed GPU launch
    draws (frames 0 .. path_frames - 1 of the camera path, cycled to
    `frames` frames), one oracle instance per thread."""
    import oracle

    m = scene.meta
    w, h = m["width"], m["height"]
    nf = m["mvps"].shape[0]
    span = max(1, min(path_frames, nf))
    sel = np.arange(frames) % span
    mvps = m["mvps"][sel]
    ords = orders[sel]
    t_mt = oracle.bench_frames(w, h, scene.occluders, mvps, ords, threads)
    single_frames = max(4, frames // 32)
    t_st = oracle.bench_frames(w, h, scene.occluders, mvps[:single_frames], ords[:single_frames], 1)
    # queries against the last path frame's buffer
    r = oracle.OracleRasterizer(w, h)
    last = nf - 1
    r.clear()
    r.setModelViewProjection(m["mvps"][last])
    r.draw_frame([scene.occluders[i] for i in orders[last]])
    nq = m["queries"].shape[0]
    t_q = oracle.bench_queries(r, m["queries"], query_reps, threads)
    q = scene.quad_count
    return {
        "value": frames * q / t_mt,
        "unit": "quads/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"frames 0-{span - 1} of the 256-frame path (the frames the timed launch draws), cycled to "
                   f"{frames} frames ({frames * q:,} quads) on {threads} threads, one oracle "
                   f"instance per thread: {t_mt:.2f} s; single thread {single_frames} frames: "
                   f"{single_frames * q / t_st:.3e} quads/s; queries: {query_reps}x{nq:,} on {threads} threads "
                   f"= {query_reps * nq / t_q:.3e} queries/s"),
        "queries_per_s": query_reps * nq / t_q,
        "single_thread_quads_per_s": single_frames * q / t_st,
    }


def views_bench(steps: int, warmup: int, seed: int = 3) -> dict:
    """The multi-view workload (64 views of mixed resolution, per-view
    occluder subsets of 1k-200k quads, 16,384 queries per view) two ways:
    batched -- one graph replay per step of the persistent multi-view launch
    (classify, draw, query every view) -- and plain -- the same views one by
    one through the single-view path (clear, setModelViewProjection,
    query-gated rz_draw_frame, then a query batch).  Units: submitted real
    quads per second and queries per second, wall time over the timed steps."""
    from rasterizer_amd import DeviceBuffer, Rasterizer, Views, scenes

    t0 = time.perf_counter()
    sc = scenes.config3(seed=seed)
    views = sc.meta["views"]
    nv = len(views)
    total_quads = sum(v["quads"] for v in views)
    total_q = sum(v["queries"].shape[0] for v in views)
    pixels = sum(v["size"][0] * v["size"][1] for v in views)
    log(f"views: {nv} views, {pixels:,} px, {total_quads:,} quads, {total_q:,} queries "
        f"({time.perf_counter() - t0:.1f} s)")
    n_mvp = 16
    mvps = [[scenes.view_mvp(v, s) for v in views] for s in range(n_mvp)]

    owner = Rasterizer(64, 64)
    V = Views(owner, [v["size"] for v in views])
    for k, v in enumerate(views):
        V.set_occluders(k, v["occluders"])
        V.set_queries(k, v["queries"])

    def batched_step(s: int) -> None:
        for k in range(nv):
            V.set_mvp(k, mvps[s % n_mvp][k])
        V.run(graph=True)

    def time_batched(what: str):
        for s in range(warmup):
            batched_step(s)
        V.sync()
        t0 = time.perf_counter()
        for s in range(steps):
            batched_step(s)
        V.sync()
        t = time.perf_counter() - t0
        gpu_ms = []
        for s in range(min(steps, 8)):
            batched_step(s)
            gpu_ms.append(V.sync())
        log(f"{what}: {t * 1e3 / steps:.3f} ms/step (GPU {np.median(gpu_ms):.3f} ms), {V.info()}")
        return t, float(np.median(gpu_ms))

    def block_counters():
        """hiZ/edge rejects of (quad, block) pairs over one step (separate, untimed)."""
        V.enable_stats(True)
        batched_step(1)  # the learned order reads the previous step's labels
        batched_step(2)
        V.sync()
        st = V.stats()
        V.enable_stats(False)
        return {**st, "reject_rate": st["rejected"] / max(st["visited"], 1)}

    # Launch configuration by the autotuner (identical bits under every
    # candidate; untimed), on this workload's first matrices.
    batched_step(0)
    V.sync()
    tuned = V.autotune(reps=3)
    log(f"views autotune: {tuned}")
    # Occluder order: each view's subset front to back for its eye.
    t_b, gpu_b = time_batched("batched, occluder order")
    cnt_b = block_counters()
    # Learned order: the trained MLP scores every batch each step (features
    # of this step's setup, with or without labels of the previous step);
    # occluders are drawn by their best batch's score.  Two models: trained
    # without the previous-label feature (so the draw tracks no labels), and
    # with it.
    from rasterizer_amd import learn

    learned = {}
    models = [("without labels", learn.WEIGHTS.with_name("order_mlp_nolabels.npz")), ("with labels", learn.WEIGHTS)]
    scene_model = learn.WEIGHTS.with_name("order_mlp_boxes3.npz")  # trained on this workload's box scene
    if scene_model.exists():
        models.insert(0, ("trained on this scene, without labels", scene_model))
    for name, path in models:
        V.set_model(learn.load(path), by_occluder=True)
        t_l, gpu_l = time_batched(f"batched, learned order ({name})")
        cnt_l = block_counters()
        log(f"reject rate: occluder order {cnt_b['reject_rate']:.4f}, learned ({name}) {cnt_l['reject_rate']:.4f}")
        V.set_model(None)
        learned[name] = {
            "granularity": "occluder (best batch score)",
            "weights": str(path.relative_to(ROOT)),
            "quads_per_s": steps * total_quads / t_l,
            "queries_per_s": steps * total_q / t_l,
            "ms_per_step": t_l * 1e3 / steps,
            "gpu_ms_per_step": gpu_l,
            "block_counters": cnt_l,
        }

    # Plain: one context per view, the single-view frame path.
    ctxs = [Rasterizer(*v["size"]) for v in views]
    handles = [ctxs[k].handle_array(v["occluders"]) for k, v in enumerate(views)]
    dq = [(DeviceBuffer(ctxs[k], v["queries"].nbytes, v["queries"]), DeviceBuffer(ctxs[k], v["queries"].shape[0]),
           DeviceBuffer(ctxs[k], v["queries"].shape[0])) for k, v in enumerate(views)]

    def plain_step(s: int) -> None:
        for k, v in enumerate(views):
            R = ctxs[k]
            R.clear()
            R.setModelViewProjection(mvps[s % n_mvp][k])
            R.draw_frame_handles(handles[k], len(v["occluders"]))
            R.query_batch_device(v["queries"].shape[0], *dq[k])

    plain_steps = max(2, steps // 4)
    for s in range(min(warmup, 2)):
        plain_step(s)
    for R in ctxs:
        R.sync()
    t0 = time.perf_counter()
    for s in range(plain_steps):
        plain_step(s)
    for R in ctxs:
        R.sync()
    t_p = time.perf_counter() - t0
    log(f"plain per-view: {t_p * 1e3 / plain_steps:.3f} ms/step")
    plain_qps = plain_steps * total_quads / t_p
    return {
        "workload": (f"{nv} views ({pixels:,} px: 16x2048^2, 24x1024^2, 16x512^2, 8x1920x1080), "
                     f"{total_quads:,} quads submitted per step from a 240,000-quad box scene, "
                     f"{total_q:,} AABB queries per step"),
        "batched_quads_per_s": steps * total_quads / t_b,
        "batched_queries_per_s": steps * total_q / t_b,
        "batched_ms_per_step": t_b * 1e3 / steps,
        "batched_gpu_ms_per_step": gpu_b,
        "batched_block_counters": cnt_b,
        "plain_quads_per_s": plain_qps,
        "plain_ms_per_step": t_p * 1e3 / plain_steps,
        "speedup_vs_plain": (steps * total_quads / t_b) / plain_qps,
        "steps": steps,
        "plain_steps": plain_steps,
        "launch_config": tuned,
        "learned_order": {name: {**d, "speedup_vs_plain": d["quads_per_s"] / plain_qps}
                          for name, d in learned.items()},
    }


def stress_bench(steps: int, warmup: int, seed: int = 5) -> dict:
    """The stress configuration: 50,000 icosphere instances (80 quads each,
    one occluder each) on a 2 km terrain, 4.26M quads, seen by a 3840x2160
    street-level view and 15 overhead 1024^2 views, each submitting the whole
    scene in front-to-back order, 65,536 queries per view.  Batched (one
    classify-then-draw graph replay per step, occluder order and learned
    order) and plain (each view through rz_draw_frame on one context per
    resolution, then its queries).  Units: submitted real quads per second."""
    from rasterizer_amd import DeviceBuffer, Rasterizer, Views, learn, scenes

    t0 = time.perf_counter()
    sc = scenes.config5(seed=seed)
    views = sc.meta["views"]
    nv = len(views)
    orders = [scenes.front_to_back(sc.occluders, v["eye"]) for v in views]
    total_quads = nv * sc.quad_count
    total_q = sum(v["queries"].shape[0] for v in views)
    log(f"stress: {len(sc.occluders)} occluders, {sc.quad_count:,} quads, {nv} views "
        f"({time.perf_counter() - t0:.1f} s)")
    sizes = [v["size"] for v in views]
    ctx = {s: Rasterizer(*s) for s in sorted(set(sizes))}
    owner = ctx[sizes[0]]
    V = Views(owner, sizes)
    for k, v in enumerate(views):
        V.set_mvp(k, v["mvp"])
        V.set_occluders(k, [sc.occluders[i] for i in orders[k]])
        V.set_queries(k, v["queries"])
    t0 = time.perf_counter()
    V.run(graph=True)
    V.sync()
    log(f"stress: plan built and run ({time.perf_counter() - t0:.1f} s), {V.info()}")
    tuned = V.autotune(reps=2)
    log(f"stress autotune: {tuned}")

    def timed(what: str):
        for _ in range(warmup):
            V.run(graph=True)
        V.sync()
        t = time.perf_counter()
        for _ in range(steps):
            V.run(graph=True)
        V.sync()
        t = time.perf_counter() - t
        log(f"stress {what}: {t * 1e3 / steps:.3f} ms/step, {steps * total_quads / t / 1e6:.1f} M quads/s")
        return t

    t_b = timed("batched, occluder order")
    t_learned = {}
    for name, path in (("without labels", learn.WEIGHTS.with_name("order_mlp_nolabels.npz")),
                       ("with labels", learn.WEIGHTS)):
        V.set_model(learn.load(path), by_occluder=True)
        t_learned[name] = timed(f"batched, learned order ({name})")
        V.set_model(None)
    t_l = t_learned["without labels"]

    handles = [ctx[sizes[k]].handle_array([sc.occluders[i] for i in orders[k]]) for k in range(nv)]
    dq = [(DeviceBuffer(ctx[sizes[k]], v["queries"].nbytes, v["queries"]),
           DeviceBuffer(ctx[sizes[k]], v["queries"].shape[0]), DeviceBuffer(ctx[sizes[k]], v["queries"].shape[0]))
          for k, v in enumerate(views)]

    def plain_step() -> None:
        for k, v in enumerate(views):
            R = ctx[sizes[k]]
            R.clear()
            R.setModelViewProjection(v["mvp"])
            R.draw_frame_handles(handles[k], len(sc.occluders))
            R.query_batch_device(v["queries"].shape[0], *dq[k])

    plain_step()
    for R in ctx.values():
        R.sync()
    plain_steps = max(1, steps // 4)
    t0 = time.perf_counter()
    for _ in range(plain_steps):
        plain_step()
    for R in ctx.values():
        R.sync()
    t_p = time.perf_counter() - t0
    log(f"stress plain per-view: {t_p * 1e3 / plain_steps:.3f} ms/step")
    plain_qps = plain_steps * total_quads / t_p
    return {
        "workload": (f"{nv} views (1x3840x2160 street level, 15x1024^2 overhead), {sc.quad_count:,} quads "
                     f"({sc.meta['instances']:,} icosphere instances + terrain) submitted per view, "
                     f"{total_q:,} AABB queries per step"),
        "batched_quads_per_s": steps * total_quads / t_b,
        "batched_ms_per_step": t_b * 1e3 / steps,
        "learned_quads_per_s": steps * total_quads / t_l,
        "learned_ms_per_step": t_l * 1e3 / steps,
        "learned_model": "rasterizer_amd/order_mlp_nolabels.npz",
        "learned_with_labels_ms_per_step": t_learned["with labels"] * 1e3 / steps,
        "plain_quads_per_s": plain_qps,
        "plain_ms_per_step": t_p * 1e3 / plain_steps,
        "speedup_vs_plain": (steps * total_quads / t_b) / plain_qps,
        "learned_speedup_vs_plain": (steps * total_quads / t_l) / plain_qps,
        "queries_per_s_batched": steps * total_q / t_b,
        "steps": steps,
        "plain_steps": plain_steps,
        "launch_config": tuned,
    }


def learned_frames_in_flight(R, scene, orders, handles, plans, F: int, tuned: dict, t_plain: float,
                             plain_value: float):
    """The frames-in-flight launches again, each frame's occluders ordered by
    the city-trained MLP (whole occluders by their best batch's score, label
    tracking off) instead of front to back, still query-gated: the batched +
    learned path at 1920x1080.  Frames 0 and F-1 are checked bit for bit
    against the single-view path drawing the occluders in the order the
    device chose, queries included.  Reports quads/s (submitted real quads),
    occluders drawn per frame against the front-to-back launches, and the
    time."""
    from rasterizer_amd import learn

    path = learn.WEIGHTS.with_name("order_mlp_city.npz")
    if not path.exists():
        return None
    model = learn.load(path)
    m = scene.meta
    W, H = m["width"], m["height"]
    nf = m["mvps"].shape[0]
    n_occ = len(scene.occluders)
    K = F * len(plans)

    def drawn(V):
        return sum(int((V.status(v) > 0).sum()) for v in range(F))

    plain_drawn = [drawn(Vj) for Vj in plans]
    for Vj in plans:
        Vj.set_model(model, by_occluder=True)
        Vj.set_config(tuned["chosen"]["waves"], tuned["chosen"]["groupsPerCU"], tuned["chosen"]["tileW"])
        Vj.run(graph=True)
        Vj.sync()
    # Check the first launch's first and last frames (each drawn from the
    # buffer its previous launch left, which the single-view path is given).
    nb = (W // 8) * (H // 8)
    V0 = plans[0]
    before = [V0.raw(v) for v in sorted({0, F - 1})]
    V0.run(graph=True)
    V0.sync()
    for (d0, h0), v in zip(before, sorted({0, F - 1})):
        order = V0.order(v)
        f = v % nf
        occ = [scene.occluders[i] for i in orders[f]]
        R.upload_raw(d0, h0)
        R.clear()
        R.setModelViewProjection(m["mvps"][f])
        st = R.draw_frame([occ[i] for i in order])
        want = np.empty_like(st)
        want[order] = st
        d, h = R.raw()
        dv, hv = V0.raw(v)
        vis, clip = R.query_batch(m["queries"])
        vv, cv = V0.queries(v)
        if not (np.array_equal(d, dv) and np.array_equal(h, hv) and np.array_equal(V0.status(v), want) and
                np.array_equal(vis, vv) and np.array_equal(clip, cv)):
            raise RuntimeError(f"batched + learned: frame {f} differs from the single-view path in the learned order")
    for Vj in plans:
        Vj.run(graph=True)
    for Vj in plans:
        Vj.sync()
    t0 = time.perf_counter()
    for Vj in plans:
        Vj.run(graph=True)
    for Vj in plans:
        Vj.sync()
    t = time.perf_counter() - t0
    learned_drawn = [drawn(Vj) for Vj in plans]
    phases = []
    for Vj in plans:
        Vj.profile(True)
        Vj.run(graph=False)
        phases.append(Vj.profile_read())
        Vj.profile(False)
    for Vj in plans:
        Vj.set_model(None)
    value = K * scene.quad_count / t
    log(f"batched + learned (city model, gated, 1080p): {K} frames {t * 1e3:.3f} ms ({value / 1e6:.1f} M quads/s); "
        f"occluders drawn {sum(learned_drawn)} vs front to back {sum(plain_drawn)}")
    return {
        "model": str(path.relative_to(ROOT)),
        "quads_per_s": value,
        "queries_per_s": K * m["queries"].shape[0] / t,
        "ms_per_frame": t * 1e3 / K,
        "front_to_back_quads_per_s": plain_value,
        "front_to_back_ms_per_frame": t_plain * 1e3 / K,
        "occluders_drawn_per_frame": sum(learned_drawn) / K,
        "front_to_back_occluders_drawn_per_frame": sum(plain_drawn) / K,
        "occluders_per_frame": n_occ,
        "raster_kernel_ms_per_launch": float(np.mean([p["raster_ms"] for p in phases])),
        "setup_phase_ms_per_launch": float(np.mean([p["setup_ms"] for p in phases])),
    }


def frames_per_launch(k: int, cap: int) -> int:
    """Largest divisor of k not above cap: K frames in equal launches."""
    f = max(1, min(k, cap))
    while k % f:
        f -= 1
    return f


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=128, help="frames timed (K)")
    ap.add_argument("--warmup", type=int, default=16)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-frames", type=int, default=16384)
    ap.add_argument("--no-views", action="store_true", help="skip the batched multi-view measurement")
    ap.add_argument("--views-steps", type=int, default=32)
    ap.add_argument("--frames-in-flight", type=int, default=128, help="frames per persistent launch (at most)")
    ap.add_argument("--no-stress", action="store_true", help="skip the stress configuration")
    ap.add_argument("--no-learned", action="store_true", help="skip the batched + learned frames-in-flight leg")
    ap.add_argument("--stress-steps", type=int, default=8)
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="first copy this many MiB device-to-device 3 times (a kernel of known traffic for "
                         "scaling the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for {args.gpus}).",
              file=sys.stderr)
        return 2

    from rasterizer_amd import DeviceBuffer, Rasterizer, Views, scenes

    if args.calibrate_copy_mib > 0:
        import torch

        a = torch.ones(args.calibrate_copy_mib << 18, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"copied {args.calibrate_copy_mib} MiB 3 times")
        del a, b

    t0 = time.perf_counter()
    scene = scenes.config2(seed=args.seed)
    m = scene.meta
    W, H = m["width"], m["height"]
    nf = m["mvps"].shape[0]
    K = args.steps
    log(f"scene: {len(scene.occluders)} occluders, {scene.quad_count} quads, {nf} frames "
        f"({time.perf_counter() - t0:.1f} s)")
    orders = np.stack([scenes.front_to_back(scene.occluders, m["eyes"][f]) for f in range(nf)])

    R = Rasterizer(W, H)
    handles = [R.handle_array([scene.occluders[i] for i in orders[f]]) for f in range(nf)]
    n_occ = len(scene.occluders)
    q = m["queries"]
    nq = q.shape[0]
    d_bounds = DeviceBuffer(R, q.nbytes, q)
    d_vis = DeviceBuffer(R, nq)
    d_clip = DeviceBuffer(R, nq)

    # --- Frames in flight: frames k = 0..K-1 (mod the path) as query-gated
    # views, F per persistent launch, each on its own 1920x1080 buffer.
    F = frames_per_launch(K, args.frames_in_flight)
    launches = K // F
    V = Views(R, [(W, H)] * F)
    V.set_query_gated(True)

    def load_launch(j: int) -> None:
        for v in range(F):
            f = (j * F + v) % nf
            V.set_occluders(v, [scene.occluders[i] for i in orders[f]])
            V.set_queries(v, q)
            V.set_mvp(v, m["mvps"][f])

    plans = []
    if launches == 1:
        load_launch(0)
        t0 = time.perf_counter()
        V.run(graph=True)
        V.sync()
        log(f"frames in flight: {F} frames per launch, plan built and run ({time.perf_counter() - t0:.1f} s), "
            f"{V.info()}")
    else:
        # Several launches: one Views per launch, so that no plan is rebuilt
        # inside the timed region.
        for j in range(launches):
            Vj = V if j == 0 else Views(R, [(W, H)] * F)
            Vj.set_query_gated(True)
            for v in range(F):
                f = (j * F + v) % nf
                Vj.set_occluders(v, [scene.occluders[i] for i in orders[f]])
                Vj.set_queries(v, q)
                Vj.set_mvp(v, m["mvps"][f])
            Vj.run(graph=True)
            Vj.sync()
            plans.append(Vj)
    plans = plans or [V]
    # Bit-exact check against the single-view path from the same (zeroed)
    # buffer: the views' first run started from zeroed buffers.
    nb = (W // 8) * (H // 8)
    for v in sorted({0, F - 1}):
        f = v % nf
        R.upload_raw(np.zeros(nb * 64, np.uint16), np.zeros(nb + 8, np.uint16))
        R.clear()
        R.setModelViewProjection(m["mvps"][f])
        R.draw_frame_handles(handles[f], n_occ)
        d, h = R.raw()
        dv, hv = plans[0].raw(v)
        vis, clip = R.query_batch(q)
        vv, cv = plans[0].queries(v)
        if not (np.array_equal(d, dv) and np.array_equal(h, hv) and np.array_equal(vis, vv) and
                np.array_equal(clip, cv)):
            raise RuntimeError(f"frames in flight: frame {f} differs from the single-view path")
    # Launch configuration by the autotuner on the first plan (identical bits
    # required; untimed), applied to every plan.
    fif_tuned = plans[0].autotune(reps=2)
    log(f"frames in flight autotune: {fif_tuned}")
    for Vj in plans[1:]:
        Vj.set_config(fif_tuned["chosen"]["waves"], fif_tuned["chosen"]["groupsPerCU"], fif_tuned["chosen"]["tileW"])
        Vj.run(graph=True)
        Vj.sync()
    for Vj in plans:  # warm-up launch of each plan (untimed)
        Vj.run(graph=True)
    for Vj in plans:
        Vj.sync()
    t0 = time.perf_counter()
    for Vj in plans:
        Vj.run(graph=True)
    for Vj in plans:
        Vj.sync()
    t_fif = time.perf_counter() - t0
    # Per-phase HIP-event timing of the same launches (separate pass).
    phases = []
    for Vj in plans:
        Vj.profile(True)
        Vj.run(graph=False)
        phases.append(Vj.profile_read())
        Vj.profile(False)
    raster_ms = float(np.mean([p["raster_ms"] for p in phases]))
    setup_ms = float(np.mean([p["setup_ms"] for p in phases]))
    query_ms = float(np.mean([p["query_ms"] for p in phases]))
    value = K * scene.quad_count / t_fif
    log(f"frames in flight: {K} frames in {launches} launch(es) of {F}: {t_fif * 1e3:.3f} ms "
        f"({value / 1e6:.1f} M quads/s, {K * nq / t_fif / 1e9:.2f} G queries/s); per launch: raster kernel "
        f"{raster_ms:.3f} ms, setup phase {setup_ms:.3f} ms, queries {query_ms:.3f} ms")

    learned_fif = (None if args.no_learned else
                   learned_frames_in_flight(R, scene, orders, handles, plans, F, fif_tuned, t_fif, value))

    # --- Single frame at a time (the plain single-view path).
    def raster_step(k: int) -> None:
        f = k % nf
        R.clear()
        R.setModelViewProjection(m["mvps"][f])
        R.draw_frame_handles(handles[f], n_occ)

    def query_step() -> None:
        R.query_batch_device(nq, d_bounds, d_vis, d_clip)

    for k in range(args.warmup):
        raster_step(k)
        query_step()
    R.sync()
    t0 = time.perf_counter()
    for k in range(K):
        raster_step(k)
    R.sync()
    t_raster = time.perf_counter() - t0
    t0 = time.perf_counter()
    for k in range(K):
        query_step()
    R.sync()
    t_query = time.perf_counter() - t0
    log(f"single frame: raster {t_raster * 1e3 / K:.3f} ms/frame, queries {t_query * 1e3 / K:.3f} ms/batch")
    R.profile(True)
    for k in range(K):
        raster_step(k)
        query_step()
    prof = R.profile_read()
    R.profile(False)
    frame_ms = prof["frame_ms"] / max(prof["draws"], 1)
    kernel_ms = {"prep": prof["prep_ms"] / max(prof["draws"], 1), "setup": prof["setup_ms"] / max(prof["draws"], 1),
                 "frame": frame_ms, "query": prof["query_ms"] / max(prof["query_batches"], 1)}
    log("single-frame kernels: " + ", ".join(f"{k} {v:.4f} ms" for k, v in kernel_ms.items()))

    traffic_bytes, traffic_source = measured_traffic("rz_views_kernel_city1080_traffic.txt", F)
    alg_frame = algorithmic_bytes_per_frame(scene.quad_count, W, H)
    alg_launch = F * alg_frame
    achieved = alg_launch / (raster_ms * 1e-3) / 1e9
    queries_per_s = K * nq / t_fif
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "quads/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": t_fif * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"city+terrain {W}x{H}: {scene.quad_count:,} quads in {n_occ} occluders, frames 0-{K - 1} "
                         f"of a {nf}-frame street-level camera path, query-gated draw in front-to-back order, "
                         f"then {nq:,} AABB queries against each frame's buffer, in the same timed launches"),
            "mode": f"plain method, {F} frames in flight per persistent launch (query-gated views)",
            "seed": args.seed,
        },
        "queries_per_s": queries_per_s,
        "frames_in_flight": {"frames_per_launch": F, "launches": launches, "wall_ms": t_fif * 1e3,
                             "launch_config": fif_tuned,
                             "raster_kernel_ms_per_launch": raster_ms, "setup_phase_ms_per_launch": setup_ms,
                             "query_kernel_ms_per_launch": query_ms},
        "single_frame": {"quads_per_s": K * scene.quad_count / t_raster, "ms_per_frame": t_raster * 1e3 / K,
                         "kernel_ms": kernel_ms,
                         "frame_kernel_hbm_gbs": alg_frame / (frame_ms * 1e-3) / 1e9,
                         "queries_per_s": K * nq / t_query, "query_ms_per_batch": t_query * 1e3 / K},
        "batched_learned_1080p": learned_fif,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_bytes,
            "traffic_source": traffic_source,
            "kernel": "rz_views_kernel (query-gated, frames in flight)",
            "algorithmic_bytes": alg_launch,
        },
        "cpu_baseline": None,
    }
    if not args.no_views:
        out["batched"] = views_bench(args.views_steps, min(args.warmup, 4))
    if not args.no_stress:
        out["stress"] = stress_bench(args.stress_steps, 2)
    if not args.no_cpu:
        threads = min(len(os.sched_getaffinity(0)), 16)
        t0 = time.perf_counter()
        out["cpu_baseline"] = cpu_baseline(scene, orders, threads, args.cpu_frames, 4, K)
        log(f"cpu baseline {time.perf_counter() - t0:.1f} s")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
