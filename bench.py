"""Benchmark of the MI355X coverage rasterizer.

Headline workload (one GPU): Rasterizer::mask of one path of 200k segments
(40% lines, 30% quads, 25% cubics, 5% elliptic arcs; seeded random walk,
random affine) into a 4096x4096 f32 mask with SignedDifferenceRasterizer
semantics, nonzero fill. One step = the whole pipeline (flatten -> bin ->
fine) with the path already resident in HBM. Also measured on the same GPU
and reported in the same JSON line: the batched 10k-path scene render next to
its plain one-pipeline-per-path form, and the reference's own many-circles
scene benchmark. Prints one JSON line.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_mask_c2_4096.json")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--rule", choices=["nonzero", "evenodd"], default="nonzero")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-scene", action="store_true", help="skip the scene measurements")
    ap.add_argument("--workload", choices=["scene", "circles", "atlas", "diff", "stroke"],
                    help="only this measurement, printed as one JSON line (for per-kernel "
                         "profiles of one workload)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy a 256 MiB device buffer 3 times (calibrates HBM counters)")
    ap.add_argument("--batched-only", action="store_true",
                    help="with --workload scene/circles: only the batched form")
    ap.add_argument("--diff-steps", type=int, default=200, help="Adam steps of the differentiable render")
    ap.add_argument("--cpu-reps", type=int, default=16)
    return ap.parse_args()


def copy_bandwidth_gbps(torch):
    """Achievable device copy bandwidth (read + write bytes / time)."""
    n = 1 << 28  # 1 GiB of f32
    a = torch.empty(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    reps = 10
    for _ in range(reps):
        b.copy_(a)
    e.record()
    torch.cuda.synchronize()
    ms = s.elapsed_time(e) / reps
    del a, b
    return 2.0 * n * 4 / (ms * 1e-3) / 1e9


def kernel_stats(lib, _abi, ctx, run, reps):
    """Per-kernel (launches, device ms) per run from events on the context's stream."""
    lib.rz_profile(ctx.h, 1)
    for _ in range(reps):
        run()
    stats = (_abi.rz_kernel_stat * 64)()
    n = C.c_size_t()
    ctx.check(lib.rz_profile_stats(ctx.h, stats, 64, C.byref(n)))
    lib.rz_profile(ctx.h, 0)
    return {stats[i].name.decode(): (stats[i].calls / reps, stats[i].total_ms / reps)
            for i in range(n.value)}


def set_graph(lib, _abi, ctx, on):
    """HIP-graph replay of repeated frames on the context (RZ_PARAM_GRAPH)."""
    ctx.check(lib.rz_set_param(ctx.h, _abi.RZ_PARAM_GRAPH, int(on)))


def settle(lib, _abi, ctx, run):
    """Run until the context's buffers fit (a first run may grow them)."""
    for _ in range(4):
        run()
        s = lib.rz_sync(ctx.h)
        if s != _abi.RZ_E_CAPACITY:
            break
    ctx.check(s)


def last_counts(lib, ctx):
    """Flattened lines and cell records of the context's last run."""
    nl, nr, ne = C.c_size_t(), C.c_size_t(), C.c_size_t()
    ctx.check(lib.rz_last_counts(ctx.h, C.byref(nl), C.byref(nr), C.byref(ne)))
    return nl.value, nr.value


def scene_measure(torch, rz, lib, _abi, name, ws, W, H, frames_b, frames_p, forms=None):
    """Batched and plain Scene::render of a generated scene into a W x H layer."""
    ds = rz.DeviceScene(ws.scene())
    ctx = ds.ctx
    out = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    view = (0.0, 0.0, float(W), float(H))
    stream = torch.cuda.ExternalStream(lib.rz_stream(ctx.h))
    ident = rz.Transform.identity()
    res = {}
    for form, batched, frames in (("batched", True, frames_b), ("plain", False, frames_p),
                                  ("graph", True, frames_b)):
        if forms is not None and form not in forms:
            continue
        set_graph(lib, _abi, ctx, form == "graph")

        def run():
            ds.render_device(ident, view, out.data_ptr(), W, batched=batched)
        settle(lib, _abi, ctx, run)
        if form == "graph":  # the first clean run settles, the next one is captured
            for _ in range(2):
                run()
                ctx.check(lib.rz_sync(ctx.h))
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(frames):
            run()
        b.record(stream)
        ctx.check(lib.rz_sync(ctx.h))
        torch.cuda.synchronize()
        res[form] = a.elapsed_time(b) / frames
    set_graph(lib, _abi, ctx, False)
    k = kernel_stats(lib, _abi, ctx, lambda: ds.render_device(ident, view, out.data_ptr(), W), 3)
    ctx.check(lib.rz_sync(ctx.h))
    n_lines, n_recs = last_counts(lib, ctx)
    ds.close()
    segs = sum(p.segments_count() for p in ws.paths)
    return {
        "workload": name,
        "paths": len(ws.paths),
        "segments": segs,
        "lines": n_lines,
        "cell_records": n_recs,
        "layer": [W, H],
        "batched_ms_per_frame": round(res["batched"], 4),
        "batched_graph_ms_per_frame": round(res["graph"], 4) if "graph" in res else None,
        "plain_ms_per_frame": round(res["plain"], 3) if "plain" in res else None,
        "batched_over_plain": round(res["plain"] / res["batched"], 1) if "plain" in res else None,
        "batched_mpix_s": round(W * H / res["batched"] / 1e3, 1),
        "batched_msegs_s": round(segs / res["batched"] / 1e3, 2),
        "batched_kernels_ms": {n: round(v[1], 4) for n, v in sorted(k.items())},
    }


def atlas_measure(torch, rz, lib, _abi, frames):
    """Glyph atlas: 65,536 glyph masks into one atlas, one pipeline per frame."""
    gs = rz.workload_glyphs(4, 65_536, 16384)
    W, H = gs.atlas_size
    dg = rz.DeviceGlyphs(gs.paths, gs.slots)
    ctx = dg.ctx
    out = torch.zeros((H, W), dtype=torch.float32, device="cuda")

    def run():
        dg.mask_device(out.data_ptr(), W)
    stream = torch.cuda.ExternalStream(lib.rz_stream(ctx.h))
    timed = {}
    for graph in (False, True):
        set_graph(lib, _abi, ctx, graph)
        settle(lib, _abi, ctx, run)
        for _ in range(2 if graph else 0):
            run()
            ctx.check(lib.rz_sync(ctx.h))
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(frames):
            run()
        b.record(stream)
        ctx.check(lib.rz_sync(ctx.h))
        torch.cuda.synchronize()
        timed[graph] = a.elapsed_time(b) / frames
    set_graph(lib, _abi, ctx, False)
    ms = timed[False]
    k = kernel_stats(lib, _abi, ctx, run, 2)
    ctx.check(lib.rz_sync(ctx.h))
    n_lines, n_recs = last_counts(lib, ctx)
    dg.close()
    px = sum(s.width * s.height for s in gs.slots)
    segs = sum(p.segments_count() for p in gs.paths)
    del out
    return {
        "workload": "glyph atlas: 65,536 glyphs of 8-200 quads (1-3 contours), masks 16-128 px, "
                    f"one {W}x{H} f32 atlas, NonZero",
        "glyphs": len(gs.paths),
        "segments": segs,
        "lines": n_lines,
        "cell_records": n_recs,
        "mask_pixels": px,
        "ms_per_frame": round(ms, 4),
        "graph_ms_per_frame": round(timed[True], 4),
        "mpix_s": round(px / ms / 1e3, 1),
        "msegs_s": round(segs / ms / 1e3, 1),
        "kernels_ms": {n: round(v[1], 4) for n, v in sorted(k.items())},
    }, gs


def stroke_measure(torch, rz, lib, _abi, reps, cpu=True):
    """Path::stroke of the 200k-segment random walk (the mask benchmark's path;
    width 2, round joins and caps) on the GPU, host path in and out as
    rz_path_stroke takes them, and the 4096x4096 mask of the stroked path:
    the reference's "stroke" and "mask stroked" benches
    (benches/rasterize_bench.rs:35-74) on that path."""
    import time
    ws = rz.workload_random_walk(2, 200_000, 4096.0)
    path, tr = ws.paths[0], ws.transform
    ctx = rz.Context(0)
    style = rz.StrokeStyle(2.0, rz.LineJoin.Round, rz.LineCap.Round)
    stroked = path.stroke(style, ctx)
    t0 = time.perf_counter()
    for _ in range(reps):
        stroked = path.stroke(style, ctx)
    ms = (time.perf_counter() - t0) * 1e3 / reps
    k = kernel_stats(lib, _abi, ctx, lambda: path.stroke(style, ctx), 2)
    W = H = 4096
    dp = rz.DevicePath(stroked, ctx)
    out = torch.zeros((H, W), dtype=torch.float32, device="cuda")

    def run():
        dp.mask_device(tr, W, H, out.data_ptr(), W)
    settle(lib, _abi, ctx, run)
    stream = torch.cuda.ExternalStream(lib.rz_stream(ctx.h))
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(reps):
        run()
    b.record(stream)
    ctx.check(lib.rz_sync(ctx.h))
    torch.cuda.synchronize()
    mask_ms = a.elapsed_time(b) / reps
    res = {
        "workload": "Path::stroke of the 200k-segment random walk (width 2, round joins and caps), "
                    "then its 4096x4096 mask (NonZero)",
        "segments": path.segments_count(),
        "stroked_segments": stroked.segments_count(),
        "stroke_ms": round(ms, 3),
        "stroke_msegs_s": round(path.segments_count() / ms / 1e3, 1),
        "stroke_kernels_ms": {n: round(v[1], 4) for n, v in sorted(k.items())},
        "mask_stroked_ms": round(mask_ms, 4),
        "mask_stroked_mpix_s": round(W * H / mask_ms / 1e3, 1),
    }
    dp.close()
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _oracle as O
        t, n = O.time_stroke(path, 2.0, "round", 4.0, "round", reps=1)
        res["cpu"] = {"stroke_ms": round(t * 1e3, 2), "cores": 1, "kind": "port",
                      "sample": "the oracle's Path::stroke of the same path, one thread, once",
                      "stroked_segments": n}
    return res


def diff_measure(torch, rz, steps):
    """Differentiable render: 256 closed paths x 32 cubics, 1024x1024 RGBA,
    L2 loss against a perturbed target, forward + adjoint backward + Adam;
    kernel by kernel, then with each step's forward and backward replayed
    from HIP graphs (the points and colours they read change every step)."""
    from rasterize_amd import _abi
    ds, tp, tc = rz.workload_diff(5, 256, 32, 1024.0)
    dev = rz.DeviceScene(ds.scene())
    L = _abi.lib()
    view = (0.0, 0.0, 1024.0, 1024.0)
    ident = rz.Transform.identity()
    idx = torch.tensor(ds.index, device="cuda")
    with torch.no_grad():
        T = rz.render_torch(dev, torch.tensor(tp, device="cuda")[idx],
                            torch.tensor(tc, dtype=torch.float32, device="cuda"), ident, view).clone()

    def run(graph):
        dev.ctx.check(L.rz_set_param(dev.ctx.h, _abi.RZ_PARAM_GRAPH, int(graph)))
        U = torch.tensor(ds.points, dtype=torch.float32, device="cuda", requires_grad=True)
        Cc = torch.tensor(ds.colors, dtype=torch.float32, device="cuda", requires_grad=True)
        opt = torch.optim.Adam([{"params": [U], "lr": 0.25}, {"params": [Cc], "lr": 0.005}])

        def step():
            out = rz.render_torch(dev, U[idx], Cc, ident, view)
            loss = ((out - T) ** 2).sum()
            opt.zero_grad()
            loss.backward()
            opt.step()
            return loss

        loss0 = float(step().detach())
        for _ in range(3):
            step()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        a.record()
        for _ in range(steps):
            loss = step()
        b.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        return wall * 1e3 / steps, a.elapsed_time(b) / steps, loss0, float(loss.detach())

    eager = run(False)
    graph = run(True)
    dev.ctx.check(L.rz_set_param(dev.ctx.h, _abi.RZ_PARAM_GRAPH, 0))
    dev.close()
    return {
        "workload": "differentiable render: 256 closed paths x 32 cubics (control points U[0,1024)^2), "
                    "solid RGBA, 1024x1024, L2 loss vs target (points + N(0,2px), colours + N(0,0.05))",
        "parameters": int(ds.points.size + ds.colors.size),
        "adam_steps": steps,
        "ms_per_step": round(eager[0], 3),
        "gpu_ms_per_step": round(eager[1], 3),
        "graph_ms_per_step": round(graph[0], 3),
        "loss_first": round(eager[2], 2),
        "loss_last": round(eager[3], 2),
        "gradient_check": "tests/test_gpu_fullsize.py::test_c5_gradients_full_size: 512 sampled "
                          "parameters vs FP64 central differences and the FP64 forward-mode gradient",
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design "
                                   "only (bands of tile rows per GPU), --gpus must be 1"}))
        sys.exit(2)
    import numpy as np
    import torch

    import rasterize_amd as rz
    from rasterize_amd import _abi

    if args.copy_calibration:
        n = 64 << 20  # 256 MiB of f32
        a = torch.ones(n, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": 4 * n}))
        return
    lib = _abi.lib()
    if args.workload == "scene":
        print(json.dumps(scene_measure(torch, rz, lib, _abi, "batched scene: 10,000 paths, 2048x2048",
                                       rz.workload_scene(3, 10_000, 2048.0), 2048, 2048,
                                       args.steps, 1, ["batched"] if args.batched_only else None)),
              flush=True)
        return
    if args.workload == "circles":
        print(json.dumps(scene_measure(torch, rz, lib, _abi, "1024 circles, 1024x1024",
                                       rz.workload_circles(1024, 1024.0), 1024, 1024, args.steps,
                                       1, ["batched"] if args.batched_only else None)), flush=True)
        return
    if args.workload == "atlas":
        print(json.dumps(atlas_measure(torch, rz, lib, _abi, args.steps)[0]), flush=True)
        return
    if args.workload == "diff":
        print(json.dumps(diff_measure(torch, rz, args.diff_steps)), flush=True)
        return
    if args.workload == "stroke":
        print(json.dumps(stroke_measure(torch, rz, lib, _abi, max(1, min(args.steps, 20)),
                                        not args.no_cpu)), flush=True)
        return

    W = H = 4096
    rule = rz.FillRule.NonZero if args.rule == "nonzero" else rz.FillRule.EvenOdd
    ws = rz.workload_random_walk(2, 200_000, 4096.0)
    path, tr = ws.paths[0], ws.transform
    n_seg = path.segments_count()
    n_pts = int(path.point_offsets[-1])
    ctx = rz.Context(0)
    dpath = C.c_void_p()
    ctx.check(lib.rz_path_upload(ctx.h, C.byref(path.abi()), C.byref(dpath)))
    out = torch.empty((H, W), dtype=torch.float32, device="cuda")

    def step():
        ctx.check(lib.rz_mask_device(ctx.h, dpath, tr.abi(), 0.05, rule, _abi.RZ_EDGE_MASK, W, H,
                                     C.c_void_p(out.data_ptr()), W))

    # warmup (sizes the context's buffers; a capacity miss is re-run)
    for _ in range(max(1, args.warmup)):
        settle(lib, _abi, ctx, step)
    nl, nr, ne = C.c_size_t(), C.c_size_t(), C.c_size_t()
    ctx.check(lib.rz_last_counts(ctx.h, C.byref(nl), C.byref(nr), C.byref(ne)))
    n_lines, n_recs = nl.value, nr.value

    stream = torch.cuda.ExternalStream(lib.rz_stream(ctx.h))

    def timed_steps():
        torch.cuda.synchronize()
        t_start = torch.cuda.Event(enable_timing=True)
        t_end = torch.cuda.Event(enable_timing=True)
        t_start.record(stream)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        t_end.record(stream)
        ctx.check(lib.rz_sync(ctx.h))
        torch.cuda.synchronize()
        return t_start.elapsed_time(t_end), time.perf_counter() - t0

    # each step launched kernel by kernel, then the same step replayed from a
    # HIP graph (captured once the step has run without growing a buffer);
    # the faster of the two is the headline
    ms_eager, wall_eager = timed_steps()
    set_graph(lib, _abi, ctx, True)
    for _ in range(3):
        step()
        ctx.check(lib.rz_sync(ctx.h))
    ms_graph, wall_graph = timed_steps()
    set_graph(lib, _abi, ctx, False)
    launch = "hip graph replay" if ms_graph < ms_eager else "kernel by kernel"
    ms_total, wall = (ms_graph, wall_graph) if ms_graph < ms_eager else (ms_eager, wall_eager)
    ms_step = ms_total / args.steps

    # per-kernel device time over a few profiled steps (events on the stream)
    kern = kernel_stats(lib, _abi, ctx, step, min(10, args.steps))
    ctx.check(lib.rz_sync(ctx.h))
    per_step = {k: v[1] for k, v in kern.items()}
    dom = max(per_step, key=per_step.get)
    dom_ms = per_step[dom] / kern[dom][0]

    # Algorithmic bytes of a mask step: the f32 mask written once and each
    # flattened line read once (16 B: f32 x0, y0, x1, y1), 4*W*H + 16*lines.
    # Intermediate buffers are not counted. `achieved` divides them by the
    # dominant kernel's average launch time (the roofline of the kernel that
    # bounds the step); `frac_step` by the whole step's time.
    alg = 4 * W * H + 16 * n_lines
    alg_note = "4*W*H (f32 mask written) + 16*lines (f32 line read), per step"
    achieved = alg / (dom_ms * 1e-3) / 1e9
    achieved_step = alg / (ms_step * 1e-3) / 1e9
    # the same bytes over each kernel's time (what each kernel alone would reach)
    per_kernel_gbs = {k: round(alg / (per_step[k] / kern[k][0] * 1e-3) / 1e9, 1) for k in per_step}

    traffic = traffic_step = None
    traffic_src = None
    if os.path.exists(TRAFFIC_FILE):
        tf = json.load(open(TRAFFIC_FILE))
        ks = tf.get("kernels", {})
        if dom in ks:
            traffic = ks[dom]["hbm_bytes"]
            traffic_src = os.path.relpath(TRAFFIC_FILE, ROOT)
        if ks:
            traffic_step = int(sum(int(v["hbm_bytes"]) * int(v.get("launches_per_step", 1))
                                   for v in ks.values()))

    copy_bw = copy_bandwidth_gbps(torch)

    mpix = W * H * args.steps / (ms_total * 1e-3) / 1e6
    msegs = n_seg * args.steps / (ms_total * 1e-3) / 1e6

    scenes = None
    ws_c3 = None
    if not args.no_scene:
        ws_c3 = rz.workload_scene(3, 10_000, 2048.0)
        scenes = [
            scene_measure(torch, rz, lib, _abi, "batched scene: 10,000 star/blob paths, 60% solid / "
                          "25% linear / 15% radial paint, 2048x2048 RGBA f32",
                          ws_c3, 2048, 2048, 10, 2),
            scene_measure(torch, rz, lib, _abi, "reference many-circles scene bench: 1024 circles, "
                          "1024x1024 RGBA f32", rz.workload_circles(1024, 1024.0), 1024, 1024, 20, 2),
        ]

    atlas = None
    glyph_set = None
    diff = None
    if not args.no_scene:
        atlas, glyph_set = atlas_measure(torch, rz, lib, _abi, 5)
        if args.diff_steps > 0:
            diff = diff_measure(torch, rz, args.diff_steps)

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _oracle as O
        p = O.opath(path)
        t = np.array(tr.m)
        reps = args.cpu_reps
        secs = O.lib().orc_time_sd_mask(C.byref(p), O._dp(t), 0.05, rule, W, H, reps, 1)
        nth = O.host_threads()
        reps_n = 2 * nth
        secs_n = O.lib().orc_time_sd_mask(C.byref(p), O._dp(t), 0.05, rule, W, H, reps_n, nth)
        cpu = {
            "value": round(W * H * reps / secs / 1e6, 3),
            "unit": "Mpix/s",
            "cores": 1,
            "kind": "port",
            "sample": f"{reps} full masks of the same 200k-segment path at 4096x4096 "
                      f"({secs:.1f} s), the oracle's C++ restatement of "
                      "SignedDifferenceRasterizer::mask, one host thread (the reference is "
                      "single-threaded)",
            "msegs_per_s": round(n_seg * reps / secs / 1e6, 4),
            "host_cores": os.cpu_count(),
            "threads": {
                "cores": nth,
                "value": round(W * H * reps_n / secs_n / 1e6, 3),
                "unit": "Mpix/s",
                "sample": f"{reps_n} masks spread over {nth} host threads ({secs_n:.1f} s); an "
                          "upper bound for the reference, which has no threading",
            },
        }
        if ws_c3 is not None and scenes:
            # the 10k-path scene on the host: one thread over the top 512 of the
            # 2048 rows (a view of the same scene), scaled to the frame; then the
            # whole frame on every host thread (bands of rows)
            fills = [(q, pt.abi(), r) for q, pt, r in zip(ws_c3.paths, ws_c3.paints, ws_c3.rules)]
            ident = rz.Transform.identity()
            t1 = time.perf_counter()
            O.scene_render_mt(fills, ident, (0.0, 0.0, 2048.0, 512.0), threads=1)
            one = (time.perf_counter() - t1) * 4.0
            t1 = time.perf_counter()
            O.scene_render_mt(fills, ident, (0.0, 0.0, 2048.0, 2048.0), threads=nth)
            many = time.perf_counter() - t1
            gpu_ms = scenes[0]["batched_ms_per_frame"]
            scenes[0]["cpu"] = {
                "ms_per_frame_1thread": round(one * 1e3, 1),
                "sample_1thread": "the oracle's Scene::render of rows 0-512 (a view), one thread, x4",
                "ms_per_frame_threads": round(many * 1e3, 1),
                "threads": nth,
                "batched_gpu_over_1thread": round(one * 1e3 / gpu_ms, 1),
                "batched_gpu_over_threads": round(many * 1e3 / gpu_ms, 1),
            }
        if glyph_set is not None:
            # the glyph atlas on the host: the first 4096 glyphs, one at a time
            t1 = time.perf_counter()
            px = 0
            for gp, sl in zip(glyph_set.paths[:4096], glyph_set.slots[:4096]):
                O.sd_mask(gp, rz.Transform.identity(), sl.width, sl.height, 0)
                px += sl.width * sl.height
            atlas["cpu_mpix_s_1thread"] = round(px / (time.perf_counter() - t1) / 1e6, 2)
            atlas["cpu_sample"] = "the oracle's SignedDifferenceRasterizer::mask of the first 4096 glyphs"

    rec = {
        "metric": "Mpix/s filled (Rasterizer::mask, 4096x4096, 200k-segment path)",
        "value": round(mpix, 2),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 flatten/bin, f32+i32 fixed-point coverage",
        "data": "synthetic (seeded Rnd LCG random walk, seed 2)",
        "config": {
            "workload": "large single path: 200k segments (40% line, 30% quad, 25% cubic, "
                        "5% arc), random affine, 4096x4096 f32 mask",
            "fill_rule": args.rule,
            "segments": n_seg,
            "lines": n_lines,
            "cell_records": n_recs,
            "width": W,
            "height": H,
            "flatness": 0.05,
        },
        "msegs_per_s": round(msegs, 3),
        "launch": launch,
        "ms_per_step_kernel_by_kernel": round(ms_eager / args.steps, 4),
        "ms_per_step_graph": round(ms_graph / args.steps, 4),
        "wall_s": round(wall, 3),
        "kernels_ms_per_step": {k: round(v, 4) for k, v in sorted(per_step.items())},
        "kernels_gb_s": per_kernel_gbs,
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": traffic,
            "traffic_step": traffic_step,
            "traffic_source": traffic_src,
            "algorithmic_bytes": alg,
            "algorithmic_bytes_def": alg_note,
            "kernel_ms": round(dom_ms, 4),
            "achieved_step": round(achieved_step, 1),
            "frac_step": round(achieved_step / HBM_PEAK_GBPS, 4),
            "copy_bw_measured": round(copy_bw, 1),
        },
        "scenes": scenes,
        "atlas": atlas,
        "differentiable": diff,
        "cpu_baseline": cpu,
    }
    print(json.dumps(rec), flush=True)
    lib.rz_path_free(ctx.h, dpath)


if __name__ == "__main__":
    main()
