#!/usr/bin/env python3
"""Benchmark: brickmap ray traversal on one MI355X.

Workload (one "step"): the 1024^3 terrain + caves world (seeded fBm heightmap,
3D value-noise caves, ~12% of bricks allocated), 1920x1080 primary rays from a
fixed camera above the terrain (their integer world origin at the camera's
sector corner; --anchor floor puts it at floor(eye)), then one cosine-distributed
diffuse bounce ray per primary hit (seeded PCG), i.e.
    vrt_cast(primary) -> vrt_bounce_generate -> vrt_cast(bounce, device count)
Rays, world and output buffers are resident in HBM before timing starts.
value = rays traced (primary + bounce) per second, in millions.

Also reported: the HBM roofline of the cast kernel (algorithmic bytes per
launch from the kernel's own counters / average launch time from HIP events
on the launch stream), the same step with the camera rays anchored at
floor(eye) (the reference's choice, CpuRenderer.cpp:447), the 256^3 terrain
(640x360 primary rays) and the 2048^3 sphere worlds (sparse and literal:
1920x1080 primary rays and 2M incoherent rays), the batched and
differentiable forms, the GLSL-path cast and the renderer frame, and the
reference CPU path (the oracle's SIMD port of CpuRenderer.cpp:172-224) timed
on this host's cores on the same ray buffers.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--tile 8x8] [--schedule grid|persistent]
                    [--primary-order|--bounce-order|--fused-order auto|forward|reverse]
                    [--anchor sector|floor] [--only all|batched|diff|c1|c3|extras]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import statistics
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per ray: ray in (6 x f32), hit out (t + 16-byte record; + 12 B of
# position for primaries), 8 B sector mask per lookup, 4 B base slot + 8 B occupancy mask per
# descent into a brick, 1 B voxel per hit
B_RAY_IN = 24
B_HIT_OUT = 20
B_POS_OUT = 12
B_ITER = 8
B_DESCENT = 12
B_HIT = 1


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes(c, with_pos: bool) -> float:
    rays, iters, desc, hits = (float(x) for x in c)
    return rays * (B_RAY_IN + B_HIT_OUT + (B_POS_OUT if with_pos else 0)) + iters * B_ITER + desc * B_DESCENT + hits * B_HIT


def _time_ms(fn, reps: int) -> float:
    import torch

    fn()
    torch.cuda.synchronize()
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        fn()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / reps


_TIMING_STREAM = None


def _graph_time_ms(launch, reps: int, replays: int = 5) -> float:
    """GPU time per call of launch(stream_handle): `reps` calls captured in one graph, the
    graph replayed `replays` times between HIP events, so host launch overhead does not enter
    small launches.  launch must issue no memset and no synchronisation."""
    import torch

    global _TIMING_STREAM
    if _TIMING_STREAM is None:
        _TIMING_STREAM = torch.cuda.Stream()  # one stream: per-stream library state is allocated once
    side = _TIMING_STREAM
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(2):
            launch(side.cuda_stream)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    # captured on the warm-up stream: per-stream state sized by the warm-up (the compact
    # schedule's survivor buffer) is the captured launches' state
    with torch.cuda.graph(g, stream=side):
        h = side.cuda_stream
        for _ in range(reps):
            launch(h)
    g.replay()
    torch.cuda.synchronize()
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(replays):
        g.replay()
    e.record()
    torch.cuda.synchronize()
    del g
    return s.elapsed_time(e) / (reps * replays)


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def roofline_of(bytes_per_launch: float, ms: float) -> dict:
    achieved = bytes_per_launch / (ms / 1e3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "algorithmic_bytes_per_launch": int(bytes_per_launch)}


# ---------------------------------------------------------------- plain casts of one ray set


def cast_section(world, ow, rb, name: str, max_iters: int, cpu_threads: int | None, cpu_runs: int,
                 reps: int = 20) -> dict:
    """One ray set on one world: the cast's rate (GPU time from a captured graph of `reps`
    launches), its algorithmic-byte roofline (the counting variant's totals), and the CPU path
    on the same rays with bit-exact parity."""
    import numpy as np
    import torch

    import voxelrt_amd as V
    from voxelrt_amd import _lib as L

    lib = V.lib()
    hits = V.HitBatch.empty(rb.n)
    r_c, h_c = rb.to_c(), hits.to_c()
    ctr = V.new_counters()
    world.cast(rb, max_iters, out=hits, counters=ctr)
    torch.cuda.synchronize()
    c = ctr.cpu().numpy()

    def timed(sched, cap):
        opts = L.CastOpts(sched, 0, 0, 0, 0, 0, L.VRT_ORDER_FORWARD, cap)

        def launch(stream):
            L.check(lib.vrt_cast_ex(world.handle, C.byref(r_c), max_iters, C.byref(h_c), C.byref(opts), stream))

        return _graph_time_ms(launch, reps)

    # the grid, and the compact schedule (a capped pass, then the survivors repacked) at four
    # first-pass caps: the fastest is the section's launch
    trials = {(L.VRT_SCHED_GRID, 0): timed(L.VRT_SCHED_GRID, 0)}
    for cap in (24, 32, 48, 64):
        trials[(L.VRT_SCHED_COMPACT, cap)] = timed(L.VRT_SCHED_COMPACT, cap)
    (sched, cap), ms = min(trials.items(), key=lambda kv: kv[1])
    world.cast(rb, max_iters, out=hits, schedule=sched, pass_cap=cap)  # the measured launch's hits
    torch.cuda.synchronize()
    res = {
        "rays": rb.n,
        "value": round(rb.n / ms / 1e3, 1),
        "unit": "Mrays/s",
        "ms_per_launch": round(ms, 5),
        "schedule": "grid" if sched == L.VRT_SCHED_GRID else f"compact, first pass {cap} steps",
        "grid_ms_per_launch": round(trials[(L.VRT_SCHED_GRID, 0)], 5),
        "iters_per_ray": round(float(c[1]) / rb.n, 3),
        "descents_per_ray": round(float(c[2]) / rb.n, 3),
        "hit_fraction": round(float(c[3]) / rb.n, 4),
        "roofline": roofline_of(algorithmic_bytes(c, False), ms),
    }
    if cpu_threads is not None:
        o, d = rb.o.cpu().numpy(), rb.d.cpu().numpy()
        ow.cast_simd(o, d, rb.world_origin, max_iters, threads=cpu_threads)  # warm-up
        times, ref = [], None
        for _ in range(cpu_runs):
            t0 = time.perf_counter()
            ref = ow.cast_simd(o, d, rb.world_origin, max_iters, threads=cpu_threads)
            times.append(time.perf_counter() - t0)
        med = statistics.median(times)
        ok = bool(np.array_equal(hits.rec.cpu().numpy(), ref["rec"]) and
                  np.array_equal(hits.t.cpu().numpy().view(np.uint32), ref["t"].view(np.uint32)))
        res["cpu_value"] = round(rb.n / med / 1e6, 2)
        res["parity_with_cpu"] = ok
    log(f"  {name}: {rb.n} rays, {res['value']} Mrays/s ({ms:.4f} ms, {res['schedule']}; grid "
        f"{res['grid_ms_per_launch']:.4f} ms), {res['iters_per_ray']} lookups/ray, "
        f"roofline {res['roofline']['achieved']} GB/s" +
        (f", CPU {res['cpu_value']} Mrays/s, parity {res['parity_with_cpu']}" if cpu_threads is not None else ""))
    return res


def bench_c1(max_iters: int, cpu_threads, cpu_runs: int) -> dict:
    """256^3 fBm terrain, 640x360 primary rays (the reference CPU path's configuration)."""
    import voxelrt_amd as V
    from voxelrt_amd import workloads as W

    log("C1: 256^3 terrain, 640x360 primary rays")
    scene = W.C1_WORLD.build()
    world = V.World.from_scene(scene)
    ow = None
    if cpu_threads is not None:
        import oracle as O

        ow = O.OracleWorld.from_stream(scene.stream())
    res = {"workload": "256^3 fBm terrain (4 octaves, freq 1/64, gain 0.5), 640x360 primary rays",
           "primary": cast_section(world, ow, W.c1_camera().rays_device(), "C1 primary", max_iters, cpu_threads,
                                   cpu_runs, reps=50)}
    del world, ow
    return res


def bench_c3(max_iters: int, cpu_threads, cpu_runs: int, literal: bool) -> dict:
    """2048^3 sphere worlds: 587 spheres (~2% fill) and the literal 20k spheres; 1920x1080 primary
    rays and 2M incoherent rays (origins uniform in the box, directions uniform on the sphere)."""
    import gc

    import voxelrt_amd as V
    from voxelrt_amd import workloads as W

    out = {}
    for sparse in (True, False) if literal else (True,):
        key = "sparse_587" if sparse else "literal_20000"
        log(f"C3 {key}: building the 2048^3 sphere world")
        t0 = time.time()
        scene = W.c3_world(sparse)
        world = V.World.from_scene(scene)
        ow = None
        if cpu_threads is not None:
            import oracle as O

            ow = O.OracleWorld.from_stream(scene.stream())
        log(f"  built in {time.time() - t0:.1f} s, {world.info().num_bricks} bricks")
        o_i, d_i = V.incoherent_rays_host(2_000_000, 2048, 2048, seed=33)
        res = {"bricks": int(world.info().num_bricks)}
        res["primary"] = cast_section(world, ow, W.c3_camera().rays_device(), f"C3 {key} primary", max_iters,
                                      cpu_threads, cpu_runs)
        res["incoherent"] = cast_section(world, ow, V.RayBatch.from_numpy(o_i, d_i), f"C3 {key} incoherent",
                                         max_iters, cpu_threads, cpu_runs)
        out[key] = res
        del world, ow, scene
        gc.collect()
    return out


# ---------------------------------------------------------------- batched


def bench_batched(world, max_iters: int, dev) -> dict:
    """64 cameras of uneven square sizes (128..1024 px) on the same world plus each camera's
    bounce set truncated to a seeded fraction U(0.25, 1): all 128 segments in one hipGraph
    launch with the autotuned tuning, against plain vrt_cast launches over the same rays."""
    import numpy as np
    import torch

    import voxelrt_amd as V
    from voxelrt_amd import workloads as W

    log("batched: 64 cameras + bounce sets")
    cams = W.c4_cameras(64, seed=4, tile=(1, 1))
    rng = np.random.default_rng(44)
    prim = [c.rays_device() for c in cams]
    parts, segs = [], []
    off = 0
    for c, rb in zip(cams, prim):
        h = world.cast(rb, max_iters, with_pos=True)
        br, _, cnt = V.bounce(rb, h, c.c.width, W.BOUNCE_SEED)
        torch.cuda.synchronize()
        nb = int(int(cnt.item()) * rng.uniform(0.25, 1.0))
        parts.append((br.o[:, :nb].clone(), br.d[:, :nb].clone()))
        segs.append((off, nb, rb.world_origin, None))
        off += nb
        del h, br
    rays = V.RayBatch(torch.cat([p[0] for p in parts], 1).contiguous(),
                      torch.cat([p[1] for p in parts], 1).contiguous(), (0, 0, 0))
    del parts
    batch = V.Batch(cams, segs, rays, max_iters)
    tuning, tune_ms = V.autotune(world, batch, None)
    plan = V.BatchPlan(world, batch, tuning)
    n_prim = sum(c.n for c in cams)
    n_total = batch.total
    ms_graph = _time_ms(plan.graph_launch, 10)
    ms_launch = _time_ms(plan.launch, 10)
    # plain per-ray kernel on the same rays: 128 launches over materialised buffers, camera rays
    # materialised in the tuned tile order (the plain kernel's best case)
    tiled = [V.CameraSpec((0, 0, 0), 0, 0, 16, 16) for _ in cams]
    for t_cam, c in zip(tiled, cams):
        t_cam.c = type(c.c).from_buffer_copy(c.c)
        t_cam.c.tile_w, t_cam.c.tile_h = tuning[0], tuning[1]
    prim_tiled = [c.rays_device() for c in tiled]
    del prim
    bounce_views = [V.RayBatch(rays.o[:, ro:ro + n].contiguous(), rays.d[:, ro:ro + n].contiguous(), Wo)
                    for (ro, n, Wo, _) in segs]
    all_rays = prim_tiled + bounce_views
    outs = [V.HitBatch.empty(r.n) for r in all_rays]

    def plain():
        for r, o in zip(all_rays, outs):
            world.cast(r, max_iters, out=o)

    ms_plain = _time_ms(plain, 5)
    # parity of the batch against the plain launches, ray by ray: bounce segments in order; a
    # camera segment holds pixel (px, py) at py * width + px, the plain buffer ray i of the tuned
    # tile order is pixel (tx * tw + i % tw, ty * th + (i % (tw th)) / tw) of tile i / (tw th)
    plan.launch()
    torch.cuda.synchronize()
    ok = all(torch.equal(batch.segment_hits(len(cams) + k)[1], outs[len(cams) + k].rec) for k in range(len(segs)))
    tw, th = tuning[0], tuning[1]
    for k, c in enumerate(cams):
        w, h = c.c.width, c.c.height
        i = torch.arange(w * h, device=dev, dtype=torch.int64)
        tile, within = i // (tw * th), i % (tw * th)
        px = (tile % (w // tw)) * tw + within % tw
        py = (tile // (w // tw)) * th + within // tw
        t_b, rec_b, _ = batch.segment_hits(k)
        ok = ok and torch.equal(rec_b[py * w + px], outs[k].rec) and torch.equal(t_b[py * w + px], outs[k].t)
    res = {
        "value": round(n_total / ms_graph / 1e3, 1),
        "unit": "Mrays/s",
        "rays": n_total,
        "primary_rays": n_prim,
        "segments": len(cams) + len(segs),
        "ms_per_launch": round(ms_graph, 4),
        "ms_per_launch_no_graph": round(ms_launch, 4),
        "tuning": {"tile": f"{tuning[0]}x{tuning[1]}", "rays_per_lane": tuning[2], "chunk": tuning[3]},
        "plain_value": round(n_total / ms_plain / 1e3, 1),
        "plain_launches": len(outs),
        "parity_with_plain": bool(ok),
    }
    log(f"  batched {res['value']} Mrays/s ({n_total} rays, tuning {tuning}), plain {res['plain_value']} Mrays/s, "
        f"parity {ok}")
    plan.close()
    return res


# ---------------------------------------------------------------- shader method and renderer


def bench_extras(world, max_iters: int) -> dict:
    """The GLSL-path cast (vrt_cast_shader, all four modes) on the C2 camera rays and the GPU
    renderer's frame (vrt_render_frame, 0/1/2 bounces) at 1920x1080 on the same world: GPU time
    per launch from captured graphs (no host transfers, no allocation in the launches)."""
    import voxelrt_amd as V
    from voxelrt_amd import _lib as L
    from voxelrt_amd import workloads as W

    log("shader method and renderer frame on the 1024^3 terrain + caves world")
    lib = V.lib()
    cam = W.c2_camera()
    rb = cam.rays_device()
    hits = V.HitBatch.empty(rb.n)
    r_c, h_c = rb.to_c(), hits.to_c()
    shader = {}
    for method in ("full", "coarse"):
        for lut in (False, True):
            o = L.ShaderOpts({"full": L.VRT_SHADER_FULL, "coarse": L.VRT_SHADER_COARSE}[method], int(lut), 0, 0)

            def launch(stream, o=o):
                L.check(lib.vrt_cast_shader(world.handle, C.byref(r_c), C.byref(o), C.byref(h_c), stream))

            ms = _graph_time_ms(launch, 20)
            key = f"{'rayCast' if method == 'full' else 'rayCastCoarse'}{', ray-cell LUT' if lut else ''}"
            shader[key] = {"ms_per_launch": round(ms, 5), "value": round(rb.n / ms / 1e3, 1), "unit": "Mrays/s"}
            log(f"  {key}: {shader[key]['value']} Mrays/s ({ms:.4f} ms, {rb.n} rays)")
    frames = {}
    g = V.GBufferImages(cam.c.width, cam.c.height)
    g_c = g.to_c()
    for bounces in (0, 1, 2):
        o = L.RenderOpts(bounces, 0, 0, 0, (C.c_float * 3)(0.6, 0.75, 1.0), 0.0, 0.0)

        def launch(stream, o=o):
            L.check(lib.vrt_render_frame(world.handle, C.byref(cam.c), C.byref(o), C.byref(g_c), stream))

        ms = _graph_time_ms(launch, 10)
        frames[f"{bounces} bounces"] = {"ms_per_frame": round(ms, 4),
                                        "pixels_per_s": round(cam.c.width * cam.c.height / ms * 1e3)}
        log(f"  renderer frame, {bounces} bounces: {ms:.4f} ms")
    return {"camera": f"{cam.c.width}x{cam.c.height}, world origin {list(rb.world_origin)}",
            "shader_method": shader, "render_frame": frames}


# ---------------------------------------------------------------- differentiable


def bench_differentiable(dev) -> dict:
    """512^3 terrain, per-voxel RGB albedo U[0,1], 16 views of 512x512, L2 loss against images
    rendered from albedo + N(0, 0.05): forward + backward through the libtorch autograd Function."""
    import torch

    import voxelrt_amd as V
    from voxelrt_amd import workloads as W

    log("differentiable: 512^3 terrain, 16 views of 512x512")
    scene = W.C5_WORLD.build()
    world = V.World.from_scene(scene)
    cams = W.c5_cameras(16, 512)
    nvox = world.info().num_bricks * 512
    g = torch.Generator(device=dev).manual_seed(5)
    albedo = torch.rand((nvox, 3), generator=g, device=dev, dtype=torch.float32)
    target, _, _ = V.render_saved(world, (albedo + 0.05 * torch.randn(albedo.shape, generator=g, device=dev)),
                                  cams, W.MAX_ITERS)
    npix = sum(c.n for c in cams)
    a = albedo.clone().requires_grad_(True)

    def step(mode):
        # the L2 loss as torch's fused form: one elementwise pass and a reduction forward, one
        # elementwise pass backward (written out as ((img - target) ** 2).sum() it takes five
        # passes over the 50 MB image, ~100 us of the step)
        a.grad = None
        img = V.render(world, a, cams, W.MAX_ITERS, grad=mode)
        loss = torch.nn.functional.mse_loss(img, target, reduction="sum")
        loss.backward()

    # the gradient's form: dense written whole (the default), dense zero-filled and scatter-added
    # with atomics, sparse COO with one entry per pixel
    ms_mode = {m: _time_ms(lambda: step(m), 10) for m in ("bricks", "atomic", "sparse")}
    ms_step = ms_mode["bricks"]
    ms_fwd = _time_ms(lambda: V.render_saved(world, albedo, cams, W.MAX_ITERS), 10)
    img, sv, sw = V.render_saved(world, albedo, cams, W.MAX_ITERS)
    gi = 2.0 * (img - target)
    ms_bwd = {m: _time_ms(lambda: V.render_backward(world, gi, sv, sw, nvox, mode=m), 10)
              for m in ("bricks", "atomic", "sparse")}
    prims = [c.rays_device() for c in cams]
    outs = [V.HitBatch.empty(r.n) for r in prims]

    def plain():
        for r, o in zip(prims, outs):
            world.cast(r, W.MAX_ITERS, out=o)

    ms_plain = _time_ms(plain, 10)
    res = {
        "value": round(npix / ms_step / 1e3, 1),
        "unit": "Mrays/s",
        "what": "forward + backward (autograd Function, the L2 loss as mse_loss(reduction=sum) and its "
                "gradient included), dense gradient written whole (vrt_render_bwd_bricks)",
        "rays": npix,
        "ms_per_step": round(ms_step, 4),
        "atomic_value": round(npix / ms_mode["atomic"] / 1e3, 1),
        "sparse_value": round(npix / ms_mode["sparse"] / 1e3, 1),
        "forward_only_value": round(npix / ms_fwd / 1e3, 1),
        "backward_ms": {m: round(t, 4) for m, t in ms_bwd.items()},
        "plain_value": round(npix / ms_plain / 1e3, 1),
        "voxels": nvox,
    }
    log(f"  fwd+bwd {res['value']} Mrays/s (atomic {res['atomic_value']}, sparse {res['sparse_value']}), "
        f"forward {res['forward_only_value']}, plain {res['plain_value']} Mrays/s, backward {res['backward_ms']} ms")
    del world
    return res


# ---------------------------------------------------------------- the metric: the C2 step


def c2_step(world, cam, args, K: int, order_arg: dict, event_every: int) -> dict:
    """The timed step for one camera: primary cast, bounce generation, bounce cast (device count).
    Grid block orders given as "auto" are chosen by timing whole steps with each pair after
    warm-up.  Returns rates, launch times (HIP events on the launch stream), counters and buffers."""
    import torch

    import voxelrt_amd as V
    from voxelrt_amd import _lib as L
    from voxelrt_amd import workloads as W

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    sched = 1 if args.schedule == "grid" else 2
    prim = cam.rays_device()
    n = prim.n
    hits_p = V.HitBatch.empty(n, with_pos=True)
    hits_b = V.HitBatch.empty(n, with_pos=False)
    bounce_rays = V.RayBatch.empty(n, prim.world_origin)
    src = torch.empty(n, dtype=torch.int32, device=dev)
    count = torch.zeros(1, dtype=torch.int32, device=dev)
    scratch = torch.empty(int(V.lib().vrt_bounce_scratch_bytes(n)) // 4 + 1, dtype=torch.int32, device=dev)
    lib = V.lib()
    r_prim, h_prim = prim.to_c(), hits_p.to_c()
    r_b, h_b = bounce_rays.to_c(), hits_b.to_c()
    bufs_b = bounce_rays.buffers()
    sh = stream.cuda_stream
    order = {k: L.VRT_ORDER_REVERSE if v == "reverse" else L.VRT_ORDER_FORWARD for k, v in order_arg.items()}

    def opts(counters=None, cnt=None, which="primary"):
        return L.CastOpts(sched, args.refill, args.chunk, 0, counters.data_ptr() if counters is not None else 0,
                          cnt.data_ptr() if cnt is not None else 0, order[which], 0)

    cur = {"primary": opts(), "bounce": opts(cnt=count, which="bounce")}  # the timed launches' options

    def cast_primary(o=None):
        o = cur["primary"] if o is None else o
        L.check(lib.vrt_cast_ex(world.handle, C.byref(r_prim), W.MAX_ITERS, C.byref(h_prim), C.byref(o), sh))

    # bounce rays in primary order, or grouped by direction octant (vrt_bounce_generate_ex): the
    # same rays, the same hit per ray; the order changes the bounce launch's wave coherence
    gen_orders = {"primary": L.VRT_BOUNCE_PRIMARY_ORDER, "direction": L.VRT_BOUNCE_DIRECTION_ORDER}
    cur["gen"] = gen_orders.get(args.bounce_generation, L.VRT_BOUNCE_DIRECTION_ORDER)

    def gen_bounce():
        L.check(lib.vrt_bounce_generate_ex(C.byref(r_prim), C.byref(h_prim), cam.c.width, W.BOUNCE_SEED, cur["gen"],
                                           C.byref(bufs_b), src.data_ptr(), count.data_ptr(), scratch.data_ptr(), sh))

    def cast_bounce(o=None):
        o = cur["bounce"] if o is None else o
        L.check(lib.vrt_cast_ex(world.handle, C.byref(r_b), W.MAX_ITERS, C.byref(h_b), C.byref(o), sh))

    for _ in range(args.warmup):
        cast_primary()
        gen_bounce()
        cast_bounce()
    torch.cuda.synchronize()
    nb = int(count.item())
    trial = {}
    auto_orders = "auto" in (order_arg["primary"], order_arg["bounce"]) and sched == L.VRT_SCHED_GRID
    if auto_orders or args.bounce_generation == "auto":
        # the grid's block order decides which waves form each launch's tail and the generation
        # order which rays share a wave; results per ray are the same: time whole steps with each
        # combination and keep the fastest
        both = (L.VRT_ORDER_FORWARD, L.VRT_ORDER_REVERSE)
        gens = tuple(gen_orders.values()) if args.bounce_generation == "auto" else (cur["gen"],)
        for go in gens:
            for po in both if auto_orders and order_arg["primary"] == "auto" else (order["primary"],):
                for bo in both if auto_orders and order_arg["bounce"] == "auto" else (order["bounce"],):
                    order["primary"], order["bounce"], cur["gen"] = po, bo, go
                    cur["primary"], cur["bounce"] = opts(), opts(cnt=count, which="bounce")
                    trial[(go, po, bo)] = _time_ms(lambda: (cast_primary(), gen_bounce(), cast_bounce()), 20)
        cur["gen"], order["primary"], order["bounce"] = min(trial, key=trial.get)
        cur["primary"], cur["bounce"] = opts(), opts(cnt=count, which="bounce")
    names = ("forward", "reverse")
    gen_names = {v: k for k, v in gen_orders.items()}

    # counters (same traversal, counting kernel variant; outside the timed region)
    cp = torch.zeros(4, dtype=torch.int64, device=dev)
    cb = torch.zeros(4, dtype=torch.int64, device=dev)
    cast_primary(opts(counters=cp))
    gen_bounce()
    cast_bounce(opts(counters=cb, cnt=count, which="bounce"))
    torch.cuda.synchronize()
    cp_h, cb_h = cp.cpu().numpy(), cb.cpu().numpy()
    assert int(cp_h[0]) == n and int(cb_h[0]) == nb, (cp_h, cb_h, n, nb)

    # timed region: K steps; every event_every-th step (and the first) has HIP events around
    # each of its launches on the launch stream (an event between two launches costs the
    # stream ~3.5 us, so events on every step would add 5% to the step)
    every = max(1, event_every)
    sampled = [k for k in range(K) if k % every == 0]
    ev = {k: [torch.cuda.Event(enable_timing=True) for _ in range(4)] for k in sampled}
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for k in range(K):
        e = ev.get(k)
        if e is None:
            cast_primary()
            gen_bounce()
            cast_bounce()
            continue
        e[0].record(stream)
        cast_primary()
        e[1].record(stream)
        gen_bounce()
        e[2].record(stream)
        cast_bounce()
        e[3].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    res = {
        "prim": prim, "n": n, "nb": nb, "hits_p": hits_p, "hits_b": hits_b, "bounce_rays": bounce_rays, "src": src,
        "wall": wall, "K": K, "cp": cp_h, "cb": cb_h,
        "step_ms": [ev[k][0].elapsed_time(ev[k][3]) for k in sampled],
        "prim_ms": [ev[k][0].elapsed_time(ev[k][1]) for k in sampled],
        "gen_ms": [ev[k][1].elapsed_time(ev[k][2]) for k in sampled],
        "bounce_ms": [ev[k][2].elapsed_time(ev[k][3]) for k in sampled],
        "order": {k: names[v] for k, v in order.items() if k != "fused"},
        "bounce_generation": gen_names[cur["gen"]],
        "order_trials": {f"{gen_names[g]}: {names[p]}/{names[b]}": round(t, 4) for (g, p, b), t in trial.items()},
    }
    res["value"] = (n + nb) * K / wall / 1e6
    res["ms_per_step"] = wall * 1e3 / K
    return res


def read_profile_json(name: str, want: dict):
    """A committed profile summary whose workload fields equal `want`, else None."""
    f = ROOT / "profiles" / name
    if not f.exists():
        return None
    try:
        data = json.loads(f.read_text())
    except (ValueError, OSError):
        return None
    return data if all(data.get(k) == v for k, v in want.items()) else None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--tile", default="8x8")
    ap.add_argument("--schedule", default="grid", choices=["grid", "persistent"])
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--refill", type=int, default=0)
    for launch in ("primary", "bounce", "fused"):
        ap.add_argument(f"--{launch}-order", default="auto", choices=["auto", "forward", "reverse"],
                        help=f"grid block order of the {launch} launch; auto: timed after warm-up")
    ap.add_argument("--bounce-generation", default="auto", choices=["auto", "primary", "direction"],
                    help="bounce rays in primary order or grouped by direction octant (vrt_bounce_generate_ex); "
                         "auto: timed after warm-up with the block orders")
    ap.add_argument("--anchor", default="sector", choices=["floor", "sector"],
                    help="integer world origin of the camera rays: the camera's sector corner (default; the "
                         "aligned-origin step) or floor(eye)")
    ap.add_argument("--event-every", type=int, default=8,
                    help="HIP events around the launches of every n-th timed step (the roofline's launch times)")
    ap.add_argument("--only", default="all", choices=["all", "batched", "diff", "c1", "c3", "extras"],
                    help="run one section alone and print its result (profiling runs); the default runs the "
                         "metric with every section")
    ap.add_argument("--cpu-runs", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the shader-method and renderer section")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-diff", action="store_true")
    ap.add_argument("--no-configs", action="store_true", help="skip the 256^3 and 2048^3 sections")
    ap.add_argument("--no-c3-literal", action="store_true",
                    help="skip the literal 20k-sphere world (its host-side build takes ~30 s)")
    ap.add_argument("--no-floor", action="store_true", help="skip the C2 step with floor(eye)-anchored rays")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="also run one device copy of this many MiB (a known-size kernel for scaling PMC byte counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported here (a multi-GPU cast "
              "would replicate the world on every GPU and split the rays between them)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import voxelrt_amd as V
    from voxelrt_amd import _lib as L
    from voxelrt_amd import workloads as W

    sys.path.insert(0, str(ROOT / "oracle"))
    cpu_threads = None
    if not args.no_cpu:
        import oracle as O  # the CPU baseline (oracle's SIMD port of the reference CPU cast)

        cpu_threads = O.num_cores()

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.only != "all":
        if args.only == "diff":
            res = bench_differentiable(dev)
        elif args.only == "c1":
            res = bench_c1(W.MAX_ITERS, cpu_threads, args.cpu_runs)
        elif args.only == "c3":
            res = bench_c3(W.MAX_ITERS, cpu_threads, args.cpu_runs, not args.no_c3_literal)
        else:
            log("building the 1024^3 terrain + caves world")
            world = V.World.from_scene(W.C2_WORLD.build())
            if args.only == "batched":
                res = bench_batched(world, W.MAX_ITERS, dev)
            else:
                res = bench_extras(world, W.MAX_ITERS)
        print(json.dumps({"section": args.only, **res}), flush=True)
        return 0

    tile = tuple(int(x) for x in args.tile.split("x"))
    log("building the 1024^3 terrain + caves world")
    t0 = time.time()
    scene = W.C2_WORLD.build()
    log(f"  generated in {time.time() - t0:.1f} s, brick fraction {W.brick_fraction(scene):.4f}")
    world = V.World.from_scene(scene)
    info = world.info()
    cam = W.c2_camera(tile=tile, anchor=args.anchor)

    if args.copy_mb:
        a = torch.empty(args.copy_mb << 18, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        a.fill_(1.0)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    order_arg = {"primary": args.primary_order, "bounce": args.bounce_order}
    st = c2_step(world, cam, args, args.steps, order_arg, args.event_every)
    if st["order_trials"]:
        log("bounce generation: block order (primary/bounce): " +
            ", ".join(f"{k} {v:.4f} ms" for k, v in st["order_trials"].items()))
    n, nb, K = st["n"], st["nb"], st["K"]
    prim, hits_p, hits_b, bounce_rays, src = st["prim"], st["hits_p"], st["hits_b"], st["bounce_rays"], st["src"]
    cp_h, cb_h = st["cp"], st["cb"]
    prim_ms, bounce_ms, gen_ms, step_ms = st["prim_ms"], st["bounce_ms"], st["gen_ms"], st["step_ms"]
    value, ms_per_step = st["value"], st["ms_per_step"]

    bytes_p = algorithmic_bytes(cp_h, True)
    bytes_b = algorithmic_bytes(cb_h, False)
    avg_launch_s = (statistics.mean(prim_ms) + statistics.mean(bounce_ms)) / 2 / 1e3
    achieved = (bytes_p + bytes_b) / 2 / avg_launch_s / 1e9
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None,
        "launch_samples": len(step_ms),
    }
    # achievable HBM bandwidth on this device, measured in the same run: a 1 GiB device-to-device
    # copy (reads and writes 1 GiB each)
    ca = torch.empty(1 << 28, dtype=torch.float32, device=dev)
    cb = torch.empty_like(ca)
    ca.fill_(1.0)
    copy_ms = _time_ms(lambda: cb.copy_(ca), 10)
    achievable = 2 * ca.numel() * 4 / (copy_ms / 1e3) / 1e9
    del ca, cb
    roofline["achievable"] = round(achievable, 1)
    roofline["frac_of_achievable"] = round(achieved / achievable, 4)
    want = {"tile": args.tile, "schedule": args.schedule, "order": st["order"], "anchor": args.anchor,
            "bounce_generation": st["bounce_generation"]}
    tr = read_profile_json("traffic_cast_c2.json", want)
    if tr:
        roofline["traffic"] = tr.get("hbm_bytes_per_launch")
    vb = read_profile_json("valu_cast_c2.json", want)
    if vb:
        # the launches' binding ceiling: VALU issue (SQ_ACTIVE_INST_VALU x 4 / (SIMDs x busy cycles))
        roofline["valu_busy"] = {"primary": vb["primary"], "bounce": vb["bounce"], "mean": vb["mean"],
                                 "source": vb["source"]}
    log(f"primary {n} rays: {statistics.mean(prim_ms):.3f} ms ({n / statistics.mean(prim_ms) / 1e3:.0f} Mrays/s), "
        f"iters/ray {cp_h[1] / n:.2f}, descents/ray {cp_h[2] / n:.2f}, hit {cp_h[3] / n:.3f}")
    log(f"bounce  {nb} rays: {statistics.mean(bounce_ms):.3f} ms ({nb / statistics.mean(bounce_ms) / 1e3:.0f} Mrays/s), "
        f"iters/ray {cb_h[1] / max(nb, 1):.2f}, descents/ray {cb_h[2] / max(nb, 1):.2f}, hit {cb_h[3] / max(nb, 1):.3f}")
    log(f"bounce generation {statistics.mean(gen_ms):.3f} ms; step {statistics.mean(step_ms):.3f} ms (events, "
        f"{len(step_ms)} sampled steps), {ms_per_step:.3f} ms (wall); cast roofline {achieved:.0f} GB/s = "
        f"{achieved / HBM_PEAK_GBS:.1%} of 8 TB/s, {achieved / achievable:.1%} of the {achievable:.0f} GB/s a 1 GiB "
        f"device copy reaches")
    launches = {
        "primary_ms": round(statistics.mean(prim_ms), 5),
        "bounce_generation_ms": round(statistics.mean(gen_ms), 5),
        "bounce_ms": round(statistics.mean(bounce_ms), 5),
        "primary_value": round(n / statistics.mean(prim_ms) / 1e3, 1),
        "bounce_value": round(nb / statistics.mean(bounce_ms) / 1e3, 1),
    }

    # the same step with the camera rays anchored at floor(eye), the reference's convention
    floor = None
    if not args.no_floor and args.anchor != "floor":
        fst = c2_step(world, W.c2_camera(tile=tile, anchor="floor"), args, args.steps, order_arg, args.event_every)
        floor = {
            "value": round(fst["value"], 1),
            "unit": "Mrays/s",
            "ms_per_step": round(fst["ms_per_step"], 4),
            "world_origin": list(fst["prim"].world_origin),
            "primary_ms": round(statistics.mean(fst["prim_ms"]), 5),
            "bounce_ms": round(statistics.mean(fst["bounce_ms"]), 5),
            "order": fst["order"],
            "bounce_generation": fst["bounce_generation"],
            "bounce_rays": fst["nb"],
        }
        log(f"floor(eye) anchor: {floor['value']} Mrays/s ({floor['ms_per_step']:.4f} ms per step)")
        del fst

    # the same step fused into one kernel (vrt_cast_frame: a lane traces its pixel's primary ray,
    # then that hit's bounce ray); reported beside the plain per-ray pipeline, checked against it
    fp = V.HitBatch.empty(n, with_pos=False)
    fb = V.HitBatch.empty(n, with_pos=False)
    V.cast_frame(world, cam, W.MAX_ITERS, W.BOUNCE_SEED, schedule=1, out=(fp, fb))
    torch.cuda.synchronize()
    src_h = src[:nb].long()
    frame_ok = bool(torch.equal(fp.rec, hits_p.rec) and torch.equal(fb.rec[src_h], hits_b.rec[:nb]))
    fused_order = L.VRT_ORDER_REVERSE if args.fused_order == "reverse" else L.VRT_ORDER_FORWARD
    if args.fused_order == "auto":
        ft = {o_try: _time_ms(lambda: V.cast_frame(world, cam, W.MAX_ITERS, W.BOUNCE_SEED, schedule=1,
                                                   out=(fp, fb), order=o_try), 20)
              for o_try in (L.VRT_ORDER_FORWARD, L.VRT_ORDER_REVERSE)}
        fused_order = min(ft, key=ft.get)
    frame_ms = _time_ms(lambda: V.cast_frame(world, cam, W.MAX_ITERS, W.BOUNCE_SEED, schedule=1, out=(fp, fb),
                                             order=fused_order), K)
    fused = {
        "value": round((n + nb) / frame_ms / 1e3, 1),
        "unit": "Mrays/s",
        "ms_per_step": round(frame_ms, 4),
        "kernel": "vrt_cast_frame (grid): primary + bounce per lane",
        "order": "reverse" if fused_order else "forward",
        "parity_with_plain": frame_ok,
    }
    log(f"fused frame {fused['value']} Mrays/s ({frame_ms:.3f} ms), parity with the plain pipeline {frame_ok}")

    cpu = None
    if cpu_threads is not None:
        log("CPU baseline: building the reference's dense world layout")
        ow = O.OracleWorld.from_stream(scene.stream())
        po = prim.o.cpu().numpy()
        pd = prim.d.cpu().numpy()
        # the CPU path traces the bounce rays in primary order, as RenderRow does per pixel
        # (CpuRenderer.cpp:389-392); the GPU's hits are compared in that order
        perm = np.argsort(src[:nb].cpu().numpy().astype(np.int64), kind="stable")
        bo = np.ascontiguousarray(bounce_rays.o[:, :nb].cpu().numpy()[:, perm])
        bd = np.ascontiguousarray(bounce_rays.d[:, :nb].cpu().numpy()[:, perm])
        ow.cast_simd(po, pd, prim.world_origin, W.MAX_ITERS, threads=cpu_threads)  # warm-up
        times = []
        for _ in range(args.cpu_runs):
            t0 = time.perf_counter()
            rp = ow.cast_simd(po, pd, prim.world_origin, W.MAX_ITERS, threads=cpu_threads)
            rb = ow.cast_simd(bo, bd, prim.world_origin, W.MAX_ITERS, threads=cpu_threads)
            times.append(time.perf_counter() - t0)
        med = statistics.median(times)
        # same rays, same answers: the GPU against the CPU path
        ok = np.array_equal(hits_p.rec.cpu().numpy(), rp["rec"]) and np.array_equal(
            hits_b.rec[:nb].cpu().numpy()[perm], rb["rec"])
        cpu = {
            "value": round((n + nb) / med / 1e6, 2),
            "unit": "Mrays/s",
            "cores": cpu_threads,
            "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"the full step: {n} primary + {nb} bounce rays, SIMD width {rp['lanes']}, "
                      f"median of {args.cpu_runs} runs, traversal only",
            "parity_with_gpu": bool(ok),
        }
        log(f"CPU {cpu['value']} Mrays/s on {cpu_threads} cores ({cpu['cpu_model']}); GPU/CPU parity {ok}")
        del ow

    del st
    batched = None if args.no_batched else bench_batched(world, W.MAX_ITERS, dev)
    extras = None if args.no_extras else bench_extras(world, W.MAX_ITERS)
    del world
    configs = None
    if not args.no_configs:
        configs = {"c1_256_terrain": bench_c1(W.MAX_ITERS, cpu_threads, args.cpu_runs),
                   "c3_2048_spheres": bench_c3(W.MAX_ITERS, cpu_threads, args.cpu_runs, not args.no_c3_literal)}
    differentiable = None if args.no_diff else bench_differentiable(dev)

    line = {
        "metric": "Mrays/s (1024^3, primary+diffuse)",
        "value": round(value, 1),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "1024^3 fBm terrain + 3D value-noise caves, 1920x1080 primary + 1 diffuse bounce per hit",
            "world_bricks": int(info.num_bricks),
            "primary_rays": n,
            "bounce_rays": nb,
            "max_iters": W.MAX_ITERS,
            "tile": args.tile,
            "schedule": args.schedule,
            "order": want["order"],
            "bounce_generation": want["bounce_generation"],
            "camera_anchor": f"{args.anchor}: world origin {list(prim.world_origin)}, eye {[round(x, 4) for x in cam.c.eye]}",
            "parallelism": "single GPU",
        },
        "roofline": roofline,
        "launches": launches,
        "floor_anchor": floor,
        "fused_frame": fused,
        "cpu_baseline": cpu,
        "batched": batched,
        "shader_and_renderer": extras,
        "configs": configs,
        "differentiable": differentiable,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
