#!/usr/bin/env python3
"""Benchmark: ReSTIR DI + 3-bounce path frames on one MI355X.

Workload (default): 1920x1080 ReSTIR DI plus a 3-bounce diffuse path
(ReSTIR PT, reconnection shift, max_depth 3) on a 256^3 voxel grid split
into 64 BLAS chunks of 64^3 under one TLAS (4,194,304 voxels at 25 % fill,
83,886 emissive voxels), camera orbiting 1 degree/frame, TAA at 0.1 blend.
A step is one frame: primary rays, RIS + visibility, temporal reuse, path
segments and temporal path reuse, spatial reuse with pairwise MIS, spatial
path reuse, shading and TAA.  The scene is resident in HBM before the timed
region.

Prints one JSON line: Mrays/s (traversal queries per second), frame ms, the
roofline of the dominant traversal kernel (and of the next one and of the
per-pixel passes), HBM bytes per launch measured by rocprofv3 counter runs
of this build (child processes started before this one touches the GPU),
and the CPU oracle timed on the host cores on a bounded band of rows of the
same frame.  Beside it: the 512^3 palette workload (ReSTIR PT with the
hybrid shift, 4 bounces), the 64-view batched comparison, the
inverse-rendering step, a .vox scene, plain path tracing, BLAS/TLAS build
and refit, and GPU voxel deletion.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cube-tracing_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def depth_of(w):
    return w.restir.max_depth if w.restir.indirect else 0


def cpu_baseline(w, sc, acc, budget_s):
    """Oracle (OpenMP, all host cores) on bands of rows of frames 0 and 1."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import ctamd
    import oracle as O
    t0 = time.time()
    w2o = np.stack([acc.world_to_object(i) for i in range(sc.info.instance_count)])
    osc = O.OScene.from_scene(sc, world_to_object=w2o)
    log(f"cpu baseline: oracle LBVH built in {time.time() - t0:.1f}s")
    fr = O.OFrame(osc, w.width, w.height)
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    # Whole frames (or, when one frame is too long, a centred band of rows)
    # until about `budget_s` seconds of CPU work have been timed.
    rows = w.height
    total_rays, total_t, sampled_rows, frames = 0, 0.0, 0, 0
    mid = w.height // 2
    for f in range(64):
        p = ctamd.frame_params(w, f)
        r0 = max(0, mid - rows // 2)
        r1 = min(w.height, r0 + rows)
        t1 = time.perf_counter()
        fr.render(p, threads=cores, rows=(r0, r1))
        dt = time.perf_counter() - t1
        st = fr.stats()
        total_rays += st.primary_rays + st.shadow_rays + st.bounce_rays
        total_t += dt
        sampled_rows += r1 - r0
        frames += 1
        if dt > budget_s / 2 and rows > 16:
            rows //= 2
        if total_t >= budget_s:
            break
    return {"value": round(total_rays / total_t / 1e6, 3), "unit": "Mrays/s", "cores": cores, "kind": "port",
            "sample": f"{frames} consecutive frames of the same workload from frame 0, {sampled_rows} rows in "
                      f"total ({total_rays} traversal queries in {total_t:.1f}s, OpenMP {cores} threads)"}


def c0_bench(sets=5):
    """The 64^3 plain workload (31,457 voxels, 8 lights, 256x256, ReSTIR DI
    + 1 bounce, frames 0-3): GPU frame time by events over `sets` runs of
    the four frames (context reset between runs), and the oracle on the
    host cores rendering the same four frames once."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ctamd
    import oracle as O
    w = ctamd.small_plain()
    sc = ctamd.Scene(ctamd.scene_desc(w))
    acc, ctx, _ = setup(sc, w)
    params = [ctamd.frame_params(w, f) for f in range(4)]
    ms, rays = [], 0
    for k in range(sets + 1):
        ctx.reset()
        ctx.set_instrumentation(4)
        for p in params:
            ctx.render(acc, p)
            st = ctx.stats()
            if k > 0:  # the first set is a warm-up
                ms.append(st.ms_total)
                rays += st.primary_rays + st.shadow_rays + st.bounce_rays
    ctx.set_instrumentation(0)
    gpu_ms = sum(ms) / len(ms)
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    fr = O.OFrame(O.OScene.from_scene(sc), w.width, w.height)
    t0 = time.perf_counter()
    crays = 0
    for p in params:
        fr.render(p, threads=cores)
        st = fr.stats()
        crays += st.primary_rays + st.shadow_rays + st.bounce_rays
    cpu_s = time.perf_counter() - t0
    out = {"workload": "small_plain_64^3_256x256_di_pt1_4frames", "voxels": sc.info.primitive_count,
           "gpu_frame_ms_mean": round(gpu_ms, 4), "gpu_mrays_s": round(rays / (sum(ms) / 1e3) / 1e6, 1),
           "cpu_frame_ms_mean": round(cpu_s * 1e3 / 4, 2), "cpu_mrays_s": round(crays / cpu_s / 1e6, 2),
           "cpu_cores": cores, "cpu_kind": "port", "rays_per_frame": crays // 4,
           "gpu_over_cpu": round((cpu_s * 1e3 / 4) / gpu_ms, 1)}
    ctx.close()
    acc.close()
    sc.close()
    return out


def batched_bench(frames=5, warmup=2, n_views=64):
    """Batched views (4 scenes, 64 views of uneven size): per-view
    ct_frame_render against one batched, graph-replayed render."""
    import ctamd
    t0 = time.time()
    scenes, views = ctamd.batched_views(n_views=n_views)
    built = []
    for sw in scenes:
        sc = ctamd.Scene(ctamd.scene_desc(sw))
        acc = ctamd.Accel(max_instances=sc.info.instance_count + 8,
                          max_primitives=sc.info.primitive_count + 32 * sc.info.instance_count + 1024,
                          max_cube_lights=sc.info.cube_light_count + 8)
        tmp = ctamd.Ctx(8, 8)
        ctamd.load_scene(sc, acc, tmp)
        built.append((sc, acc))
    ctx_plain, ctx_batch = [], []
    for k, w in views:
        for lst in (ctx_plain, ctx_batch):
            c = ctamd.Ctx(w.width, w.height)
            ctamd.bind_scene(built[k][0], c)
            lst.append(c)
    accs = [built[k][1] for k, _ in views]
    b = ctamd.Batch(ctx_batch, accs)
    log(f"batched: {len(views)} views over {len(scenes)} scenes "
        f"({sum(sc.info.primitive_count for sc, _ in built)} voxels, "
        f"{sum(w.width * w.height for _, w in views)} pixels) set up in {time.time() - t0:.1f}s")
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    params = [[ctamd.frame_params(w, f) for _, w in views] for f in range(warmup + frames)]
    # the batch's own launch shape (2 view groups, 5 traversal blocks per CU,
    # far rays from 2 units first); the launch-shape tuner is switched off
    log("batched: default launch shape: 2 view groups, traversal grid 5 blocks/CU, far rays from 2 units")
    # plain: one ct_frame_render per view and frame, issued back to back (the
    # views' own streams overlap on the GPU), then the same with each view
    # finished before the next starts, as a renderer drawing one view per
    # frame does (draw(), main.cpp:1436), then the same renders enqueued on
    # one stream.  Every loop starts from the same history: the later ones
    # render the same frame numbers on copies.
    t_plain, t_serial, t_batch, rays = 0.0, 0.0, 0.0, 0
    for f in range(warmup + frames):
        hip.hipDeviceSynchronize()
        t1 = time.perf_counter()
        for (k, w), c, p in zip(views, ctx_plain, params[f]):
            c.render(built[k][1], p)
        hip.hipDeviceSynchronize()
        if f >= warmup:
            t_plain += time.perf_counter() - t1
    ctx_serial = []
    for k, w in views:
        c = ctamd.Ctx(w.width, w.height)
        ctamd.bind_scene(built[k][0], c)
        ctx_serial.append(c)
    for f in range(warmup + frames):
        hip.hipDeviceSynchronize()
        t1 = time.perf_counter()
        for (k, w), c, p in zip(views, ctx_serial, params[f]):
            c.render(built[k][1], p)
            hip.hipDeviceSynchronize()
        if f >= warmup:
            t_serial += time.perf_counter() - t1
    # plain on one stream: the 64 renders enqueued one after another on a
    # single stream with no host synchronisation between them (each view's
    # launches follow the previous view's, as one command queue runs them)
    import torch
    one = torch.cuda.Stream()
    ctx_one = []
    for k, w in views:
        c = ctamd.Ctx(w.width, w.height)
        ctamd.bind_scene(built[k][0], c)
        ctx_one.append(c)
    t_one = 0.0
    for f in range(warmup + frames):
        hip.hipDeviceSynchronize()
        t1 = time.perf_counter()
        for (k, w), c, p in zip(views, ctx_one, params[f]):
            c.render(built[k][1], p, stream=one.cuda_stream)
        hip.hipDeviceSynchronize()
        if f >= warmup:
            t_one += time.perf_counter() - t1
    for c in ctx_one:
        c.close()
    for f in range(warmup + frames):
        hip.hipDeviceSynchronize()
        t1 = time.perf_counter()
        b.render(params[f], use_graph=True)
        hip.hipDeviceSynchronize()
        if f >= warmup:
            t_batch += time.perf_counter() - t1
            st = b.stats()
            rays += st.primary_rays + st.shadow_rays + st.bounce_rays
    return {"views": len(views), "scenes": len(scenes), "pixels": sum(w.width * w.height for _, w in views),
            "indirect_bounces": views[0][1].restir.max_depth, "frames": frames,
            "plain_ms_per_frame": round(1000 * t_plain / frames, 3),
            "plain_serial_ms_per_frame": round(1000 * t_serial / frames, 3),
            "plain_one_stream_ms_per_frame": round(1000 * t_one / frames, 3),
            "speedup_vs_one_stream": round(t_one / t_batch, 2),
            "batched_ms_per_frame": round(1000 * t_batch / frames, 3),
            "speedup": round(t_plain / t_batch, 2), "speedup_vs_serial": round(t_serial / t_batch, 2), "batched_mrays_s": round(rays / t_batch / 1e6, 1),
            "graph": True, "blocks_per_cu": 5, "groups": 2, "far_ray": 2.0, "tuned": False}


def frame_adjoint_bench(w, acc, ctx, sc, f, reps=10):
    """The frame adjoint (ct_frame_adjoint) on the 1080p workload: one frame
    rendered, then the gradient of sum(g * shading output) with its final
    reservoirs held fixed, with respect to every material's albedo and
    emission (one material per voxel), timed over `reps` calls."""
    import torch
    import ctamd
    from ctamd.adjoint import frame_adjoint
    p = ctamd.frame_params(w, f)
    ctx.render(acc, p)
    ctx.stats()
    nm = sc.info.material_count
    g = torch.randn((w.width * w.height, 4), dtype=torch.float32, device="cuda")
    gd, ge = frame_adjoint(ctx, acc, p, g, nm)  # untimed: allocation of the gradients
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(reps):
        gd, ge = frame_adjoint(ctx, acc, p, g, nm)
    torch.cuda.synchronize()
    t = (time.perf_counter() - t1) / reps
    return {"workload": w.name, "materials": nm, "pixels": w.width * w.height, "ms_per_call": round(1000 * t, 3),
            "nonzero_albedo_grads": int(torch.count_nonzero(gd).item()),
            "note": "includes zeroing the two [materials][3] gradient tensors and a device sync per call"}


def adjoint_bench(steps=3, fit_steps=200):
    """Inverse rendering on 128^3 voxels, 8 views of 512^2 at 16 spp, 3
    bounces: forward render vs forward + path-replay adjoint + Adam step
    (after one untimed step), then the 200-step Adam fit from the initial
    parameters (learning rates ramped over 10 steps) with its loss curve."""
    import torch
    from ctamd import adjoint as AD
    P = AD.c4_problem()
    w, r, target = P["workload"], P["renderer"], P["target"]
    a, e = P["albedo"], P["emission"]
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(steps):
        with torch.no_grad():
            r.forward(a, e)
    torch.cuda.synchronize()
    t_fwd = (time.perf_counter() - t1) / steps
    a2 = a.detach().clone().requires_grad_(True)
    e2 = e.detach().clone().requires_grad_(True)
    AD.fit(r, target, a2, e2, 1)  # untimed: one-time autograd and allocator set-up
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    AD.fit(r, target, a2, e2, steps)
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t1) / steps
    t1 = time.perf_counter()
    hist = AD.fit(r, target, a, e, fit_steps)
    torch.cuda.synchronize()
    fit_s = time.perf_counter() - t1
    paths = w.views * w.width * w.height * w.spp
    curve = {str(i): round(hist[i], 6) for i in sorted(set(list(range(0, fit_steps, 20)) + [1, 2, fit_steps - 1]))}
    return {"voxels": P["scene"].info.primitive_count, "views": w.views, "resolution": [w.width, w.height],
            "spp": w.spp, "bounces": w.max_depth, "paths_per_render": paths, "forward_ms": round(1000 * t_fwd, 2),
            "adam_step_ms": round(1000 * t_step, 2), "step_over_forward": round(t_step / t_fwd, 2),
            "forward_mpaths_s": round(paths / t_fwd / 1e6, 1), "fit_steps": fit_steps, "fit_s": round(fit_s, 2),
            "fit_loss_curve": curve, "fit_loss_ratio": round(hist[-1] / hist[0], 6),
            "fit_first_steps_decrease": all(hist[i + 1] < hist[i] for i in range(10))}


def vox_bench(frames=8, warmup=2, n=128, chunk=64):
    """A .vox scene end to end: a procedural n^3 MagicaVoxel model (stepped
    tower, about 1 % emissive voxels) parsed by the library's loader
    (host), loaded (staging, BLAS/TLAS build, upload), then ReSTIR DI + a
    3-bounce path at 1920x1080 on it (two-stream frames, GPU events)."""
    import ctamd
    from ctamd.voxfile import monument
    data = monument(n=n)
    t0 = time.perf_counter()
    sc = ctamd.Scene.from_vox(data, chunk=(chunk, chunk, chunk))
    parse_s = time.perf_counter() - t0
    acc = ctamd.Accel(max_instances=sc.info.instance_count + 8,
                      max_primitives=sc.info.primitive_count + 32 * sc.info.instance_count + 1024,
                      max_cube_lights=max(sc.info.cube_light_count, 1) + 8)
    ctx = ctamd.Ctx(1920, 1080)
    t0 = time.perf_counter()
    ctamd.load_scene(sc, acc, ctx)
    load_s = time.perf_counter() - t0
    w = ctamd.vox_view(sc, frames=warmup + frames)
    for f in range(warmup):
        ctx.render(acc, ctamd.frame_params(w, f))
    ctx.stats()
    ctx.set_instrumentation(4)
    ms, rays = [], 0
    for f in range(warmup, warmup + frames):
        ctx.render(acc, ctamd.frame_params(w, f))
        st = ctx.stats()
        ms.append(st.ms_total)
        rays += st.primary_rays + st.shadow_rays + st.bounce_rays
    med = sorted(ms)[len(ms) // 2]
    nv = sc.info.primitive_count
    return {"file_bytes": len(data), "grid": [n, n, n], "voxels": nv, "blas": sc.info.instance_count,
            "materials": sc.info.material_count, "lights": sc.info.cube_light_count,
            "parse_ms": round(parse_s * 1e3, 2), "parse_mvoxels_s": round(nv / parse_s / 1e6, 1),
            "load_ms": round(load_s * 1e3, 1), "frame_ms_median": round(med, 3),
            "mrays_s": round(rays / (sum(ms) / 1e3) / 1e6, 1)}


def path_trace_bench(acc, sc, w, spp=1, depth=3, reps=3):
    """Plain path tracing with the material callables on the 1080p scene:
    its materials with every 4th voxel metal, dielectric and constant
    medium in turn (the rest Lambertian), one 1920x1080 view from the
    workload's first camera, `spp` samples and `depth` bounces per pixel
    (after one untimed call that uploads the materials)."""
    import ctamd
    from ctamd import adjoint as AD
    import numpy as np
    n = sc.info.material_count
    mats = (ctamd.Material * n).from_buffer_copy(sc.materials)
    u = np.frombuffer(mats, np.uint32).reshape(n, 21)  # MATERIAL as 21 words
    f = u.view(np.float32)
    kind = np.arange(n, dtype=np.uint32) % 4
    u[:, 0] = kind
    metal, dielectric = kind == ctamd.CT_MATERIAL_METAL, kind == ctamd.CT_MATERIAL_DIELECTRIC
    f[metal, 7:10] = f[metal, 4:7]  # specular = diffuse
    f[metal, 17] = 0.3              # roughness
    f[dielectric, 18] = 1.5         # ior
    p = ctamd.frame_params(w, 0)
    v = ctamd.AdjointView()
    v.width, v.height = w.width, w.height
    v.inv_view[:] = p.inv_view[:]
    v.inv_proj[:] = p.inv_proj[:]
    import torch
    AD.path_trace(acc, [v], mats, spp=spp, max_depth=depth)
    torch.cuda.synchronize()  # untimed first call
    t0 = time.perf_counter()
    for _ in range(reps):
        img = AD.path_trace(acc, [v], None, spp=spp, max_depth=depth)  # materials kept in the accel
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / reps * 1e3
    samples = w.width * w.height * spp
    return {"resolution": [w.width, w.height], "spp": spp, "bounces": depth, "ms": round(ms, 2),
            "msamples_s": round(samples / ms / 1e3, 1), "mean_radiance": round(float(img.mean()), 5)}


def brush_bench(ctx, acc, w, f, window=16):
    """GPU voxel deletion on the 1080p scene (last, as it edits the scene):
    one squared-brush stroke of half-size `window` at the image centre of
    the last frame, then processing (device compaction of the marked
    instance, task queuing) and update_scene (BLAS rebuild); then two
    bulk deletions of 1 % of the voxels of every BLAS marked from a list,
    each followed by the settling update_scene.  Wall times of each call."""
    import numpy as np
    import ctamd
    ctx.render(acc, ctamd.frame_params(w, f))
    ctx.stats()
    acc.update_scene(True)  # idle
    res = {}
    t0 = time.perf_counter()
    ctamd.brush_apply(ctx, acc, w.width // 2, w.height // 2, window)
    _, marked = acc.brush_counters()
    t1 = time.perf_counter()
    tasks, vox = acc.process_voxel_modifications()
    t2 = time.perf_counter()
    acc.update_scene(True)
    t3 = time.perf_counter()
    acc.update_scene(True)  # settle
    res["stroke"] = {"window": window, "voxels": vox, "instances": tasks, "mark_ms": round((t1 - t0) * 1e3, 3),
                     "process_ms": round((t2 - t1) * 1e3, 3), "update_scene_ms": round((t3 - t2) * 1e3, 3)}
    # Two bulk deletions in a row: the first also allocates the pinned
    # staging the host mirror's refresh goes through.
    for rep, key in enumerate(("bulk_1pct", "bulk_1pct_again")):
        inst = acc.read_instances()
        rng = np.random.default_rng(3 + rep)
        marks = []
        for i, r in enumerate(inst):
            if r.primitive_count > 100:
                k = r.primitive_count // 100
                marks += [(i, int(p)) for p in rng.choice(r.primitive_count, k, replace=False)]
        marks = np.array(marks, np.uint32)
        t0 = time.perf_counter()
        acc.brush_mark(marks)
        t1 = time.perf_counter()
        tasks, vox = acc.process_voxel_modifications()
        t2 = time.perf_counter()
        acc.update_scene(True)
        t3 = time.perf_counter()
        acc.update_scene(True)  # settle
        t4 = time.perf_counter()
        res[key] = {"voxels": vox, "instances": tasks, "mark_ms": round((t1 - t0) * 1e3, 3),
                    "process_ms": round((t2 - t1) * 1e3, 3), "update_scene_ms": round((t3 - t2) * 1e3, 3),
                    "settle_ms": round((t4 - t3) * 1e3, 3)}
    return res


def update_bench(acc, n_inst, reps=5):
    """Scene updates on the loaded accel (after every other measurement, as
    they move nothing but still rewrite it): UPDATE_BLAS_FROM_CPU tasks with
    an identity transform on 1 and on all BLAS, then update_scene(synchronize)
    -- settle, task processing, BLAS refits, TLAS rebuild and table uploads.
    Median wall time per call."""
    import numpy as np
    ident = np.eye(4, dtype=np.float32).reshape(16)
    res = {}
    for k in (1, n_inst):
        times = []
        for _ in range(reps):
            for i in range(k):
                acc.update_blas(i, ident)
            t0 = time.perf_counter()
            acc.update_scene(True)
            times.append(time.perf_counter() - t0)
        res[f"refit_{k}_blas_ms"] = round(sorted(times)[len(times) // 2] * 1e3, 3)
    return res


def setup(sc, w):
    """Accel sized for the scene, a frame context of the workload's size,
    and the timed BLAS/TLAS build + upload."""
    import ctamd
    acc = ctamd.Accel(max_instances=sc.info.instance_count + 8,
                      max_primitives=sc.info.primitive_count + 32 * sc.info.instance_count + 1024,
                      max_cube_lights=sc.info.cube_light_count + 8)
    ctx = ctamd.Ctx(w.width, w.height)
    t0 = time.perf_counter()
    ctamd.load_scene(sc, acc, ctx)
    return acc, ctx, time.perf_counter() - t0


# Algorithmic bytes of the per-pixel passes: the pixel buffers each pass
# must read and write (structure-of-arrays records: DI 64 B, reservoir
# 36 B, neighbour record 128 B, spatial reuse's per-neighbour result 48 B,
# hit 16 B, shadow-ray record 36 B) plus
# the scene records it reads per pixel (AABB 24 B, instance 48 B,
# primitive 8 B, previous transform 64 B, material 32 B, light 48 B per RIS
# candidate).
RIS_HIT_BYTES = 16 + 24 + 48 + 8 + 64 + 32 + 36 + 8 + 36 + 64 + 4 + 4 + 32  # + temporal reuse's two targets
RIS_MISS_BYTES = 16 + 16 + 16 + 64 + 4
# spatial reuse: own records in both passes, the 16 B header of the
# per-neighbour results written by the emit pass and read by the finish pass
SPATIAL_HIT_BYTES = (64 + 32 + 36 + 4) + (64 + 32 + 36 + 4 + 16 + 36 + 64) + 2 * 16
# per neighbour: its 128 B record gathered once, its 48 B result written and read
SPATIAL_NEIGHBOUR_BYTES = 128 + 2 * 48
RAY_BYTES = 36


def frame_bench(w, acc, ctx, K, Wm, traffic):
    """K timed frames of workload w (two halves on two streams, as the
    library renders a large frame) after Wm warm-up frames and one
    instrumented frame; then K frames on one stream with events around
    every pass for the per-kernel rooflines."""
    import ctamd
    import numpy as np
    f = 0
    for _ in range(Wm):
        ctx.render(acc, ctamd.frame_params(w, f))
        f += 1
    ctx.stats()
    # One instrumented frame (not timed): traversal node/leaf counters.
    ctx.set_instrumentation(2)
    ctx.render(acc, ctamd.frame_params(w, f))
    f += 1
    ist = ctx.stats()
    hits = int((np.frombuffer(bytes(ctx.read(ctamd.CT_BUF_DI, w.width, w.height)), np.float32)
                .reshape(-1, 15)[:, 3] > 0).sum())
    # Algorithmic bytes per query: a shadow ray reads its 36 B record, a
    # path segment reads 32 B and writes a 16 B hit; every BVH2 node
    # fetched is 64 B, every leaf 28 B (centre + id + instance record
    # share), every instance entry 64 B.
    n_any = ist.shadow_rays_ris + ist.shadow_rays_temporal + ist.shadow_rays_pt + ist.shadow_rays_spatial
    bytes_any = 36 * n_any + 64 * ist.nodes + 28 * ist.leaves + 64 * ist.instances
    bytes_seg = 48 * ist.bounce_rays + 64 * ist.pt_nodes + 28 * ist.pt_leaves + 64 * ist.pt_instances
    b_any = bytes_any / max(n_any, 1)
    b_seg = bytes_seg / max(ist.bounce_rays, 1)
    npix = w.width * w.height
    n_c = w.restir.max_ris_cube_samples
    algo_frame = {
        "k_primary": 16 * npix + 64 * ist.closest_nodes + 28 * ist.closest_leaves + 64 * ist.closest_instances,
        "k_ris": hits * (RIS_HIT_BYTES + 48 * n_c) + (npix - hits) * RIS_MISS_BYTES
                 + RAY_BYTES * (ist.shadow_rays_ris + ist.shadow_rays_temporal),
        "k_spatial": hits * (SPATIAL_HIT_BYTES + SPATIAL_NEIGHBOUR_BYTES * w.restir.max_neighbors)
                     + RAY_BYTES * ist.shadow_rays_spatial,
    }
    log(f"{w.name}: instrumented frame: {hits} hit pixels, {n_any} shadow rays ({ist.shadow_rays_ris} RIS, "
        f"{ist.shadow_rays_temporal} temporal, {ist.shadow_rays_pt} path reuse, {ist.shadow_rays_spatial} "
        f"spatial) {b_any:.0f} B/ray; {ist.bounce_rays} path segments {b_seg:.0f} B/segment; "
        f"closest: {ist.closest_nodes} nodes")
    # Timed region: K frames as the library runs them, with only the
    # frame's start/end events.
    ctx.set_instrumentation(4)
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipDeviceSynchronize()
    t_start = time.perf_counter()
    rays = 0
    ms_frames = []
    for _ in range(K):
        ctx.render(acc, ctamd.frame_params(w, f))
        f += 1
        st = ctx.stats()  # drains the stream, reads the frame's counters and events
        rays += st.primary_rays + st.shadow_rays + st.bounce_rays
        ms_frames.append(st.ms_total)
    hip.hipDeviceSynchronize()
    wall = time.perf_counter() - t_start
    # Kernel durations for the roofline: K more frames run whole on one
    # stream with events around every pass (outside the timed region), so
    # each traversal launch covers the whole frame's rays on its own.
    ctx.set_instrumentation(1)
    ms = {"k_trace_any": 0.0, "k_trace_segments": 0.0, "k_primary": 0.0, "k_ris": 0.0, "k_spatial": 0.0}
    n_any_t = n_seg_t = 0
    ms_pt, ms_whole = [], []
    for _ in range(K):
        ctx.render(acc, ctamd.frame_params(w, f))
        f += 1
        st = ctx.stats()
        ms["k_trace_any"] += st.ms_trace_any
        ms["k_trace_segments"] += st.ms_pt_segments
        ms["k_primary"] += st.ms_primary
        ms["k_ris"] += st.ms_ris
        ms["k_spatial"] += st.ms_spatial
        n_any_t += st.shadow_rays_ris + st.shadow_rays_temporal + st.shadow_rays_pt + st.shadow_rays_spatial
        n_seg_t += st.bounce_rays
        ms_pt.append(st.ms_pt)
        ms_whole.append(st.ms_total)
    ctx.set_instrumentation(0)
    depth = depth_of(w)
    algo = {"k_trace_any": b_any * n_any_t / K, "k_trace_segments": b_seg * n_seg_t / K, **algo_frame}
    # any-hit launches: RIS and temporal visibility, path-reuse visibility
    # (with paths), spatial visibility
    launches = {"k_trace_any": 3 if depth else 2,
                "k_trace_segments": max(depth, 1), "k_primary": 1, "k_ris": 1, "k_spatial": 2}
    desc = {"k_trace_any": "k_trace_any_frames (shadow-ray any-hit traversal)",
            "k_trace_segments": "k_trace_segments (path-segment closest-hit traversal)",
            "k_primary": "k_primary (camera-ray closest-hit traversal)",
            "k_ris": "k_ris (hit shading, motion vector, RIS candidates, visibility-ray emission)",
            "k_spatial": "k_spatial (spatial reuse emit + finish, shading)"}

    def roof(k):
        per_frame_ms = ms[k] / K
        achieved = algo[k] / (per_frame_ms / 1e3) / 1e9 if per_frame_ms > 0 else 0.0
        t = traffic.get(k)
        r = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": t,
             "algorithmic_bytes_per_launch": round(algo[k] / launches[k]),
             "kernel": desc[k], "ms_per_frame": round(per_frame_ms, 3), "launches_per_frame": launches[k]}
        if t:
            r["hbm_gbs_physical"] = round(t * launches[k] / (per_frame_ms / 1e3) / 1e9, 1)
            r["frac_physical"] = round(r["hbm_gbs_physical"] / HBM_PEAK_GBS, 4)
        return r

    order = sorted(("k_trace_any", "k_trace_segments"), key=lambda k: -ms[k])
    med = lambda v: sorted(v)[len(v) // 2]
    return {"mrays": round(rays / wall / 1e6, 2), "ms_per_step": round(1000.0 * wall / K, 3),
            "next_frame": f, "roofline": roof(order[0]), "roofline_next_kernel": roof(order[1]),
            "frame_kernels": {k: roof(k) for k in ("k_primary", "k_ris", "k_spatial")},
            "config": {"frame_ms_gpu_events_median": round(med(ms_frames), 3),
                       "frame_ms_one_stream_median": round(med(ms_whole), 3),
                       "pt_ms_gpu_events_median": round(med(ms_pt), 3), "rays_per_frame": rays // K}}


def palette_bench(K=8, Wm=3, traffic=None):
    """The 512^3 palette workload: 16,106,127 voxels in 512 chunks, a
    256-entry palette (3 emissive entries, 187,748 cube lights), ReSTIR DI
    plus ReSTIR PT with the hybrid shift, 4 bounces, M-cap 20, 3 spatial
    neighbours, 1920x1080, orbiting camera, TAA."""
    import ctamd
    w = ctamd.palette_pt(frames=Wm + 2 * K + 2)
    t0 = time.time()
    sc = ctamd.Scene(ctamd.scene_desc(w))
    gen_s = time.time() - t0
    acc, ctx, load_s = setup(sc, w)
    log(f"palette: {sc.info.primitive_count} voxels in {sc.info.instance_count} chunks, "
        f"{sc.info.cube_light_count} lights; generated in {gen_s:.1f}s, built + uploaded in {load_s:.2f}s")
    r = frame_bench(w, acc, ctx, K, Wm, traffic or {})
    out = {"workload": f"restir_pt_hybrid4_{w.width}x{w.height}_grid512_palette256_mcap20_n3_taa",
           "voxels": sc.info.primitive_count, "blas": sc.info.instance_count,
           "lights": sc.info.cube_light_count, "palette_entries": sc.info.material_count,
           "frames": K, "mrays_s": r["mrays"], "ms_per_frame": r["ms_per_step"], **r["config"],
           "load_ms": round(load_s * 1e3, 1), "roofline": r["roofline"],
           "roofline_next_kernel": r["roofline_next_kernel"], "roofline_frame_kernels": r["frame_kernels"]}
    ctx.close()
    acc.close()
    sc.close()
    return out


def traffic_probe(which, grid, width, height, depth):
    """Child of a rocprofv3 counter run: a 1 GiB read + 1 GiB write
    elementwise pass (the counters' calibration), a separator pass, then
    one warm-up frame and two whole frames on one stream of the
    workload."""
    import torch
    import ctamd
    x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.mul(x, 2.0, out=y)
    torch.cuda.synchronize()
    del x, y
    if which == "palette":
        w = ctamd.palette_pt(frames=4)
    else:
        w = ctamd.hd_di(frames=4, width=width, height=height, grid=grid, indirect=depth > 0,
                        max_depth=max(depth, 1))
    sc = ctamd.Scene(ctamd.scene_desc(w))
    acc, ctx, _ = setup(sc, w)
    ctx.render(acc, ctamd.frame_params(w, 0))
    ctx.stats()
    s = torch.ones(1 << 10, device="cuda")
    torch.mul(s, 2.0, out=s)  # separator: the frames counted follow this pass
    torch.cuda.synchronize()
    ctx.set_instrumentation(1)
    for f in (1, 2):
        ctx.render(acc, ctamd.frame_params(w, f))
        ctx.stats()
    print(json.dumps({"probe": which, "frames": 2}), flush=True)


TRAFFIC_KERNELS = {"k_trace_any": "k_trace_any_frames", "k_trace_segments": "k_trace_segments",
                   "k_primary": "k_primary", "k_ris": "k_ris", "k_spatial": "k_spatial"}


def _counter_rows(folder):
    """Per-dispatch counter sums of a rocprofv3 --pmc run, in dispatch
    order: its CSV output, else its SQLite database."""
    import csv
    import glob
    rows = {}
    for path in glob.glob(os.path.join(folder, "**", "*counter_collection.csv"), recursive=True):
        with open(path) as fh:
            for r in csv.DictReader(fh):
                ent = rows.setdefault(int(r["Dispatch_Id"]), {"name": r["Kernel_Name"], "value": 0.0})
                ent["value"] += float(r["Counter_Value"])
    if not rows:
        import sqlite3
        for path in glob.glob(os.path.join(folder, "**", "*.db"), recursive=True):
            con = sqlite3.connect(path)
            try:
                for did, name, val in con.execute("select dispatch_id, name, counter_value from pmc_events"):
                    ent = rows.setdefault(int(did), {"name": name, "value": 0.0})
                    ent["value"] += float(val)
            finally:
                con.close()
    return [rows[k] for k in sorted(rows)]


def measure_traffic(probes, argv_extra, timeout_s=240):
    """HBM bytes per launch of the frame's kernels, measured by this build:
    per workload, two rocprofv3 counter runs (FETCH_SIZE, then WRITE_SIZE)
    over a child process of this script.  Each counter is scaled by what it
    reads for the child's calibration pass (1 GiB read and 1 GiB written),
    and averaged over the launches of the two frames after the separator.
    Runs before this process touches the GPU."""
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        return {"note": "rocprofv3 not found: traffic not measured"}
    out = {}
    notes = []
    tmp = tempfile.mkdtemp(prefix="ct_traffic_", dir="/tmp")
    env = dict(os.environ, TMPDIR="/tmp")
    for which in probes:
        res = {}
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            folder = os.path.join(tmp, f"{which}_{counter}")
            cmd = [prof, "--pmc", counter, "--output-format", "csv", "-d", folder, "-o", "c", "--", sys.executable, os.path.abspath(__file__),
                   "--traffic-probe", which] + argv_extra
            t0 = time.time()
            p = None
            try:
                p = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                   timeout=timeout_s, start_new_session=True)
                ok = p.returncode == 0 or b'"probe"' in p.stdout
            except subprocess.TimeoutExpired:
                ok = False
            rows = _counter_rows(folder) if ok else []
            log(f"traffic: {which} {counter}: {len(rows)} dispatches ({time.time() - t0:.1f}s)")
            if not rows:
                if ok or p is not None:
                    log("traffic: counter run output:\n" + p.stdout.decode(errors="replace")[-3000:])
                notes.append(f"{which} {counter}: counter run failed")
                break
            sep = max(i for i, r in enumerate(rows) if "elementwise" in r["name"])
            calib = max(r["value"] for r in rows[:sep] if "elementwise" in r["name"])
            scale = float(1 << 30) / (calib * 1024.0)
            for k, pat in TRAFFIC_KERNELS.items():
                vals = [r["value"] for r in rows[sep + 1:] if pat in r["name"]]
                if vals:
                    res.setdefault(k, 0.0)
                    res[k] += scale * 1024.0 * sum(vals) / len(vals)
            res[f"scale_{counter}"] = round(scale, 4)
        else:
            out[which] = {k: (int(v) if k.startswith("k_") else v) for k, v in res.items()}
    shutil.rmtree(tmp, ignore_errors=True)
    out["note"] = ("HBM bytes per launch = FETCH_SIZE x scale + WRITE_SIZE x scale (KiB x 1024), each from its "
                   "own rocprofv3 --pmc run of this build over two whole frames on one stream; scale = 1 GiB / "
                   "the counter's reading for a 1 GiB elementwise pass in the same run"
                   + ("; " + "; ".join(notes) if notes else ""))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--depth", type=int, default=3, help="path bounces (ReSTIR PT max_depth); 0 = DI only")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU baseline work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-view batched comparison")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the inverse-rendering step timing")
    ap.add_argument("--no-vox", action="store_true",
                    help="skip the .vox scene and material path-tracing timings")
    ap.add_argument("--no-brush", action="store_true", help="skip the GPU voxel-deletion timing")
    ap.add_argument("--no-palette", action="store_true", help="skip the 512^3 palette workload")
    ap.add_argument("--no-c0", action="store_true", help="skip the 64^3 plain workload (GPU and oracle)")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 counter runs that measure HBM traffic per launch")
    ap.add_argument("--traffic-probe", choices=("hd", "palette"), help=argparse.SUPPRESS)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device copy (rocprof counter calibration)")
    args = ap.parse_args()
    if args.traffic_probe:
        traffic_probe(args.traffic_probe, args.grid, args.width, args.height, args.depth)
        return
    if args.calibrate_copy:
        import torch
        x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": x.numel() * 4, "copy_bytes_written": y.numel() * 4, "launches": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the per-frame path is single-GPU "
              f"and multi-GPU sharding is not part of this build", file=sys.stderr)
        sys.exit(2)
    traffic = {}
    if not args.no_traffic:
        # counter runs first: each is a child process, started before this
        # process initialises the GPU
        probes = ["hd"] + ([] if args.no_palette else ["palette"])
        traffic = measure_traffic(probes, ["--grid", str(args.grid), "--width", str(args.width),
                                           "--height", str(args.height), "--depth", str(args.depth)])
        log(f"traffic: {traffic}")
    import ctamd
    if not ctamd.device_available():
        raise SystemExit("bench: no gfx950 device visible")
    K, Wm = args.steps, args.warmup
    w = ctamd.hd_di(frames=Wm + K + 1, width=args.width, height=args.height, grid=args.grid,
                    indirect=args.depth > 0, max_depth=max(args.depth, 1))
    t0 = time.time()
    sc = ctamd.Scene(ctamd.scene_desc(w))
    log(f"scene: {sc.info.instance_count} instances, {sc.info.primitive_count} voxels, "
        f"{sc.info.cube_light_count} lights ({time.time() - t0:.1f}s)")
    acc, ctx, load_s = setup(sc, w)
    log(f"BLAS/TLAS build + upload: {load_s:.3f}s")
    hd = frame_bench(w, acc, ctx, K, Wm, traffic.get("hd", {}))
    f = hd["next_frame"]
    palette = None
    if not args.no_palette:
        palette = palette_bench(traffic=traffic.get("palette"))
        log(f"palette: {palette}")
    c0 = None
    if not args.no_c0:
        c0 = c0_bench()
        log(f"c0: {c0}")
    batched = adjoint = None
    if not args.no_batched:
        batched = batched_bench()
        log(f"batched: {batched}")
    frame_adj = None
    if not args.no_adjoint:
        frame_adj = frame_adjoint_bench(w, acc, ctx, sc, f)
        f += 1
        log(f"frame adjoint: {frame_adj}")
        adjoint = adjoint_bench()
        log(f"adjoint: {adjoint}")
    vox = None
    if not args.no_vox:
        vox = vox_bench()
        log(f"vox: {vox}")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(w, sc, acc, args.cpu_budget)
    build = {"voxels": sc.info.primitive_count, "blas": sc.info.instance_count,
             "load_ms": round(load_s * 1e3, 1),
             "load_mvoxels_s": round(sc.info.primitive_count / load_s / 1e6, 1)}
    build.update(update_bench(acc, sc.info.instance_count))
    log(f"build: {build}")
    ptrace = None
    if not args.no_vox:
        ptrace = path_trace_bench(acc, sc, w)
        log(f"path_trace: {ptrace}")
    brush = None
    if not args.no_brush:
        brush = brush_bench(ctx, acc, w, f)
        log(f"brush: {brush}")
    depth = depth_of(w)
    out = {
        "metric": "Mrays/s + 1080p frame ms (ReSTIR DI/PT), 1 MI355X; batched/adjoint vs plain",
        "value": hd["mrays"], "unit": "Mrays/s", "n_gpus": 1, "steps": K, "warmup": Wm,
        "ms_per_step": hd["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"restir_di_pt{depth}_{w.width}x{w.height}_grid{args.grid}_64blas_taa",
                   "width": w.width, "height": w.height, "grid": args.grid, "blas": sc.info.instance_count,
                   "voxels": sc.info.primitive_count, "lights": sc.info.cube_light_count,
                   **hd["config"], "indirect_bounces": depth},
        "roofline": hd["roofline"],
        "roofline_next_kernel": hd["roofline_next_kernel"],
        "roofline_frame_kernels": hd["frame_kernels"],
        "traffic_measurement": traffic.get("note"),
        "cpu_baseline": cpu,
        "palette": palette,
        "c0": c0,
        "batched": batched,
        "adjoint": adjoint,
        "frame_adjoint": frame_adj,
        "build": build,
        "vox": vox,
        "path_trace": ptrace,
        "brush": brush,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
