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
roofline of the any-hit traversal kernel and the CPU oracle timed on the
host cores on a bounded band of rows of the same frame.
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


TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "hbm_traffic_trace_hd_di_pt3.md")


def _traffic_summary():
    """Measured HBM bytes per launch of the traversal kernels on the default
    workload: the `traffic_bytes_per_launch:` line of the counter summary."""
    try:
        with open(TRAFFIC_SUMMARY) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch:"):
                    return {k: int(v) for k, v in (kv.split("=") for kv in line.split(":", 1)[1].split())}
    except OSError:
        pass
    return {}


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
    tune = b.autotune(params[0], repeats=2)
    log(f"batched: autotuned {tune.groups} view groups, traversal grid {tune.blocks_per_cu} blocks/CU ({tune.ms:.2f} ms)")
    # plain: one ct_frame_render per view and frame, issued back to back (the
    # views' own streams overlap on the GPU), then the same with each view
    # finished before the next starts, as a renderer drawing one view per
    # frame does (draw(), main.cpp:1436).  Both loops start from the same
    # history: the serial one renders the same frame numbers on copies.
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
            "batched_ms_per_frame": round(1000 * t_batch / frames, 3),
            "speedup": round(t_plain / t_batch, 2), "speedup_vs_serial": round(t_serial / t_batch, 2), "batched_mrays_s": round(rays / t_batch / 1e6, 1),
            "graph": True, "tuned_blocks_per_cu": tune.blocks_per_cu, "tuned_groups": tune.groups}


def adjoint_bench(steps=3):
    """Inverse-rendering step on 128^3 voxels, 8 views of 512^2 at 16 spp,
    3 bounces: forward render vs forward + path-replay adjoint (after one
    untimed step; `losses` lists every step)."""
    import torch
    import ctamd
    from ctamd import adjoint as AD
    w = AD.AdjointWorkload()
    sc, alb, emi = AD.build_scene(w)
    acc = ctamd.Accel(max_instances=sc.info.instance_count + 8,
                      max_primitives=sc.info.primitive_count + 32 * sc.info.instance_count + 1024,
                      max_cube_lights=max(sc.info.cube_light_count, 1) + 8)
    AD._load_geometry(sc, acc)
    slots = AD.VoxelSlots(sc, acc)
    extent = w.grid * ctamd.CT_VOXEL_EXTENT
    views = AD.orbit_views(w.views, w.width, w.height, (0.0, 0.0, 0.0), 1.6 * extent, 0.4 * extent)
    r = AD.DiffRenderer(acc, views, w.spp, w.max_depth, w.target_seed)
    dev = torch.device("cuda")
    target = r.forward(torch.from_numpy(slots.scatter(alb)).to(dev), torch.from_numpy(slots.scatter(emi)).to(dev))
    a0, e0 = AD.initial_parameters(w, alb, emi)
    a = torch.from_numpy(slots.scatter(a0)).to(dev).requires_grad_(True)
    e = torch.from_numpy(slots.scatter(e0)).to(dev).requires_grad_(True)
    opt = torch.optim.Adam([{"params": [a], "lr": 0.02}, {"params": [e], "lr": 0.5}])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(steps):
        with torch.no_grad():
            r.forward(a, e)
    torch.cuda.synchronize()
    t_fwd = (time.perf_counter() - t1) / steps
    losses = []

    def adam_step():
        opt.zero_grad()
        loss = ((AD.render(r, a, e) - target) ** 2).mean()
        loss.backward()
        opt.step()
        losses.append(float(loss.detach()))

    adam_step()  # untimed: the process's first backward pays one-time autograd and allocator set-up
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(steps):
        adam_step()
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t1) / steps
    paths = w.views * w.width * w.height * w.spp
    return {"voxels": sc.info.primitive_count, "views": w.views, "resolution": [w.width, w.height], "spp": w.spp,
            "bounces": w.max_depth, "paths_per_render": paths, "forward_ms": round(1000 * t_fwd, 2),
            "adam_step_ms": round(1000 * t_step, 2), "step_over_forward": round(t_step / t_fwd, 2),
            "forward_mpaths_s": round(paths / t_fwd / 1e6, 1), "losses": [round(x, 6) for x in losses]}


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
    instance, task queuing) and update_scene (BLAS rebuild); then a
    bulk deletion of 1 % of the voxels of every BLAS marked from a list.
    Wall times of each call."""
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
    inst = acc.read_instances()
    rng = np.random.default_rng(3)
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
    res["bulk_1pct"] = {"voxels": vox, "instances": tasks, "mark_ms": round((t1 - t0) * 1e3, 3),
                        "process_ms": round((t2 - t1) * 1e3, 3), "update_scene_ms": round((t3 - t2) * 1e3, 3)}
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
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device copy (rocprof counter calibration)")
    args = ap.parse_args()
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
    acc = ctamd.Accel(max_instances=sc.info.instance_count + 8,
                      max_primitives=sc.info.primitive_count + 32 * sc.info.instance_count + 1024,
                      max_cube_lights=sc.info.cube_light_count + 8)
    ctx = ctamd.Ctx(w.width, w.height)
    t0 = time.perf_counter()
    ctamd.load_scene(sc, acc, ctx)
    load_s = time.perf_counter() - t0
    log(f"BLAS/TLAS build + upload: {load_s:.3f}s")
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
    # Algorithmic bytes per query: a shadow ray reads its
    # 36 B record, a path segment reads 32 B and writes a 16 B hit; every BVH2
    # node fetched is 64 B, every leaf 28 B (centre + id + instance record
    # share), every instance entry 64 B.
    n_any = ist.shadow_rays_ris + ist.shadow_rays_temporal + ist.shadow_rays_pt + ist.shadow_rays_spatial
    bytes_any = 36 * n_any + 64 * ist.nodes + 28 * ist.leaves + 64 * ist.instances
    bytes_seg = 48 * ist.bounce_rays + 64 * ist.pt_nodes + 28 * ist.pt_leaves + 64 * ist.pt_instances
    b_any = bytes_any / max(n_any, 1)
    b_seg = bytes_seg / max(ist.bounce_rays, 1)
    log(f"instrumented frame: {n_any} shadow rays ({ist.shadow_rays_ris} RIS, {ist.shadow_rays_temporal} temporal, "
        f"{ist.shadow_rays_pt} path reuse, {ist.shadow_rays_spatial} spatial) {b_any:.0f} B/ray; "
        f"{ist.bounce_rays} path segments {b_seg:.0f} B/segment; closest: {ist.closest_nodes} nodes")
    # Timed region: K frames as the library runs them (two halves of the
    # image on two streams), with only the frame's start/end events.
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
    ms_per_step = 1000.0 * wall / K
    mrays = rays / wall / 1e6
    # Kernel durations for the roofline: K more frames run whole on one
    # stream with events around every pass (outside the timed region), so
    # each traversal launch covers the whole frame's rays on its own.
    ctx.set_instrumentation(1)
    ms_any = ms_seg = 0.0
    n_any_t = n_seg_t = 0
    ms_pt, ms_whole = [], []
    for _ in range(K):
        ctx.render(acc, ctamd.frame_params(w, f))
        f += 1
        st = ctx.stats()
        ms_any += st.ms_trace_any
        ms_seg += st.ms_pt_segments
        n_any_t += st.shadow_rays_ris + st.shadow_rays_temporal + st.shadow_rays_pt + st.shadow_rays_spatial
        n_seg_t += st.bounce_rays
        ms_pt.append(st.ms_pt)
        ms_whole.append(st.ms_total)
    kern = {
        "k_trace_any": {"desc": "k_trace_any (shadow-ray any-hit traversal)", "ms": ms_any,
                        "achieved": b_any * n_any_t / (ms_any / 1000.0) / 1e9 if ms_any > 0 else 0.0},
        "k_trace_segments": {"desc": "k_trace_segments (path-segment closest-hit traversal)", "ms": ms_seg,
                             "achieved": b_seg * n_seg_t / (ms_seg / 1000.0) / 1e9 if ms_seg > 0 else 0.0},
    }
    order = sorted(kern, key=lambda k: -kern[k]["ms"])
    # HBM bytes per launch from the committed FETCH_SIZE/WRITE_SIZE counter
    # runs of this workload (rocprofv3 cannot run inside the timed process);
    # the algorithmic bytes per launch of this run are reported beside them.
    measured = _traffic_summary() if (args.width, args.height, args.grid, args.depth) == (1920, 1080, 256, 3) else {}
    # any-hit launches per frame: RIS and temporal visibility, path-reuse visibility (with paths), spatial
    launches = {"k_trace_any": 3 if depth_of(w) else 2, "k_trace_segments": max(depth_of(w), 1)}
    algo = {"k_trace_any": b_any * n_any_t / K, "k_trace_segments": b_seg * n_seg_t / K}

    def roof(k):
        v = kern[k]
        return {"bound": "hbm", "achieved": round(v["achieved"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(v["achieved"] / HBM_PEAK_GBS, 4), "traffic": measured.get(k),
                "algorithmic_bytes_per_launch": round(algo[k] / launches[k]),
                "kernel": v["desc"], "ms_per_frame": round(v["ms"] / K, 3)}

    batched = adjoint = None
    if not args.no_batched:
        batched = batched_bench()
        log(f"batched: {batched}")
    if not args.no_adjoint:
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
    ms_sorted = sorted(ms_frames)
    depth = w.restir.max_depth if w.restir.indirect else 0
    out = {
        "metric": "Mrays/s + 1080p frame ms (ReSTIR DI/PT), 1 MI355X; batched/adjoint vs plain",
        "value": round(mrays, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": K, "warmup": Wm,
        "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"restir_di_pt{depth}_{w.width}x{w.height}_grid{args.grid}_64blas_taa",
                   "width": w.width, "height": w.height, "grid": args.grid, "blas": sc.info.instance_count,
                   "voxels": sc.info.primitive_count, "lights": sc.info.cube_light_count,
                   "frame_ms_gpu_events_median": round(ms_sorted[len(ms_sorted) // 2], 3),
                   "frame_ms_one_stream_median": round(sorted(ms_whole)[len(ms_whole) // 2], 3),
                   "pt_ms_gpu_events_median": round(sorted(ms_pt)[len(ms_pt) // 2], 3),
                   "rays_per_frame": rays // K, "indirect_bounces": depth},
        "roofline": roof(order[0]),
        "roofline_next_kernel": roof(order[1]),
        "cpu_baseline": cpu,
        "batched": batched,
        "adjoint": adjoint,
        "build": build,
        "vox": vox,
        "path_trace": ptrace,
        "brush": brush,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
