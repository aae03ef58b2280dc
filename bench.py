#!/usr/bin/env python3
"""Primary-ray throughput of the HIP raymarch on one MI355X.

A step renders every view of the workload once (one plain launch per view,
primary rays generated on the device, outputs normal + t + leaf index written
to HBM). Inputs (the octree) are resident in HBM before the timed region.

Workloads:
  ico2048 (default) 2048^3 octree built on the GPU from a seeded fBm-displaced
          icosphere (frequency 316 -> 1,997,120 triangles, displacement 5% of
          the radius, 8 threads' sub-block partition, 1 GiB budget), 8 seeded
          orbit cameras at 1920x1080 (16,588,800 rays per step).
  dragon  the reference's 256^3 dragon octree, same 8 cameras at 1920x1080.

Prints one JSON line (rank 0). The CPU baseline is the oracle (a plain-C
restatement of the reference raymarch) on the host cores, on the same rays.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "primary Grays/s @2048^3 1080p; batched-vs-plain rays/s + p99 latency, 1 GPU"
HBM_PEAK_GBS = 8000.0


def _pkg():
    from __graft_entry__ import _import_package
    return _import_package()


def progress(msg):
    """A line on stderr per finished stage (long runs show they are moving)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def profiled_traffic(workload, width, height):
    """HBM bytes per primary launch measured by rocprofv3 (FETCH_SIZE and
    WRITE_SIZE, calibrated on a 1 GiB copy) for this workload, from the
    committed profile summary; None when no profile covers it."""
    path = os.path.join(ROOT, "profiles", "raymarch_primary_ico2048_1080p_pmc.txt")
    if workload != "icosphere2048_8x1080p" or (width, height) != (1920, 1080) or not os.path.exists(path):
        return None
    for line in open(path):
        if line.strip().startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None



def build_roofline(st):
    """Construction against HBM bandwidth: algorithmic bytes 36*T + 24*K + 4*W
    over the device phase of the build (upload through last scatter), for a
    second build in the process (the first also loads the builder's kernels)."""
    sec = st["device_ms"] / 1e3
    achieved = st["algorithmic_bytes"] / sec / 1e9 if sec > 0 else 0.0
    return {"bound": "hbm", "device_ms": round(st["device_ms"], 2), "triangles": st["triangles"],
            "candidates": st["candidates"], "leaves": st["leaves"], "words": st["words"],
            "algorithmic_bytes": st["algorithmic_bytes"], "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
            "leaf_voxels_per_s": round(st["leaves"] / sec, 1) if sec > 0 else None}


def make_workload(pkg, name, cache_dir, no_cpu=False):
    if name == "dragon":
        tree = pkg.VoxelOctree(os.path.join(ROOT, "tests", "golden", "dragon.oct"))
        return tree, {"workload": "dragon256_8x1080p", "octree": "reference dragon 256^3",
                      "resolution": 256}
    if name == "ico2048":
        ply = os.path.join(cache_dir, "ico316_seed2048.ply")
        if not os.path.exists(ply):
            os.makedirs(cache_dir, exist_ok=True)
            v, f = pkg.scenes.icosphere(316, seed=2048, displacement=0.05)
            pkg.scenes.write_ply(ply, v, f)
        # the library's kernels loaded up front (svo_load_device_code), so the
        # first build does not pay the runtime's code loading
        t0 = time.time()
        pkg.svo.load_device_code()
        load_s = time.time() - t0
        t0 = time.time()
        tree = pkg.VoxelOctree.build_from_ply(ply, 2048, threads=8, memory_budget=1 << 30)
        build_s = time.time() - t0
        # a second build in the same process: the first one also pays the HIP
        # runtime's one-time loading of every builder kernel
        t0 = time.time()
        warm = pkg.VoxelOctree.build_from_ply(ply, 2048, threads=8, memory_budget=1 << 30)
        warm_s = time.time() - t0
        bst = warm.build_stats()
        del warm
        # the built tree must be the reference builder's, byte for byte
        import hashlib
        with open(os.path.join(ROOT, "tests", "golden", "builds.json")) as fp:
            ref = json.load(fp)["ico2048"]
        md5 = hashlib.md5(tree.encode()).hexdigest()
        # the reference's own builder on this host, same PLY, partition (8 threads) and budget
        model, ncpu = host_cpu()
        rb = reference_build(ply, 2048, 8, 1 << 30, cache_dir) if not no_cpu else None
        build_cpu = None
        if rb is not None:
            total = rb["parse"] + rb["build"] + rb["save"]
            build_cpu = {"value": round(ref["leaves"] / (rb["parse"] + rb["build"]) / 1e6, 3), "unit": "M leaf voxels/s",
                         "cores": rb["threads"], "kind": "reference", "cpu_model": model, "host_cpu_count": ncpu,
                         "seconds_without_save": round(rb["parse"] + rb["build"], 3),
                         "seconds_with_lz4_save": round(total, 3),
                         "phases_s": {"ply_parse": rb["parse"], "voxelize_and_build": rb["build"], "lz4_save": rb["save"]},
                         "oct_md5_equals_gpu_build": rb["md5"] == md5,
                         "sample": "the whole 2048^3 build: oracle/_ref/svo_ref build-ply (the reference's PlyLoader, "
                                   "VoxelData, VoxelOctree, save) with the 8-thread partition and a 1 GiB budget",
                         "gpu_speedup_without_save": round((rb["parse"] + rb["build"]) / warm_s, 1)}
        return tree, {"workload": "icosphere2048_8x1080p", "octree": "GPU-built 2048^3 displaced icosphere",
                      "triangles": 1997120, "resolution": 2048, "device_code_load_seconds": round(load_s, 3),
                      "build_seconds_first_build": round(build_s, 3),
                      "build_seconds": round(warm_s, 3),
                      "build_leaf_voxels_per_s": round(ref["leaves"] / warm_s, 1),
                      "build_note": "PLY file -> octree words in host memory, second build of the process",
                      "build_roofline": build_roofline(bst),
                      "build_oct_md5_matches_reference": md5 == ref["md5"],
                      "build_cpu_baseline": build_cpu}
    raise SystemExit(f"unknown workload {name}")


def run_mesh_big(pkg, cache_dir):
    """The headline mesh built at 4096^3 and 8192^3 (8 threads' partition,
    1 GiB budget): end-to-end seconds of a second build in the process, the
    device phase, the construction roofline and the .oct md5 against the
    reference builder's (tests/golden/builds.json)."""
    import hashlib
    with open(os.path.join(ROOT, "tests", "golden", "builds.json")) as fp:
        golden = json.load(fp)
    ply = os.path.join(cache_dir, "ico316_seed2048.ply")
    if not os.path.exists(ply):
        os.makedirs(cache_dir, exist_ok=True)
        v, f = pkg.scenes.icosphere(316, seed=2048, displacement=0.05)
        pkg.scenes.write_ply(ply, v, f)
    pkg.svo.load_device_code()
    out = {}
    for key in ("ico4096", "ico8192"):
        g = golden.get(key)
        if g is None:
            continue
        res = g["resolution"]
        t0 = time.time()
        tree = pkg.VoxelOctree.build_from_ply(ply, res, threads=8, memory_budget=1 << 30)
        first_s = time.time() - t0
        md5 = hashlib.md5(tree.encode()).hexdigest()
        del tree
        t0 = time.time()
        warm = pkg.VoxelOctree.build_from_ply(ply, res, threads=8, memory_budget=1 << 30)
        warm_s = time.time() - t0
        st = warm.build_stats()
        del warm
        out[key] = {"resolution": res, "build_seconds": round(warm_s, 3), "build_seconds_first_build": round(first_s, 3),
                    "leaves": st["leaves"], "words": st["words"], "candidates": st["candidates"],
                    "peak_bytes": st["peak_bytes"], "build_roofline": build_roofline(st),
                    "leaf_voxels_per_s": round(st["leaves"] / warm_s, 1),
                    "oct_md5_matches_reference": md5 == g["md5"]}
        progress(f"mesh build {res}^3 done")
    return out


def host_cpu():
    """(model name, os.cpu_count()) of this host."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


def reference_build(ply, resolution, threads, mem, cache_dir):
    """The reference builder's in-memory mode (oracle/_ref/svo_ref build-ply:
    PlyLoader -> VoxelData -> VoxelOctree -> save, src/Main.cpp:322-327) on
    `threads` host threads, each phase timed: {"parse", "build", "save",
    "md5", "threads"}, or None when that build is not present."""
    import hashlib
    import subprocess
    exe = os.path.join(ROOT, "oracle", "_ref", "svo_ref")
    if not os.access(exe, os.X_OK):
        return None
    os.makedirs(cache_dir, exist_ok=True)
    out = os.path.join(cache_dir, "ref_build.oct")
    try:
        r = subprocess.run([exe, "build-ply", ply, str(resolution), str(threads), str(mem), out], capture_output=True,
                           text=True, timeout=900)
    except (OSError, subprocess.SubprocessError):
        return None
    for line in r.stdout.splitlines():
        if line.startswith("build_seconds"):
            f = line.split()
            return {"parse": float(f[2]), "build": float(f[4]), "save": float(f[6]), "threads": int(f[8]),
                    "md5": hashlib.md5(open(out, "rb").read()).hexdigest()}
    return None


def reference_rays(tree, rays7, threads, cache_dir):
    """The reference's own VoxelOctree::raymarch (oracle/_ref/svo_ref, built
    from the reference sources) on rays (n x {o, d, rayScale}) with `threads`
    host threads: {"rate", "rays", "threads", "out" (hit, normal, t bits;
    0 where unwritten)}, or None when that build is not present."""
    import subprocess
    exe = os.path.join(ROOT, "oracle", "_ref", "svo_ref")
    if not os.access(exe, os.X_OK):
        return None
    os.makedirs(cache_dir, exist_ok=True)
    oct_path = os.path.join(cache_dir, "bench_tree.oct")
    rays_path = os.path.join(cache_dir, "bench_rays.f32")
    out_path = os.path.join(cache_dir, "bench_rays.out")
    tree.save(oct_path)
    np.ascontiguousarray(rays7, np.float32).tofile(rays_path)
    try:
        r = subprocess.run([exe, "raymarch", oct_path, rays_path, out_path, str(threads)], capture_output=True,
                           text=True, timeout=900)
    except (OSError, subprocess.SubprocessError):
        return None
    for line in r.stdout.splitlines():
        if line.startswith("raymarch_seconds"):
            f = line.split()
            out = np.fromfile(out_path, np.uint32).reshape(-1, 3)
            return {"rate": int(f[3]) / float(f[1]) / 1e9, "rays": int(f[3]), "threads": int(f[5]), "out": out}
    return None


def reference_cpu_rate(tree, cam, width, height, threads, cache_dir):
    """reference_rays on one view's primary rays."""
    import oracle_py as orc
    d = orc.primary_dirs(np.frombuffer(cam.tform, np.float32), width, height)
    rays = np.zeros((len(d), 7), np.float32)
    rays[:, 0:3] = np.frombuffer(cam.pos, np.float32)
    rays[:, 3:6] = d
    return reference_rays(tree, rays, threads, cache_dir)


def run_mixed(pkg, torch, stream, steps, warmup, waves_per_cu, refill):
    """Mixed batch: 64 GPU-built octrees (128^3..2048^3, SDF scenes and
    displaced meshes) and 256 views (320x240..1920x1080): 256 plain per-view
    launches vs one batched launch (form picked by the library's autotuner), same rays."""
    t0 = time.time()
    trees = [pkg.scenes.build_mixed_octree(pkg, sp) for sp in pkg.scenes.mixed_octrees()]
    build_s = time.time() - t0
    for t in trees:
        t.upload()
    specs = pkg.scenes.mixed_views()
    dev = torch.device("cuda", 0)
    sptr = stream.cuda_stream
    outs, cams = [], []
    for v in specs:
        n = v["width"] * v["height"]
        outs.append((torch.empty(n, dtype=torch.int32, device=dev), torch.empty(n, dtype=torch.float32, device=dev),
                     torch.empty(n, dtype=torch.int64, device=dev)))
        cams.append(trees[v["tree"]].orbit_camera(v["pitch"], v["yaw"], 1.0))
    rays = sum(v["width"] * v["height"] for v in specs)
    ren = pkg.svo.Renderer(len(specs))
    bviews = [ren.view(trees[v["tree"]], cams[i], v["width"], v["height"], None, outs[i][0].data_ptr(),
                       outs[i][1].data_ptr(), outs[i][2].data_ptr()) for i, v in enumerate(specs)]

    def plain():
        for i, v in enumerate(specs):
            trees[v["tree"]].render_primary_device(cams[i], v["width"], v["height"], None, outs[i][0].data_ptr(),
                                                   outs[i][1].data_ptr(), outs[i][2].data_ptr(), sptr)

    def batched(w=waves_per_cu, r=refill, f=None, tw=0):
        ren.render_device(bviews, sptr, w, r, f, tw)

    def timed(fn, k):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(k):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / k

    # launch form and shape: the library's autotuner on this workload
    tune = {}
    form = None
    tile_w = 0
    if not waves_per_cu and not refill:
        (form, waves_per_cu, refill, tile_w), tune = ren.autotune(bviews, sptr)
    plain_ms = timed(plain, steps)
    batched_ms = timed(lambda: batched(waves_per_cu, refill, form, tile_w), steps)
    # the same launch captured once in a HIP graph and replayed
    graph = ren.capture(bviews, sptr, waves_per_cu, refill, form, tile_w)
    graph_ms = timed(lambda: graph.launch(sptr), steps)
    # the grid form with the tiles of all 256 views in a batch plan's cost order
    planned = None
    if form == pkg.svo.BATCH_GRID:
        bplan = pkg.svo.TilePlan.for_batch(bviews, tile_w=tile_w)
        planned_ms = timed(lambda: ren.render_device(bviews, sptr, 0, 0, form, tile_w, plan=bplan), steps)
        planned = {"grays_s": round(rays / (planned_ms / 1e3) / 1e9, 4), "ms_per_step": round(planned_ms, 3),
                   "vs_plain": round(plain_ms / planned_ms, 3),
                   "form": "the autotuned grid launch through a batch tile plan (svo_render_batch_planned_device)"}
    # parity: the graph-replayed batched launch reproduces the plain launches bit for bit
    plain()
    torch.cuda.synchronize()
    snap = [(a.clone(), b.clone(), c.clone()) for a, b, c in outs]
    for a, b, c in outs:
        a.fill_(-1)
        b.fill_(-1.0)
        c.fill_(-1)
    graph.launch(sptr)
    torch.cuda.synchronize()
    graph.close()
    mism = sum(int((a != sa).sum() + (b.view(torch.int32) != sb.view(torch.int32)).sum() + (c != sc).sum())
               for (a, b, c), (sa, sb, sc) in zip(outs, snap))
    if planned is not None:
        for a, b, c in outs:
            a.fill_(-1)
            b.fill_(-1.0)
            c.fill_(-1)
        ren.render_device(bviews, sptr, 0, 0, form, tile_w, plan=bplan)
        torch.cuda.synchronize()
        planned["mismatched_fields_vs_plain"] = sum(
            int((a != sa).sum() + (b.view(torch.int32) != sb.view(torch.int32)).sum() + (c != sc).sum())
            for (a, b, c), (sa, sb, sc) in zip(outs, snap))
        mism += planned["mismatched_fields_vs_plain"]
        bplan.close()
    run_mixed.trees = trees
    return {"workload": "mixed_64octrees_256views", "rays_per_step": rays, "octree_build_seconds": round(build_s, 2),
            "plain_grays_s": round(rays / (plain_ms / 1e3) / 1e9, 4), "plain_ms_per_step": round(plain_ms, 3),
            "batched_grays_s": round(rays / (batched_ms / 1e3) / 1e9, 4), "batched_ms_per_step": round(batched_ms, 3),
            "batched_vs_plain": round(plain_ms / batched_ms, 3),
            "graph_replay": {"grays_s": round(rays / (graph_ms / 1e3) / 1e9, 4), "ms_per_step": round(graph_ms, 3),
                             "vs_plain": round(plain_ms / graph_ms, 3),
                             "form": "the autotuned batched launch captured in one HIP graph, replayed per step"},
            "planned": planned,
            "launch": {"form": "grid" if form == pkg.svo.BATCH_GRID else "persistent", "waves_per_cu": waves_per_cu,
                       "refill": refill, "tile": f"{tile_w or 8}x{64 // (tile_w or 8)}", "autotune_ms": tune},
            "mismatched_fields_vs_plain": mism}


def run_terrain(pkg, torch, stream, steps, warmup, cpu_threads, cache_dir, n=8192, seed=8192, width=3840, height=2160):
    """Large terrain: the 8192^3 exposed shell of a seeded fBm heightfield,
    built on the GPU column by column (512^3 cache blocks under a 1 GiB
    budget), then 3840x2160 primary rays plus one hard-shadow ray per hit
    toward a seeded directional light. A step is one primary pass and one
    shadow pass; rays counted = primary rays + shadow rays traced."""
    import oracle_py as orc
    torch.cuda.synchronize()
    t0 = time.time()
    tree = pkg.VoxelOctree.build_from_terrain(n, seed, memory_budget=1 << 30)
    build_s = time.time() - t0
    t0 = time.time()
    warm = pkg.VoxelOctree.build_from_terrain(n, seed, memory_budget=1 << 30)
    warm_s = time.time() - t0
    warm_stats = warm.build_stats()
    del warm
    tree.upload()
    words = tree.words()
    st = orc.octree_stats(words)
    pitch, yaw, light = pkg.scenes.terrain_view(seed)
    cam = tree.orbit_camera(pitch, yaw, 1.0)
    eps = 0.25 / n
    m = width * height
    dev = torch.device("cuda", 0)
    hit = torch.empty(m, dtype=torch.uint8, device=dev)
    nrm = torch.empty(m, dtype=torch.int32, device=dev)
    tt = torch.empty(m, dtype=torch.float32, device=dev)
    lf = torch.empty(m, dtype=torch.int64, device=dev)
    sh = torch.empty(m, dtype=torch.uint8, device=dev)
    sptr = stream.cuda_stream

    def step():
        tree.render_shadow_device(cam, width, height, light, eps, hit.data_ptr(), tt.data_ptr(), sh.data_ptr(),
                                  nrm.data_ptr(), lf.data_ptr(), sptr)

    def primary_only():
        tree.render_primary_device(cam, width, height, hit.data_ptr(), nrm.data_ptr(), tt.data_ptr(), lf.data_ptr(),
                                   sptr)

    def timed(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(steps):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / steps

    prim_ms = timed(primary_only)
    # the same step through a tile plan (each launch starts the tiles the
    # previous one found slowest, primary plus shadow passes); its outputs
    # must equal the unplanned step's
    plan = pkg.svo.TilePlan(width, height)
    planned_out = [torch.empty_like(x) for x in (hit, nrm, tt, lf, sh)]

    def planned_step():
        ph, pn, pt, pl, ps = planned_out
        tree.render_shadow_device(cam, width, height, light, eps, ph.data_ptr(), pt.data_ptr(), ps.data_ptr(),
                                  pn.data_ptr(), pl.data_ptr(), sptr, plan=plan)

    planned_ms = timed(planned_step)
    ms = timed(step)
    planned_mism = sum(int(torch.count_nonzero(a.view(torch.uint8) != b.view(torch.uint8)))
                       for a, b in zip(planned_out, (hit, nrm, tt, lf, sh)))
    plan.close()
    g_hit = hit.cpu().numpy()
    g_sh = sh.cpu().numpy()
    shadow_rays = int((g_hit == 1).sum())
    rays = m + shadow_rays
    # parity and CPU baseline on a sample of pixels: the oracle's primary and shadow rays
    rng = np.random.default_rng(7)
    sample = np.sort(rng.choice(m, 200000, replace=False))
    tf = np.frombuffer(cam.tform, np.float32)
    pos = np.frombuffer(cam.pos, np.float32).copy()
    d = orc.primary_dirs(tf, width, height)[sample]
    c0 = time.perf_counter()
    prim = orc.raymarch(words, pos[None, :], d, 0.0, threads=cpu_threads)
    hits = prim["hit"] == 1
    back = (prim["t"][hits] - np.float32(eps)).astype(np.float32)
    o = (pos[None, :] + d[hits] * back[:, None]).astype(np.float32)
    shr = orc.raymarch(words, o, np.broadcast_to(light, o.shape).copy(), 0.0, threads=cpu_threads)
    cpu_s = time.perf_counter() - c0
    ref_sh = np.zeros(len(sample), np.uint8)
    ref_sh[hits] = (shr["hit"] != 0).astype(np.uint8)
    g_t = tt.cpu().numpy()[sample]
    g_n = nrm.cpu().numpy().view(np.uint32)[sample]
    mism = int(np.count_nonzero(g_hit[sample] != prim["hit"]) + np.count_nonzero(g_n != prim["normal"])
               + np.count_nonzero(g_t.view(np.uint32) != prim["t"].view(np.uint32))
               + np.count_nonzero(g_sh[sample] != ref_sh))
    cpu_rays = len(sample) + int(hits.sum())
    counts = prim["counts"] + shr["counts"]
    b_ray = 4.0 * float(counts.sum()) / cpu_rays + 16.0
    cpu_baseline = {"value": round(cpu_rays / cpu_s / 1e9, 6), "unit": "Grays/s", "cores": cpu_threads, "kind": "port",
                    "sample": f"{len(sample)} sampled pixels' primary rays + their {int(hits.sum())} shadow rays "
                              f"through the oracle"}
    # the same rays through the reference's own raymarch, when its build is present
    rays7 = np.zeros((cpu_rays, 7), np.float32)
    rays7[:len(sample), 0:3] = pos
    rays7[:len(sample), 3:6] = d
    rays7[len(sample):, 0:3] = o
    rays7[len(sample):, 3:6] = light
    ref = reference_rays(tree, rays7, cpu_threads, cache_dir)
    ref_parity = None
    if ref is not None:
        ro = ref["out"]
        ref_parity = {"rays": cpu_rays, "mismatched_fields": int(
            np.count_nonzero(g_n != ro[:len(sample), 1]) + np.count_nonzero(g_t.view(np.uint32) != ro[:len(sample), 2])
            + np.count_nonzero(g_sh[sample][hits] != (ro[len(sample):, 0] != 0)))}
        mism += ref_parity["mismatched_fields"]
        cpu_baseline = {"value": round(ref["rate"], 6), "unit": "Grays/s", "cores": ref["threads"], "kind": "reference",
                        "sample": f"{len(sample)} sampled pixels' primary rays + their {int(hits.sum())} shadow rays "
                                  f"through the reference's own VoxelOctree::raymarch (oracle/_ref/svo_ref, "
                                  f"{ref['threads']} threads); the oracle port: {cpu_rays / cpu_s / 1e9:.6f} Grays/s"}
    return {"workload": "terrain8192_4k_shadow", "octree": f"GPU-built {n}^3 fBm heightfield shell (seed {seed})",
            "build_seconds": round(build_s, 3), "leaf_voxels": int(st["leaves"]), "words": int(len(words)),
            "build_leaf_voxels_per_s": round(st["leaves"] / build_s, 1),
            "build_seconds_second_build": round(warm_s, 3),
            "build_roofline": build_roofline(warm_stats),
            "width": width, "height": height, "primary_rays": m, "shadow_rays": shadow_rays,
            "grays_s": round(rays / (ms / 1e3) / 1e9, 4), "ms_per_step": round(ms, 3),
            "primary_only_grays_s": round(m / (prim_ms / 1e3) / 1e9, 4), "primary_only_ms": round(prim_ms, 3),
            "shadow_pass_ms": round(ms - prim_ms, 3),
            "planned": {"grays_s": round(rays / (planned_ms / 1e3) / 1e9, 4), "ms_per_step": round(planned_ms, 3),
                        "mismatched_bytes_vs_unplanned": planned_mism,
                        "form": "svo_render_shadow_planned_device: the same launch through a tile plan"},
            "bytes_per_ray": round(b_ray, 2),
            "hbm_frac": round(b_ray * rays / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 5),
            "cpu_baseline": cpu_baseline,
            "parity": {"rays_checked": cpu_rays, "mismatched_fields": mism, "vs_reference_binary": ref_parity}}


def run_frames(pkg, torch, stream, tree, cams, width, height, steps, warmup, cpu_check=True, angles=None,
               threads=16, cache_dir=None):
    """Viewer frames (src/Main.cpp renderBatch/renderTile/shade): per frame a
    coarse LOD pass over 8x8-tile corners and a fine pass over tiles that may
    hit, shaded to RGBA, 16 stripes, stride 1. One step renders every camera's
    frame once; the rate counts coarse + traced fine rays."""
    import oracle_py as orc
    dev = torch.device("cuda", 0)
    m = width * height
    nd = pkg.svo.frame_depth_count(width, height, 16)
    rgba = [torch.empty(m, dtype=torch.int32, device=dev) for _ in cams]
    depth = [torch.empty(nd, dtype=torch.float32, device=dev) for _ in cams]
    sptr = stream.cuda_stream

    def step():
        for k, c in enumerate(cams):
            tree.render_frame_device(c, width, height, rgba[k].data_ptr(), depth[k].data_ptr(), 16, 1, sptr)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(steps):
        step()
    b.record(stream)
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / steps
    # rays traced: every coarse corner, plus every pixel of a traced tile (those
    # carry an opaque colour; pixels of skipped tiles stay 0)
    traced = 0
    for k in range(len(cams)):
        img = rgba[k].cpu().numpy().view(np.uint32)
        traced += nd + int(np.count_nonzero(img))
    out = {"workload": f"viewer_frames_{width}x{height}", "frames_per_step": len(cams),
           "ms_per_frame": round(ms / len(cams), 4), "frames_per_s": round(len(cams) / (ms / 1e3), 1),
           "rays_per_step": traced, "grays_s": round(traced / (ms / 1e3) / 1e9, 4),
           "fine_pixels_traced_frac": round((traced - nd * len(cams)) / (m * len(cams)), 4)}
    # The same frames in one svo_render_frames_device call per step: up to 16
    # frames share one coarse launch and one fine launch on this one stream, so
    # a frame's coarse pass (a few hundred waves) and its fine pass's tail run
    # beside the other frames' work. Images and depth tables must equal the
    # one-frame-per-call ones.
    rgba2 = [torch.empty_like(x) for x in rgba]
    depth2 = [torch.empty_like(x) for x in depth]
    rptrs, dptrs = [x.data_ptr() for x in rgba2], [x.data_ptr() for x in depth2]

    def step2():
        tree.render_frames_device(cams, width, height, rptrs, dptrs, 16, 1, sptr)

    for _ in range(max(warmup, 1)):
        step2()
    torch.cuda.synchronize()
    a.record(stream)
    for _ in range(steps):
        step2()
    b.record(stream)
    torch.cuda.synchronize()
    ms2 = a.elapsed_time(b) / steps
    diff = sum(int(torch.count_nonzero(rgba2[k] != rgba[k])) + int(torch.count_nonzero(
        depth2[k].view(torch.int32) != depth[k].view(torch.int32))) for k in range(len(cams)))
    out["frames_call"] = {"ms_per_frame": round(ms2 / len(cams), 4), "frames_per_s": round(len(cams) / (ms2 / 1e3), 1),
                          "vs_one_frame_per_call": round(ms / ms2, 3), "mismatched_vs_one_frame_per_call": diff,
                          "form": f"one svo_render_frames_device call for the {len(cams)} frames"}
    # One frame per call through a tile plan per camera (a viewer's next frame
    # of the same view): the fine pass starts the tiles the previous frame
    # found slowest first. Same images and depth tables.
    plans = [pkg.svo.TilePlan.for_frame(width, height, 16) for _ in cams]
    rgba3 = [torch.empty_like(x) for x in rgba]
    depth3 = [torch.empty_like(x) for x in depth]

    def step3():
        for k, c in enumerate(cams):
            tree.render_frame_device(c, width, height, rgba3[k].data_ptr(), depth3[k].data_ptr(), 16, 1, sptr,
                                     plan=plans[k])

    for _ in range(max(warmup, 1)):
        step3()
    torch.cuda.synchronize()
    a.record(stream)
    for _ in range(steps):
        step3()
    b.record(stream)
    torch.cuda.synchronize()
    ms3 = a.elapsed_time(b) / steps
    diff3 = sum(int(torch.count_nonzero(rgba3[k] != rgba[k])) + int(torch.count_nonzero(
        depth3[k].view(torch.int32) != depth[k].view(torch.int32))) for k in range(len(cams)))
    # the frames call through a plan made for its frame count (the fine pass
    # over all frames' tiles in the previous call's cost order)
    fplan = pkg.svo.TilePlan.for_frames(width, height, len(cams), 16) if len(cams) <= 16 else None
    if fplan is not None:
        rgba4 = [torch.empty_like(x) for x in rgba]
        depth4 = [torch.empty_like(x) for x in depth]
        rptrs4, dptrs4 = [x.data_ptr() for x in rgba4], [x.data_ptr() for x in depth4]
        for _ in range(max(warmup, 1)):
            tree.render_frames_device(cams, width, height, rptrs4, dptrs4, 16, 1, sptr, plan=fplan)
        torch.cuda.synchronize()
        a.record(stream)
        for _ in range(steps):
            tree.render_frames_device(cams, width, height, rptrs4, dptrs4, 16, 1, sptr, plan=fplan)
        b.record(stream)
        torch.cuda.synchronize()
        ms4 = a.elapsed_time(b) / steps
        diff4 = sum(int(torch.count_nonzero(rgba4[k] != rgba[k])) + int(torch.count_nonzero(
            depth4[k].view(torch.int32) != depth[k].view(torch.int32))) for k in range(len(cams)))
        out["frames_call"]["planned"] = {"ms_per_frame": round(ms4 / len(cams), 4),
                                         "frames_per_s": round(len(cams) / (ms4 / 1e3), 1),
                                         "mismatched_vs_one_frame_per_call": diff4}
        fplan.close()
    out["planned"] = {"ms_per_frame": round(ms3 / len(cams), 4), "frames_per_s": round(len(cams) / (ms3 / 1e3), 1),
                      "vs_one_frame_per_call": round(ms / ms3, 3), "mismatched_vs_one_frame_per_call": diff3,
                      "form": "one frame per call, the fine pass in a tile plan's cost order (one plan per camera)"}
    for pl in plans:
        pl.close()
    if cpu_check:
        words = tree.words()
        tf = np.frombuffer(cams[0].tform, np.float32)
        c0 = time.perf_counter()
        ref_rgba, ref_depth = orc.render_frame(words, tf, tree.center(), width, height, 16, 1)
        cpu_s = time.perf_counter() - c0
        g_rgba = rgba[0].cpu().numpy().view(np.uint32).reshape(height, width)
        g_depth = depth[0].cpu().numpy().view(np.uint32)
        out["parity"] = {"frames_checked": 1, "mismatched_pixels": int(np.count_nonzero(g_rgba != ref_rgba)),
                         "mismatched_depths": int(np.count_nonzero(g_depth != ref_depth.view(np.uint32)))}
        out["cpu_baseline"] = {"value": round(1.0 / cpu_s, 3), "unit": "frames/s", "cores": 1, "kind": "port",
                               "sample": f"one {width}x{height} frame through the oracle's frame loop (one thread; "
                                         f"the reference viewer splits it over 16 threads)"}
        ref = reference_frame(tree, angles[0] if angles else None, width, height, threads, cache_dir)
        if ref is not None:
            nd1 = len(ref_depth)
            out["parity"]["vs_reference_classes"] = {
                "mismatched_pixels": int(np.count_nonzero(g_rgba.reshape(-1) != ref["out"][:m])),
                "mismatched_depths": int(np.count_nonzero(g_depth != ref["out"][m:m + nd1]))}
            out["cpu_baseline"] = {"value": round(1.0 / ref["seconds"], 3), "unit": "frames/s", "cores": threads,
                                   "kind": "reference",
                                   "sample": f"one {width}x{height} frame: the viewer's renderBatch/renderTile/shade "
                                             f"loop on the reference's own VoxelOctree::raymarch, Mat4, Vec3 "
                                             f"(oracle/_ref/svo_ref frame; src/Main.cpp needs SDL), one thread per "
                                             f"stripe on {threads} threads; the oracle port on one thread: "
                                             f"{1.0 / cpu_s:.3f} frames/s"}
    return out


def reference_frame(tree, angle, width, height, threads, cache_dir):
    """oracle/_ref/svo_ref's frame loop (the reference's classes) on one
    orbit camera (pitch, yaw; radius 1): {"seconds", "out"} or None."""
    import subprocess
    exe = os.path.join(ROOT, "oracle", "_ref", "svo_ref")
    if angle is None or cache_dir is None or not os.access(exe, os.X_OK):
        return None
    os.makedirs(cache_dir, exist_ok=True)
    oct_path = os.path.join(cache_dir, "bench_tree.oct")
    cam_path = os.path.join(cache_dir, "bench_cam.f32")
    out_path = os.path.join(cache_dir, "bench_frame.out")
    tree.save(oct_path)
    np.float32([[angle[0], angle[1], 1.0]]).tofile(cam_path)
    try:
        r = subprocess.run([exe, "frame", oct_path, cam_path, str(width), str(height), "1", out_path, str(threads)],
                           capture_output=True, text=True, timeout=600)
    except (OSError, subprocess.SubprocessError):
        return None
    for line in r.stdout.splitlines():
        if line.startswith("frame_seconds"):
            return {"seconds": float(line.split()[1]), "out": np.fromfile(out_path, np.uint32)}
    return None


def run_served(pkg, trees, rates, seconds, seed=2000, full=False, waves_per_cu=0):
    """Served stream over the 64 resident octrees: seeded Poisson request
    streams, each request (octree U{0..63}, seeded orbit camera, resolution as
    the mixed views). For each mean rate the live server (svo_server_*:
    requests run as their own launches while few are in flight, the others
    together in batched launches; outputs left in device memory), the same
    server with its persistent launch (requests beyond four in flight folded
    into one running launch) and the one-request-per-launch baseline (plain
    kernel, outputs in device memory) replay the same schedule; latency is
    scheduled arrival -> completion seen on the host. Requests per rate:
    min(5000, rate * seconds), or 5000 with `full`. A final burst (every
    request at t=0) measures saturated throughput."""
    rng = np.random.default_rng(seed)
    out = {"rates": [], "latency": "scheduled arrival -> completion seen by the host (outputs resident in HBM)"}

    def make(n):
        reqs = []
        for v in pkg.scenes.mixed_views(seed=int(rng.integers(1 << 30)), count=n, trees=len(trees)):
            t = trees[v["tree"]]
            reqs.append((v["tree"], t.orbit_camera(v["pitch"], v["yaw"], 1.0), v["width"], v["height"]))
        return reqs

    def stats(lat, mk, reqs):
        rays = sum(w * h for _, _, w, h in reqs)
        return {"p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
                "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
                "grays_s": round(rays / mk / 1e9, 4), "requests": len(reqs)}

    srv = pkg.svo.Server(trees, output_budget=4 << 30)
    # the same server with the persistent launch instead of batched launches
    srvp = pkg.svo.Server(trees, output_budget=4 << 30, waves_per_cu=waves_per_cu, persistent_launch=True)
    # first use of the serving kernels, outside the measurements
    warm = make(8)
    srv.run_schedule(warm, np.zeros(len(warm)))
    srvp.run_schedule(warm, np.zeros(len(warm)))
    for lam in rates:
        n = 5000 if full else int(max(200, min(5000, lam * seconds)))
        reqs = make(n)
        arr = np.cumsum(rng.exponential(1.0 / lam, n))
        arr -= arr[0]
        lat_s, mk_s, _ = srv.run_schedule(reqs, arr)
        progress(f"served {lam:g} req/s: live server, {n} requests")
        lat_b, mk_b, _ = srvp.run_schedule(reqs, arr)
        progress(f"served {lam:g} req/s: live server with the persistent launch")
        lat_p, mk_p = pkg.svo.serve_replay(trees, reqs, arr, streams=1, one_per_launch=True, device_outputs=True)
        progress(f"served {lam:g} req/s: one per launch")
        out["rates"].append({"req_per_s": lam, "live_server": stats(lat_s, mk_s, reqs),
                             "live_server_persistent_launch": stats(lat_b, mk_b, reqs),
                             "one_per_launch": stats(lat_p, mk_p, reqs)})
    reqs = make(1000)
    arr = np.zeros(len(reqs))
    lat_s, mk_s, _ = srv.run_schedule(reqs, arr)
    st = srv.stats()
    srv.close()
    lat_b, mk_b, _ = srvp.run_schedule(reqs, arr)
    stp = srvp.stats()
    srvp.close()
    lat_p, mk_p = pkg.svo.serve_replay(trees, reqs, arr, streams=1, one_per_launch=True, device_outputs=True)
    lat_r, mk_r = pkg.svo.serve_replay(trees, reqs, arr, streams=4, device_outputs=True)
    s_, p_, r_ = stats(lat_s, mk_s, reqs), stats(lat_p, mk_p, reqs), stats(lat_r, mk_r, reqs)
    out["burst"] = {"live_server": s_, "live_server_persistent_launch": stats(lat_b, mk_b, reqs),
                    "one_per_launch": p_, "relaunch_batched_4_streams": r_,
                    "live_vs_one_per_launch": round(s_["grays_s"] / p_["grays_s"], 3)}
    out["server"] = {"direct_requests": st["direct"], "batched_requests": st["batched"],
                     "persistent_launches": st["launches"]}
    out["persistent_server"] = {"direct_requests": stp["direct"], "persistent_launches": stp["launches"],
                                "waves_per_cu": waves_per_cu or 24}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="ico2048", choices=["ico2048", "dragon"])
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--views", type=int, default=8)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and parity leg (profiling runs)")
    ap.add_argument("--raster", action="store_true",
                    help="headline launches in raster tile order (no tile plans; the cost-ordered form is then the "
                         "comparison)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="run one plain device copy of this size first (HBM counter calibration)")
    ap.add_argument("--cache", default=os.path.join(os.environ.get("TMPDIR", "/tmp"), "svo_bench_cache"))
    ap.add_argument("--waves-per-cu", type=int, default=0, help="batched form: persistent waves per CU (0 = default)")
    ap.add_argument("--refill", type=int, default=0, help="batched form: idle lanes that trigger a refill (0 = default)")
    ap.add_argument("--no-mixed", action="store_true", help="skip the 64-octree / 256-view batched workload")
    ap.add_argument("--no-serve", action="store_true", help="skip the served (request stream) workload")
    ap.add_argument("--serve-rates", default="100,200,500,1000,2000", help="mean request rates (req/s) to replay")
    ap.add_argument("--serve-seconds", type=float, default=4.0, help="schedule length per rate (at most 5000 requests)")
    ap.add_argument("--serve-full", action="store_true", help="5000 requests at every rate")
    ap.add_argument("--serve-waves-per-cu", type=int, default=0,
                    help="live server with the persistent launch: wave slots per CU it holds (0 = the library's "
                         "default, 24)")
    ap.add_argument("--no-terrain", action="store_true", help="skip the 8192^3 terrain + shadow-ray workload")
    ap.add_argument("--no-frames", action="store_true", help="skip the viewer-frame workload")
    ap.add_argument("--only", choices=["terrain", "frames", "build", "meshbig"], default=None,
                    help="run just that workload and print its JSON (profiling runs); build: the headline "
                         "octree's two GPU builds from its PLY file; meshbig: the same mesh at 4096^3 and 8192^3")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: measures one GPU only (no multi-GPU path); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import torch

    pkg = _pkg()
    if pkg.svo.device_count() < 1 or not torch.cuda.is_available():
        raise SystemExit("bench.py: no GPU visible")
    torch.cuda.set_device(0)
    if args.only == "meshbig":
        print(json.dumps(run_mesh_big(pkg, args.cache)), flush=True)
        return
    if args.only == "terrain":
        print(json.dumps(run_terrain(pkg, torch, torch.cuda.current_stream(), args.steps, args.warmup,
                                     args.cpu_threads, args.cache)), flush=True)
        return
    tree, config = make_workload(pkg, args.workload, args.cache, no_cpu=args.no_cpu)
    if args.only == "build":
        print(json.dumps({k: v for k, v in config.items() if k.startswith(("workload", "octree", "triangles",
                                                                           "resolution", "device_code", "build"))}),
              flush=True)
        return
    tree.upload()
    W, H, V = args.width, args.height, args.views
    n = W * H
    angles = pkg.scenes.orbit_angles(1080, V)
    cams = [tree.orbit_camera(p, y, 1.0) for p, y in angles]

    if args.only == "frames":
        print(json.dumps(run_frames(pkg, torch, torch.cuda.current_stream(), tree, cams, W, H, args.steps,
                                    args.warmup, cpu_check=not args.no_cpu, angles=angles, threads=args.cpu_threads,
                                    cache_dir=args.cache)), flush=True)
        return
    dev = torch.device("cuda", 0)
    normal = [torch.empty(n, dtype=torch.int32, device=dev) for _ in range(V)]
    tbuf = [torch.empty(n, dtype=torch.float32, device=dev) for _ in range(V)]
    leaf = [torch.empty(n, dtype=torch.int64, device=dev) for _ in range(V)]
    stream = torch.cuda.current_stream()
    sptr = stream.cuda_stream

    # one tile plan per view: a view's render starts the tiles its previous
    # render found slowest first (the same rays, results and launch count)
    plans = [pkg.svo.TilePlan(W, H) for _ in range(V)]

    def step(events=None, raster=args.raster):
        for v in range(V):
            if events is not None:
                events[v][0].record(stream)
            if raster:
                tree.render_primary_device(cams[v], W, H, None, normal[v].data_ptr(), tbuf[v].data_ptr(),
                                           leaf[v].data_ptr(), sptr)
            else:
                plans[v].render_device(tree, cams[v], None, normal[v].data_ptr(), tbuf[v].data_ptr(),
                                       leaf[v].data_ptr(), sptr)
            if events is not None:
                events[v][1].record(stream)

    # batched form: every view of the step in one launch
    renderer = pkg.svo.Renderer(max(V, 1))
    bviews = [renderer.view(tree, cams[v], W, H, None, normal[v].data_ptr(), tbuf[v].data_ptr(), leaf[v].data_ptr())
              for v in range(V)]

    # the grid form through a batch tile plan: each launch starts the tiles the
    # previous one found slowest (over all views) first
    bplan = pkg.svo.TilePlan.for_batch(bviews) if not (args.waves_per_cu or args.refill) else None

    def batched_step(ordered=True):
        renderer.render_device(bviews, sptr, args.waves_per_cu, args.refill, plan=bplan if ordered else None)

    def time_graph():
        g = renderer.capture(bviews, sptr, args.waves_per_cu, args.refill)
        for _ in range(args.warmup):
            g.launch(sptr)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(args.steps):
            g.launch(sptr)
        b.record(stream)
        torch.cuda.synchronize()
        g.close()
        return a.elapsed_time(b) / args.steps

    def time_batched(ordered=True):
        for _ in range(args.warmup):
            batched_step(ordered)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(args.steps):
            batched_step(ordered)
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / args.steps

    if args.calibrate_copy_mib:
        a = torch.ones(args.calibrate_copy_mib << 18, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    ev = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(V)]
          for _ in range(args.steps)]
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record(stream)
    for k in range(args.steps):
        step(ev[k])
    stop.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    total_ms = start.elapsed_time(stop)
    launch_ms = [a.elapsed_time(b) for row in ev for a, b in row]
    avg_launch_s = float(np.mean(launch_ms)) / 1e3
    rays = n * V * args.steps
    value = rays / (total_ms / 1e3) / 1e9
    # the same views with every launch in raster tile order (no plan), for comparison
    other = not args.raster
    for _ in range(args.warmup):
        step(raster=other)
    ev2 = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(V)]
           for _ in range(args.steps)]
    torch.cuda.synchronize()
    start.record(stream)
    for k in range(args.steps):
        step(ev2[k], raster=other)
    stop.record(stream)
    torch.cuda.synchronize()
    other_ms = start.elapsed_time(stop)
    other_launch = float(np.mean([a.elapsed_time(b) for row in ev2 for a, b in row]))
    compare = {"order": "raster" if other else "cost-ordered", "value": round(rays / (other_ms / 1e3) / 1e9, 4),
               "ms_per_step": round(other_ms / args.steps, 4), "avg_launch_ms": round(other_launch, 4)}
    batched_ms = time_batched()
    batched_raster_ms = time_batched(False) if bplan is not None else None
    graph_ms = time_graph()
    batched = {"value": round(n * V / (batched_ms / 1e3) / 1e9, 4), "unit": "Grays/s",
               "ms_per_step": round(batched_ms, 4), "vs_plain": round((total_ms / args.steps) / batched_ms, 3),
               "form": (f"one launch of every tile of the {V} views (one wave per 8x8 tile), tiles in the cost order "
                        "of the previous launch (svo_render_batch_planned_device)"
                        if bplan is not None else
                        f"one persistent launch for all {V} views (waves/CU {args.waves_per_cu}, refill {args.refill})"),
               "graph_replay": {"value": round(n * V / (graph_ms / 1e3) / 1e9, 4), "ms_per_step": round(graph_ms, 4),
                                "order": "the views' tile order (a graph replays one launch's parameters)"}}
    if batched_raster_ms is not None:
        batched["other_order"] = {"order": "the views' tile order", "value": round(n * V / (batched_raster_ms / 1e3) / 1e9, 4),
                                  "ms_per_step": round(batched_raster_ms, 4)}

    if args.no_cpu:
        print(json.dumps({"metric": METRIC, "value": round(value, 4), "unit": "Grays/s", "n_gpus": 1,
                          "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(total_ms / args.steps, 4),
                          "avg_launch_ms": round(avg_launch_s * 1e3, 4), "config": config, "batched": batched,
                          "other_order": compare, "profiling_run": True}),
              flush=True)
        return

    mixed = None
    if not args.no_mixed:
        mixed = run_mixed(pkg, torch, stream, max(1, min(args.steps, 5)), 1, args.waves_per_cu, args.refill)
        progress("mixed 64-octree batch done")
    served = None
    if not args.no_mixed and not args.no_serve:
        served = run_served(pkg, run_mixed.trees, [float(x) for x in args.serve_rates.split(",")], args.serve_seconds,
                            full=args.serve_full, waves_per_cu=args.serve_waves_per_cu)
        progress("served stream done")

    frames = None
    if not args.no_frames:
        frames = run_frames(pkg, torch, stream, tree, cams, W, H, max(1, min(args.steps, 5)), 1, angles=angles,
                            threads=args.cpu_threads, cache_dir=args.cache)
        progress("viewer frames done")
    terrain = None
    if not args.no_terrain:
        terrain = run_terrain(pkg, torch, stream, max(1, min(args.steps, 5)), 1, args.cpu_threads, args.cache)
        progress("terrain done")

    # batched outputs must equal the plain kernel's: snapshot plain, run batched, compare
    step()
    torch.cuda.synchronize()
    snap = [(normal[v].clone(), tbuf[v].clone(), leaf[v].clone()) for v in range(V)]
    for v in range(V):
        normal[v].fill_(-1)
        tbuf[v].fill_(-1.0)
        leaf[v].fill_(-1)
    batched_step()
    torch.cuda.synchronize()
    batched_mismatch = sum(int((normal[v] != snap[v][0]).sum() + (tbuf[v].view(torch.int32) != snap[v][1].view(torch.int32)).sum()
                               + (leaf[v] != snap[v][2]).sum()) for v in range(V))
    batched["mismatched_fields_vs_plain"] = batched_mismatch

    # parity + fetch counts: the oracle on the same rays of every view
    import oracle_py as orc
    words = tree.words()
    cpu_rays = 0
    cpu_time = 0.0
    counts = np.zeros(3, np.uint64)
    mismatches = 0
    for v in range(V):
        tf = np.frombuffer(cams[v].tform, np.float32)
        pos = np.frombuffer(cams[v].pos, np.float32)
        d = orc.primary_dirs(tf, W, H)
        c0 = time.perf_counter()
        ref = orc.raymarch(words, pos[None, :], d, 0.0, threads=args.cpu_threads)
        cpu_time += time.perf_counter() - c0
        cpu_rays += n
        counts += ref["counts"]
        g_leaf = leaf[v].cpu().numpy().view(np.uint64)
        g_nrm = normal[v].cpu().numpy().view(np.uint32)
        g_t = tbuf[v].cpu().numpy().view(np.uint32)
        mismatches += int(np.count_nonzero(g_leaf != ref["leaf"]) + np.count_nonzero(g_nrm != ref["normal"])
                          + np.count_nonzero(g_t != ref["t"].view(np.uint32)))
    reads_per_ray = float(counts.sum()) / cpu_rays
    b_ray = 4.0 * reads_per_ray + 16.0
    achieved = b_ray * n / avg_launch_s / 1e9
    # the batched step is one launch of every view's tiles: its time is the launch's
    batched_gbs = b_ray * n * V / (batched_ms / 1e3) / 1e9
    batched["roofline"] = {"bound": "hbm", "achieved": round(batched_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(batched_gbs / HBM_PEAK_GBS, 5), "bytes_per_ray": round(b_ray, 2)}
    port_rate = cpu_rays / cpu_time / 1e9
    cpu_baseline = {"value": round(port_rate, 6), "unit": "Grays/s", "cores": args.cpu_threads, "kind": "port",
                    "sample": f"{cpu_rays} primary rays ({cpu_rays // n} of the {V} views, full {W}x{H}) through "
                              f"the oracle raymarch"}
    # every view through the reference's own raymarch, timed on this host
    ref_views, ref_mism, ref_sec, ref_rays = 0, 0, 0.0, 0
    for v in range(V):
        ref = reference_cpu_rate(tree, cams[v], W, H, args.cpu_threads, args.cache)
        if ref is None:
            break
        # normal and t bits of every ray (0 where a miss leaves them unwritten)
        g_nrm = normal[v].cpu().numpy().view(np.uint32)
        g_t = tbuf[v].cpu().numpy().view(np.uint32)
        ref_mism += int(np.count_nonzero(g_nrm != ref["out"][:, 1]) + np.count_nonzero(g_t != ref["out"][:, 2]))
        ref_sec += ref["rays"] / (ref["rate"] * 1e9)
        ref_rays += ref["rays"]
        ref_views += 1
        ref_threads = ref["threads"]
    ref_parity = None
    model, ncpu = host_cpu()
    cpu_baseline["cpu_model"], cpu_baseline["host_cpu_count"] = model, ncpu
    if ref_views:
        ref_parity = {"rays": ref_rays, "views": ref_views, "mismatched_fields": ref_mism}
        mismatches += ref_mism
        cpu_baseline = {"value": round(ref_rays / ref_sec / 1e9, 6), "unit": "Grays/s", "cores": ref_threads,
                        "kind": "reference", "cpu_model": model, "host_cpu_count": ncpu,
                        "sample": f"{ref_rays} primary rays (all {ref_views} views, {W}x{H}) through the reference's "
                                  f"own VoxelOctree::raymarch (oracle/_ref/svo_ref, {ref_threads} threads = this "
                                  f"host's CPU share for one GPU, 16 = the reference viewer's NumThreads, "
                                  f"src/Main.cpp:57; raymarch loop only); the oracle port on all {cpu_rays} rays: "
                                  f"{port_rate:.6f} Grays/s"}

    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Grays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total_ms / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": dict(config, views=V, width=W, height=H, rays_per_step=n * V,
                       kernel=("plain, one launch per view, raster tile order" if args.raster else
                               "plain, one launch per view, tiles started in the order of the view's previous "
                               "render's per-tile costs (svo_render_primary_planned_device)"),
                       other_order=compare),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": profiled_traffic(config["workload"], W, H),
                     "bytes_per_ray": round(b_ray, 2), "avg_launch_ms": round(avg_launch_s * 1e3, 4)},
        "cpu_baseline": cpu_baseline,
        "batched": batched,
        "mixed_batched": mixed,
        "served": served,
        "large_terrain": terrain,
        "viewer_frames": frames,
        "parity": {"rays_checked": cpu_rays, "mismatched_fields": mismatches, "vs_reference_binary": ref_parity,
                   "note": "plain-kernel outputs vs the oracle on every ray and vs the reference's own raymarch on "
                           "view 0; batched outputs vs plain below"},
        "wall_s": round(wall, 3),
    }
    print(json.dumps(out), flush=True)
    if mismatches or batched_mismatch or config.get("build_oct_md5_matches_reference") is False or \
            (mixed and mixed["mismatched_fields_vs_plain"]) or (terrain and (terrain["parity"]["mismatched_fields"] or terrain["planned"]["mismatched_bytes_vs_unplanned"])) or \
            (frames and (frames["parity"]["mismatched_pixels"] or frames["parity"]["mismatched_depths"]
                         or frames["frames_call"]["mismatched_vs_one_frame_per_call"]
                         or frames["frames_call"].get("planned", {}).get("mismatched_vs_one_frame_per_call")
                         or frames["planned"]["mismatched_vs_one_frame_per_call"]
                         or any(frames["parity"].get("vs_reference_classes", {}).values()))):
        sys.exit(3)


if __name__ == "__main__":
    main()
