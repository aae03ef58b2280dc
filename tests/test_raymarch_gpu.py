"""GPU parity of the HIP raymarch (libsvo_hip.so through its C ABI).

Bar: bit-exact. hit code, normal word, t (compared as bits) and leaf index
against (a) the reference's own outputs for the same rays (golden npz made
by the reference build) and (b) the oracle on camera frames and random rays
at sizes the oracle finishes in seconds. The other launch paths (batched
renders in the grid and persistent forms, HIP-graph replays, the request
loop, several viewer frames per call, octrees past 2^32 words) must equal
the plain per-view kernel or single-frame renders bit for bit.
"""
import os

import numpy as np
import pytest

import oracle_py as orc

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def _same(mine, ref, what=""):
    assert np.array_equal(mine["hit"], ref["hit"]), what
    assert np.array_equal(mine["normal"], ref["normal"]), what
    assert np.array_equal(_bits(mine["t"]), _bits(ref["t"])), what
    assert np.array_equal(mine["leaf"], ref["leaf"]), what


@pytest.fixture(scope="module")
def trees(svo, golden):
    return {
        "dragon": svo.VoxelOctree(os.path.join(golden, "dragon.oct")),
        "sdf128": svo.VoxelOctree(os.path.join(golden, "build_sdf128.oct")),
    }


@pytest.mark.parametrize("name", ["dragon", "sdf128"])
def test_batch_matches_reference_goldens(trees, golden, name):
    g = np.load(os.path.join(golden, f"raymarch_{name}.npz"))
    rays = g["rays"]
    print(f"[gpu] raymarch batch vs reference goldens: {name}, {len(rays)} rays", flush=True)
    r = trees[name].raymarch_batch(rays[:, 0:3], rays[:, 3:6], rays[:, 6])
    assert np.array_equal(r["hit"] != 0, g["hit"] != 0)
    leaf = r["hit"] == 1
    assert np.array_equal(r["normal"][leaf], g["normal"][leaf])
    hit = r["hit"] != 0
    assert np.array_equal(_bits(r["t"][hit]), _bits(g["t"][hit]))
    assert np.array_equal(r["leaf"][leaf], g["leaf"][leaf].astype(np.uint64))
    # and the full output record against the oracle
    _same(r, orc.raymarch(trees[name].words(), rays[:, 0:3], rays[:, 3:6], rays[:, 6]), name)


def test_extreme_and_nan_rays(trees, golden, svo):
    """Extreme components against the reference's outputs; rays whose
    coefficients are NaN (the reference never ends them) come back as misses
    from the batch, device and single-ray entries."""
    g = np.load(os.path.join(golden, "raymarch_extremes.npz"))
    rays = g["rays"]
    t = trees["dragon"]
    print(f"[gpu] extreme rays: {len(rays)}", flush=True)
    r = t.raymarch_batch(rays[:, 0:3], rays[:, 3:6], rays[:, 6])
    assert np.array_equal(r["hit"] != 0, g["hit"] != 0)
    leaf = r["hit"] == 1
    assert np.array_equal(r["normal"][leaf], g["normal"][leaf])
    hit = r["hit"] != 0
    assert np.array_equal(_bits(r["t"][hit]), _bits(g["t"][hit]))
    assert np.array_equal(r["leaf"][leaf], g["leaf"][leaf].astype(np.uint64))
    _same(r, orc.raymarch(t.words(), rays[:, 0:3], rays[:, 3:6], rays[:, 6]), "extremes")
    o = np.float32([[np.nan, 1.5, 1.5], [1.5, 1.5, 1.5], [np.inf, 0.5, 0.5], [1.2, 1.3, np.nan]])
    d = np.float32([[0.0, 0.0, 1.0], [np.nan, 1.0, 0.0], [-np.inf, 0.1, 0.2], [np.nan, np.nan, np.nan]])
    print("[gpu] NaN rays", flush=True)
    r = t.raymarch_batch(o, d, 0.0)
    assert not r["hit"].any() and not r["normal"].any() and not r["t"].any()
    assert (r["leaf"] == svo.svo.LEAF_NONE).all()
    for i in range(len(o)):
        assert t.raymarch(o[i], d[i], 0.0)[0] == 0


@pytest.mark.parametrize("name,w,h,pitch,yaw", [
    ("dragon", 320, 180, 20.0, 35.0),
    ("dragon", 257, 131, -55.0, 301.5),
    ("sdf128", 320, 240, 10.0, 120.0),
    ("sdf128", 1, 1, 0.0, 0.0),
    ("dragon", 9, 7, 89.0, 10.0),
])
def test_render_primary_matches_oracle(trees, name, w, h, pitch, yaw):
    t = trees[name]
    print(f"[gpu] render primary {name} {w}x{h} pitch={pitch} yaw={yaw}", flush=True)
    cam = t.orbit_camera(pitch, yaw, 1.0)
    mine = t.render_primary(cam, w, h)
    tf = np.frombuffer(cam.tform, np.float32)
    pos = np.frombuffer(cam.pos, np.float32)
    d = orc.primary_dirs(tf, w, h)
    ref = orc.raymarch(t.words(), pos[None, :], d, 0.0)
    _same(mine, ref, f"{name} {w}x{h}")


def test_concurrent_first_use_of_a_handle(svo, golden):
    """The reference's viewer raymarches one tree from 16 threads: here 8
    threads render and trace single rays on a fresh (not yet uploaded) handle
    at once, and each result equals a sequential one."""
    import threading
    t = svo.VoxelOctree(os.path.join(golden, "dragon.oct"))
    cams = [t.orbit_camera(10.0 * k, 25.0 * k, 1.0) for k in range(8)]
    rng = np.random.default_rng(9)
    o = rng.uniform(-0.5, 3.5, (8, 3)).astype(np.float32)
    d = (rng.uniform(0.9, 2.1, (8, 3)).astype(np.float32) - o).astype(np.float32)
    renders, rays, errors = [None] * 8, [None] * 8, []

    def work(k):
        try:
            renders[k] = t.render_primary(cams[k], 96, 64)
            rays[k] = t.raymarch(o[k], d[k], 0.0)
        except Exception as e:  # reported below
            errors.append(repr(e))

    print("[gpu] 8 threads on a fresh handle", flush=True)
    threads = [threading.Thread(target=work, args=(k,)) for k in range(8)]
    for x in threads:
        x.start()
    for x in threads:
        x.join()
    assert not errors, errors
    for k in range(8):
        _same(renders[k], t.render_primary(cams[k], 96, 64), f"camera {k}")
        assert rays[k] == t.raymarch(o[k], d[k], 0.0)


def test_octree_beyond_2_32_words(svo, golden):
    """A tree of more than 2^32 words runs the 64-bit-index kernels: the
    dragon moved to word 2^32 + 8 behind a far root (16 GiB, mostly zero
    pages) gives the dragon's hits, normals and t, with leaf indices moved by
    the same amount, for arbitrary rays, primary renders, viewer frames and a
    batched launch."""
    import torch
    from test_format import relocated
    c, w = orc.read_oct(os.path.join(golden, "dragon.oct"))
    base = (1 << 32) + 8
    print("[gpu] 2^32+ word octree: building the relocated array", flush=True)
    moved = relocated(w, base)
    big = svo.VoxelOctree.from_words(moved, c)
    assert big.word_count() == base + len(w) and big.depth() == 9
    small = svo.VoxelOctree(os.path.join(golden, "dragon.oct"))
    shift = np.uint64(base)

    def shifted(r):
        r = dict(r)
        leaf = r["hit"] == 1
        r["leaf"] = np.where(leaf, r["leaf"] + shift, r["leaf"])
        return r

    rng = np.random.default_rng(8)
    o = rng.uniform(-0.5, 3.5, (20000, 3)).astype(np.float32)
    d = (rng.uniform(0.9, 2.1, (20000, 3)).astype(np.float32) - o).astype(np.float32)
    print("[gpu] 2^32+ word octree: rays", flush=True)
    mine = big.raymarch_batch(o, d, 0.0)
    assert (mine["hit"] == 1).sum() > 2000
    _same(mine, shifted(small.raymarch_batch(o, d, 0.0)), "rays")
    _same({k: v[:500] for k, v in mine.items()}, orc.raymarch(moved, o[:500], d[:500], 0.0), "rays vs oracle")
    del moved
    cam = small.orbit_camera(20.0, 35.0, 1.0)
    print("[gpu] 2^32+ word octree: primary, frame, batch", flush=True)
    _same(big.render_primary(cam, 320, 180), shifted(small.render_primary(cam, 320, 180)), "primary")
    rgba, depth = big.render_frame(cam, 320, 176, 16, 1)
    ref_rgba, ref_depth = small.render_frame(cam, 320, 176, 16, 1)
    assert np.array_equal(rgba, ref_rgba) and np.array_equal(_bits(depth), _bits(ref_depth))
    dev = torch.device("cuda", 0)
    ren = svo.svo.Renderer(4)
    for form in (svo.svo.BATCH_GRID, svo.svo.BATCH_PERSISTENT):
        o4 = {k: torch.zeros(96 * 64, dtype=dt, device=dev)
              for k, dt in (("hit", torch.uint8), ("normal", torch.int32), ("t", torch.float32), ("leaf", torch.int64))}
        ren.render_device([ren.view(big, cam, 96, 64, o4["hit"].data_ptr(), o4["normal"].data_ptr(),
                                    o4["t"].data_ptr(), o4["leaf"].data_ptr())], None, 0, 0, form)
        torch.cuda.synchronize()
        got = {"hit": o4["hit"].cpu().numpy(), "normal": o4["normal"].cpu().numpy().view(np.uint32),
               "t": o4["t"].cpu().numpy(), "leaf": o4["leaf"].cpu().numpy().view(np.uint64)}
        _same(got, shifted(small.render_primary(cam, 96, 64)), f"batched form {form}")


def test_single_ray_entry_matches_batch(trees):
    t = trees["dragon"]
    rng = np.random.default_rng(5)
    o = rng.uniform(-0.5, 3.5, (40, 3)).astype(np.float32)
    d = (rng.uniform(1.0, 2.0, (40, 3)) - o).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    print("[gpu] single-ray entry", flush=True)
    b = t.raymarch_batch(o, d, 0.01)
    for i in range(len(o)):
        hit, normal, tt = t.raymarch(o[i], d[i], 0.01)
        assert hit == b["hit"][i]
        if hit == 1:
            assert normal == b["normal"][i]
        if hit:
            assert np.float32(tt).view(np.uint32) == _bits(b["t"][i:i + 1])[0]


def test_far_pointers_and_empty_batch(trees, svo):
    """Trees with far pointers (the dragon has 24 far words) and a zero-ray
    call; rays aimed at the far-pointer subtrees are part of the frames above."""
    t = trees["dragon"]
    r = t.raymarch_batch(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32))
    assert len(r["hit"]) == 0


def test_batched_renderer_matches_plain_kernel(svo, trees, golden):
    """One launch over views of different octrees, sizes (edge tiles, sizes
    below one tile) and LOD scales equals the plain per-view kernel bit for
    bit, in the grid form and in the persistent form for every refill
    threshold / wave count tried."""
    import torch
    dev = torch.device("cuda", 0)
    specs = [("dragon", 320, 180, 20.0, 35.0, 0.0), ("sdf128", 257, 131, -30.0, 200.0, 0.0),
             ("dragon", 7, 5, 10.0, 10.0, 0.0), ("sdf128", 640, 360, 45.0, 300.0, 0.0),
             ("dragon", 200, 100, -60.0, 90.0, 0.02)]
    ren = svo.svo.Renderer(16)
    for waves, refill, form in [(0, 0, svo.svo.BATCH_GRID), (0, 0, svo.svo.BATCH_PERSISTENT), (8, 1, None),
                                (32, 64, None), (16, 40, None)]:
        outs, views = [], []
        for name, w, h, p, y, rs in specs:
            t = trees[name]
            cam = t.orbit_camera(p, y, 1.0)
            o = {"hit": torch.full((w * h,), 9, dtype=torch.uint8, device=dev),
                 "normal": torch.zeros(w * h, dtype=torch.int32, device=dev),
                 "t": torch.zeros(w * h, dtype=torch.float32, device=dev),
                 "leaf": torch.zeros(w * h, dtype=torch.int64, device=dev)}
            outs.append((t, cam, w, h, rs, o))
            views.append(ren.view(t, cam, w, h, o["hit"].data_ptr(), o["normal"].data_ptr(), o["t"].data_ptr(),
                                  o["leaf"].data_ptr(), ray_scale=rs))
        print(f"[gpu] batched renderer waves={waves} refill={refill} form={form}", flush=True)
        ren.render_device(views, None, waves, refill, form)
        torch.cuda.synchronize()
        for t, cam, w, h, rs, o in outs:
            if rs == 0.0:
                ref = t.render_primary(cam, w, h)
            else:
                d = orc.primary_dirs(np.frombuffer(cam.tform, np.float32), w, h)
                ref = orc.raymarch(t.words(), np.frombuffer(cam.pos, np.float32)[None, :], d, rs)
            mine = {"hit": o["hit"].cpu().numpy(), "normal": o["normal"].cpu().numpy().view(np.uint32),
                    "t": o["t"].cpu().numpy(), "leaf": o["leaf"].cpu().numpy().view(np.uint64)}
            _same(mine, ref, f"{w}x{h} rs={rs}")


def test_batched_renderer_with_a_deep_octree(svo, trees):
    """A batch whose deepest octree has depth 13 (4096^3 terrain) keeps the
    top three stack entries of every view in registers; each view, the
    shallow dragon's included, equals the plain kernel, and the terrain's
    plain render equals the oracle on a sample of rays."""
    import torch
    dev = torch.device("cuda", 0)
    deep = svo.VoxelOctree.build_from_terrain(4096, 41, memory_budget=1 << 30)
    assert deep.depth() == 13
    pitch, yaw, _ = svo.scenes.terrain_view(41)
    specs = [(deep, 320, 180, pitch, yaw), (trees["dragon"], 200, 120, 20.0, 35.0), (deep, 77, 45, 30.0, 100.0)]
    ren = svo.svo.Renderer(8)
    outs, views = [], []
    for t, w, h, p, y in specs:
        cam = t.orbit_camera(p, y, 1.0)
        o = {"hit": torch.full((w * h,), 9, dtype=torch.uint8, device=dev),
             "normal": torch.zeros(w * h, dtype=torch.int32, device=dev),
             "t": torch.zeros(w * h, dtype=torch.float32, device=dev),
             "leaf": torch.zeros(w * h, dtype=torch.int64, device=dev)}
        outs.append((t, cam, w, h, o))
        views.append(ren.view(t, cam, w, h, o["hit"].data_ptr(), o["normal"].data_ptr(), o["t"].data_ptr(),
                              o["leaf"].data_ptr()))
    for form in (svo.svo.BATCH_GRID, svo.svo.BATCH_PERSISTENT):
        print(f"[gpu] batched renderer with a depth-13 octree, form {form}", flush=True)
        for _, _, _, _, o in outs:
            o["hit"].fill_(9)
        ren.render_device(views, None, 0, 0, form)
        torch.cuda.synchronize()
        for t, cam, w, h, o in outs:
            mine = {"hit": o["hit"].cpu().numpy(), "normal": o["normal"].cpu().numpy().view(np.uint32),
                    "t": o["t"].cpu().numpy(), "leaf": o["leaf"].cpu().numpy().view(np.uint64)}
            _same(mine, t.render_primary(cam, w, h), f"{w}x{h} form {form}")
    t, cam, w, h, _ = outs[0]
    plain = t.render_primary(cam, w, h)
    sample = np.arange(0, w * h, 7)
    d = orc.primary_dirs(np.frombuffer(cam.tform, np.float32), w, h)[sample]
    ref = orc.raymarch(t.words(), np.frombuffer(cam.pos, np.float32)[None, :], d, 0.0)
    _same({k: v[sample] for k, v in plain.items()}, ref, "terrain 4096 vs oracle")
    assert (plain["hit"] == 1).mean() > 0.2


def _batch_setup(svo, trees, ren, dev, specs):
    import torch
    outs, views = [], []
    for name, w, h, p, y, rs in specs:
        t = trees[name]
        cam = t.orbit_camera(p, y, 1.0)
        o = {"hit": torch.full((w * h,), 9, dtype=torch.uint8, device=dev),
             "normal": torch.zeros(w * h, dtype=torch.int32, device=dev),
             "t": torch.zeros(w * h, dtype=torch.float32, device=dev),
             "leaf": torch.zeros(w * h, dtype=torch.int64, device=dev)}
        outs.append((t, cam, w, h, rs, o))
        views.append(ren.view(t, cam, w, h, o["hit"].data_ptr(), o["normal"].data_ptr(), o["t"].data_ptr(),
                              o["leaf"].data_ptr(), ray_scale=rs))
    return outs, views


def _check_batch(outs, refs):
    for (t, cam, w, h, rs, o), ref in zip(outs, refs):
        mine = {"hit": o["hit"].cpu().numpy(), "normal": o["normal"].cpu().numpy().view(np.uint32),
                "t": o["t"].cpu().numpy(), "leaf": o["leaf"].cpu().numpy().view(np.uint64)}
        _same(mine, ref, f"{w}x{h} rs={rs}")


@pytest.mark.parametrize("form", [0, 1])
def test_batch_graph_replays_match_eager(svo, trees, form):
    """A batched launch captured in a HIP graph, replayed several times with
    the program's own eager work and synchronisations in between (outputs
    cleared by eager fills each time): every replay equals the plain kernel."""
    import torch
    dev = torch.device("cuda", 0)
    specs = [("dragon", 320, 180, 20.0, 35.0, 0.0), ("sdf128", 257, 131, -30.0, 200.0, 0.0),
             ("dragon", 200, 100, -60.0, 90.0, 0.02)]
    ren = svo.svo.Renderer(8)
    outs, views = _batch_setup(svo, trees, ren, dev, specs)
    refs = []
    for t, cam, w, h, rs, o in outs:
        if rs == 0.0:
            refs.append(t.render_primary(cam, w, h))
        else:
            d = orc.primary_dirs(np.frombuffer(cam.tform, np.float32), w, h)
            refs.append(orc.raymarch(t.words(), np.frombuffer(cam.pos, np.float32)[None, :], d, rs))
    stream = torch.cuda.Stream()
    print(f"[gpu] batch graph capture, form {form}", flush=True)
    g = ren.capture(views, stream.cuda_stream, form=form)
    for k in range(4):
        with torch.cuda.stream(stream):
            for _, _, _, _, _, o in outs:
                o["hit"].fill_(9)
                o["normal"].fill_(-1)
                o["t"].fill_(-1.0)
                o["leaf"].fill_(-1)
            # eager work of the program between replays, on the same stream
            trees["dragon"].render_primary_device(outs[0][1], 16, 16, None, None, None, None, stream.cuda_stream)
        print(f"[gpu] batch graph replay {k}", flush=True)
        g.launch(stream.cuda_stream)
        stream.synchronize()
        _check_batch(outs, refs)
    g.close()


def test_autotune_picks_a_shape_and_keeps_results(svo, trees):
    import torch
    dev = torch.device("cuda", 0)
    specs = [("dragon", 320, 180, 20.0, 35.0, 0.0), ("sdf128", 640, 360, 45.0, 300.0, 0.0)]
    ren = svo.svo.Renderer(4)
    outs, views = _batch_setup(svo, trees, ren, dev, specs)
    print("[gpu] batch autotune", flush=True)
    (form, w, r, tw), times = ren.autotune(views)
    torch.cuda.synchronize()
    assert len(times) == 13 and all(v > 0 for v in times.values())
    if form == svo.svo.BATCH_GRID:
        assert w == 0 and r == 0 and tw in (4, 8, 16, 32)
        assert times[f"grid{tw}x{64 // tw}"] == min(times.values())
    else:
        assert form == svo.svo.BATCH_PERSISTENT and w in (16, 24, 32) and r in (16, 32, 64) and tw == 8
        assert times[f"{w}x{r}"] == min(times.values())
    ren.render_device(views, None, w, r, form, tw)
    torch.cuda.synchronize()
    with pytest.raises(svo.SvoError) as e:  # no such launch form
        ren.render_device(views, None, 0, 0, 7)
    assert e.value.code == svo.svo.ERR_INVALID
    with pytest.raises(svo.SvoError) as e:  # no such tile shape
        ren.render_device(views, None, 0, 0, svo.svo.BATCH_GRID, 12)
    assert e.value.code == svo.svo.ERR_INVALID
    _check_batch(outs, [t.render_primary(cam, ww, hh) for t, cam, ww, hh, _, _ in outs])


@pytest.mark.parametrize("tile_w", [4, 16, 32])
def test_grid_tile_shapes_match_plain_kernel(svo, trees, tile_w):
    """The grid form with 4x16, 16x4 and 32x2 tiles (edge tiles, views below
    one tile, LOD scale) equals the plain 8x8 kernel bit for bit."""
    import torch
    dev = torch.device("cuda", 0)
    specs = [("dragon", 320, 180, 20.0, 35.0, 0.0), ("sdf128", 257, 131, -30.0, 200.0, 0.0),
             ("dragon", 7, 5, 10.0, 10.0, 0.0), ("dragon", 200, 100, -60.0, 90.0, 0.02)]
    ren = svo.svo.Renderer(8)
    outs, views = _batch_setup(svo, trees, ren, dev, specs)
    print(f"[gpu] grid form tile width {tile_w}", flush=True)
    ren.render_device(views, None, 0, 0, svo.svo.BATCH_GRID, tile_w)
    torch.cuda.synchronize()
    refs = []
    for t, cam, w, h, rs, o in outs:
        if rs == 0.0:
            refs.append(t.render_primary(cam, w, h))
        else:
            d = orc.primary_dirs(np.frombuffer(cam.tform, np.float32), w, h)
            refs.append(orc.raymarch(t.words(), np.frombuffer(cam.pos, np.float32)[None, :], d, rs))
    _check_batch(outs, refs)


@pytest.mark.parametrize("one_per_launch,streams,budget,form", [(False, 4, 1 << 28, 0), (False, 2, 1 << 24, 0),
                                                               (True, 1, 1 << 26, 0), (False, 4, 1 << 28, 1)])
def test_served_results_equal_alone_and_in_the_mix(svo, trees, one_per_launch, streams, budget, form):
    """Requests arriving while others run are folded into the next launch;
    each served result equals the same request rendered alone."""
    rng = np.random.default_rng(17)
    names = ["dragon", "sdf128"]
    tl = [trees[k] for k in names]
    reqs = []
    for i in range(48):
        ti = int(rng.integers(0, 2))
        w = int(rng.integers(8, 400))
        h = int(rng.integers(8, 300))
        cam = tl[ti].orbit_camera(float(rng.uniform(-60, 60)), float(rng.uniform(0, 360)), 1.0)
        reqs.append((ti, cam, w, h))
    arrivals = np.cumsum(rng.exponential(1.0 / 3000.0, len(reqs)))
    outs = [(np.zeros(w * h, np.uint32), np.zeros(w * h, np.float32), np.zeros(w * h, np.uint64))
            for _, _, w, h in reqs]
    print(f"[gpu] serve replay one_per_launch={one_per_launch} streams={streams} budget={budget} form={form}",
          flush=True)
    lat, makespan = svo.svo.serve_replay(tl, reqs, arrivals, streams=streams, output_budget=budget,
                                         one_per_launch=one_per_launch, outputs=outs, batch_form=form)
    assert np.all(lat > 0) and makespan > 0
    for (ti, cam, w, h), (nrm, t, leaf) in zip(reqs, outs):
        ref = tl[ti].render_primary(cam, w, h)
        assert np.array_equal(nrm, ref["normal"])
        assert np.array_equal(t.view(np.uint32), ref["t"].view(np.uint32))
        assert np.array_equal(leaf, ref["leaf"])


def test_served_request_larger_than_budget_is_refused(svo, trees):
    t = trees["dragon"]
    cam = t.orbit_camera(0.0, 0.0, 1.0)
    with pytest.raises(svo.SvoError) as e:
        svo.svo.serve_replay([t], [(0, cam, 1920, 1080)], [0.0], streams=2, output_budget=1 << 20)
    assert e.value.code == svo.svo.ERR_LIMIT


@pytest.mark.parametrize("name", ["dragon", "sdf128"])
def test_frame_matches_reference_goldens(trees, golden, name):
    """The viewer frame (coarse LOD pass + per-tile fine pass + shading)
    against the reference driver's frames: RGBA and coarse depths bit-equal."""
    g = np.load(os.path.join(golden, f"frame_{name}.npz"))
    w, h = (int(v) for v in g["size"])
    t = trees[name]
    for i, (pitch, yaw, radius) in enumerate(g["cams"]):
        cam = t.orbit_camera(float(pitch), float(yaw), float(radius))
        for stride in (1, 3):
            print(f"[gpu] frame {name} cam {i} stride {stride} {w}x{h}", flush=True)
            rgba, depth = t.render_frame(cam, w, h, 16, stride)
            assert np.array_equal(rgba, g[f"rgba_s{stride}"][i]), (name, i, stride)
            assert np.array_equal(_bits(depth), _bits(g[f"depth_s{stride}"][i])), (name, i, stride)


@pytest.mark.parametrize("name,w,h,stripes,stride,pitch,yaw", [
    ("dragon", 1280, 720, 16, 1, 25.0, 40.0),
    ("dragon", 1280, 720, 16, 3, -10.0, 200.0),
    ("sdf128", 641, 359, 7, 1, 5.0, 75.0),
    ("sdf128", 8, 8, 1, 2, 0.0, 0.0),
    ("dragon", 1, 1, 1, 1, 0.0, 0.0),
])
def test_frame_matches_oracle(trees, name, w, h, stripes, stride, pitch, yaw):
    t = trees[name]
    print(f"[gpu] frame vs oracle {name} {w}x{h} stripes={stripes} stride={stride}", flush=True)
    cam = t.orbit_camera(pitch, yaw, 1.0)
    rgba, depth = t.render_frame(cam, w, h, stripes, stride)
    tf = np.frombuffer(cam.tform, np.float32)
    ref_rgba, ref_depth = orc.render_frame(t.words(), tf, t.center(), w, h, stripes, stride)
    assert np.array_equal(rgba, ref_rgba)
    assert np.array_equal(_bits(depth), _bits(ref_depth))


def test_frame_device_entry_and_errors(svo, trees):
    import torch
    t = trees["dragon"]
    cam = t.orbit_camera(30.0, 10.0, 1.0)
    w, h = 320, 176
    print("[gpu] frame device entry", flush=True)
    ref, _ = t.render_frame(cam, w, h, 16, 1)
    out = torch.full((h, w), 0x12345678, dtype=torch.int32, device="cuda")
    t.render_frame_device(cam, w, h, out.data_ptr(), None, 16, 1, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy().view(np.uint32), ref)
    with pytest.raises(svo.SvoError):
        t.render_frame(cam, w, 100, 16, 1)  # 16 stripes of 7 rows leave the last one empty
    with pytest.raises(svo.SvoError):
        t.render_frame(cam, w, h, 16, 9)


def test_frames_in_one_call_equal_single_frames(svo, trees):
    """svo_render_frames_device (up to 16 frames per coarse launch and per
    fine launch, on the caller's one stream) gives each frame's RGBA and depth
    table exactly as a single-frame render, with per-frame NULL depths, an odd
    frame count and stride 3, and for 18 frames (two launch groups, 16 + 2);
    bad cameras or null images are refused before any frame is enqueued."""
    import torch
    t = trees["dragon"]
    w, h, stripes = 320, 176, 16
    nd = svo.svo.frame_depth_count(w, h, stripes)
    angles = [(30.0, 10.0), (20.0, 35.0), (-15.0, 120.0), (45.0, 200.0), (5.0, 300.0)]
    cams = [t.orbit_camera(p, y, 1.0) for p, y in angles]
    stream = torch.cuda.Stream()
    for stride in (1, 3):
        rgba = [torch.full((h * w,), 0x12345678, dtype=torch.int32, device="cuda") for _ in cams]
        depth = [torch.full((nd,), -1.0, dtype=torch.float32, device="cuda") for _ in cams]
        dptrs = [d.data_ptr() if k != 2 else None for k, d in enumerate(depth)]
        print(f"[gpu] frames in one call, stride {stride}", flush=True)
        t.render_frames_device(cams, w, h, [x.data_ptr() for x in rgba], dptrs, stripes, stride, stream.cuda_stream)
        stream.synchronize()
        for k, cam in enumerate(cams):
            ref_rgba, ref_depth = t.render_frame(cam, w, h, stripes, stride)
            assert np.array_equal(rgba[k].cpu().numpy().view(np.uint32), ref_rgba.reshape(-1)), (stride, k)
            if k != 2:
                assert np.array_equal(depth[k].cpu().numpy().view(np.uint32), ref_depth.view(np.uint32)), (stride, k)
            else:
                assert (depth[k].cpu().numpy() == -1.0).all()
    # 18 frames: two launch groups (16 + 2)
    w2, h2 = 72, 40
    cams18 = [t.orbit_camera(-40.0 + 5.0 * k, 20.0 * k, 1.0) for k in range(18)]
    rgba18 = [torch.zeros(h2 * w2, dtype=torch.int32, device="cuda") for _ in cams18]
    print("[gpu] frames in one call, 18 frames", flush=True)
    t.render_frames_device(cams18, w2, h2, [x.data_ptr() for x in rgba18], None, 4, 1, stream.cuda_stream)
    stream.synchronize()
    for k, cam in enumerate(cams18):
        ref_rgba, _ = t.render_frame(cam, w2, h2, 4, 1)
        assert np.array_equal(rgba18[k].cpu().numpy().view(np.uint32), ref_rgba.reshape(-1)), k
    bad = t.orbit_camera(0.0, 0.0, 1.0)
    bad.pos[0] = float("nan")
    out = [torch.full((h * w,), 7, dtype=torch.int32, device="cuda") for _ in range(2)]
    for cs, ptrs in (([cams[0], bad], [x.data_ptr() for x in out]), ([cams[0], cams[1]], [out[0].data_ptr(), None])):
        with pytest.raises(svo.SvoError) as e:
            t.render_frames_device(cs, w, h, ptrs, None, stripes, 1, stream.cuda_stream)
        assert e.value.code == svo.svo.ERR_INVALID
    stream.synchronize()
    assert all(bool((x == 7).all()) for x in out)  # nothing was enqueued


def test_planned_frames_equal_unplanned_frames(svo, trees):
    """svo_render_frame_planned_device: a viewer's frames through one tile plan
    (the first in raster order, the next ones with the fine pass in the cost
    order the previous frame filed, over seven camera moves that take every
    counter set and id region through filling, reading and clearing twice)
    give RGBA and depth tables bit-equal to unplanned frames, at stride 1 and
    3 and with a NULL depth table; a plan made for another size or stripe
    split is refused before anything is enqueued."""
    import torch
    t = trees["dragon"]
    w, h, stripes = 1280, 720, 16
    nd = svo.svo.frame_depth_count(w, h, stripes)
    stream = torch.cuda.Stream()
    for stride in (1, 3):
        plan = svo.svo.TilePlan.for_frame(w, h, stripes)
        for k in range(7):
            cam = t.orbit_camera(20.0 - 6.0 * k, 30.0 + 15.0 * k, 1.0)
            rgba = torch.full((h * w,), 0x12345678, dtype=torch.int32, device="cuda")
            depth = torch.full((nd,), -1.0, dtype=torch.float32, device="cuda")
            dptr = depth.data_ptr() if k != 4 else None
            print(f"[gpu] planned frame {k} stride {stride} {w}x{h}", flush=True)
            t.render_frame_device(cam, w, h, rgba.data_ptr(), dptr, stripes, stride, stream.cuda_stream, plan=plan)
            stream.synchronize()
            ref_rgba, ref_depth = t.render_frame(cam, w, h, stripes, stride)
            assert np.array_equal(rgba.cpu().numpy().view(np.uint32), ref_rgba.reshape(-1)), (stride, k)
            if dptr is not None:
                assert np.array_equal(depth.cpu().numpy().view(np.uint32), ref_depth.view(np.uint32)), (stride, k)
    cam = t.orbit_camera(0.0, 0.0, 1.0)
    out = torch.full((h * w,), 7, dtype=torch.int32, device="cuda")
    for pw, ph, ps in ((w, h + 8, stripes), (w, h, 8)):
        other = svo.svo.TilePlan.for_frame(pw, ph, ps)
        print(f"[gpu] planned frame refused: plan {pw}x{ph}/{ps}", flush=True)
        with pytest.raises(svo.SvoError) as e:
            t.render_frame_device(cam, w, h, out.data_ptr(), None, stripes, 1, stream.cuda_stream, plan=other)
        assert e.value.code == svo.svo.ERR_INVALID
    stream.synchronize()
    assert bool((out == 7).all())


def test_planned_frame_calls_equal_unplanned_calls(svo, trees):
    """svo_render_frames_planned_device: 8 frames per call through one plan
    (the fine pass over all 8 frames' tiles in the cost order the previous
    call filed), over five calls of moving cameras at stride 1 and 3, give
    the same RGBA and depth tables as svo_render_frames_device; a plan made
    for another frame count is refused before anything is enqueued."""
    import torch
    t = trees["dragon"]
    w, h, stripes, nf = 640, 360, 16, 8
    nd = svo.svo.frame_depth_count(w, h, stripes)
    stream = torch.cuda.Stream()
    for stride in (1, 3):
        plan = svo.svo.TilePlan.for_frames(w, h, nf, stripes)
        for call in range(5):
            cams = [t.orbit_camera(15.0 - 4.0 * call + 2.0 * k, 20.0 * call + 40.0 * k, 1.0) for k in range(nf)]
            rgba = [torch.full((h * w,), 0x12345678, dtype=torch.int32, device="cuda") for _ in cams]
            depth = [torch.full((nd,), -1.0, dtype=torch.float32, device="cuda") for _ in cams]
            ref_rgba = [torch.zeros_like(x) for x in rgba]
            ref_depth = [torch.zeros_like(x) for x in depth]
            print(f"[gpu] planned frames call {call} stride {stride} {nf}x{w}x{h}", flush=True)
            t.render_frames_device(cams, w, h, [x.data_ptr() for x in rgba], [x.data_ptr() for x in depth], stripes,
                                   stride, stream.cuda_stream, plan=plan)
            t.render_frames_device(cams, w, h, [x.data_ptr() for x in ref_rgba], [x.data_ptr() for x in ref_depth],
                                   stripes, stride, stream.cuda_stream)
            stream.synchronize()
            for k in range(nf):
                assert torch.equal(rgba[k], ref_rgba[k]), (stride, call, k)
                assert torch.equal(depth[k].view(torch.int32), ref_depth[k].view(torch.int32)), (stride, call, k)
    cams = [t.orbit_camera(0.0, 30.0 * k, 1.0) for k in range(nf)]
    out = [torch.full((h * w,), 7, dtype=torch.int32, device="cuda") for _ in cams]
    other = svo.svo.TilePlan.for_frames(w, h, nf - 1, stripes)
    print("[gpu] planned frames call refused: plan for 7 frames", flush=True)
    with pytest.raises(svo.SvoError) as e:
        t.render_frames_device(cams, w, h, [x.data_ptr() for x in out], None, stripes, 1, stream.cuda_stream,
                               plan=other)
    assert e.value.code == svo.svo.ERR_INVALID
    stream.synchronize()
    assert all(bool((x == 7).all()) for x in out)


@pytest.mark.parametrize("name,w,h", [("dragon", 1, 1), ("dragon", 9, 7), ("empty", 64, 48), ("one", 33, 17),
                                       ("sdf128", 520, 8)])
def test_tile_plans_on_small_and_empty_images(svo, trees, golden, name, w, h):
    """Tile plans at the edges: a 1x1 image (one tile: a plan of one
    replica), images of fewer tiles than the 64 replicas, an octree without
    leaves and a one-voxel octree, and a one-tile-high strip: three planned
    primary renders (raster, then the filed order) equal the plain kernel."""
    import torch
    dev = torch.device("cuda", 0)
    if name in trees:
        t = trees[name]
    else:
        t = svo.VoxelOctree(os.path.join(golden, "build_grid_empty16.oct" if name == "empty" else "build_grid_one16.oct"))
    plan = svo.svo.TilePlan(w, h)
    for k, (pitch, yaw) in enumerate([(20.0, 35.0), (20.0, 35.0), (-30.0, 200.0)]):
        cam = t.orbit_camera(pitch, yaw, 1.0)
        o = {"hit": torch.full((w * h,), 9, dtype=torch.uint8, device=dev),
             "normal": torch.zeros(w * h, dtype=torch.int32, device=dev),
             "t": torch.zeros(w * h, dtype=torch.float32, device=dev),
             "leaf": torch.zeros(w * h, dtype=torch.int64, device=dev)}
        print(f"[gpu] planned render {name} {w}x{h} #{k}", flush=True)
        plan.render_device(t, cam, o["hit"].data_ptr(), o["normal"].data_ptr(), o["t"].data_ptr(), o["leaf"].data_ptr())
        torch.cuda.synchronize()
        mine = {"hit": o["hit"].cpu().numpy(), "normal": o["normal"].cpu().numpy().view(np.uint32),
                "t": o["t"].cpu().numpy(), "leaf": o["leaf"].cpu().numpy().view(np.uint64)}
        _same(mine, t.render_primary(cam, w, h), f"planned render {name} {k}")


def test_cost_ordered_renders_match_plain_kernel(svo, trees):
    """svo_render_primary_planned_device: the first render of a plan (raster
    order), the next ones (tiles in the order the previous render filed them
    by cost) and renders of another camera through the same plan (a stale
    order) all equal the plain kernel bit for bit over seven renders, which
    take every counter set and id region of the plan through filling, reading
    and clearing twice; every pixel is written (outputs start from a fill no
    result has); a plan refuses another image size."""
    import torch
    dev = torch.device("cuda", 0)
    t = trees["dragon"]
    w, h = 333, 197
    plan = svo.svo.TilePlan(w, h)
    a, b = t.orbit_camera(20.0, 35.0, 1.0), t.orbit_camera(-40.0, 250.0, 1.0)
    cams = [a, a, b, a, b, b, a]
    for k, cam in enumerate(cams):
        o = {"hit": torch.full((w * h,), 9, dtype=torch.uint8, device=dev),
             "normal": torch.zeros(w * h, dtype=torch.int32, device=dev),
             "t": torch.zeros(w * h, dtype=torch.float32, device=dev),
             "leaf": torch.zeros(w * h, dtype=torch.int64, device=dev)}
        print(f"[gpu] cost-ordered render {k}", flush=True)
        plan.render_device(t, cam, o["hit"].data_ptr(), o["normal"].data_ptr(), o["t"].data_ptr(), o["leaf"].data_ptr())
        torch.cuda.synchronize()
        mine = {"hit": o["hit"].cpu().numpy(), "normal": o["normal"].cpu().numpy().view(np.uint32),
                "t": o["t"].cpu().numpy(), "leaf": o["leaf"].cpu().numpy().view(np.uint64)}
        _same(mine, t.render_primary(cam, w, h), f"planned render {k}")
    other = svo.svo.TilePlan(w + 1, h)
    with pytest.raises(svo.SvoError) as e:
        lib = svo.svo.lib
        import ctypes
        svo.svo._check(lib.svo_render_primary_planned_device(t.handle, other._p, ctypes.byref(cams[0]), w, h,
                                                             None, None, None, None, None))
    assert e.value.code == svo.svo.ERR_INVALID
    plan.close()
    other.close()


def test_cost_ordered_batches_match_unordered(svo, trees):
    """svo_render_batch_planned_device: five launches of one batch (views of
    two octrees, edge tiles, a view below one tile, a LOD view) through one
    batch plan, in the 8x8 and 16x4 tile shapes, equal the unordered grid
    launch on every ray (outputs start from a fill no result has); the plan
    refuses another batch, the persistent form and a plain image plan of
    another tile count."""
    import torch
    dev = torch.device("cuda", 0)
    specs = [("dragon", 320, 180, 20.0, 35.0, 0.0), ("sdf128", 257, 131, -30.0, 200.0, 0.0),
             ("dragon", 7, 5, 10.0, 10.0, 0.0), ("dragon", 200, 100, -60.0, 90.0, 0.02)]
    ren = svo.svo.Renderer(8)

    def batch():
        outs, views = [], []
        for name, w, h, p, y, rs in specs:
            t = trees[name]
            o = {"hit": torch.full((w * h,), 9, dtype=torch.uint8, device=dev),
                 "normal": torch.zeros(w * h, dtype=torch.int32, device=dev),
                 "t": torch.zeros(w * h, dtype=torch.float32, device=dev),
                 "leaf": torch.full((w * h,), -2, dtype=torch.int64, device=dev)}
            outs.append(o)
            views.append(ren.view(t, t.orbit_camera(p, y, 1.0), w, h, o["hit"].data_ptr(), o["normal"].data_ptr(),
                                  o["t"].data_ptr(), o["leaf"].data_ptr(), ray_scale=rs))
        return outs, views

    for tile_w in (8, 16):
        ref, rviews = batch()
        ren.render_device(rviews, None, tile_w=tile_w)
        torch.cuda.synchronize()
        outs, views = batch()
        plan = svo.svo.TilePlan.for_batch(views, tile_w=tile_w)
        for k in range(5):
            for o in outs:
                o["hit"].fill_(9)
                o["leaf"].fill_(-2)
            print(f"[gpu] planned batch tile_w={tile_w} launch {k}", flush=True)
            ren.render_device(views, None, tile_w=tile_w, plan=plan)
            torch.cuda.synchronize()
            for a, b in zip(outs, ref):
                assert all(bool(torch.equal(a[f], b[f])) for f in ("hit", "normal", "t", "leaf")), (tile_w, k)
        with pytest.raises(svo.SvoError) as e:
            ren.render_device(views[:2], None, tile_w=tile_w, plan=plan)
        assert e.value.code == svo.svo.ERR_INVALID
        with pytest.raises(svo.SvoError) as e:
            ren.render_device(views, None, form=svo.svo.BATCH_PERSISTENT, plan=plan)
        assert e.value.code == svo.svo.ERR_INVALID
        plan.close()
    image_plan = svo.svo.TilePlan(64, 64)
    with pytest.raises(svo.SvoError) as e:
        ren.render_device(views, None, plan=image_plan)
    assert e.value.code == svo.svo.ERR_INVALID
    image_plan.close()
