"""GPU parity of the full-size render workloads against the oracle.

Bar: bit-exact (hit code, normal word, t bits, leaf index).

- icosphere 2048^3, 8 seeded 1920x1080 views: the GPU-built octree (its .oct
  hashes to the reference builder's) rendered by the plain kernel equals the
  oracle on every ray; the batched grid form and its HIP-graph replay equal
  the plain kernel on every ray.
- mixed batch: 64 GPU-built octrees (128^3..2048^3 SDF scenes and displaced
  meshes) and 256 views (320x240..1920x1080): the batched launch in the grid
  form, the persistent form and a HIP-graph replay equal the 256 plain
  launches on every ray, and every view equals the oracle on every 97th pixel.
- served stream: a seeded Poisson stream (2000 req/s, 500 requests, then a
  burst of 100) over the 64 octrees through the live server; every result
  equals the request rendered alone on every ray and the oracle on every
  97th pixel.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle_py as orc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _host(o):
    return {"hit": o["hit"].cpu().numpy(), "normal": o["normal"].cpu().numpy().view(np.uint32),
            "t": o["t"].cpu().numpy(), "leaf": o["leaf"].cpu().numpy().view(np.uint64)}


def _same(mine, ref, what):
    for k in ("hit", "normal", "leaf"):
        bad = np.flatnonzero(mine[k] != ref[k])
        assert len(bad) == 0, (what, k, len(bad), bad[:5])
    bad = np.flatnonzero(mine["t"].view(np.uint32) != ref["t"].view(np.uint32))
    assert len(bad) == 0, (what, "t", len(bad), bad[:5])


def _outputs(torch, n, dev):
    return {"hit": torch.full((n,), 9, dtype=torch.uint8, device=dev),
            "normal": torch.full((n,), -1, dtype=torch.int32, device=dev),
            "t": torch.full((n,), -1.0, dtype=torch.float32, device=dev),
            "leaf": torch.full((n,), -2, dtype=torch.int64, device=dev)}


def _ptrs(o):
    return o["hit"].data_ptr(), o["normal"].data_ptr(), o["t"].data_ptr(), o["leaf"].data_ptr()


def _equal_on_device(torch, a, b):
    return all(bool(torch.equal(a[k], b[k])) for k in ("hit", "normal", "t", "leaf"))


@pytest.fixture(scope="module")
def ico2048(svo, tmp_path_factory):
    b = json.load(open(os.path.join(ROOT, "tests", "golden", "builds.json")))["ico2048"]
    v, f = svo.scenes.icosphere(b["frequency"], seed=b["seed"], displacement=b["displacement"])
    ply = str(tmp_path_factory.mktemp("ico") / "ico2048.ply")
    svo.scenes.write_ply(ply, v, f)
    print("[gpu] building the 2048^3 icosphere", flush=True)
    t = svo.VoxelOctree.build_from_ply(ply, b["resolution"], threads=b["threads"], memory_budget=b["mem"])
    assert hashlib.md5(t.encode()).hexdigest() == b["md5"]
    return t


def test_ico2048_eight_views_match_oracle(svo, ico2048):
    """The headline workload's render half: 8 seeded 1920x1080 orbit views of
    the GPU-built 2048^3 tree, plain kernel vs the oracle on all 16,588,800
    rays; the batched grid form, a graph replay, the cost-ordered renders of
    the bench's headline (one tile plan per view, three renders each: raster
    order, then the tiles filed by the previous render's costs) and three
    cost-ordered batched launches (one plan over the 8 views' tiles) vs the
    plain kernel."""
    import torch
    dev = torch.device("cuda", 0)
    w, h = 1920, 1080
    words = ico2048.words()
    angles = svo.scenes.orbit_angles(1080, 8)
    cams = [ico2048.orbit_camera(p, y, 1.0) for p, y in angles]
    plain = []
    for k, cam in enumerate(cams):
        o = _outputs(torch, w * h, dev)
        print(f"[gpu] ico2048 view {k} plain render", flush=True)
        ico2048.render_primary_device(cam, w, h, *_ptrs(o))
        torch.cuda.synchronize()
        mine = _host(o)
        d = orc.primary_dirs(np.frombuffer(cam.tform, np.float32), w, h)
        ref = orc.raymarch(words, np.frombuffer(cam.pos, np.float32)[None, :], d, 0.0, threads=16)
        _same(mine, ref, f"view {k} vs oracle")
        assert (mine["hit"] == 1).mean() > 0.1
        plain.append(o)
    ren = svo.svo.Renderer(8)
    outs = [_outputs(torch, w * h, dev) for _ in cams]
    views = [ren.view(ico2048, cam, w, h, *_ptrs(o)) for cam, o in zip(cams, outs)]
    print("[gpu] ico2048 batched grid form", flush=True)
    ren.render_device(views, None, 0, 0, svo.svo.BATCH_GRID)
    torch.cuda.synchronize()
    assert all(_equal_on_device(torch, a, b) for a, b in zip(outs, plain))
    stream = torch.cuda.Stream()
    g = ren.capture(views, stream.cuda_stream)
    for rep in range(2):
        for o in outs:
            o["hit"].fill_(9)
            o["leaf"].fill_(-2)
        torch.cuda.synchronize()
        print(f"[gpu] ico2048 graph replay {rep}", flush=True)
        g.launch(stream.cuda_stream)
        stream.synchronize()
        assert all(_equal_on_device(torch, a, b) for a, b in zip(outs, plain))
    g.close()
    plans = [svo.svo.TilePlan(w, h) for _ in cams]
    for rep in range(3):
        for o in outs:
            o["hit"].fill_(9)
            o["leaf"].fill_(-2)
        torch.cuda.synchronize()
        print(f"[gpu] ico2048 cost-ordered renders, pass {rep}", flush=True)
        for plan, cam, o in zip(plans, cams, outs):
            plan.render_device(ico2048, cam, *_ptrs(o))
        torch.cuda.synchronize()
        assert all(_equal_on_device(torch, a, b) for a, b in zip(outs, plain))
    for plan in plans:
        plan.close()
    bplan = svo.svo.TilePlan.for_batch(views)
    for rep in range(3):
        for o in outs:
            o["hit"].fill_(9)
            o["leaf"].fill_(-2)
        torch.cuda.synchronize()
        print(f"[gpu] ico2048 cost-ordered batch, launch {rep}", flush=True)
        ren.render_device(views, None, 0, 0, svo.svo.BATCH_GRID, plan=bplan)
        torch.cuda.synchronize()
        assert all(_equal_on_device(torch, a, b) for a, b in zip(outs, plain))
    bplan.close()


@pytest.fixture(scope="module")
def mixed(svo):
    print("[gpu] building the 64 mixed octrees", flush=True)
    trees = [svo.scenes.build_mixed_octree(svo, sp) for sp in svo.scenes.mixed_octrees()]
    return trees, svo.scenes.mixed_views()


def test_mixed_batch_matches_plain_and_oracle(svo, mixed):
    """64 octrees x 256 views (147,812,595 rays): grid form, persistent form
    and a graph replay of the batched launch equal the plain per-view
    launches on every ray; every view equals the oracle on every 97th pixel."""
    import torch
    dev = torch.device("cuda", 0)
    trees, specs = mixed
    assert len(trees) == 64 and len(specs) == 256
    cams = [trees[v["tree"]].orbit_camera(v["pitch"], v["yaw"], 1.0) for v in specs]
    plain = []
    for i, v in enumerate(specs):
        o = _outputs(torch, v["width"] * v["height"], dev)
        trees[v["tree"]].render_primary_device(cams[i], v["width"], v["height"], *_ptrs(o))
        plain.append(o)
    torch.cuda.synchronize()
    print("[gpu] mixed: plain launches done; oracle on every 97th pixel", flush=True)
    words = {}
    for i, v in enumerate(specs):
        ti, w, h = v["tree"], v["width"], v["height"]
        if ti not in words:
            words[ti] = trees[ti].words()
        sample = np.arange(i % 97, w * h, 97)
        d = orc.primary_dirs(np.frombuffer(cams[i].tform, np.float32), w, h)[sample]
        ref = orc.raymarch(words[ti], np.frombuffer(cams[i].pos, np.float32)[None, :], d, 0.0, threads=16)
        mine = {k: x[sample] for k, x in _host(plain[i]).items()}
        _same(mine, ref, f"mixed view {i} ({w}x{h}, tree {ti}) vs oracle")
    ren = svo.svo.Renderer(256)
    outs = [_outputs(torch, v["width"] * v["height"], dev) for v in specs]
    views = [ren.view(trees[v["tree"]], cams[i], v["width"], v["height"], *_ptrs(outs[i]))
             for i, v in enumerate(specs)]
    for form, waves, refill in [(svo.svo.BATCH_GRID, 0, 0), (svo.svo.BATCH_PERSISTENT, 32, 64),
                                (svo.svo.BATCH_PERSISTENT, 24, 16)]:
        for o in outs:
            o["hit"].fill_(9)
            o["leaf"].fill_(-2)
        print(f"[gpu] mixed batched form={form} waves={waves} refill={refill}", flush=True)
        ren.render_device(views, None, waves, refill, form)
        torch.cuda.synchronize()
        bad = [i for i in range(len(specs)) if not _equal_on_device(torch, outs[i], plain[i])]
        assert not bad, (form, waves, refill, bad[:5])
    stream = torch.cuda.Stream()
    g = ren.capture(views, stream.cuda_stream)
    for o in outs:
        o["hit"].fill_(9)
        o["leaf"].fill_(-2)
    torch.cuda.synchronize()
    print("[gpu] mixed graph replay", flush=True)
    g.launch(stream.cuda_stream)
    stream.synchronize()
    bad = [i for i in range(len(specs)) if not _equal_on_device(torch, outs[i], plain[i])]
    assert not bad, ("graph", bad[:5])
    g.close()


@pytest.mark.parametrize("persistent", [False, True])
def test_served_stream_over_the_mixed_octrees(svo, mixed, persistent):
    """The served workload through the live server (its default direct and
    batched launches, or with persistent_launch): a seeded Poisson stream at
    2000 req/s of 500 requests (octree U{0..63}, seeded orbit camera,
    resolution as the mixed views) plus a burst of 100 requests at once.
    Every served result equals the request rendered alone by the plain kernel
    on every ray and the oracle on every 97th pixel."""
    trees, _ = mixed
    rng = np.random.default_rng(2000)
    specs = svo.scenes.mixed_views(seed=2000, count=600, trees=len(trees))
    reqs = [(v["tree"], trees[v["tree"]].orbit_camera(v["pitch"], v["yaw"], 1.0), v["width"], v["height"])
            for v in specs]
    arr = np.cumsum(rng.exponential(1.0 / 2000.0, 500))
    arr -= arr[0]
    arr = np.concatenate([arr, np.full(100, arr[-1] + 0.01)])
    budget = sum(w * h for _, _, w, h in reqs) * 16 + (1 << 26)
    srv = svo.svo.Server(trees, output_budget=budget, persistent_launch=persistent)
    print(f"[gpu] served stream: 600 requests, arena {budget >> 20} MiB, persistent={persistent}", flush=True)
    lat, mk, tickets = srv.run_schedule(reqs, arr, keep=True)
    assert np.all(lat > 0) and mk > 0
    words = {}
    for i, ((ti, cam, w, h), tk) in enumerate(zip(reqs, tickets)):
        out = srv.copy_outputs(int(tk), w, h)
        ref = trees[ti].render_primary(cam, w, h)
        _same({"hit": ref["hit"], **out}, ref, f"served request {i} vs alone")
        if ti not in words:
            words[ti] = trees[ti].words()
        sample = np.arange(i % 97, w * h, 97)
        d = orc.primary_dirs(np.frombuffer(cam.tform, np.float32), w, h)[sample]
        o = orc.raymarch(words[ti], np.frombuffer(cam.pos, np.float32)[None, :], d, 0.0, threads=16)
        _same({"hit": o["hit"], **{k: v[sample] for k, v in out.items()}}, o, f"served request {i} vs oracle")
        srv.release(int(tk))
    srv.close()
