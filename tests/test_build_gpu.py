"""GPU octree construction parity (libsvo_hip.so through its C ABI).

Bar: byte-identical. The word array (and the .oct bytes written from it)
must equal the reference builder's for the same (mesh or volume, resolution,
memory budget, thread count): checked against the reference's own .oct
outputs (golden fixtures) and, at sizes without fixtures, against the oracle.
"""
import json
import os

import numpy as np
import pytest

import oracle_py as orc

pytestmark = pytest.mark.gpu


def _builds(golden):
    with open(os.path.join(golden, "builds.json")) as fp:
        return json.load(fp)


def _mesh_cases():
    with open(os.path.join(os.path.dirname(__file__), "golden", "builds.json")) as fp:
        return [(b["mesh"], b["resolution"], b["threads"], b["mem"], b["oct"]) for b in json.load(fp)["mesh_builds"]]


def _depth_as_walked(svo, t):
    """A built tree with leaves takes its depth from the builder (no host walk
    of the words); it must equal what the host walk (octree_validate, run for
    trees made from words) reports for the same words."""
    assert t.depth() == svo.VoxelOctree.from_words(t.words(), t.center()).depth()


@pytest.mark.parametrize("mesh,res,threads,mem,octfile", _mesh_cases())
def test_mesh_build_matches_reference(svo, golden, mesh, res, threads, mem, octfile):
    print(f"[gpu] build {mesh} R={res} threads={threads} mem={mem}", flush=True)
    t = svo.VoxelOctree.build_from_ply(os.path.join(golden, f"mesh_{mesh}.ply"), res, threads=threads,
                                       memory_budget=mem)
    raw = open(os.path.join(golden, octfile), "rb").read()
    gc, gw = orc.decode_oct_bytes(raw)
    w = t.words()
    assert len(w) == len(gw)
    assert np.array_equal(w, gw)
    assert np.array_equal(t.center().view(np.uint32), gc.view(np.uint32))
    assert t.encode() == raw
    _depth_as_walked(svo, t)


def test_grid_builds_match_reference(svo, golden):
    for b in _builds(golden)["grid_builds"]:
        print(f"[gpu] grid build {b['name']} mem={b['mem']}", flush=True)
        vol = np.load(os.path.join(golden, f"grid_{b['name']}.npz"))["vol"]
        t = svo.VoxelOctree.build_from_grid(vol, memory_budget=b["mem"])
        raw = open(os.path.join(golden, b["oct"]), "rb").read()
        assert t.encode() == raw, b["name"]
        _depth_as_walked(svo, t)


def test_edge_volumes_match_reference(svo, golden):
    vols = np.load(os.path.join(golden, "grid_edges.npz"))
    for b in _builds(golden)["grid_edges"]:
        print(f"[gpu] edge volume {b['name']}", flush=True)
        t = svo.VoxelOctree.build_from_grid(vols[b["name"]], memory_budget=1 << 30)
        assert t.encode() == open(os.path.join(golden, b["oct"]), "rb").read(), b["name"]
        _depth_as_walked(svo, t)
    empty = svo.VoxelOctree.build_from_grid(vols["empty16"])
    r = empty.render_primary(empty.orbit_camera(20.0, 30.0, 1.0), 64, 48)
    assert not r["hit"].any()
    with pytest.raises(svo.SvoError) as e:
        svo.VoxelOctree.build_from_grid(np.full((1, 1, 1), 7, np.uint32))
    assert e.value.code == svo.svo.ERR_UNSUPPORTED


def test_convert_to_volume_matches_reference(svo, golden, tmp_path):
    """PlyLoader::convertToVolume on the GPU: the .voxel bytes equal the
    reference's for single-slab and multi-slab budgets (the latter carry the
    slab buffer between slabs); a budget below one slice is refused and
    leaves no file."""
    import hashlib
    for c in _builds(golden)["convert"]:
        print(f"[gpu] convert {c}", flush=True)
        out = tmp_path / f"{c['mesh']}_{c['mem']}.voxel"
        svo.svo.convert_to_volume(os.path.join(golden, f"mesh_{c['mesh']}.ply"), c["resolution"], str(out),
                                  threads=c["threads"], memory_budget=c["mem"])
        raw = out.read_bytes()
        assert len(raw) == c["size"] and hashlib.md5(raw).hexdigest() == c["md5"], c
    bad = tmp_path / "small.voxel"
    with pytest.raises(svo.SvoError) as e:
        svo.svo.convert_to_volume(os.path.join(golden, "mesh_ico10.ply"), 64, str(bad), threads=4, memory_budget=1000)
    assert e.value.code == svo.svo.ERR_LIMIT and not bad.exists()


def test_sdf_scene_build_matches_reference(svo, golden):
    b = _builds(golden)["sdf128"]
    print("[gpu] sdf128 grid build", flush=True)
    vol = svo.scenes.sdf_volume(b["n"], seed=b["seed"], compress=orc.compress_material_batch)
    t = svo.VoxelOctree.build_from_grid(vol, memory_budget=b["mem"])
    assert t.encode() == open(os.path.join(golden, "build_sdf128.oct"), "rb").read()


@pytest.mark.parametrize("freq,res,threads,mem", [
    (40, 256, 8, 1 << 30),     # 32,000 triangles, one cache block
    (40, 256, 16, 1 << 23),    # 8 MiB budget: 64^3 cache blocks, top-table nodes
    (24, 512, 8, 1 << 26),     # 64 MiB budget: 128^3 cache blocks
])
def test_mesh_build_matches_oracle(svo, tmp_path, freq, res, threads, mem):
    v, f = svo.scenes.icosphere(freq, seed=freq + res, displacement=0.05)
    ply = str(tmp_path / "m.ply")
    svo.scenes.write_ply(ply, v, f)
    print(f"[gpu] oracle-vs-GPU build ico{freq} R={res} threads={threads} mem={mem}", flush=True)
    c, words, info = orc.Mesh.from_ply(ply).build(res, threads, mem)
    t = svo.VoxelOctree.build_from_ply(ply, res, threads=threads, memory_budget=mem)
    w = t.words()
    assert len(w) == len(words)
    assert np.array_equal(w, words)
    assert np.array_equal(t.center(), c)


def test_built_tree_renders_like_oracle(svo, tmp_path):
    """The GPU-built tree is consumed by the GPU raymarch with the same
    results as the oracle raymarch over the same words."""
    v, f = svo.scenes.icosphere(30, seed=9, displacement=0.05)
    ply = str(tmp_path / "m.ply")
    svo.scenes.write_ply(ply, v, f)
    t = svo.VoxelOctree.build_from_ply(ply, 512, threads=8)
    cam = t.orbit_camera(25.0, 140.0, 1.0)
    out = t.render_primary(cam, 320, 240)
    d = orc.primary_dirs(np.frombuffer(cam.tform, np.float32), 320, 240)
    ref = orc.raymarch(t.words(), np.frombuffer(cam.pos, np.float32)[None, :], d, 0.0)
    for k in ("hit", "normal", "leaf"):
        assert np.array_equal(out[k], ref[k])
    assert (out["hit"] == 1).sum() > 10000


def test_build_errors(svo, tmp_path):
    with pytest.raises(svo.SvoError) as e:
        svo.VoxelOctree.build_from_ply(str(tmp_path / "none.ply"), 64)
    assert e.value.code == svo.svo.ERR_IO
    bad = tmp_path / "bad.ply"
    bad.write_bytes(b"ply\nformat binary_little_endian 1.0\nelement vertex 3\nproperty float x\n"
                    b"property float y\nproperty float z\nend_header\n" + bytes(36))
    with pytest.raises(svo.SvoError) as e:
        svo.VoxelOctree.build_from_ply(str(bad), 64)
    assert e.value.code == svo.svo.ERR_FORMAT
    with pytest.raises(svo.SvoError) as e:
        svo.VoxelOctree.build_from_ply(str(bad), 64, memory_budget=1000)
    v, f = svo.scenes.icosphere(4, seed=1)
    ply = str(tmp_path / "m.ply")
    svo.scenes.write_ply(ply, v, f)
    with pytest.raises(svo.SvoError) as e:
        svo.VoxelOctree.build_from_ply(ply, 64, memory_budget=1000)
    assert e.value.code == svo.svo.ERR_LIMIT


def test_ico2048_full_size_build_matches_reference(svo, golden, tmp_path):
    """Full-size mesh build: 1,997,120-triangle displaced icosphere at
    2048^3, 8 threads' partition, 1 GiB budget; the .oct must hash to the
    reference builder's output for the same PLY."""
    import hashlib
    b = _builds(golden)["ico2048"]
    v, f = svo.scenes.icosphere(b["frequency"], seed=b["seed"], displacement=b["displacement"])
    ply = str(tmp_path / "ico2048.ply")
    svo.scenes.write_ply(ply, v, f)
    assert hashlib.md5(open(ply, "rb").read()).hexdigest() == b["ply_md5"]
    print("[gpu] full-size mesh build 2048^3", flush=True)
    t = svo.VoxelOctree.build_from_ply(ply, b["resolution"], threads=b["threads"], memory_budget=b["mem"])
    w = t.words()
    assert len(w) == b["words"]
    st = orc.octree_stats(w)
    assert st["desc_per_level"] == b["desc_per_level"]
    assert st["leaves"] == b["leaves"] and st["far_words"] == b["far_words"]
    data = t.encode()
    assert len(data) == b["size"]
    assert hashlib.md5(data).hexdigest() == b["md5"]


@pytest.mark.parametrize("key", ["ico4096", "ico8192"])
def test_large_mesh_builds_match_reference(svo, golden, tmp_path, key):
    """The same 1,997,120-triangle mesh built at 4096^3 and 8192^3 (8 threads'
    partition, 1 GiB budget; at 8192^3 about 16x the 2048^3 build's 39 M
    (voxel, triangle) candidates, in one global sort): the .oct must hash to
    the reference builder's output, and the builder's depth must be the
    host walk's."""
    import hashlib
    b = _builds(golden)[key]
    v, f = svo.scenes.icosphere(b["frequency"], seed=b["seed"], displacement=b["displacement"])
    ply = str(tmp_path / "ico.ply")
    svo.scenes.write_ply(ply, v, f)
    print(f"[gpu] large mesh build {b['resolution']}^3", flush=True)
    t = svo.VoxelOctree.build_from_ply(ply, b["resolution"], threads=b["threads"], memory_budget=b["mem"])
    st = t.build_stats()
    assert st["words"] == b["words"] and st["leaves"] == b["leaves"]
    data = t.encode()
    assert len(data) == b["size"]
    assert hashlib.md5(data).hexdigest() == b["md5"]
    w = t.words()
    assert t.depth() == svo.VoxelOctree.from_words(w, t.center()).depth()


@pytest.mark.parametrize("n,seed,mem", [(64, 3, 1 << 30), (256, 4, 1 << 30), (256, 6, 1 << 22)])
def test_sdf_source_matches_oracle_build(svo, n, seed, mem):
    """GPU SDF scene source: its octree equals the oracle's file-source build of
    the very voxels it generates (dense copy from svo_sdf_volume)."""
    prims = svo.scenes.sdf_primitives(seed)
    print(f"[gpu] SDF source n={n} seed={seed} mem={mem}", flush=True)
    vol = svo.svo.sdf_volume(prims, n)
    assert (vol != 0).sum() > 1000
    c, words, _ = orc.build_grid(vol, mem)
    t = svo.VoxelOctree.build_from_sdf(prims, n, memory_budget=mem)
    assert np.array_equal(t.words(), words)
    assert np.array_equal(t.center(), c)


def test_sdf256_build_and_render(svo):
    """SDF scene: 256^3 octree of a seeded 16-primitive SDF scene (built
    through the raw-volume route), 1280x720 primary rays from a seeded orbit
    camera: GPU build == oracle build, GPU render == oracle raymarch."""
    vol = svo.scenes.sdf_volume(256, seed=101, compress=orc.compress_material_batch)
    c, words, _ = orc.build_grid(vol)
    print("[gpu] SDF 256^3 grid build + 1280x720 render", flush=True)
    t = svo.VoxelOctree.build_from_grid(vol)
    assert np.array_equal(t.words(), words)
    (pitch, yaw), = svo.scenes.orbit_angles(101, 1)
    cam = t.orbit_camera(pitch, yaw, 1.0)
    out = t.render_primary(cam, 1280, 720)
    d = orc.primary_dirs(np.frombuffer(cam.tform, np.float32), 1280, 720)
    ref = orc.raymarch(words, np.frombuffer(cam.pos, np.float32)[None, :], d, 0.0)
    for k in ("hit", "normal", "leaf"):
        assert np.array_equal(out[k], ref[k])
    assert np.array_equal(out["t"].view(np.uint32), ref["t"].view(np.uint32))


# ---- fBm heightfield terrain source -------------------------------------------

@pytest.mark.parametrize("n,seed", [(128, 11), (256, 12)])
def test_terrain_source_matches_definition(svo, n, seed):
    """GPU heights and voxels equal the numpy statement of the terrain."""
    print(f"[gpu] terrain source n={n} seed={seed}", flush=True)
    H, h = svo.svo.terrain_heights(n, seed)
    Hr, _, hr = svo.scenes.terrain_heights(n, seed)
    assert np.array_equal(H, Hr)
    assert np.array_equal(h.view(np.uint32), hr.view(np.uint32))
    vol = svo.svo.terrain_volume(n, seed)
    assert np.array_equal(vol, svo.scenes.terrain_volume(n, seed, compress=orc.compress_material_batch))


def test_terrain_builds_match_reference(svo, golden):
    import hashlib
    for b in _builds(golden)["terrain"]:
        print(f"[gpu] terrain build vs reference {b['oct']}", flush=True)
        t = svo.VoxelOctree.build_from_terrain(b["n"], b["seed"], memory_budget=b["mem"])
        c, words = orc.read_oct(os.path.join(golden, b["oct"]))
        assert np.array_equal(t.words(), words), b["oct"]
        assert np.array_equal(t.center().view(np.uint32), c.view(np.uint32))
        assert hashlib.md5(t.encode()).hexdigest() == b["md5"]


@pytest.mark.parametrize("n,seed,mem", [(512, 21, 1 << 24), (512, 22, 1 << 30), (64, 23, 1 << 16)])
def test_terrain_build_matches_oracle(svo, n, seed, mem):
    print(f"[gpu] terrain build vs oracle n={n} seed={seed} mem={mem}", flush=True)
    vol = svo.svo.terrain_volume(n, seed)
    c, words, _ = orc.build_grid(vol, mem)
    del vol
    t = svo.VoxelOctree.build_from_terrain(n, seed, memory_budget=mem)
    assert np.array_equal(t.words(), words)
    assert np.array_equal(t.center(), c)


def _shadow_reference(words, cam, w, h, light, eps, sample):
    """Oracle primary + shadow rays on the pixel indices `sample`."""
    tf = np.frombuffer(cam.tform, np.float32)
    pos = np.frombuffer(cam.pos, np.float32).copy()
    d = orc.primary_dirs(tf, w, h)[sample]
    prim = orc.raymarch(words, pos[None, :], d, 0.0)
    hit = prim["hit"] == 1
    back = (prim["t"][hit] - np.float32(eps)).astype(np.float32)
    o = (pos[None, :] + d[hit] * back[:, None]).astype(np.float32)
    sh = orc.raymarch(words, o, np.broadcast_to(light, o.shape).copy(), 0.0)
    shadow = np.zeros(len(sample), np.uint8)
    shadow[hit] = (sh["hit"] != 0).astype(np.uint8)
    return prim, shadow


# 1283x723: 14,651 tiles, partial edge tiles on both axes
@pytest.mark.parametrize("n,seed,w,h", [(256, 31, 320, 180), (1024, 32, 641, 359), (1024, 33, 1283, 723)])
def test_shadow_rays_match_oracle(svo, n, seed, w, h):
    print(f"[gpu] shadow rays terrain n={n} {w}x{h}", flush=True)
    t = svo.VoxelOctree.build_from_terrain(n, seed)
    pitch, yaw, light = svo.scenes.terrain_view(seed)
    cam = t.orbit_camera(pitch, yaw, 1.0)
    eps = 0.25 / n
    r = t.render_shadow(cam, w, h, light, eps)
    sample = np.arange(w * h)
    prim, shadow = _shadow_reference(t.words(), cam, w, h, light, eps, sample)
    assert np.array_equal(r["hit"], prim["hit"])
    assert np.array_equal(r["normal"], prim["normal"])
    assert np.array_equal(r["t"].view(np.uint32), prim["t"].view(np.uint32))
    assert np.array_equal(r["leaf"], prim["leaf"])
    assert np.array_equal(r["shadow"], shadow)
    hits = r["hit"] == 1
    assert hits.mean() > 0.3 and 0.02 < r["shadow"][hits].mean() < 0.98
    for bad_eps in (float("nan"), float("inf")):  # would move the shadow origins off to NaN/inf
        with pytest.raises(svo.SvoError) as e:
            t.render_shadow(cam, 16, 16, light, bad_eps)
        assert e.value.code == svo.svo.ERR_INVALID


def test_planned_shadow_renders_equal_unplanned(svo):
    """svo_render_shadow_planned_device: primary + shadow renders through one
    tile plan (the first in raster order, the next ones in the cost order the
    previous launch filed from its primary plus shadow passes), over five
    camera moves, give hit, normal, t, leaf and shadow arrays equal to the
    unplanned renders; a plan for another size or render entry is refused."""
    import torch
    n, seed, w, h = 1024, 33, 1283, 723
    t = svo.VoxelOctree.build_from_terrain(n, seed)
    pitch, yaw, light = svo.scenes.terrain_view(seed)
    eps = 0.25 / n
    m = w * h
    plan = svo.svo.TilePlan(w, h)
    stream = torch.cuda.Stream()
    bufs = [[torch.full((m,), fill, dtype=dt, device="cuda") for dt, fill in
             ((torch.uint8, 7), (torch.int32, 7), (torch.float32, -1.0), (torch.int64, 7), (torch.uint8, 7))]
            for _ in range(2)]
    for k in range(5):
        cam = t.orbit_camera(pitch + 2.0 * k, yaw + 9.0 * k, 1.0)
        print(f"[gpu] planned shadow render {k} {w}x{h}", flush=True)
        for j, pl in enumerate((plan, None)):
            hit, nrm, tt, lf, sh = bufs[j]
            t.render_shadow_device(cam, w, h, light, eps, hit.data_ptr(), tt.data_ptr(), sh.data_ptr(), nrm.data_ptr(),
                                   lf.data_ptr(), stream.cuda_stream, plan=pl)
        stream.synchronize()
        for a, b in zip(bufs[0], bufs[1]):
            assert torch.equal(a.view(torch.uint8), b.view(torch.uint8)), k
    cam = t.orbit_camera(pitch, yaw, 1.0)
    for other in (svo.svo.TilePlan(w + 8, h), svo.svo.TilePlan.for_frame(w, h, 16)):
        with pytest.raises(svo.SvoError) as e:
            t.render_shadow_device(cam, w, h, light, eps, bufs[0][0].data_ptr(), bufs[0][2].data_ptr(),
                                   bufs[0][4].data_ptr(), stream=stream.cuda_stream, plan=other)
        assert e.value.code == svo.svo.ERR_INVALID


def test_terrain_full_size(svo):
    """8192^3 terrain (exposed shell of a seeded fBm heightfield, 512^3 cache
    blocks under a 1 GiB budget), built on the GPU; 3840x2160 primary rays plus
    one shadow ray per hit. Size-independent checks: the leaf count equals the
    shell count implied by the GPU heights; sampled heights equal the numpy
    definition; a short vertical probe above sampled columns hits the column's
    top voxel with its word; sampled primary and shadow rays equal the
    oracle's on the GPU-built words."""
    import time
    n, seed = 8192, 8192
    print("[gpu] terrain 8192^3 build", flush=True)
    t0 = time.time()
    tree = svo.VoxelOctree.build_from_terrain(n, seed, memory_budget=1 << 30)
    print(f"[gpu] built in {time.time() - t0:.2f} s", flush=True)
    H, hh = svo.svo.terrain_heights(n, seed)
    rng = np.random.default_rng(5)
    xs, zs = rng.integers(0, n, 4096), rng.integers(0, n, 4096)
    Hr, _, hr = svo.scenes.terrain_heights(n, seed, x=xs, z=zs)
    assert np.array_equal(H[zs, xs], Hr) and np.array_equal(hh[zs, xs].view(np.uint32), hr.view(np.uint32))
    words = tree.words()
    st = orc.octree_stats(words)
    assert st["leaves"] == svo.scenes.terrain_shell_count(H)
    assert len(st["desc_per_level"]) == 13
    # vertical probes: start two voxels above the column top
    f32 = np.float32
    inv = f32(1.0) / f32(n)
    o = np.stack([f32(1.0) + (xs.astype(f32) + f32(0.5)) * inv, f32(1.0) + (Hr + 2).astype(f32) * inv,
                  f32(1.0) + (zs.astype(f32) + f32(0.5)) * inv], axis=1).astype(f32)
    d = np.broadcast_to(np.float32([0.0, -1.0, 0.0]), o.shape).copy()
    g = tree.raymarch_batch(o, d, np.zeros(len(o), np.float32))
    ref = orc.raymarch(words, o, d, 0.0)
    assert np.array_equal(g["hit"], ref["hit"]) and np.array_equal(g["normal"], ref["normal"])
    assert np.all(g["hit"] == 1)
    vol_words = svo.scenes.terrain_volume_words(n, seed, xs, zs, compress=orc.compress_material_batch)
    assert np.array_equal(g["normal"], vol_words)
    # primary + shadow at 3840x2160, sampled against the oracle
    pitch, yaw, light = svo.scenes.terrain_view(seed)
    cam = tree.orbit_camera(pitch, yaw, 1.0)
    w, h = 3840, 2160
    print("[gpu] terrain 4K primary + shadow", flush=True)
    r = tree.render_shadow(cam, w, h, light, 0.25 / n)
    sample = np.sort(rng.choice(w * h, 60000, replace=False))
    prim, shadow = _shadow_reference(words, cam, w, h, light, 0.25 / n, sample)
    assert np.array_equal(r["hit"][sample], prim["hit"])
    assert np.array_equal(r["normal"][sample], prim["normal"])
    assert np.array_equal(r["t"][sample].view(np.uint32), prim["t"].view(np.uint32))
    assert np.array_equal(r["shadow"][sample], shadow)
    assert (r["hit"] == 1).mean() > 0.3


# ---- PLY ingest: device decode of binary bodies, host parser otherwise ---------

def _write_custom_ply(path, verts, faces, vtypes, count_type="uchar", idx_type="int", extra_vertex=None,
                      extra_face=None, extra_elem=None):
    """Binary little-endian PLY with chosen property types. verts: (n, k)
    columns named by vtypes [(name, type)]; faces: list of index lists."""
    np_t = {"char": "i1", "uchar": "u1", "short": "<i2", "ushort": "<u2", "int": "<i4", "uint": "<u4",
            "float": "<f4", "double": "<f8"}
    hdr = ["ply", "format binary_little_endian 1.0"]
    if extra_elem:
        hdr += [f"element {extra_elem[0]} {extra_elem[1]}", "property int tag"]
    hdr.append(f"element vertex {len(verts)}")
    vfields = list(vtypes) + ([extra_vertex] if extra_vertex else [])
    hdr += [f"property {t} {n}" for n, t in vfields]
    hdr.append(f"element face {len(faces)}")
    if extra_face:
        hdr.append(f"property {extra_face[1]} {extra_face[0]}")
    hdr += [f"property list {count_type} {idx_type} vertex_indices", "end_header"]
    body = bytearray()
    if extra_elem:
        body += np.arange(extra_elem[1], dtype="<i4").tobytes()
    rec = np.zeros(len(verts), dtype=[(n, np_t[t]) for n, t in vfields])
    for j, (n, t) in enumerate(vtypes):
        rec[n] = verts[:, j]
    body += rec.tobytes()
    for k, f in enumerate(faces):
        if extra_face:
            body += np.array([k], np_t[extra_face[1]]).tobytes()
        body += np.array([len(f)], np_t[count_type]).tobytes() + np.array(f, np_t[idx_type]).tobytes()
    with open(path, "wb") as fp:
        fp.write(("\n".join(hdr) + "\n").encode() + bytes(body))


@pytest.mark.parametrize("case", ["types", "quads", "extra"])
def test_ply_ingest_matches_oracle(svo, tmp_path, case):
    """Binary PLY variants through svo_build_from_ply against the oracle's own
    parse + build: other vertex property types and an extra vertex property
    (device decode), faces of mixed sizes (host parser), an element before the
    vertices and a face property before the index list (device decode)."""
    v, f = svo.scenes.icosphere(20, seed=77, displacement=0.05)
    v = v.astype(np.float64)
    rng = np.random.default_rng(1)
    faces = [list(x) for x in f]
    if case == "types":
        cols = np.concatenate([v, rng.integers(0, 256, (len(v), 3))], axis=1)
        vt = [("x", "double"), ("y", "double"), ("z", "double"), ("red", "uchar"), ("green", "uchar"), ("blue", "uchar")]
        kw = dict(count_type="uchar", idx_type="uint", extra_vertex=("confidence", "float"))
    elif case == "quads":
        cols, vt = v, [("x", "float"), ("y", "float"), ("z", "float")]
        for k in range(0, len(faces) - 1, 7):  # merge some triangle pairs into 4- and 5-gons
            faces[k] = faces[k] + [faces[k + 1][2]] + ([faces[k + 1][1]] if k % 2 else [])
        kw = dict(count_type="int", idx_type="int")
    else:
        nrm = v / np.linalg.norm(v, axis=1, keepdims=True)
        cols = np.concatenate([v, nrm], axis=1)
        vt = [("x", "float"), ("y", "float"), ("z", "float"), ("nx", "float"), ("ny", "float"), ("nz", "float")]
        kw = dict(count_type="uchar", idx_type="int", extra_face=("flags", "ushort"), extra_elem=("camera", 3))
    ply = str(tmp_path / f"{case}.ply")
    _write_custom_ply(ply, cols, faces, vt, **kw)
    print(f"[gpu] PLY ingest {case}", flush=True)
    c, words, _ = orc.Mesh.from_ply(ply).build(256, 8, 1 << 30)
    t = svo.VoxelOctree.build_from_ply(ply, 256, threads=8)
    assert np.array_equal(t.words(), words)
    assert np.array_equal(t.center(), c)


def test_ply_ingest_rejects_bad_indices(svo, tmp_path):
    v, f = svo.scenes.icosphere(4, seed=1)
    faces = [list(x) for x in f]
    faces[5][1] = len(v) + 3
    ply = str(tmp_path / "bad.ply")
    _write_custom_ply(ply, v, faces, [("x", "float"), ("y", "float"), ("z", "float")])
    with pytest.raises(svo.SvoError) as e:
        svo.VoxelOctree.build_from_ply(ply, 64)
    assert e.value.code == svo.svo.ERR_FORMAT


def test_build_stats_are_consistent(svo, tmp_path):
    """The counters the construction roofline is computed from."""
    v, f = svo.scenes.icosphere(12, seed=3, displacement=0.05)
    ply = str(tmp_path / "m.ply")
    svo.scenes.write_ply(ply, v, f)
    print("[gpu] build stats", flush=True)
    t = svo.VoxelOctree.build_from_ply(ply, 128, threads=8)
    st = t.build_stats()
    words = t.words()
    kinds = orc.octree_stats(words)
    assert st["triangles"] == len(f) and st["words"] == len(words) and st["leaves"] == kinds["leaves"]
    assert st["candidates"] >= st["leaves"] > 0 and st["device_ms"] > 0 and st["peak_bytes"] > 0
    assert st["algorithmic_bytes"] == 36 * st["triangles"] + 24 * st["candidates"] + 4 * st["words"]
    g = svo.VoxelOctree.build_from_terrain(64, 5)
    gs = g.build_stats()
    assert gs["triangles"] == 0 and gs["candidates"] == gs["leaves"] == orc.octree_stats(g.words())["leaves"]
    loaded = svo.VoxelOctree.from_words(words, t.center())
    assert loaded.build_stats()["words"] == 0 and loaded.build_stats()["peak_bytes"] == 0
