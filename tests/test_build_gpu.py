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


def test_grid_builds_match_reference(svo, golden):
    for b in _builds(golden)["grid_builds"]:
        print(f"[gpu] grid build {b['name']} mem={b['mem']}", flush=True)
        vol = np.load(os.path.join(golden, f"grid_{b['name']}.npz"))["vol"]
        t = svo.VoxelOctree.build_from_grid(vol, memory_budget=b["mem"])
        raw = open(os.path.join(golden, b["oct"]), "rb").read()
        assert t.encode() == raw, b["name"]


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
