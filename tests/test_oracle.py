"""Pins the oracle (oracle/svo_oracle.c) before anything is checked against it.

What pins what:
- decode + word layout: the shipped dragon fixture's decoded facts (header,
  119,887 words, descriptors per level, leaves, far words).
- raymarch (hit, t, normal) and its leaf index: reference outputs for the same
  rays on the dragon and on a reference-built SDF octree (golden npz; the leaf
  index comes from a "tagged" copy whose leaf words are their own indices).
- raymarch, independently of the reference: a closed-form slab entry on a
  one-level solid octree, a brute-force voxel ray cast on small random grids,
  misses outside the root box, no LOD exit at rayScale 0, thread determinism.
- camera matrices, material codec, triBoxOverlap: reference outputs (golden npz).
- builder (PLY mesh and raw-volume sources, incl. multi-block budgets, odd
  depths and far pointers): reference .oct outputs for fixed (resolution,
  threads, budget); plus the invariant that the emitted leaves are exactly
  the voxelization when no hazard applies.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle_py as orc


def _eq_bits(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


def test_dragon_layout(golden):
    c, w = orc.read_oct(os.path.join(golden, "dragon.oct"))
    assert list(c) == [0.5, 0.2265625, 0.33203125]
    assert len(w) == 119887
    st = orc.octree_stats(w)
    assert st["desc_per_level"] == [1, 3, 15, 68, 289, 1237, 5352, 22191]
    assert st["leaves"] == 90707
    assert st["far_words"] == 24
    assert st["depth"] == 9


@pytest.mark.parametrize("name", ["dragon", "sdf128"])
def test_raymarch_matches_reference(golden, name):
    g = np.load(os.path.join(golden, f"raymarch_{name}.npz"))
    path = os.path.join(golden, "dragon.oct" if name == "dragon" else "build_sdf128.oct")
    _, words = orc.read_oct(path)
    rays = g["rays"]
    r = orc.raymarch(words, rays[:, 0:3], rays[:, 3:6], rays[:, 6])
    assert np.array_equal(r["hit"] != 0, g["hit"] != 0)
    leaf = r["hit"] == 1
    assert leaf.sum() > 1000 and (r["hit"] == 2).sum() > 10
    assert np.array_equal(r["normal"][leaf], g["normal"][leaf])
    hit = r["hit"] != 0
    assert _eq_bits(r["t"][hit], g["t"][hit])
    assert np.array_equal(r["leaf"][leaf], g["leaf"][leaf].astype(np.uint64))


def _nan_rays():
    """Rays the reference never ends: a NaN origin or direction component, or
    an infinite origin on an axis whose direction is infinite (-0 * inf)."""
    o = np.float32([[np.nan, 1.5, 1.5], [1.5, 1.5, 1.5], [np.inf, 0.5, 0.5], [0.2, -np.inf, 0.3],
                    [1.2, 1.3, np.nan]])
    d = np.float32([[0.0, 0.0, 1.0], [np.nan, 1.0, 0.0], [-np.inf, 0.1, 0.2], [0.1, np.inf, 0.0],
                    [np.nan, np.nan, np.nan]])
    return o, d


def test_extreme_rays_match_reference(golden):
    """Signed zeros, tiny, huge and infinite components and rayScales (golden
    from the reference binary); rays with a NaN coefficient are misses."""
    g = np.load(os.path.join(golden, "raymarch_extremes.npz"))
    _, words = orc.read_oct(os.path.join(golden, "dragon.oct"))
    rays = g["rays"]
    r = orc.raymarch(words, rays[:, 0:3], rays[:, 3:6], rays[:, 6])
    assert np.array_equal(r["hit"] != 0, g["hit"] != 0)
    leaf = r["hit"] == 1
    assert leaf.sum() > 100
    assert np.array_equal(r["normal"][leaf], g["normal"][leaf])
    hit = r["hit"] != 0
    assert _eq_bits(r["t"][hit], g["t"][hit])
    assert np.array_equal(r["leaf"][leaf], g["leaf"][leaf].astype(np.uint64))
    o, d = _nan_rays()
    r = orc.raymarch(words, o, d, 0.0)
    assert not r["hit"].any()


@pytest.mark.parametrize("name", ["dragon", "sdf128"])
def test_frame_matches_reference(golden, name):
    """Viewer frames (src/Main.cpp:81-202, 16 stripes) against the reference
    driver's frames, which run on the reference's own raymarch, Mat4, Vec3,
    invSqrt and decompressMaterial."""
    g = np.load(os.path.join(golden, f"frame_{name}.npz"))
    center, words = orc.read_oct(os.path.join(golden, "dragon.oct" if name == "dragon" else "build_sdf128.oct"))
    w, h = (int(v) for v in g["size"])
    for i, (pitch, yaw, radius) in enumerate(g["cams"]):
        tf, _ = orc.orbit_camera(pitch, yaw, radius, center)
        for stride in (1, 3):
            rgba, depth = orc.render_frame(words, tf, center, w, h, 16, stride)
            assert np.array_equal(rgba, g[f"rgba_s{stride}"][i]), (name, i, stride)
            assert _eq_bits(depth, g[f"depth_s{stride}"][i]), (name, i, stride)


def test_frame_invariants(golden):
    """Properties of the frame the formulas fix: every traced pixel is opaque
    grey; stride-3 pixels copy the stride corner of their 8x8 tile, and those
    corners carry the stride-1 value of the same ray; tiles whose four corner
    depths all miss stay 0; the stride does not change the coarse pass."""
    center, words = orc.read_oct(os.path.join(golden, "dragon.oct"))
    w, h, stripes = 203, 97, 4
    tf, _ = orc.orbit_camera(15.0, 70.0, 1.0, center)
    s1, d1 = orc.render_frame(words, tf, center, w, h, stripes, 1)
    s3, d3 = orc.render_frame(words, tf, center, w, h, stripes, 3)
    assert np.array_equal(d1.view(np.uint32), d3.view(np.uint32))
    for img in (s1, s3):
        lit = img != 0
        assert lit.mean() > 0.2
        px = img[lit]
        assert np.all(px >> 24 == 0xFF)
        assert np.all((px & 0xFF) == ((px >> 8) & 0xFF)) and np.all((px & 0xFF) == ((px >> 16) & 0xFF))
    sh = (h - 1) // stripes + 1
    ys, xs = np.mgrid[0:h, 0:w]
    y0 = (ys // sh) * sh
    ty0 = y0 + ((ys - y0) // 8) * 8
    tx0 = (xs // 8) * 8
    cy = ys - (ys - ty0) % 3
    cx = xs - (xs - tx0) % 3
    assert np.array_equal(s3, s3[cy, cx])
    corner = (cy == ys) & (cx == xs)
    assert np.array_equal(s3[corner], s1[corner])
    # coarse corner rays: depth table row-major per stripe, tilesX = (w-1)//8 + 2
    tiles_x = (w - 1) // 8 + 2
    base = 0
    f = np.float32
    for i in range(stripes):
        sy0, sy1 = i * sh, min((i + 1) * sh, h)
        tiles_y = (sy1 - sy0 - 1) // 8 + 2
        d = d1[base:base + tiles_x * tiles_y].reshape(tiles_y, tiles_x)
        for y in range(1, tiles_y):
            for x in range(1, tiles_x):
                quad = d[y - 1:y + 1, x - 1:x + 1]
                t0, t1 = (y - 1) * 8 + sy0, (x - 1) * 8
                if np.all(quad == f(1e10)):
                    assert not s1[t0:min(t0 + 8, sy1), t1:min(t1 + 8, w)].any()
        base += tiles_x * tiles_y
    assert base == len(d1)
    assert 0 < (d1 < f(1e10)).mean() < 1


def test_convert_to_volume_matches_reference(golden):
    """PlyLoader::convertToVolume's .voxel bytes (reference driver: the
    function's slab loop on the reference's own processBlock). The small
    budgets split the volume into 9-12-slice slabs; their bytes depend on the
    slab buffer and counts being reused without clearing."""
    for c in _builds(golden)["convert"]:
        m = orc.Mesh.from_ply(os.path.join(golden, f"mesh_{c['mesh']}.ply"))
        data, slice_z = m.convert_to_volume(c["resolution"], c["threads"], c["mem"])
        assert len(data) == c["size"], c
        assert hashlib.md5(data).hexdigest() == c["md5"], c
        assert (slice_z < int(np.frombuffer(data[8:12], np.int32)[0])) == (c["mem"] < 1 << 30), c


def test_edge_volumes_match_reference(golden):
    """No voxel, one voxel, a voxel only on the unscanned last slice of an odd
    depth, a full volume; a one-voxel volume (on which the reference crashes)
    is refused."""
    vols = np.load(os.path.join(golden, "grid_edges.npz"))
    for b in _builds(golden)["grid_edges"]:
        c, words, _ = orc.build_grid(vols[b["name"]], 1 << 30)
        rc, rwords = orc.read_oct(os.path.join(golden, b["oct"]))
        assert np.array_equal(words, rwords) and _eq_bits(c, rc), b["name"]
    assert len(orc.read_oct(os.path.join(golden, "build_grid_empty16.oct"))[1]) == 1
    with pytest.raises(RuntimeError):
        orc.build_grid(np.full((1, 1, 1), 7, np.uint32))


def test_terrain_builds_match_reference(golden, svo):
    """fBm terrain volumes (exposed shell) through the oracle's file-source
    builder equal the reference builder's .oct for the same volume."""
    for b in _builds(golden)["terrain"]:
        vol = svo.scenes.terrain_volume(b["n"], b["seed"], compress=orc.compress_material_batch)
        assert hashlib.md5(vol.tobytes()).hexdigest() == b["volume_md5"]
        assert int((vol != 0).sum()) == b["voxels"]
        c, words, _ = orc.build_grid(vol, b["mem"])
        rc, rwords = orc.read_oct(os.path.join(golden, b["oct"]))
        assert np.array_equal(words, rwords), b["oct"]
        assert _eq_bits(c, rc)


def test_terrain_shell_invariants(svo):
    """The exposed shell is what the solid heightfield implies: every solid
    voxel with an empty 6-neighbour (outside the grid counts as solid) is in
    the shell, and nothing else is."""
    n = 64
    H, _, _ = svo.scenes.terrain_heights(n, 5)
    vol = svo.scenes.terrain_volume(n, 5)
    y = np.arange(n)[None, :, None]
    solid = y <= H[:, None, :]  # [z, y, x]
    pad = np.pad(solid, 1, constant_values=True)
    exposed = solid & ~(pad[2:, 1:-1, 1:-1] & pad[:-2, 1:-1, 1:-1] & pad[1:-1, 2:, 1:-1] & pad[1:-1, :-2, 1:-1]
                        & pad[1:-1, 1:-1, 2:] & pad[1:-1, 1:-1, :-2])
    assert np.array_equal(vol != 0, exposed)
    assert 0 < H.min() and H.max() < n


def test_camera_matches_reference(golden):
    g = np.load(os.path.join(golden, "camera.npz"))
    for inp, tf in zip(g["inputs"], g["tform"]):
        mine, _ = orc.orbit_camera(inp[0], inp[1], inp[2], inp[3:6])
        assert _eq_bits(mine, tf)


def test_material_codec_matches_reference(golden):
    g = np.load(os.path.join(golden, "material.npz"))
    inp = g["inputs"]
    w = orc.compress_material_batch(inp[:, 0:3], inp[:, 3])
    assert np.array_equal(w, g["word"])
    n, s = orc.decompress_material_batch(w)
    dec = np.concatenate([n, s[:, None]], axis=1)
    assert _eq_bits(dec, g["decoded"])


def test_tribox_matches_reference(golden):
    g = np.load(os.path.join(golden, "tribox.npz"))
    assert np.array_equal(orc.tribox_batch(g["inputs"]), g["overlap"])
    assert 0.1 < g["overlap"].mean() < 0.9


def _builds(golden):
    with open(os.path.join(golden, "builds.json")) as fp:
        return json.load(fp)


def _mesh_cases():
    with open(os.path.join(os.path.dirname(__file__), "golden", "builds.json")) as fp:
        return [(b["mesh"], b["resolution"], b["threads"], b["mem"], b["oct"]) for b in json.load(fp)["mesh_builds"]]


@pytest.mark.parametrize("mesh,res,threads,mem,octfile", _mesh_cases())
def test_mesh_build_matches_reference(golden, mesh, res, threads, mem, octfile):
    m = orc.Mesh.from_ply(os.path.join(golden, f"mesh_{mesh}.ply"))
    c, words, info = m.build(res, threads, mem)
    gc, gw = orc.read_oct(os.path.join(golden, octfile))
    assert _eq_bits(c, gc)
    assert len(words) == len(gw)
    assert np.array_equal(words, gw)


def test_grid_builds_match_reference(golden):
    for b in _builds(golden)["grid_builds"]:
        vol = np.load(os.path.join(golden, f"grid_{b['name']}.npz"))["vol"]
        c, words, info = orc.build_grid(vol, b["mem"])
        gc, gw = orc.read_oct(os.path.join(golden, b["oct"]))
        assert _eq_bits(c, gc), b
        assert np.array_equal(words, gw), b
        # the fixtures put voxels on the last slice of odd-depth volumes
        if vol.shape[0] % 2 == 1:
            assert info["hazard_voxels"] > 0


def test_sdf_scene_build_matches_reference(golden, svo):
    import hashlib
    b = _builds(golden)["sdf128"]
    vol = svo.scenes.sdf_volume(b["n"], seed=b["seed"], compress=orc.compress_material_batch)
    assert hashlib.md5(vol.tobytes()).hexdigest() == b["volume_md5"]
    c, words, _ = orc.build_grid(vol, b["mem"])
    gc, gw = orc.read_oct(os.path.join(golden, "build_sdf128.oct"))
    assert np.array_equal(words, gw)
    # the numpy packer used by the scene generator agrees with the oracle codec
    rng = np.random.default_rng(3)
    n = rng.normal(size=(5000, 3)).astype(np.float32)
    assert np.array_equal(svo.scenes.compress_material_np(n, 0.8), orc.compress_material_batch(n, 0.8))


def test_mesh_leaves_equal_voxelization(golden):
    """Invariant (no hazard applies to these meshes): the leaves the DFS emits
    are exactly the non-zero voxels of the block voxelization."""
    for mesh, res, threads in [("ico10", 64, 8), ("ico16n", 128, 2), ("ico6a", 64, 8)]:
        m = orc.Mesh.from_ply(os.path.join(golden, f"mesh_{mesh}.ply"))
        vol = m.voxelize(res, threads)
        _, words, info = m.build(res, threads)
        kind = orc.octree_kinds(words)
        leaves = np.sort(words[kind == 3])
        assert np.array_equal(leaves, np.sort(vol[vol != 0]))


# ---- closed forms / brute force ------------------------------------------------

def test_solid_single_level_octree_slab_entry():
    """Root with 8 leaf children (a solid 2^3 grid over [1,2]^3): t is the slab
    entry of the root box and the leaf is the octant holding the entry point."""
    words = np.zeros(9, np.uint32)
    words[0] = (0xFF << 8) | (1 << 18)
    words[1:] = np.arange(1, 9, dtype=np.uint32) * 100  # storage order = ascending Morton
    rng = np.random.default_rng(0)
    n = 20000
    o = rng.uniform(-2.0, 5.0, (n, 3)).astype(np.float32)
    tgt = rng.uniform(1.05, 1.95, (n, 3)).astype(np.float32)
    d = (tgt - o).astype(np.float64)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d.astype(np.float32)
    out = orc.raymarch(words, o, d)
    o64, d64 = o.astype(np.float64), d.astype(np.float64)
    t0 = (1.0 - o64) / d64
    t1 = (2.0 - o64) / d64
    tent = np.maximum(np.minimum(t0, t1).max(axis=1), 0.0)
    outside = np.any((o < 1) | (o > 2), axis=1)
    sel = outside
    assert np.all(out["hit"][sel] == 1)
    assert np.allclose(out["t"][sel], tent[sel], rtol=2e-5, atol=2e-5)
    p = o64 + d64 * tent[:, None]
    m = ((p[:, 0] >= 1.5).astype(int) | ((p[:, 1] >= 1.5).astype(int) << 1) | ((p[:, 2] >= 1.5).astype(int) << 2))
    away = np.all(np.abs(p - 1.5) > 1e-4, axis=1) & sel
    assert np.array_equal(out["normal"][away], (m[away] + 1).astype(np.uint32) * 100)


def test_misses_outside_root_box():
    _, words = orc.read_oct(os.path.join(os.path.dirname(__file__), "golden", "dragon.oct"))
    rng = np.random.default_rng(1)
    n = 5000
    o = rng.uniform(3.0, 4.0, (n, 3)).astype(np.float32)
    d = rng.uniform(0.1, 1.0, (n, 3)).astype(np.float32)  # pointing away from [1,2]^3
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    assert np.all(orc.raymarch(words, o, d)["hit"] == 0)


def test_brute_force_voxel_cast():
    """Random occupancy on 8^3 grids: the first voxel hit equals a float64
    slab-test brute force over all occupied voxels (rays with a near tie
    between two voxel entries are skipped)."""
    rng = np.random.default_rng(2)
    for trial in range(4):
        n = 8
        occ = rng.random((n, n, n)) < 0.15
        zz, yy, xx = np.nonzero(occ)
        vol = np.zeros((n, n, n), np.uint32)
        vol[zz, yy, xx] = (xx + n * (yy + n * zz) + 1).astype(np.uint32)
        _, words, _ = orc.build_grid(vol)
        k = 3000
        o = rng.uniform(-1.0, 4.0, (k, 3)).astype(np.float32)
        tgt = rng.uniform(1.0, 2.0, (k, 3)).astype(np.float32)
        d = (tgt - o).astype(np.float64)
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        d = d.astype(np.float32)
        out = orc.raymarch(words, o, d)
        lo = np.stack([xx, yy, zz], 1) / n + 1.0
        hi = lo + 1.0 / n
        o64, d64 = o.astype(np.float64), d.astype(np.float64)
        checked = 0
        for i in range(k):
            with np.errstate(divide="ignore", invalid="ignore"):
                a = (lo - o64[i]) / d64[i]
                b = (hi - o64[i]) / d64[i]
            tn = np.minimum(a, b).max(axis=1)
            tf = np.maximum(a, b).min(axis=1)
            ok = (tn <= tf) & (tf >= 0)
            if not ok.any():
                assert out["hit"][i] == 0
                checked += 1
                continue
            te = np.where(ok, np.maximum(tn, 0), np.inf)
            order = np.argsort(te)
            if len(order) > 1 and te[order[1]] - te[order[0]] < 1e-5:
                continue
            assert out["hit"][i] == 1
            assert out["normal"][i] == vol[zz[order[0]], yy[order[0]], xx[order[0]]]
            assert abs(out["t"][i] - te[order[0]]) < 1e-4
            checked += 1
        assert checked > 0.9 * k


def test_no_lod_exit_at_zero_scale_and_thread_determinism(golden):
    _, words = orc.read_oct(os.path.join(golden, "dragon.oct"))
    for pitch, yaw in [(20.0, 35.0), (-40.0, 200.0)]:
        tf, pos = orc.orbit_camera(pitch, yaw, 1.0, [0.5, 0.2265625, 0.33203125])
        d = orc.primary_dirs(tf, 160, 90)
        a = orc.raymarch(words, pos[None, :], d, 0.0, threads=1)
        b = orc.raymarch(words, pos[None, :], d, 0.0, threads=7)
        assert not np.any(a["hit"] == 2)
        for k in ("hit", "normal", "leaf"):
            assert np.array_equal(a[k], b[k])
        assert _eq_bits(a["t"], b["t"])
        assert np.array_equal(a["counts"], b["counts"])


def test_inv_sqrt_error_bound():
    """The one-step Quake estimate stays within its published 0.175% bound."""
    x = np.float32(np.geomspace(1e-6, 1e6, 4001))
    y = np.array([orc.inv_sqrt(v) for v in x], np.float64)
    rel = np.abs(y * np.sqrt(x.astype(np.float64)) - 1.0)
    assert rel.max() < 1.8e-3
