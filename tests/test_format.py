"""CPU tests of the product library's host side: the C ABI loads and exports
every declared symbol, .oct decode/encode is byte-identical to the reference
(every reference-written .oct fixture re-encodes to its own bytes; a 2-slice
stream matches the reference encoder's md5), errors are reported, not fatal,
and the host camera matches the reference's matrices."""
import ctypes
import hashlib
import json
import os
import re

import numpy as np
import pytest

import oracle_py as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))



def _device_count_or_none():
    """Visible HIP devices as the library counts them (None if it cannot load)."""
    try:
        import conftest
        return conftest.load_package().svo.device_count()
    except Exception:
        return None

def test_library_exports_every_declared_symbol(svo):
    header = open(os.path.join(ROOT, "include", "svo.h")).read()
    declared = set(re.findall(r"\b(svo_[a-z_0-9]+)\s*\(", header))
    assert declared == set(svo.svo.EXPORTS)
    lib = ctypes.CDLL(svo.svo.LIB_PATH)
    for name in declared:
        assert hasattr(lib, name), name


def test_dragon_reencode_matches_reference(svo, golden):
    lz = json.load(open(os.path.join(golden, "lz4.json")))
    t = svo.VoxelOctree(os.path.join(golden, "dragon.oct"))
    data = t.encode()
    assert len(data) == lz["dragon_resave"]["size"]
    assert hashlib.md5(data).hexdigest() == lz["dragon_resave"]["md5"]
    c, w = orc.read_oct(os.path.join(golden, "dragon.oct"))
    assert np.array_equal(t.words(), w)
    assert np.array_equal(t.center(), c)
    assert t.depth() == 9


def test_reference_octs_reencode_to_same_bytes(svo, golden, tmp_path):
    files = [f for f in os.listdir(golden) if f.startswith("build_") and f.endswith(".oct")]
    assert len(files) >= 10
    for f in files:
        raw = open(os.path.join(golden, f), "rb").read()
        t = svo.VoxelOctree.from_bytes(raw)
        assert t.encode() == raw, f
        out = tmp_path / f
        t.save(str(out))
        assert out.read_bytes() == raw


def _stream_words():
    n = 17 * (1 << 20) + 12345
    i = np.arange(n, dtype=np.uint64)
    w = (i * np.uint64(2654435761)) >> np.uint64(7)
    w = np.where((i % np.uint64(5)) == 0, w, i // np.uint64(3))
    w = np.where((i % np.uint64(97)) < 40, np.uint64(0), w)
    return (w & np.uint64(0xFFFFFFFF)).astype(np.uint32)


def test_two_slice_stream_matches_reference_encoder(svo, golden):
    """A word array spanning two 64 MiB slices: the second slice's matches
    reach into the first and the encoder state carries across slices."""
    lz = json.load(open(os.path.join(golden, "lz4.json")))["two_chunk_stream"]
    words = _stream_words()
    assert len(words) == lz["words"]
    enc = svo.svo.encode_words(words, (0.5, 0.25, 0.125))
    assert len(enc) == lz["size"]
    assert hashlib.md5(enc).hexdigest() == lz["md5"]
    c, w = orc.decode_oct_bytes(enc)
    assert np.array_equal(w, words)
    # the library's reader decodes the slices in parallel
    c2, w2 = svo.svo.decode_words(enc)
    assert np.array_equal(w2, words) and np.array_equal(c2, c)


def test_parallel_decode_of_dictionary_chains(svo):
    """Slices decoded before their dictionary is final: (a) a periodic stream
    whose every match in slices 1..3 chains back into the previous slice (the
    whole slice is re-decoded); (b) noise with periodic islands that straddle
    slice boundaries (short chains)."""
    n = 3 * (1 << 24) + 777  # four slices
    i = np.arange(n, dtype=np.uint64)
    per = ((i % np.uint64(3)) * np.uint64(0x9E3779B1) + np.uint64(12345)) & np.uint64(0xFFFFFFFF)
    per = per.astype(np.uint32)
    rng = np.random.default_rng(3)
    noise = rng.integers(0, 1 << 32, n, dtype=np.uint32)
    islands = ((i >> np.uint64(10)) % np.uint64(7)) == 0
    for words in (per, np.where(islands, per, noise).astype(np.uint32)):
        enc = svo.svo.encode_words(words, (0.5, 0.5, 0.5))
        c, w = svo.svo.decode_words(enc)
        assert np.array_equal(w, words)
        _, wo = orc.decode_oct_bytes(enc)
        assert np.array_equal(wo, words)


def test_load_errors_are_reported(svo, golden, tmp_path):
    with pytest.raises(svo.SvoError) as e:
        svo.VoxelOctree(str(tmp_path / "missing.oct"))
    assert e.value.code == svo.svo.ERR_IO
    raw = open(os.path.join(golden, "dragon.oct"), "rb").read()
    for bad in (raw[:10], raw[:30], raw[:-100], raw[:24] + bytes(len(raw) - 24)):
        with pytest.raises(svo.SvoError) as e:
            svo.VoxelOctree.from_bytes(bad)
        assert e.value.code == svo.svo.ERR_FORMAT
    # a child reference past the end of the array is rejected before any GPU use
    w = np.array([(0xFF << 8) | (0xFF) | (1 << 18)] + [(100 << 18) | (1 << 8) | 1] * 8, np.uint32)
    with pytest.raises(svo.SvoError) as e:
        svo.VoxelOctree.from_words(w, [0.5, 0.5, 0.5])
    assert e.value.code == svo.svo.ERR_FORMAT



def test_corrupt_streams_load_valid_trees_or_fail_cleanly(svo, golden):
    """Random corruptions of a reference .oct (byte flips in the header, the
    slice header and the LZ4 body, overwritten runs, truncations) either
    decode to a tree the structural check accepts or are refused with
    ERR_FORMAT; none may crash the process."""
    raw = np.frombuffer(open(os.path.join(golden, "dragon.oct"), "rb").read(), np.uint8)
    rng = np.random.default_rng(2024)
    loaded = refused = 0
    for i in range(300):
        b = raw.copy()
        kind = i % 4
        if kind == 0:  # a few flipped bits in the first 4 KiB
            for _ in range(int(rng.integers(1, 5))):
                b[int(rng.integers(0, 4096))] ^= np.uint8(1 << int(rng.integers(0, 8)))
        elif kind == 1:  # flipped bits anywhere
            for _ in range(int(rng.integers(1, 9))):
                b[int(rng.integers(0, len(b)))] ^= np.uint8(1 << int(rng.integers(0, 8)))
        elif kind == 2:  # a random run overwritten
            at, k = int(rng.integers(0, len(b) - 64)), int(rng.integers(1, 65))
            b[at:at + k] = rng.integers(0, 256, k, dtype=np.uint8)
        else:  # truncated
            b = b[:int(rng.integers(0, len(b)))]
        try:
            t = svo.VoxelOctree.from_bytes(b.tobytes())
        except svo.SvoError as e:
            assert e.code == svo.svo.ERR_FORMAT, (i, e.code, str(e))
            refused += 1
            continue
        assert 1 <= t.depth() <= 23 and t.word_count() >= 1
        # an accepted tree is one the traversal walks without leaving the array
        orc.raymarch(t.words(), *_probe_rays(rng, 64))
        loaded += 1
    assert refused > 100 and loaded > 10, (loaded, refused)


def _probe_rays(rng, n):
    """n rays from random points around the unit cube [1,2]^3 toward random
    points inside it (origins, directions)."""
    o = rng.uniform(0.2, 2.8, (n, 3)).astype(np.float32)
    tgt = rng.uniform(1.0, 2.0, (n, 3)).astype(np.float32)
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o, d.astype(np.float32)


def test_descriptor_masks_are_checked_where_the_traversal_reads():
    """The traversal finds a non-leaf child m at base + popc(non-leaf bits below
    m) * (1 + large) and a leaf child at base + popc(valid bits below m)
    (src/VoxelOctree.cpp:282,290-293). Masks the reference builder never writes
    (non-leaf neither 0 nor equal to valid, non-leaf bits without a valid bit)
    are checked at those addresses: here both send the traversal far outside
    the array and must be refused."""
    import conftest
    svo = conftest.load_package()
    # child 1 (non-leaf) sits at base + 0, the leaf word 0xFFFFFFFF, read as a
    # far descriptor pointing ~2^46 words away
    with pytest.raises(svo.SvoError) as e:
        svo.VoxelOctree.from_words(np.uint32([0x00040302, 0xFFFFFFFF, 0x00040100, 0x00808080]), [0.5] * 3)
    assert e.value.code == svo.svo.ERR_FORMAT
    # a non-leaf bit without its valid bit shifts child 1 to base + 1 (word 2, a leaf word)
    with pytest.raises(svo.SvoError) as e:
        svo.VoxelOctree.from_words(np.uint32([(1 << 18) | (0x02 << 8) | 0x03, (1 << 18) | (0x01 << 8), 0x00808080]),
                                   [0.5] * 3)
    assert e.value.code == svo.svo.ERR_FORMAT
    # the same trees with the masks the reference writes are accepted
    assert svo.VoxelOctree.from_words(np.uint32([(1 << 18) | (0x02 << 8) | 0x02, (1 << 18) | (0x01 << 8), 0x00808080]),
                                      [0.5] * 3).depth() == 3


def test_mutated_descriptor_masks_load_traceable_trees_or_fail_cleanly(golden):
    """Random bit flips in the masks and flags of the dragon's descriptors
    (non-leaf, valid, large, far): every tree the check accepts traces in the
    bounds-checked oracle without leaving the word array."""
    import conftest
    svo = conftest.load_package()
    words = svo.VoxelOctree(os.path.join(golden, "dragon.oct")).words()
    rng = np.random.default_rng(355)
    # descriptors of the top levels: the words reached from the root within 3 levels
    top, frontier = [0], [0]
    for _ in range(3):
        nxt = []
        for i in frontier:
            w = int(words[i])
            off, valid, nonleaf, per = w >> 18, (w >> 8) & 0xFF, w & 0xFF, 1 + ((w >> 16) & 1)
            nxt += [i + off + bin(nonleaf & ((1 << m) - 1)).count("1") * per
                    for m in range(8) if (valid >> m) & (nonleaf >> m) & 1]
        top += nxt
        frontier = nxt
    loaded = refused = 0
    o, d = _probe_rays(rng, 256)
    for _ in range(300):
        w = words.copy()
        for _ in range(int(rng.integers(1, 4))):
            i = int(top[int(rng.integers(0, len(top)))])
            w[i] ^= np.uint32(1 << int(rng.integers(0, 18)))
        try:
            t = svo.VoxelOctree.from_words(w, [0.5] * 3)
        except svo.SvoError as e:
            assert e.code == svo.svo.ERR_FORMAT, str(e)
            refused += 1
            continue
        orc.raymarch(t.words(), o, d)
        loaded += 1
    assert loaded > 20 and refused > 20, (loaded, refused)


def test_word_count_claim_is_bounded_by_the_stream(svo):
    """A header claiming 2^38 words with one empty slice record per 64 MiB
    (128 KiB of file) is refused before the word array is allocated: LZ4
    writes at most 255 bytes per input byte."""
    n = 1 << 38
    slices = (n * 4) >> 26
    data = np.zeros(3, np.float32).tobytes() + int(n).to_bytes(8, "little") + bytes(8 * slices)
    with pytest.raises(svo.SvoError) as e:
        svo.VoxelOctree.from_bytes(data)
    assert e.value.code == svo.svo.ERR_FORMAT and "too short" in str(e.value)


def test_shared_subtrees_are_refused_at_once(svo):
    """Twenty levels of descriptor blocks where all eight children of a node
    point at the same next block, then one block of leaves (a DAG the
    reference never writes: each non-leaf child gets its own block), make
    8^20 root-to-leaf paths out of 169 words, every pointer in range. The
    check refuses a slot's second parent instead of walking the paths."""
    import time
    levels = 20

    def chain(nonleaf):
        w = np.full(1 + 8 * (levels + 1), 0x00808080, np.uint32)  # last block: leaf material words
        w[0] = (1 << 18) | (0xFF << 8) | nonleaf
        for j in range(levels):
            blk, nxt = 1 + 8 * j, 1 + 8 * (j + 1)
            mask = 0 if j == levels - 1 else nonleaf
            for k in range(8 if nonleaf == 0xFF else 1):
                w[blk + k] = ((nxt - (blk + k)) << 18) | (0xFF << 8) | mask
        return w

    t0 = time.perf_counter()
    with pytest.raises(svo.SvoError) as e:
        svo.VoxelOctree.from_words(chain(0xFF), [0.5, 0.5, 0.5])
    assert e.value.code == svo.svo.ERR_FORMAT
    assert time.perf_counter() - t0 < 1.0
    # one non-leaf child per node (the other seven are leaf words): a tree
    assert svo.VoxelOctree.from_words(chain(0x01), [0.5, 0.5, 0.5]).depth() == levels + 2


def _ply(tmp_path, name, header, body=b""):
    path = tmp_path / name
    path.write_bytes(("\n".join(["ply"] + header + ["end_header"]) + "\n").encode() + body)
    return str(path)


def test_crafted_ply_counts_are_refused_before_sizing(svo, tmp_path):
    """Element counts the body cannot hold are refused from the header, before
    anything is sized by them: a property-less vertex element whose count
    times 9 floats wraps to 2 (a heap overflow in an unchecked parser), counts
    beyond the body of binary and ascii files, a list length of 1e30 and a
    face index beyond int range."""
    face = ["element face 1", "property list uchar int vertex_indices"]
    xyz = ["property float x", "property float y", "property float z"]
    cases = [
        _ply(tmp_path, "wrap.ply", ["format ascii 1.0", "element vertex 2049638230412172402"] + face, b"3 0 0 0\n"),
        _ply(tmp_path, "bin.ply", ["format binary_little_endian 1.0", "element vertex 1000000"] + xyz + face,
             bytes(64)),
        _ply(tmp_path, "asc.ply", ["format ascii 1.0", "element vertex 1000000"] + xyz + face, b"0 0 0\n3 0 0 0\n"),
        _ply(tmp_path, "list.ply", ["format ascii 1.0", "element vertex 3"] + xyz +
             ["element face 1", "property list float int vertex_indices"], b"0 0 0\n1 0 0\n0 1 0\n1e30 0 1 2\n"),
        _ply(tmp_path, "idx.ply", ["format ascii 1.0", "element vertex 3"] + xyz +
             ["element face 1", "property list uchar float vertex_indices"], b"0 0 0\n1 0 0\n0 1 0\n3 0 1 1e12\n"),
    ]
    for path in cases:
        with pytest.raises(svo.SvoError) as e:
            svo.VoxelOctree.build_from_ply(path, 64)
        assert e.value.code == svo.svo.ERR_FORMAT, (path, str(e.value))


@pytest.mark.skipif(_device_count_or_none() != 0, reason="host-parse fuzz: with a device the mutated meshes would build")
def test_corrupt_ply_files_fail_cleanly(svo, golden, tmp_path):
    """Random corruptions of an ascii and a binary PLY (header and body bytes,
    truncations) are refused or reach the device step (ERR_DEVICE where no
    device is present); none may crash the process."""
    scenes = svo.scenes
    v, f = scenes.icosphere(4, seed=7, displacement=0.05)
    srcs = []
    for ascii in (True, False):
        pth = str(tmp_path / f"src_{int(ascii)}.ply")
        scenes.write_ply(pth, v, f, normals=v, colors=np.full((len(v), 3), 200, np.uint8), ascii=ascii)
        srcs.append(np.frombuffer(open(pth, "rb").read(), np.uint8))
    rng = np.random.default_rng(99)
    allowed = {svo.svo.ERR_FORMAT, svo.svo.ERR_UNSUPPORTED, svo.svo.ERR_DEVICE, svo.svo.ERR_LIMIT}
    seen = set()
    for i in range(400):
        b = srcs[i % 2].copy()
        hdr = int(bytes(b).index(b"end_header")) + 11
        for _ in range(int(rng.integers(1, 6))):
            k = int(rng.integers(0, hdr)) if rng.random() < 0.5 else int(rng.integers(0, len(b)))
            b[k] = np.uint8(rng.integers(0, 256)) if rng.random() < 0.5 else np.uint8(b"0123456789 \n-e."[rng.integers(0, 15)])
        if rng.random() < 0.2:
            b = b[:int(rng.integers(0, len(b)))]
        pth = tmp_path / "m.ply"
        pth.write_bytes(b.tobytes())
        try:
            svo.VoxelOctree.build_from_ply(str(pth), 32)
            raise AssertionError("built without a device")
        except svo.SvoError as e:
            assert e.code in allowed, (i, e.code, str(e))
            seen.add(e.code)
    assert {svo.svo.ERR_FORMAT, svo.svo.ERR_DEVICE} <= seen


def test_voxel_header_is_checked_against_the_file(svo, tmp_path):
    """A .voxel header whose w*h*d exceeds the words after it (a 12-byte
    file claiming 2048^3, or sides whose product wraps 64 bits) is refused
    before the volume is allocated."""
    import time
    for dims in ((2048, 2048, 2048), (2**31 - 1,) * 3, (4, 4, 5)):
        path = tmp_path / "v.voxel"
        path.write_bytes(np.int32(dims).tobytes() + np.zeros(4 * 4 * 4, np.uint32).tobytes())
        t0 = time.perf_counter()
        with pytest.raises(svo.SvoError) as e:
            svo.VoxelOctree.build_from_voxels(str(path))
        assert e.value.code == svo.svo.ERR_FORMAT and "exceeds the file" in str(e.value), dims
        assert time.perf_counter() - t0 < 1.0

def relocated(words, base):
    """The tree with every word but the root moved up by `base` and reached
    through a far pointer in the root (a root the reference never writes, but
    one a crafted file can hold and the traversal follows)."""
    n = len(words)
    root = int(words[0])
    assert not root & 0x20000
    target = base + (root >> 18)
    out = np.zeros(base + n, np.uint32)
    out[base + 1:] = words[1:]
    out[0] = ((target >> 32) << 18) | (1 << 17) | (root & 0x1FFFF)
    out[1] = target & 0xFFFFFFFF
    return out


def test_far_root_is_validated_like_any_descriptor(svo, golden):
    c, w = orc.read_oct(os.path.join(golden, "dragon.oct"))
    base = 1000
    moved = relocated(w, base)
    t = svo.VoxelOctree.from_words(moved, c)
    assert t.depth() == 9
    rng = np.random.default_rng(3)
    o = rng.uniform(-0.5, 3.5, (3000, 3)).astype(np.float32)
    d = (rng.uniform(0.9, 2.1, (3000, 3)).astype(np.float32) - o).astype(np.float32)
    ref = orc.raymarch(w, o, d, 0.0)
    got = orc.raymarch(moved, o, d, 0.0)
    leaf = ref["hit"] == 1
    assert leaf.sum() > 300
    assert np.array_equal(got["hit"], ref["hit"]) and np.array_equal(got["normal"], ref["normal"])
    assert np.array_equal(got["t"].view(np.uint32), ref["t"].view(np.uint32))
    assert np.array_equal(got["leaf"][leaf], ref["leaf"][leaf] + np.uint64(base))
    # past the end; offset 0; and a far root whose high bits name a valid
    # one-leaf block at word 1 while the traversal's offset is (1 << 32) | w[1]
    crafted = np.array([(1 << 18) | (1 << 17) | (0x01 << 8), 0x1234], np.uint32)
    for bad in (moved[:base], np.array([moved[0], 0] + [0] * 8, np.uint32), crafted):
        with pytest.raises(svo.SvoError) as e:
            svo.VoxelOctree.from_words(bad, c)
        assert e.value.code == svo.svo.ERR_FORMAT


def test_host_camera_matches_reference(svo, golden):
    g = np.load(os.path.join(golden, "camera.npz"))
    for inp, tf in zip(g["inputs"], g["tform"]):
        t = svo.VoxelOctree.from_words(np.array([1 << 18], np.uint32), inp[3:6])
        cam = t.orbit_camera(inp[0], inp[1], inp[2])
        assert np.array_equal(np.frombuffer(cam.tform, np.float32).view(np.uint32), tf.view(np.uint32))
        _, pos = orc.orbit_camera(inp[0], inp[1], inp[2], inp[3:6])
        assert np.array_equal(np.frombuffer(cam.pos, np.float32).view(np.uint32), pos.view(np.uint32))


def test_frame_depth_count_matches_oracle_and_refuses_empty_stripes(svo):
    for w, h, stripes in [(1920, 1080, 16), (320, 176, 16), (641, 359, 7), (64, 48, 1), (64, 64, 0)]:
        assert svo.svo.frame_depth_count(w, h, stripes) == \
            orc.lib().oracle_frame_depth_count(w, h, stripes if stripes > 0 else 16), (w, h, stripes)
    for w, h, stripes in [(320, 100, 16), (0, 10, 1), (10, 0, 1), (40000, 8, 1), (8, 8, 9)]:
        with pytest.raises(svo.SvoError) as e:
            svo.svo.frame_depth_count(w, h, stripes)
        assert e.value.code == svo.svo.ERR_INVALID


def test_tile_plans_refuse_bad_arguments_before_device_use(svo, golden):
    """svo_tile_plan_create / _create_frame / _create_batch check the image
    size, the frame's stripe split and the batch's views and tile shape before
    any device call; the planned render entries refuse a null plan first."""
    m = svo.svo
    for w, h in [(0, 8), (8, 0), (65537, 8)]:
        with pytest.raises(svo.SvoError) as e:
            m.TilePlan(w, h)
        assert e.value.code == m.ERR_INVALID, (w, h)
    for w, h, stripes in [(320, 100, 16), (0, 10, 1), (8, 8, 9)]:
        with pytest.raises(svo.SvoError) as e:
            m.TilePlan.for_frame(w, h, stripes)
        assert e.value.code == m.ERR_INVALID, (w, h, stripes)
    for nframes in (0, 17):  # a multi-frame plan covers one launch group of 1 to 16 frames
        with pytest.raises(svo.SvoError) as e:
            m.TilePlan.for_frames(64, 64, nframes)
        assert e.value.code == m.ERR_INVALID, nframes
    out = ctypes.c_void_p()
    assert m.lib.svo_tile_plan_create_frame(320, 100, 16, ctypes.byref(out)) == m.ERR_INVALID and not out.value
    t = svo.VoxelOctree(os.path.join(golden, "dragon.oct"))
    cam = t.orbit_camera(30.0, 10.0, 1.0)
    for views, tile_w in [([], 0), ([m.Renderer.view(t, cam, 64, 64)], 5), ([m.Renderer.view(t, cam, 0, 64)], 0)]:
        with pytest.raises(svo.SvoError) as e:
            m.TilePlan.for_batch(views, tile_w)
        assert e.value.code == m.ERR_INVALID, (len(views), tile_w)
    assert m.lib.svo_render_primary_planned_device(t.handle, None, ctypes.byref(cam), 64, 64, None, None, None, None,
                                                   None) == m.ERR_INVALID
    assert m.lib.svo_render_frame_planned_device(t.handle, None, ctypes.byref(cam), 64, 64, 16, 1, None, None,
                                                 None) == m.ERR_INVALID
    cams = (m.Camera * 1)(cam)
    assert m.lib.svo_render_frames_planned_device(t.handle, None, cams, 1, 64, 64, 16, 1, None, None,
                                                  None) == m.ERR_INVALID
    light = np.array([0.0, 1.0, 0.0], np.float32)
    assert m.lib.svo_render_shadow_planned_device(t.handle, None, ctypes.byref(cam), 64, 64,
                                                  light.ctypes.data_as(ctypes.c_void_p), 0.001, None, None, None,
                                                  None, None, None) == m.ERR_INVALID


def test_server_refuses_bad_configuration_before_device_use(svo, golden):
    """svo_server_create checks waves_per_cu, ring_slots and max_batch before
    it looks for a device or allocates anything."""
    t = svo.VoxelOctree(os.path.join(golden, "dragon.oct"))
    for kw in ({"waves_per_cu": 3}, {"waves_per_cu": 33}, {"ring_slots": (1 << 20) + 1}, {"max_batch": 257},
               {"max_batch": -1}):
        with pytest.raises(svo.SvoError) as e:
            svo.svo.Server([t], **kw)
        assert e.value.code == svo.svo.ERR_INVALID, kw


def test_frames_call_refuses_bad_arguments_before_device_use(svo, golden):
    """svo_render_frames_device checks the frame count, the layout, the stride,
    every image pointer and every camera before anything touches a device."""
    t = svo.VoxelOctree(os.path.join(golden, "dragon.oct"))
    cam = t.orbit_camera(30.0, 10.0, 1.0)
    bad = t.orbit_camera(30.0, 10.0, 1.0)
    bad.pos[1] = float("inf")
    fake = 0x1000  # never dereferenced: every case is refused first
    cases = [([], 64, 64, [], 16, 1),                       # no frames
             ([cam], 64, 100, [fake], 16, 1),               # 16 stripes of 7 rows leave the last one empty
             ([cam], 64, 64, [fake], 16, 9),                # stride out of range
             ([cam, cam], 64, 64, [fake, None], 16, 1),     # null image
             ([cam, bad], 64, 64, [fake, fake], 16, 1)]     # camera entry not finite
    for cams, w, h, ptrs, stripes, stride in cases:
        with pytest.raises(svo.SvoError) as e:
            t.render_frames_device(cams, w, h, ptrs, None, stripes, stride)
        assert e.value.code == svo.svo.ERR_INVALID, (len(cams), w, h, stride)
