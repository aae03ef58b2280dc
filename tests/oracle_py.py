"""ctypes access to the oracle (oracle/liboracle_svo.so) for tests and the
CPU-baseline leg of bench.py. The oracle is the checker; nothing in the
product imports this module.

Also holds a small PLY reader for the oracle side of the mesh tests: it reads
the subset of PLY that src/third-party/plyfile.c accepts with PlyLoader's
property handling (src/PlyLoader.cpp:122-224): old type names only, binary
little-endian or ASCII, vertex x/y/z (+ nx/ny/nz, red/green/blue), face
vertex_indices lists.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle_svo.so")
LEAF_NONE = np.uint64(0xFFFFFFFFFFFFFFFF)

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                           stdout=subprocess.DEVNULL)
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        L.oracle_inv_sqrt.restype = ctypes.c_float
        L.oracle_inv_sqrt.argtypes = [ctypes.c_float]
        L.oracle_raymarch_batch.argtypes = [P, ctypes.c_uint64, ctypes.c_int64, P, P, P, ctypes.c_float,
                                            P, P, P, P, P, ctypes.c_int]
        L.oracle_orbit_camera.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float, P, P, P]
        L.oracle_primary_dirs.argtypes = [P, ctypes.c_int, ctypes.c_int, P]
        L.oracle_oct_decode.argtypes = [P, ctypes.c_uint64, P, P, P]
        L.oracle_octree_walk.argtypes = [P, ctypes.c_uint64, P, P, P, P, P]
        L.oracle_mesh_create.restype = P
        L.oracle_mesh_create.argtypes = [P, ctypes.c_int64, ctypes.c_int, P, P, ctypes.c_int64]
        L.oracle_mesh_free.argtypes = [P]
        L.oracle_mesh_tri_count.restype = ctypes.c_int64
        L.oracle_mesh_tri_count.argtypes = [P]
        L.oracle_mesh_tris.argtypes = [P, P]
        L.oracle_mesh_bounds.argtypes = [P, P, P]
        L.oracle_build_mesh.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, P, P, P, P]
        L.oracle_build_grid.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                        P, P, P, P]
        L.oracle_voxelize_mesh.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, P, P]
        L.oracle_convert_to_volume.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, P, P, P]
        L.oracle_compress_material_batch.argtypes = [ctypes.c_int64, P, P, P]
        L.oracle_decompress_material_batch.argtypes = [ctypes.c_int64, P, P, P]
        L.oracle_tribox_batch.argtypes = [ctypes.c_int64, P, P]
        L.oracle_free.argtypes = [P]
        L.oracle_render_frame.argtypes = [P, ctypes.c_uint64, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, P, P]
        L.oracle_frame_depth_count.restype = ctypes.c_int64
        L.oracle_frame_depth_count.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class BuildInfo(ctypes.Structure):
    _fields_ = [("data_dims", ctypes.c_int * 3), ("side_length", ctypes.c_int),
                ("cache_size", ctypes.c_int), ("sub_block", ctypes.c_int * 3),
                ("grid", ctypes.c_int * 3), ("bin_entries", ctypes.c_uint64),
                ("leaves", ctypes.c_uint64), ("hazard_voxels", ctypes.c_uint64)]

    def as_dict(self):
        return {"data_dims": list(self.data_dims), "side_length": self.side_length,
                "cache_size": self.cache_size, "sub_block": list(self.sub_block),
                "grid": list(self.grid), "bin_entries": self.bin_entries, "leaves": self.leaves,
                "hazard_voxels": self.hazard_voxels}


# ---- arithmetic ------------------------------------------------------------

def inv_sqrt(x):
    return lib().oracle_inv_sqrt(float(x))


def compress_material_batch(normals, shade):
    n = np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3)
    s = np.broadcast_to(np.asarray(shade, dtype=np.float32), (len(n),)).copy()
    out = np.zeros(len(n), np.uint32)
    lib().oracle_compress_material_batch(len(n), _p(n), _p(s), _p(out))
    return out


def decompress_material_batch(words):
    w = np.ascontiguousarray(words, dtype=np.uint32)
    n = np.zeros((len(w), 3), np.float32)
    s = np.zeros(len(w), np.float32)
    lib().oracle_decompress_material_batch(len(w), _p(w), _p(n), _p(s))
    return n, s


def tribox_batch(records):
    r = np.ascontiguousarray(records, dtype=np.float32).reshape(-1, 15)
    out = np.zeros(len(r), np.int32)
    lib().oracle_tribox_batch(len(r), _p(r), _p(out))
    return out


# ---- camera / rays -----------------------------------------------------------

def orbit_camera(pitch, yaw, radius, center):
    tf = np.zeros(16, np.float32)
    pos = np.zeros(3, np.float32)
    c = np.asarray(center, dtype=np.float32)
    lib().oracle_orbit_camera(float(pitch), float(yaw), float(radius), _p(c), _p(tf), _p(pos))
    return tf, pos


def primary_dirs(tform, width, height):
    tf = np.ascontiguousarray(tform, dtype=np.float32)
    d = np.zeros((width * height, 3), np.float32)
    lib().oracle_primary_dirs(_p(tf), int(width), int(height), _p(d))
    return d


def render_frame(words, tform, center, width, height, stripes=16, stride=1):
    """Viewer frame (src/Main.cpp:81-202) on the CPU: (rgba [H, W], coarse depths)."""
    words = np.ascontiguousarray(words, np.uint32)
    tform = np.ascontiguousarray(tform, np.float32).reshape(16)
    center = np.ascontiguousarray(center, np.float32)
    rgba = np.zeros(width * height, np.uint32)
    depth = np.zeros(lib().oracle_frame_depth_count(width, height, stripes), np.float32)
    lib().oracle_render_frame(_p(words), len(words), _p(tform), _p(center), width, height, stripes, stride,
                              _p(rgba), _p(depth))
    return rgba.reshape(height, width), depth


def raymarch(words, o, d, ray_scale=0.0, threads=8):
    """Returns dict(hit u8, normal u32, t f32, leaf u64, counts [desc, far, leaf])."""
    w = np.ascontiguousarray(words, dtype=np.uint32)
    o = np.ascontiguousarray(o, dtype=np.float32).reshape(-1, 3)
    d = np.ascontiguousarray(d, dtype=np.float32).reshape(-1, 3)
    if len(o) == 1 and len(d) > 1:
        o = np.ascontiguousarray(np.broadcast_to(o, d.shape))
    n = len(d)
    rs = None
    s0 = 0.0
    if np.ndim(ray_scale) == 0:
        s0 = float(ray_scale)
    else:
        rs = np.ascontiguousarray(ray_scale, dtype=np.float32)
    hit = np.zeros(n, np.uint8)
    normal = np.zeros(n, np.uint32)
    t = np.zeros(n, np.float32)
    leaf = np.zeros(n, np.uint64)
    counts = np.zeros(3, np.uint64)
    r = lib().oracle_raymarch_batch(_p(w), len(w), n, _p(o), _p(d), _p(rs), s0, _p(hit), _p(normal),
                                    _p(t), _p(leaf), _p(counts), int(threads))
    if r != 0:
        raise RuntimeError("oracle raymarch left the word array (malformed octree)")
    return {"hit": hit, "normal": normal, "t": t, "leaf": leaf, "counts": counts}


# ---- .oct ---------------------------------------------------------------------

def decode_oct_bytes(data):
    buf = np.frombuffer(data, dtype=np.uint8)
    c = np.zeros(3, np.float32)
    wp = ctypes.c_void_p()
    n = ctypes.c_uint64()
    r = lib().oracle_oct_decode(_p(buf), len(buf), _p(c), ctypes.byref(wp), ctypes.byref(n))
    if r != 0:
        raise ValueError("corrupt .oct")
    words = np.ctypeslib.as_array(ctypes.cast(wp, ctypes.POINTER(ctypes.c_uint32)), (n.value,)).copy() \
        if n.value else np.zeros(0, np.uint32)
    lib().oracle_free(wp)
    return c, words


def read_oct(path):
    with open(path, "rb") as fp:
        return decode_oct_bytes(fp.read())


def octree_kinds(words):
    w = np.ascontiguousarray(words, dtype=np.uint32)
    kind = np.zeros(len(w), np.uint8)
    if lib().oracle_octree_walk(_p(w), len(w), _p(kind), None, None, None, None) != 0:
        raise ValueError("malformed octree")
    return kind


def octree_stats(words):
    w = np.ascontiguousarray(words, dtype=np.uint32)
    dpl = np.zeros(24, np.uint64)
    lv = ctypes.c_uint64()
    fw = ctypes.c_uint64()
    dep = ctypes.c_int()
    if lib().oracle_octree_walk(_p(w), len(w), None, _p(dpl), ctypes.byref(lv), ctypes.byref(fw),
                                ctypes.byref(dep)) != 0:
        raise ValueError("malformed octree")
    levels = [int(x) for x in dpl]
    while levels and levels[-1] == 0:
        levels.pop()
    return {"desc_per_level": levels, "leaves": lv.value,
            "far_words": fw.value, "depth": dep.value}


# ---- PLY (oracle side) -------------------------------------------------------------

_PLY_TYPES = {"char": "i1", "uchar": "u1", "short": "<i2", "ushort": "<u2", "int": "<i4",
              "uint": "<u4", "float": "<f4", "double": "<f8"}


def read_ply(path):
    """-> (vertex table (n,9) float32, has_normals, face_counts int32, face_idx int32)."""
    with open(path, "rb") as fp:
        data = fp.read()
    end = data.index(b"end_header") + len(b"end_header")
    while data[end:end + 1] in (b"\r", b"\n"):
        end += 1
        if data[end - 1:end] == b"\n":
            break
    header = data[:end].decode().split("\n")
    fmt = None
    elems = []
    for line in header:
        tok = line.split()
        if not tok:
            continue
        if tok[0] == "format":
            fmt = tok[1]
        elif tok[0] == "element":
            elems.append([tok[1], int(tok[2]), []])
        elif tok[0] == "property":
            if tok[1] == "list":
                elems[-1][2].append((tok[4], "list", tok[2], tok[3]))
            else:
                elems[-1][2].append((tok[2], tok[1]))
    body = data[end:]
    names9 = ["x", "y", "z", "nx", "ny", "nz", "red", "green", "blue"]
    defaults = [0, 0, 0, 0, 0, 0, 255, 255, 255]
    verts = faces_counts = faces_idx = None
    has_normals = False
    if fmt == "ascii":
        toks = body.split()
        pos = 0
        for name, count, props in elems:
            if name == "vertex":
                table = np.zeros((count, 9), np.float32)
                table[:] = np.float32(defaults)
                for i in range(count):
                    for p in props:
                        if p[1] == "list":
                            k = int(toks[pos]); pos += 1 + k
                            continue
                        v = float(toks[pos]); pos += 1
                        if p[0] in names9:
                            table[i, names9.index(p[0])] = np.float32(v)
                verts = table
                has_normals = all(any(p[0] == n for p in props) for n in ("nx", "ny", "nz"))
            elif name == "face":
                cnts, idx = [], []
                for i in range(count):
                    for p in props:
                        if p[1] == "list":
                            k = int(toks[pos]); pos += 1
                            vals = [int(float(t)) for t in toks[pos:pos + k]]; pos += k
                            if p[0] == "vertex_indices":
                                cnts.append(k); idx += vals
                        else:
                            pos += 1
                faces_counts = np.array(cnts, np.int32)
                faces_idx = np.array(idx, np.int32)
            else:
                for i in range(count):
                    for p in props:
                        if p[1] == "list":
                            k = int(toks[pos]); pos += 1 + k
                        else:
                            pos += 1
        return verts, has_normals, faces_counts, faces_idx
    if fmt != "binary_little_endian":
        raise ValueError("unsupported PLY format " + str(fmt))
    off = 0
    for name, count, props in elems:
        has_list = any(p[1] == "list" for p in props)
        if not has_list:
            dt = np.dtype([(p[0], _PLY_TYPES[p[1]]) for p in props])
            rec = np.frombuffer(body, dtype=dt, count=count, offset=off)
            off += dt.itemsize * count
            if name == "vertex":
                table = np.zeros((count, 9), np.float32)
                table[:] = np.float32(defaults)
                for k, n in enumerate(names9):
                    if n in dt.names:
                        table[:, k] = rec[n].astype(np.float32)
                verts = table
                has_normals = all(n in dt.names for n in ("nx", "ny", "nz"))
            continue
        cnts, idx = [], []
        for i in range(count):
            for p in props:
                if p[1] == "list":
                    ct = np.dtype(_PLY_TYPES[p[2]])
                    it = np.dtype(_PLY_TYPES[p[3]])
                    k = int(np.frombuffer(body, ct, 1, off)[0]); off += ct.itemsize
                    vals = np.frombuffer(body, it, k, off); off += it.itemsize * k
                    if name == "face" and p[0] == "vertex_indices":
                        cnts.append(k); idx.extend(int(v) for v in vals)
                else:
                    off += np.dtype(_PLY_TYPES[p[1]]).itemsize
        if name == "face":
            faces_counts = np.array(cnts, np.int32)
            faces_idx = np.array(idx, np.int32)
    return verts, has_normals, faces_counts, faces_idx


# ---- construction ----------------------------------------------------------------

class Mesh:
    def __init__(self, verts, has_normals, face_counts, face_idx):
        self._v = np.ascontiguousarray(verts, dtype=np.float32)
        self._c = np.ascontiguousarray(face_counts, dtype=np.int32)
        self._i = np.ascontiguousarray(face_idx, dtype=np.int32)
        self.h = lib().oracle_mesh_create(_p(self._v), len(self._v), int(bool(has_normals)), _p(self._c),
                                          _p(self._i), len(self._c))

    @classmethod
    def from_ply(cls, path):
        return cls(*read_ply(path))

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_mesh_free(self.h)
            self.h = None

    def tris(self):
        n = lib().oracle_mesh_tri_count(self.h)
        out = np.zeros((n, 33), np.float32)
        lib().oracle_mesh_tris(self.h, _p(out))
        return out

    def bounds(self):
        lo = np.zeros(3, np.float32)
        hi = np.zeros(3, np.float32)
        lib().oracle_mesh_bounds(self.h, _p(lo), _p(hi))
        return lo, hi

    def build(self, resolution, threads, mem=1 << 30):
        wp = ctypes.c_void_p()
        n = ctypes.c_uint64()
        c = np.zeros(3, np.float32)
        info = BuildInfo()
        r = lib().oracle_build_mesh(self.h, int(resolution), int(threads), int(mem), ctypes.byref(wp),
                                    ctypes.byref(n), _p(c), ctypes.byref(info))
        if r != 0:
            raise RuntimeError(f"oracle mesh build failed ({r})")
        words = np.ctypeslib.as_array(ctypes.cast(wp, ctypes.POINTER(ctypes.c_uint32)), (n.value,)).copy()
        lib().oracle_free(wp)
        return c, words, info.as_dict()

    def voxelize(self, resolution, threads, mem=1 << 30):
        info_dims = np.zeros(3, np.int32)
        lo, hi = self.bounds()
        dims = [int(np.float32(hi[k] - lo[k]) * np.float32(resolution - 2)) + 2 for k in range(3)]
        out = np.zeros(dims[0] * dims[1] * dims[2], np.uint32)
        r = lib().oracle_voxelize_mesh(self.h, int(resolution), int(threads), int(mem), _p(out), _p(info_dims))
        if r != 0:
            raise RuntimeError("oracle voxelize failed")
        return out.reshape(dims[2], dims[1], dims[0])


    def convert_to_volume(self, max_size, threads, mem=1 << 30):
        """PlyLoader::convertToVolume: the .voxel file's bytes (w, h, d int32
        header, then the uint32 volume) and the slices per slab."""
        wp = ctypes.c_void_p()
        dims = np.zeros(3, np.int32)
        sz = ctypes.c_int()
        r = lib().oracle_convert_to_volume(self.h, int(max_size), int(threads), int(mem), ctypes.byref(wp),
                                           _p(dims), ctypes.byref(sz))
        if r != 0:
            raise RuntimeError(f"oracle convert failed ({r})")
        n = int(dims[0]) * int(dims[1]) * int(dims[2])
        body = np.ctypeslib.as_array(ctypes.cast(wp, ctypes.POINTER(ctypes.c_uint32)), (n,)).copy()
        lib().oracle_free(wp)
        return dims.tobytes() + body.tobytes(), sz.value


def build_grid(vol, mem=1 << 30):
    """vol: uint32 [d, h, w] (x fastest)."""
    v = np.ascontiguousarray(vol, dtype=np.uint32)
    d, h, w = v.shape
    wp = ctypes.c_void_p()
    n = ctypes.c_uint64()
    c = np.zeros(3, np.float32)
    info = BuildInfo()
    r = lib().oracle_build_grid(_p(v), w, h, d, int(mem), ctypes.byref(wp), ctypes.byref(n), _p(c),
                                ctypes.byref(info))
    if r != 0:
        raise RuntimeError(f"oracle grid build failed ({r})")
    words = np.ctypeslib.as_array(ctypes.cast(wp, ctypes.POINTER(ctypes.c_uint32)), (n.value,)).copy()
    lib().oracle_free(wp)
    return c, words, info.as_dict()
