"""Python binding of libsvo_hip.so (ctypes over the C ABI in include/svo.h).

Mirrors the reference's VoxelOctree interface (src/VoxelOctree.hpp:37-58):
``VoxelOctree(path)`` loads an .oct, ``save(path)`` writes one,
``raymarch(o, d, rayScale)`` traces one ray, ``center()`` returns the stored
center; builders mirror ``VoxelOctree(VoxelData*)`` fed by a PLY mesh or a raw
volume. All tracing and construction run in the HIP library on the GPU; this
module only marshals arguments. There is no CPU fallback: if the library is
missing, importing raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsvo_hip.so")
LEAF_NONE = np.uint64(0xFFFFFFFFFFFFFFFF)

OK, ERR_INVALID, ERR_IO, ERR_FORMAT, ERR_NOMEM, ERR_DEVICE, ERR_UNSUPPORTED, ERR_LIMIT = range(8)

# every symbol include/svo.h declares
EXPORTS = [
    "svo_last_error", "svo_version", "svo_device_count", "svo_load", "svo_load_memory", "svo_save",
    "svo_encode", "svo_encode_words", "svo_free_buffer", "svo_from_words", "svo_center", "svo_word_count", "svo_copy_words",
    "svo_depth", "svo_free", "svo_upload", "svo_raymarch", "svo_raymarch_batch", "svo_raymarch_batch_device",
    "svo_orbit_camera", "svo_render_primary", "svo_render_primary_device", "svo_build_from_ply",
    "svo_convert_to_volume",
    "svo_build_from_mesh", "svo_build_from_voxels", "svo_build_from_grid",
    "svo_renderer_create", "svo_renderer_free", "svo_render_batch_device",
    "svo_build_from_sdf", "svo_sdf_volume", "svo_serve_replay",
    "svo_render_frame", "svo_render_frame_device", "svo_render_frames_device", "svo_frame_depth_count",
    "svo_build_from_terrain", "svo_terrain_volume", "svo_terrain_heights",
    "svo_render_shadow", "svo_render_shadow_device", "svo_render_shadow_planned_device", "svo_decode_words",
    "svo_renderer_autotune", "svo_batch_graph_create", "svo_batch_graph_launch", "svo_batch_graph_free",
    "svo_server_create", "svo_server_submit", "svo_server_poll", "svo_server_wait", "svo_server_outputs",
    "svo_server_copy_outputs", "svo_server_release", "svo_server_stats_get", "svo_server_run_schedule",
    "svo_server_destroy", "svo_tile_plan_create", "svo_tile_plan_free", "svo_render_primary_planned_device",
    "svo_render_batch_noalloc", "svo_load_device_code", "svo_tile_plan_create_batch",
    "svo_render_batch_planned_device", "svo_tile_plan_create_frame", "svo_render_frame_planned_device",
    "svo_tile_plan_create_frames", "svo_render_frames_planned_device",
    "svo_build_stats_get",
]


class SvoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"svo error {code}: {msg}")
        self.code = code


class Camera(ctypes.Structure):
    _fields_ = [("tform", ctypes.c_float * 16), ("pos", ctypes.c_float * 3)]


class View(ctypes.Structure):
    _fields_ = [("tree", ctypes.c_void_p), ("camera", Camera), ("width", ctypes.c_int), ("height", ctypes.c_int),
                ("ray_scale", ctypes.c_float), ("hit", ctypes.c_void_p), ("normal", ctypes.c_void_p),
                ("t", ctypes.c_void_p), ("leaf", ctypes.c_void_p)]


BATCH_GRID, BATCH_PERSISTENT = 0, 1


class BatchOpts(ctypes.Structure):
    _fields_ = [("waves_per_cu", ctypes.c_int), ("refill_threshold", ctypes.c_int), ("form", ctypes.c_int),
                ("tile_w", ctypes.c_int), ("reserved", ctypes.c_int * 4)]


def _batch_opts(waves_per_cu, refill_threshold, form, tile_w=0):
    """Options of a batched launch: the grid form unless a form or a
    persistent launch shape is given; tile_w picks the grid form's tile shape
    (tile_w x 64/tile_w pixels; 0 = 8x8)."""
    o = BatchOpts()
    o.waves_per_cu = int(waves_per_cu)
    o.refill_threshold = int(refill_threshold)
    if form is None:
        form = BATCH_PERSISTENT if (waves_per_cu or refill_threshold) else BATCH_GRID
    o.form = int(form)
    o.tile_w = int(tile_w)
    return o


AUTOTUNE_SHAPES = ["grid8x8", "grid16x4", "grid4x16", "grid32x2"] + [f"{w}x{r}" for w in (16, 24, 32)
                                                                       for r in (16, 32, 64)]


class BuildStats(ctypes.Structure):
    _fields_ = [("triangles", ctypes.c_uint64), ("candidates", ctypes.c_uint64), ("leaves", ctypes.c_uint64),
                ("words", ctypes.c_uint64), ("device_ms", ctypes.c_double), ("peak_bytes", ctypes.c_uint64),
                ("reserved", ctypes.c_uint64 * 2)]


class SdfPrimitive(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("center", ctypes.c_float * 3), ("r1", ctypes.c_float),
                ("r2", ctypes.c_float), ("axis", ctypes.c_float * 3)]


def sdf_primitives(prims):
    """scenes.sdf_primitives(...) tuples -> ctypes array of svo_sdf_primitive."""
    arr = (SdfPrimitive * len(prims))()
    for i, (kind, c, r1, r2, axis) in enumerate(prims):
        arr[i].kind = int(kind)
        arr[i].center[:] = [float(x) for x in c]
        arr[i].r1, arr[i].r2 = float(r1), float(r2)
        arr[i].axis[:] = [float(x) for x in axis]
    return arr


def sdf_volume(prims, resolution, shade=0.8):
    """Dense uint32 [n,n,n] ([z,y,x]) of the GPU SDF source's voxels."""
    arr = sdf_primitives(prims)
    n = int(resolution)
    out = np.zeros((n, n, n), np.uint32)
    _check(lib.svo_sdf_volume(arr, len(arr), n, float(shade), _p(out)))
    return out


def terrain_volume(resolution, seed):
    """Dense uint32 [n,n,n] ([z,y,x]) of the GPU terrain source's voxels."""
    n = int(resolution)
    out = np.zeros((n, n, n), np.uint32)
    _check(lib.svo_terrain_volume(n, int(seed) & 0xFFFFFFFF, _p(out)))
    return out


def terrain_heights(resolution, seed):
    """(H int32 [n,n], h float32 [n,n]) per column, indexed [z, x]."""
    n = int(resolution)
    H = np.zeros((n, n), np.int32)
    h = np.zeros((n, n), np.float32)
    _check(lib.svo_terrain_heights(n, int(seed) & 0xFFFFFFFF, _p(H), _p(h)))
    return H, h


class Request(ctypes.Structure):
    _fields_ = [("tree", ctypes.c_int), ("camera", Camera), ("width", ctypes.c_int), ("height", ctypes.c_int)]


class ServerOpts(ctypes.Structure):
    _fields_ = [("streams", ctypes.c_int), ("output_budget", ctypes.c_uint64), ("max_batch", ctypes.c_int),
                ("one_per_launch", ctypes.c_int), ("waves_per_cu", ctypes.c_int), ("refill_threshold", ctypes.c_int),
                ("batch_form", ctypes.c_int), ("device_outputs", ctypes.c_int), ("reserved", ctypes.c_int * 4)]


def serve_replay(trees, requests, arrivals, streams=4, output_budget=1 << 30, max_batch=256, one_per_launch=False,
                 waves_per_cu=0, refill_threshold=0, outputs=None, batch_form=None, device_outputs=False):
    """Request loop over a known arrival schedule (see svo_serve_replay).
    requests: list of (tree index, Camera, width, height). outputs: optional
    list of (normal, t, leaf) numpy arrays (or None) per request. Returns
    (latencies seconds ndarray, makespan seconds)."""
    n = len(requests)
    tarr = (ctypes.c_void_p * len(trees))(*[t.handle.value for t in trees])
    rarr = (Request * n)()
    for i, (ti, cam, w, h) in enumerate(requests):
        rarr[i].tree, rarr[i].camera, rarr[i].width, rarr[i].height = int(ti), cam, int(w), int(h)
    arr = np.ascontiguousarray(arrivals, dtype=np.float64)
    lat = np.zeros(n, np.float64)
    mk = ctypes.c_double()
    o = ServerOpts()
    o.streams, o.output_budget, o.max_batch = int(streams), int(output_budget), int(max_batch)
    o.one_per_launch, o.waves_per_cu, o.refill_threshold = int(bool(one_per_launch)), int(waves_per_cu), int(refill_threshold)
    o.batch_form = _batch_opts(waves_per_cu, refill_threshold, batch_form).form
    o.device_outputs = int(bool(device_outputs))
    nout = tout = lout = None
    if outputs is not None:
        nout = (ctypes.c_void_p * n)(*[None if x is None else x[0].ctypes.data for x in outputs])
        tout = (ctypes.c_void_p * n)(*[None if x is None else x[1].ctypes.data for x in outputs])
        lout = (ctypes.c_void_p * n)(*[None if x is None else x[2].ctypes.data for x in outputs])
    _check(lib.svo_serve_replay(tarr, len(trees), rarr, _p(arr), n, ctypes.byref(o), _p(lat), ctypes.byref(mk),
                                nout, tout, lout))
    return lat, mk.value


class ServerConfig(ctypes.Structure):
    _fields_ = [("output_budget", ctypes.c_uint64), ("ring_slots", ctypes.c_int), ("waves_per_cu", ctypes.c_int),
                ("host_outputs", ctypes.c_int), ("idle_stop_us", ctypes.c_int), ("persistent_only", ctypes.c_int),
                ("persistent_launch", ctypes.c_int), ("max_batch", ctypes.c_int), ("reserved", ctypes.c_int * 3)]


class ServerStats(ctypes.Structure):
    _fields_ = [("launches", ctypes.c_uint64), ("published", ctypes.c_uint64), ("in_flight", ctypes.c_uint64),
                ("waiting", ctypes.c_uint64), ("live", ctypes.c_int), ("reserved0", ctypes.c_int),
                ("direct", ctypes.c_uint64), ("batched", ctypes.c_uint64), ("reserved", ctypes.c_int * 2)]


REQ_WAITING, REQ_IN_FLIGHT, REQ_DONE = 0, 1, 2


def _requests(requests):
    rarr = (Request * len(requests))()
    for i, (ti, cam, w, h) in enumerate(requests):
        rarr[i].tree, rarr[i].camera, rarr[i].width, rarr[i].height = int(ti), cam, int(w), int(h)
    return rarr


class Server:
    """Live request server (svo_server): requests submitted at any time run
    as their own launches while few are in flight, the others together in
    batched launches (persistent_launch: in one running persistent launch
    instead; persistent_only: every request there). requests are (tree
    index, Camera, width, height) over the octrees given here."""

    def __init__(self, trees, output_budget=0, ring_slots=0, waves_per_cu=0, host_outputs=False, idle_stop_us=0,
                 persistent_only=False, persistent_launch=False, max_batch=0):
        self._trees = list(trees)
        tarr = (ctypes.c_void_p * len(self._trees))(*[t.handle.value for t in self._trees])
        c = ServerConfig()
        c.output_budget, c.ring_slots, c.waves_per_cu = int(output_budget), int(ring_slots), int(waves_per_cu)
        c.host_outputs, c.idle_stop_us = int(bool(host_outputs)), int(idle_stop_us)
        c.persistent_only = int(bool(persistent_only))
        c.persistent_launch = int(bool(persistent_launch))
        c.max_batch = int(max_batch)
        self._s = ctypes.c_void_p()
        _check(lib.svo_server_create(tarr, len(self._trees), ctypes.byref(c), ctypes.byref(self._s)))

    def submit(self, tree, camera, width, height):
        r = Request()
        r.tree, r.camera, r.width, r.height = int(tree), camera, int(width), int(height)
        t = ctypes.c_uint64()
        _check(lib.svo_server_submit(self._s, ctypes.byref(r), ctypes.byref(t)))
        return t.value

    def poll(self, ticket):
        """(state, latency seconds or 0)."""
        st, lat = ctypes.c_int(), ctypes.c_double()
        _check(lib.svo_server_poll(self._s, int(ticket), ctypes.byref(st), ctypes.byref(lat)))
        return st.value, lat.value

    def wait(self, ticket, timeout=0.0):
        lat = ctypes.c_double()
        _check(lib.svo_server_wait(self._s, int(ticket), float(timeout), ctypes.byref(lat)))
        return lat.value

    def outputs(self, ticket):
        """(normal, t, leaf) pointers (ints) of a completed request."""
        n, t, l = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        _check(lib.svo_server_outputs(self._s, int(ticket), ctypes.byref(n), ctypes.byref(t), ctypes.byref(l)))
        return n.value, t.value, l.value

    def copy_outputs(self, ticket, width, height):
        """Host copies {"normal", "t", "leaf"} of a completed request."""
        n = int(width) * int(height)
        out = {"normal": np.zeros(n, np.uint32), "t": np.zeros(n, np.float32), "leaf": np.zeros(n, np.uint64)}
        _check(lib.svo_server_copy_outputs(self._s, int(ticket), _p(out["normal"]), _p(out["t"]), _p(out["leaf"])))
        return out

    def release(self, ticket):
        _check(lib.svo_server_release(self._s, int(ticket)))

    def stats(self):
        st = ServerStats()
        _check(lib.svo_server_stats_get(self._s, ctypes.byref(st)))
        return {k: getattr(st, k) for k in ("launches", "published", "in_flight", "waiting", "live", "direct",
                                            "batched")}

    def run_schedule(self, requests, arrivals, keep=False):
        """Submits request i at arrivals[i] seconds; returns (latencies,
        makespan, tickets or None). With keep the outputs stay until released."""
        n = len(requests)
        rarr = _requests(requests)
        arr = np.ascontiguousarray(arrivals, dtype=np.float64)
        lat = np.zeros(n, np.float64)
        mk = ctypes.c_double()
        tk = np.zeros(n, np.uint64) if keep else None
        _check(lib.svo_server_run_schedule(self._s, rarr, _p(arr), n, _p(tk), _p(lat), ctypes.byref(mk)))
        return lat, mk.value, tk

    def close(self):
        if lib is None:  # interpreter shutdown
            return
        if getattr(self, "_s", None) and self._s.value:
            lib.svo_server_destroy(self._s)
            self._s = ctypes.c_void_p()

    def __del__(self):
        self.close()


class BuildOpts(ctypes.Structure):
    _fields_ = [("memory_budget", ctypes.c_uint64), ("threads", ctypes.c_int), ("reserved", ctypes.c_int * 7)]


def _load():
    # torch's wheel ships its own libamdhip64.so.7 and only initialises with it;
    # both libraries carry that soname, so whichever loads first serves the
    # whole process. Load torch's first when it is installed, so device
    # buffers and streams from torch and from this library share one runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `make -C sparse-voxel-octrees_amd` "
                          "or __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    P, I64, U64, F = ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint64, ctypes.c_float
    L.svo_last_error.restype = ctypes.c_char_p
    sig = {
        "svo_load": [ctypes.c_char_p, P], "svo_load_memory": [P, U64, P], "svo_save": [P, ctypes.c_char_p],
        "svo_encode": [P, P, P], "svo_encode_words": [P, U64, P, P, P], "svo_free_buffer": [P], "svo_from_words": [P, U64, P, P],
        "svo_center": [P, P], "svo_word_count": [P, P], "svo_copy_words": [P, P, U64], "svo_depth": [P, P],
        "svo_free": [P], "svo_upload": [P, P], "svo_raymarch": [P, P, P, F, P, P, P],
        "svo_raymarch_batch": [P, I64, P, P, P, F, P, P, P, P],
        "svo_raymarch_batch_device": [P, I64, P, P, P, F, P, P, P, P, P],
        "svo_orbit_camera": [P, F, F, F, P],
        "svo_render_primary": [P, P, ctypes.c_int, ctypes.c_int, P, P, P, P],
        "svo_render_primary_device": [P, P, ctypes.c_int, ctypes.c_int, P, P, P, P, P],
        "svo_build_from_ply": [ctypes.c_char_p, ctypes.c_int, P, P],
        "svo_convert_to_volume": [ctypes.c_char_p, ctypes.c_int, P, ctypes.c_char_p],
        "svo_build_from_mesh": [P, I64, ctypes.c_int, P, P, I64, ctypes.c_int, P, P],
        "svo_build_from_voxels": [ctypes.c_char_p, P, P],
        "svo_build_from_grid": [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P],
        "svo_renderer_create": [ctypes.c_int, P], "svo_renderer_free": [P],
        "svo_render_batch_device": [P, P, ctypes.c_int, P, P],
        "svo_build_from_sdf": [P, ctypes.c_int, ctypes.c_int, F, P, P],
        "svo_sdf_volume": [P, ctypes.c_int, ctypes.c_int, F, P],
        "svo_serve_replay": [P, ctypes.c_int, P, P, ctypes.c_int, P, P, P, P, P, P],
        "svo_render_frame": [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P],
        "svo_render_frame_device": [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P, P],
        "svo_tile_plan_create_frame": [ctypes.c_int, ctypes.c_int, ctypes.c_int, P],
        "svo_tile_plan_create_frames": [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, P],
        "svo_render_frames_planned_device": [P, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, P, P, P],
        "svo_render_frame_planned_device": [P, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P, P],
        "svo_render_frames_device": [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P, P],
        "svo_frame_depth_count": [ctypes.c_int, ctypes.c_int, ctypes.c_int],
        "svo_build_from_terrain": [ctypes.c_int, ctypes.c_uint32, P, P],
        "svo_terrain_volume": [ctypes.c_int, ctypes.c_uint32, P],
        "svo_terrain_heights": [ctypes.c_int, ctypes.c_uint32, P, P],
        "svo_decode_words": [P, U64, P, P, P],
        "svo_renderer_autotune": [P, P, ctypes.c_int, P, P, P],
        "svo_build_stats_get": [P, P],
        "svo_batch_graph_create": [P, P, ctypes.c_int, P, P, P],
        "svo_batch_graph_launch": [P, P],
        "svo_batch_graph_free": [P],
        "svo_render_shadow": [P, P, ctypes.c_int, ctypes.c_int, P, F, P, P, P, P, P],
        "svo_render_shadow_device": [P, P, ctypes.c_int, ctypes.c_int, P, F, P, P, P, P, P, P],
        "svo_render_shadow_planned_device": [P, P, P, ctypes.c_int, ctypes.c_int, P, F, P, P, P, P, P, P],
        "svo_server_create": [P, ctypes.c_int, P, P], "svo_server_submit": [P, P, P],
        "svo_server_poll": [P, U64, P, P], "svo_server_wait": [P, U64, ctypes.c_double, P],
        "svo_server_outputs": [P, U64, P, P, P], "svo_server_release": [P, U64],
        "svo_server_copy_outputs": [P, U64, P, P, P],
        "svo_server_stats_get": [P, P], "svo_server_run_schedule": [P, P, P, ctypes.c_int, P, P, P],
        "svo_server_destroy": [P],
        "svo_tile_plan_create": [ctypes.c_int, ctypes.c_int, P], "svo_tile_plan_free": [P],
        "svo_render_primary_planned_device": [P, P, P, ctypes.c_int, ctypes.c_int, P, P, P, P, P],
        "svo_render_batch_noalloc": [P, ctypes.c_int, P, P],
        "svo_tile_plan_create_batch": [P, ctypes.c_int, P, P],
        "svo_render_batch_planned_device": [P, P, ctypes.c_int, P, P, P],
        "svo_load_device_code": [],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = None if name in ("svo_free", "svo_free_buffer", "svo_renderer_free",
                                      "svo_batch_graph_free", "svo_server_destroy",
                                      "svo_tile_plan_free") else ctypes.c_int
    L.svo_frame_depth_count.restype = I64
    return L


lib = _load()


def _check(code):
    if code != OK:
        raise SvoError(code, lib.svo_last_error().decode(errors="replace"))


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def decode_words(data):
    """(center, words) of an .oct image, without tree validation."""
    buf = np.frombuffer(data, dtype=np.uint8)
    p = ctypes.POINTER(ctypes.c_uint32)()
    n = ctypes.c_uint64()
    c = np.zeros(3, np.float32)
    _check(lib.svo_decode_words(_p(buf), len(buf), ctypes.byref(p), ctypes.byref(n), _p(c)))
    try:
        return c, np.ctypeslib.as_array(p, (n.value,)).copy() if n.value else np.zeros(0, np.uint32)
    finally:
        lib.svo_free_buffer(p)


def encode_words(words, center):
    """.oct image of an arbitrary word array (VoxelOctree::save's writer)."""
    w = np.ascontiguousarray(words, dtype=np.uint32)
    c = np.ascontiguousarray(center, dtype=np.float32)
    p = ctypes.c_void_p()
    n = ctypes.c_uint64()
    _check(lib.svo_encode_words(_p(w), len(w), _p(c), ctypes.byref(p), ctypes.byref(n)))
    try:
        return ctypes.string_at(p, n.value)
    finally:
        lib.svo_free_buffer(p)


def load_device_code():
    """Loads all of the library's kernels onto the current device now."""
    _check(lib.svo_load_device_code())


def device_count():
    return lib.svo_device_count()


def _opts(memory_budget=0, threads=0):
    o = BuildOpts()
    o.memory_budget = int(memory_budget)
    o.threads = int(threads)
    return o


class VoxelOctree:
    """Handle over one octree (host words + lazily uploaded device copy)."""

    def __init__(self, path=None, *, _handle=None):
        self._h = ctypes.c_void_p()
        if _handle is not None:
            self._h = _handle
        elif path is not None:
            _check(lib.svo_load(os.fsencode(path), ctypes.byref(self._h)))
        else:
            raise ValueError("path required")

    # -- constructors mirroring the reference's build inputs --
    @classmethod
    def from_bytes(cls, data):
        h = ctypes.c_void_p()
        buf = np.frombuffer(data, dtype=np.uint8)
        _check(lib.svo_load_memory(_p(buf), len(buf), ctypes.byref(h)))
        return cls(_handle=h)

    @classmethod
    def from_words(cls, words, center):
        h = ctypes.c_void_p()
        w = np.ascontiguousarray(words, dtype=np.uint32)
        c = np.ascontiguousarray(center, dtype=np.float32)
        _check(lib.svo_from_words(_p(w), len(w), _p(c), ctypes.byref(h)))
        return cls(_handle=h)

    @classmethod
    def build_from_ply(cls, path, resolution, threads=0, memory_budget=0):
        h = ctypes.c_void_p()
        o = _opts(memory_budget, threads)
        _check(lib.svo_build_from_ply(os.fsencode(path), int(resolution), ctypes.byref(o), ctypes.byref(h)))
        return cls(_handle=h)

    @classmethod
    def build_from_mesh(cls, verts9, has_normals, face_counts, face_idx, resolution, threads=0, memory_budget=0):
        h = ctypes.c_void_p()
        v = np.ascontiguousarray(verts9, dtype=np.float32)
        c = np.ascontiguousarray(face_counts, dtype=np.int32)
        i = np.ascontiguousarray(face_idx, dtype=np.int32)
        o = _opts(memory_budget, threads)
        _check(lib.svo_build_from_mesh(_p(v), len(v), int(bool(has_normals)), _p(c), _p(i), len(c), int(resolution),
                                       ctypes.byref(o), ctypes.byref(h)))
        return cls(_handle=h)

    @classmethod
    def build_from_grid(cls, vol, memory_budget=0):
        """vol: uint32 [d, h, w] (x fastest), as the raw .voxel layout."""
        h = ctypes.c_void_p()
        v = np.ascontiguousarray(vol, dtype=np.uint32)
        d_, h_, w_ = v.shape
        o = _opts(memory_budget, 0)
        _check(lib.svo_build_from_grid(_p(v), w_, h_, d_, ctypes.byref(o), ctypes.byref(h)))
        return cls(_handle=h)

    @classmethod
    def build_from_sdf(cls, prims, resolution, shade=0.8, memory_budget=0):
        """GPU signed-distance scene source (prims: scenes.sdf_primitives tuples)."""
        h = ctypes.c_void_p()
        arr = sdf_primitives(prims)
        o = _opts(memory_budget, 0)
        _check(lib.svo_build_from_sdf(arr, len(arr), int(resolution), float(shade), ctypes.byref(o), ctypes.byref(h)))
        return cls(_handle=h)

    @classmethod
    def build_from_terrain(cls, resolution, seed, memory_budget=0):
        """GPU fBm heightfield terrain source (exposed shell, y up)."""
        h = ctypes.c_void_p()
        o = _opts(memory_budget, 0)
        _check(lib.svo_build_from_terrain(int(resolution), int(seed) & 0xFFFFFFFF, ctypes.byref(o), ctypes.byref(h)))
        return cls(_handle=h)

    @classmethod
    def build_from_voxels(cls, path, memory_budget=0):
        h = ctypes.c_void_p()
        o = _opts(memory_budget, 0)
        _check(lib.svo_build_from_voxels(os.fsencode(path), ctypes.byref(o), ctypes.byref(h)))
        return cls(_handle=h)

    def __del__(self):
        self.close()

    def close(self):
        if lib is None:  # interpreter shutdown
            return
        if getattr(self, "_h", None) and self._h.value:
            lib.svo_free(self._h)
            self._h = ctypes.c_void_p()

    @property
    def handle(self):
        return self._h

    # -- reference members --
    def save(self, path):
        _check(lib.svo_save(self._h, os.fsencode(path)))

    def encode(self):
        p = ctypes.c_void_p()
        n = ctypes.c_uint64()
        _check(lib.svo_encode(self._h, ctypes.byref(p), ctypes.byref(n)))
        try:
            return ctypes.string_at(p, n.value)
        finally:
            lib.svo_free_buffer(p)

    def build_stats(self):
        """Counters of the GPU build that made this tree and its algorithmic
        bytes 36*triangles + 24*candidates + 4*words."""
        st = BuildStats()
        _check(lib.svo_build_stats_get(self._h, ctypes.byref(st)))
        d = {k: getattr(st, k) for k in ("triangles", "candidates", "leaves", "words", "device_ms", "peak_bytes")}
        d["algorithmic_bytes"] = 36 * st.triangles + 24 * st.candidates + 4 * st.words
        return d

    def center(self):
        c = np.zeros(3, np.float32)
        _check(lib.svo_center(self._h, _p(c)))
        return c

    def word_count(self):
        n = ctypes.c_uint64()
        _check(lib.svo_word_count(self._h, ctypes.byref(n)))
        return n.value

    def words(self):
        n = ctypes.c_uint64()
        _check(lib.svo_word_count(self._h, ctypes.byref(n)))
        w = np.zeros(n.value, np.uint32)
        _check(lib.svo_copy_words(self._h, _p(w), n.value))
        return w

    def depth(self):
        d = ctypes.c_int()
        _check(lib.svo_depth(self._h, ctypes.byref(d)))
        return d.value

    def upload(self):
        p = ctypes.c_void_p()
        _check(lib.svo_upload(self._h, ctypes.byref(p)))
        return p.value

    def raymarch(self, o, d, ray_scale=0.0):
        """VoxelOctree::raymarch: returns (hit code, normal or None, t or None)."""
        o = np.ascontiguousarray(o, dtype=np.float32)
        d = np.ascontiguousarray(d, dtype=np.float32)
        normal = ctypes.c_uint32(0)
        t = ctypes.c_float(0.0)
        hit = ctypes.c_int(0)
        _check(lib.svo_raymarch(self._h, _p(o), _p(d), float(ray_scale), ctypes.byref(normal), ctypes.byref(t),
                                ctypes.byref(hit)))
        return hit.value, (normal.value if hit.value == 1 else None), (t.value if hit.value else None)

    def raymarch_batch(self, o, d, ray_scale=0.0):
        o = np.ascontiguousarray(o, dtype=np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(d, dtype=np.float32).reshape(-1, 3)
        if len(o) == 1 and len(d) > 1:
            o = np.ascontiguousarray(np.broadcast_to(o, d.shape))
        n = len(d)
        rs = None
        rs0 = 0.0
        if np.ndim(ray_scale) == 0:
            rs0 = float(ray_scale)
        else:
            rs = np.ascontiguousarray(ray_scale, dtype=np.float32)
        hit = np.zeros(n, np.uint8)
        normal = np.zeros(n, np.uint32)
        t = np.zeros(n, np.float32)
        leaf = np.zeros(n, np.uint64)
        _check(lib.svo_raymarch_batch(self._h, n, _p(o), _p(d), _p(rs), rs0, _p(hit), _p(normal), _p(t), _p(leaf)))
        return {"hit": hit, "normal": normal, "t": t, "leaf": leaf}

    def raymarch_batch_device(self, n, o_ptr, d_ptr, rs_ptr=None, rs0=0.0, hit_ptr=None, normal_ptr=None,
                              t_ptr=None, leaf_ptr=None, stream=None):
        _check(lib.svo_raymarch_batch_device(self._h, int(n), o_ptr, d_ptr, rs_ptr, float(rs0), hit_ptr, normal_ptr,
                                             t_ptr, leaf_ptr, stream))

    def orbit_camera(self, pitch, yaw, radius=1.0):
        cam = Camera()
        _check(lib.svo_orbit_camera(self._h, float(pitch), float(yaw), float(radius), ctypes.byref(cam)))
        return cam

    def render_primary(self, cam, width, height):
        n = width * height
        hit = np.zeros(n, np.uint8)
        normal = np.zeros(n, np.uint32)
        t = np.zeros(n, np.float32)
        leaf = np.zeros(n, np.uint64)
        _check(lib.svo_render_primary(self._h, ctypes.byref(cam), int(width), int(height), _p(hit), _p(normal),
                                      _p(t), _p(leaf)))
        return {"hit": hit, "normal": normal, "t": t, "leaf": leaf}

    def render_primary_device(self, cam, width, height, hit_ptr=None, normal_ptr=None, t_ptr=None, leaf_ptr=None,
                              stream=None):
        _check(lib.svo_render_primary_device(self._h, ctypes.byref(cam), int(width), int(height), hit_ptr, normal_ptr,
                                             t_ptr, leaf_ptr, stream))

    def render_shadow(self, cam, width, height, light, eps):
        """Primary rays plus one hard-shadow ray per leaf hit toward `light`."""
        n = width * height
        out = {"hit": np.zeros(n, np.uint8), "normal": np.zeros(n, np.uint32), "t": np.zeros(n, np.float32),
               "leaf": np.zeros(n, np.uint64), "shadow": np.zeros(n, np.uint8)}
        lt = np.ascontiguousarray(light, np.float32)
        _check(lib.svo_render_shadow(self._h, ctypes.byref(cam), int(width), int(height), _p(lt), float(eps),
                                     _p(out["hit"]), _p(out["normal"]), _p(out["t"]), _p(out["leaf"]),
                                     _p(out["shadow"])))
        return out

    def render_shadow_device(self, cam, width, height, light, eps, hit_ptr, t_ptr, shadow_ptr, normal_ptr=None,
                             leaf_ptr=None, stream=None, plan=None):
        """Primary + shadow rays on the device; with `plan` (a TilePlan of the
        image size, kept for shadow renders) in the plan's cost order."""
        lt = np.ascontiguousarray(light, np.float32)
        if plan is None:
            _check(lib.svo_render_shadow_device(self._h, ctypes.byref(cam), int(width), int(height), _p(lt),
                                                float(eps), hit_ptr, normal_ptr, t_ptr, leaf_ptr, shadow_ptr, stream))
        else:
            _check(lib.svo_render_shadow_planned_device(self._h, plan._p, ctypes.byref(cam), int(width), int(height),
                                                        _p(lt), float(eps), hit_ptr, normal_ptr, t_ptr, leaf_ptr,
                                                        shadow_ptr, stream))

    def render_frame(self, cam, width, height, stripes=16, stride=1):
        """The viewer's shaded frame (src/Main.cpp:81-202): (rgba uint32
        [height, width], coarse depth table float32)."""
        rgba = np.zeros(width * height, np.uint32)
        depth = np.zeros(frame_depth_count(width, height, stripes), np.float32)
        _check(lib.svo_render_frame(self._h, ctypes.byref(cam), int(width), int(height), int(stripes), int(stride),
                                    _p(rgba), _p(depth)))
        return rgba.reshape(height, width), depth

    def render_frame_device(self, cam, width, height, rgba_ptr, depth_ptr=None, stripes=16, stride=1, stream=None,
                            plan=None):
        """One viewer frame on the device; with `plan` (TilePlan.for_frame) the
        fine pass runs in the cost order the previous frame through it filed."""
        if plan is None:
            _check(lib.svo_render_frame_device(self._h, ctypes.byref(cam), int(width), int(height), int(stripes),
                                               int(stride), rgba_ptr, depth_ptr, stream))
        else:
            _check(lib.svo_render_frame_planned_device(self._h, plan._p, ctypes.byref(cam), int(width), int(height),
                                                       int(stripes), int(stride), rgba_ptr, depth_ptr, stream))

    def render_frames_device(self, cams, width, height, rgba_ptrs, depth_ptrs=None, stripes=16, stride=1,
                             stream=None, plan=None):
        """Several frames in one call (svo_render_frames_device): device
        pointers per frame; depth_ptrs may be None or hold None entries. With
        `plan` (TilePlan.for_frames, same frame count) the fine pass runs in
        the plan's cost order."""
        n = len(cams)
        carr = (Camera * n)(*cams)
        rarr = (ctypes.c_void_p * n)(*rgba_ptrs)
        darr = None if depth_ptrs is None else (ctypes.c_void_p * n)(*depth_ptrs)
        if plan is None:
            _check(lib.svo_render_frames_device(self._h, carr, n, int(width), int(height), int(stripes), int(stride),
                                                rarr, darr, stream))
        else:
            _check(lib.svo_render_frames_planned_device(self._h, plan._p, carr, n, int(width), int(height),
                                                        int(stripes), int(stride), rarr, darr, stream))


def convert_to_volume(ply_path, resolution, voxel_path, threads=0, memory_budget=0):
    """PlyLoader::convertToVolume on the GPU: writes the dense .voxel file."""
    o = _opts(memory_budget, threads)
    _check(lib.svo_convert_to_volume(os.fsencode(ply_path), int(resolution), ctypes.byref(o),
                                     os.fsencode(voxel_path)))


def frame_depth_count(width, height, stripes=16):
    """Entries of the viewer frame's coarse depth table; raises SvoError for
    a size or stripe split the renderer refuses."""
    n = int(lib.svo_frame_depth_count(int(width), int(height), int(stripes)))
    if n < 0:
        raise SvoError(ERR_INVALID, lib.svo_last_error().decode(errors="replace"))
    return n


def render_batch_noalloc(views, stream=None, waves_per_cu=0, refill_threshold=0, form=None, tile_w=0):
    """svo_render_batch_noalloc: a batched launch whose device tables live in
    the library's fixed scratch array (at most 512 views; octrees resident)."""
    arr = (View * len(views))(*views)
    o = _batch_opts(waves_per_cu, refill_threshold, form, tile_w)
    _check(lib.svo_render_batch_noalloc(arr, len(views), ctypes.byref(o), stream))


class TilePlan:
    """The tile cost order of one image size, or of one batch, on the device
    (svo_tile_plan): renders through it start the slowest tiles first."""

    def __init__(self, width, height, _batch=None, _frame=None):
        self.width, self.height = int(width), int(height)
        self._p = ctypes.c_void_p()
        if _frame is not None:
            _check(lib.svo_tile_plan_create_frames(*_frame, ctypes.byref(self._p)))
        elif _batch is None:
            _check(lib.svo_tile_plan_create(self.width, self.height, ctypes.byref(self._p)))
        else:
            views, opts = _batch
            arr = (View * len(views))(*views)
            _check(lib.svo_tile_plan_create_batch(arr, len(views), ctypes.byref(opts), ctypes.byref(self._p)))

    @classmethod
    def for_frame(cls, width, height, stripes=16):
        """A plan for VoxelOctree.render_frame_device(..., plan=...) at this
        frame size and stripe split."""
        return cls(0, 0, _frame=(int(width), int(height), int(stripes), 1))

    @classmethod
    def for_frames(cls, width, height, nframes, stripes=16):
        """A plan for VoxelOctree.render_frames_device(..., plan=...) over
        nframes (1 to 16) frames of this size and stripe split per call."""
        return cls(0, 0, _frame=(int(width), int(height), int(stripes), int(nframes)))

    @classmethod
    def for_batch(cls, views, tile_w=0):
        """A plan for Renderer.render_device(..., plan=...) over views of these
        sizes in the grid form with this tile width."""
        return cls(0, 0, _batch=(views, _batch_opts(0, 0, None, tile_w)))

    def render_device(self, tree, cam, hit_ptr=None, normal_ptr=None, t_ptr=None, leaf_ptr=None, stream=None):
        _check(lib.svo_render_primary_planned_device(tree.handle, self._p, ctypes.byref(cam), self.width, self.height,
                                                     hit_ptr, normal_ptr, t_ptr, leaf_ptr, stream))

    def close(self):
        if lib is None:  # interpreter shutdown
            return
        if getattr(self, "_p", None) and self._p.value:
            lib.svo_tile_plan_free(self._p)
            self._p = ctypes.c_void_p()

    def __del__(self):
        self.close()


class Renderer:
    """Batched persistent renderer (svo_renderer): many views of many octrees
    per launch. Views reference device output arrays (pointers as ints)."""

    def __init__(self, max_views=256):
        self._r = ctypes.c_void_p()
        _check(lib.svo_renderer_create(int(max_views), ctypes.byref(self._r)))
        self._keep = []

    def __del__(self):
        self.close()

    def close(self):
        if lib is None:  # interpreter shutdown
            return
        if getattr(self, "_r", None) and self._r.value:
            lib.svo_renderer_free(self._r)
            self._r = ctypes.c_void_p()

    @staticmethod
    def view(tree, camera, width, height, hit=None, normal=None, t=None, leaf=None, ray_scale=0.0):
        v = View()
        v.tree = tree.handle.value
        v.camera = camera
        v.width, v.height = int(width), int(height)
        v.ray_scale = float(ray_scale)
        v.hit, v.normal, v.t, v.leaf = hit, normal, t, leaf
        return v

    def render_device(self, views, stream=None, waves_per_cu=0, refill_threshold=0, form=None, tile_w=0, plan=None):
        """One batched launch; with `plan` (TilePlan.for_batch) the grid form's
        tiles run in the plan's cost order."""
        arr = (View * len(views))(*views)
        o = _batch_opts(waves_per_cu, refill_threshold, form, tile_w)
        if plan is None:
            _check(lib.svo_render_batch_device(self._r, arr, len(views), ctypes.byref(o), stream))
        else:
            _check(lib.svo_render_batch_planned_device(self._r, arr, len(views), ctypes.byref(o), plan._p, stream))

    def autotune(self, views, stream=None):
        """((form, waves_per_cu, refill, tile_w), {"grid8x8": ms, ..., "WxR": ms})
        of the fastest launch shape (waves and refill are 0 for the grid form,
        tile_w 8 for the persistent form)."""
        arr = (View * len(views))(*views)
        best = BatchOpts()
        times = (ctypes.c_float * len(AUTOTUNE_SHAPES))()
        _check(lib.svo_renderer_autotune(self._r, arr, len(views), stream, ctypes.byref(best), times))
        return ((best.form, best.waves_per_cu, best.refill_threshold, best.tile_w or 8),
                {k: round(times[i], 3) for i, k in enumerate(AUTOTUNE_SHAPES)})

    def capture(self, views, stream=None, waves_per_cu=0, refill_threshold=0, form=None, tile_w=0):
        """BatchGraph replaying one batched launch of these views."""
        return BatchGraph(self, views, stream, waves_per_cu, refill_threshold, form, tile_w)


class BatchGraph:
    """A batched launch captured in a HIP graph (svo_batch_graph)."""

    def __init__(self, renderer, views, stream=None, waves_per_cu=0, refill_threshold=0, form=None, tile_w=0):
        self._g = ctypes.c_void_p()
        arr = (View * len(views))(*views)
        o = _batch_opts(waves_per_cu, refill_threshold, form, tile_w)
        _check(lib.svo_batch_graph_create(renderer._r, arr, len(views), ctypes.byref(o), stream, ctypes.byref(self._g)))

    def launch(self, stream=None):
        _check(lib.svo_batch_graph_launch(self._g, stream))

    def __del__(self):
        self.close()

    def close(self):
        if lib is None:  # interpreter shutdown
            return
        if getattr(self, "_g", None) and self._g.value:
            lib.svo_batch_graph_free(self._g)
            self._g = ctypes.c_void_p()
