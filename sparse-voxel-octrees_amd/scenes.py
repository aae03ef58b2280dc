"""Seeded synthetic inputs for the octree builder and the renderer.

These generators make the workloads' inputs: displaced icosphere meshes (as
binary little-endian PLY, the format PlyLoader reads, src/PlyLoader.cpp:66-224)
and signed-distance scenes voxelized into the raw ``.voxel`` layout that
VoxelData reads (``int32 w,h,d`` then ``uint32[w*h*d]`` x-fastest,
src/VoxelData.cpp:37-48). They are input synthesis, not part of the traced
path: arithmetic here is float64 numpy and only its float32 results matter.
"""
from __future__ import annotations

import math
import struct

import numpy as np

# ---------------------------------------------------------------------------
# noise
# ---------------------------------------------------------------------------


def _hash3(ix, iy, iz, seed):
    """Integer lattice hash -> float64 in [-1, 1) (wrapping uint64 arithmetic)."""
    s = np.full(np.shape(ix), (seed & 0xFFFFFFFF) * 0x27D4EB2F165667C5 % (1 << 64), dtype=np.uint64)
    with np.errstate(over="ignore"):
        h = (ix.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)
             ^ iy.astype(np.uint64) * np.uint64(0xC2B2AE3D27D4EB4F)
             ^ iz.astype(np.uint64) * np.uint64(0x165667B19E3779F9)
             ^ s)
        h ^= h >> np.uint64(29)
        h *= np.uint64(0xBF58476D1CE4E5B9)
        h ^= h >> np.uint64(32)
    return (h >> np.uint64(11)).astype(np.float64) / float(1 << 52) - 1.0


def value_noise3(p, seed):
    """Trilinear value noise with smoothstep weights; p: (..., 3) float64."""
    f = np.floor(p)
    t = p - f
    w = t * t * (3.0 - 2.0 * t)
    i = f.astype(np.int64)
    out = 0.0
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                v = _hash3(i[..., 0] + dx, i[..., 1] + dy, i[..., 2] + dz, seed)
                wx = w[..., 0] if dx else 1.0 - w[..., 0]
                wy = w[..., 1] if dy else 1.0 - w[..., 1]
                wz = w[..., 2] if dz else 1.0 - w[..., 2]
                out = out + v * wx * wy * wz
    return out


def fbm3(p, seed, octaves=6, gain=0.5, lacunarity=2.0, base_freq=4.0):
    amp, freq, total, norm = 1.0, base_freq, 0.0, 0.0
    for o in range(octaves):
        total = total + amp * value_noise3(p * freq, seed + 101 * o)
        norm += amp
        amp *= gain
        freq *= lacunarity
    return total / norm


# ---------------------------------------------------------------------------
# meshes
# ---------------------------------------------------------------------------


def _icosahedron():
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return v / np.linalg.norm(v, axis=1, keepdims=True), f


def icosphere(frequency, seed=0, displacement=0.05, octaves=6):
    """Class-I geodesic icosphere with 20*frequency^2 triangles, radially
    displaced by fBm noise of relative amplitude `displacement`.

    Returns (vertices float32 (n,3), faces int32 (m,3))."""
    f = int(frequency)
    base_v, base_f = _icosahedron()
    # barycentric lattice of one face: points (i, j) with i + j <= f
    ii, jj = np.meshgrid(np.arange(f + 1), np.arange(f + 1), indexing="ij")
    keep = ii + jj <= f
    ii, jj = ii[keep], jj[keep]
    npts = ii.size
    lut = -np.ones((f + 1, f + 1), dtype=np.int64)
    lut[ii, jj] = np.arange(npts)
    # triangles within a face lattice
    tris = []
    for i in range(f):
        for j in range(f - i):
            a, b, c = lut[i, j], lut[i + 1, j], lut[i, j + 1]
            tris.append((a, b, c))
            if j < f - i - 1:
                tris.append((b, lut[i + 1, j + 1], c))
    tris = np.array(tris, dtype=np.int64)
    A = base_v[base_f[:, 0]][:, None, :]
    B = base_v[base_f[:, 1]][:, None, :]
    C = base_v[base_f[:, 2]][:, None, :]
    P = A + (B - A) * (ii[None, :, None] / f) + (C - A) * (jj[None, :, None] / f)
    P = P.reshape(-1, 3)
    P /= np.linalg.norm(P, axis=1, keepdims=True)
    faces = (tris[None, :, :] + (np.arange(20) * npts)[:, None, None]).reshape(-1, 3)
    # merge duplicated edge/corner vertices
    q = np.round(P * (1 << 22)).astype(np.int64)
    _, first, inv = np.unique(q, axis=0, return_index=True, return_inverse=True)
    P = P[first]
    faces = inv.reshape(-1)[faces]
    if displacement:
        r = 1.0 + displacement * fbm3(P * 1.0 + 10.0, seed, octaves=octaves)
        P = P * r[:, None]
    return P.astype(np.float32), faces.astype(np.int32)


def write_ply(path, verts, faces, normals=None, colors=None, ascii=False):
    """Write a PLY that src/third-party/plyfile.c parses (old type names)."""
    verts = np.ascontiguousarray(verts, dtype=np.float32)
    faces = np.ascontiguousarray(faces, dtype=np.int32)
    nv, nf = len(verts), len(faces)
    hdr = ["ply", "format ascii 1.0" if ascii else "format binary_little_endian 1.0",
           f"element vertex {nv}", "property float x", "property float y", "property float z"]
    if normals is not None:
        hdr += ["property float nx", "property float ny", "property float nz"]
    if colors is not None:
        hdr += ["property uchar red", "property uchar green", "property uchar blue"]
    hdr += [f"element face {nf}", "property list uchar int vertex_indices", "end_header"]
    with open(path, "wb") as fp:
        fp.write(("\n".join(hdr) + "\n").encode())
        if ascii:
            for i in range(nv):
                row = [repr(float(x)) for x in verts[i]]
                if normals is not None:
                    row += [repr(float(x)) for x in normals[i]]
                if colors is not None:
                    row += [str(int(x)) for x in colors[i]]
                fp.write((" ".join(row) + "\n").encode())
            for i in range(nf):
                fp.write(("3 " + " ".join(str(int(x)) for x in faces[i]) + "\n").encode())
            return
        fields = [("p", "<f4", (3,))]
        if normals is not None:
            fields.append(("n", "<f4", (3,)))
        if colors is not None:
            fields.append(("c", "u1", (3,)))
        vrec = np.zeros(nv, dtype=fields)
        vrec["p"] = verts
        if normals is not None:
            vrec["n"] = normals
        if colors is not None:
            vrec["c"] = colors
        fp.write(vrec.tobytes())
        frec = np.zeros(nf, dtype=[("k", "u1"), ("i", "<i4", (3,))])
        frec["k"] = 3
        frec["i"] = faces
        fp.write(frec.tobytes())


def mesh_vertex_table(verts, faces, normals=None, colors=None):
    """The per-vertex table PlyLoader::readVertices builds (9 floats each:
    pos, normal (0 if absent), colour (255 if absent)) plus CSR faces."""
    nv = len(verts)
    tab = np.zeros((nv, 9), dtype=np.float32)
    tab[:, 0:3] = verts
    if normals is not None:
        tab[:, 3:6] = normals
    tab[:, 6:9] = 255.0 if colors is None else np.asarray(colors, dtype=np.float32)
    counts = np.full(len(faces), 3, dtype=np.int32)
    return tab, normals is not None, counts, np.ascontiguousarray(faces, dtype=np.int32).reshape(-1)


# ---------------------------------------------------------------------------
# signed-distance scenes (raw .voxel volumes)
# ---------------------------------------------------------------------------


def sdf_primitives(seed, count=16):
    """16 spheres/tori, centres U[0,1]^3, radii U[0.05,0.2] of the extent."""
    rng = np.random.default_rng(seed)
    prims = []
    for _ in range(count):
        kind = int(rng.integers(0, 2))
        c = rng.uniform(0.0, 1.0, 3)
        r1 = float(rng.uniform(0.05, 0.2))
        r2 = float(rng.uniform(0.05, 0.2))
        axis = rng.normal(size=3)
        axis /= np.linalg.norm(axis)
        prims.append((kind, c, r1, min(r2, r1), axis))
    return prims


def sdf_eval(prims, p):
    d = np.full(p.shape[:-1], np.inf)
    for kind, c, r1, r2, axis in prims:
        q = p - c
        if kind == 0:
            di = np.linalg.norm(q, axis=-1) - r1
        else:
            h = q @ axis
            radial = np.sqrt(np.maximum(np.sum(q * q, axis=-1) - h * h, 0.0))
            di = np.sqrt((radial - r1) ** 2 + h * h) - r2
        d = np.minimum(d, di)
    return d


def sdf_volume(n, seed, shade=0.8, compress=None, brick=8):
    """Solid-shell voxelization of the seeded SDF scene on an n^3 grid.

    Occupied: |sdf(c)| <= 0.5*sqrt(3)/n at voxel centres c = (v+0.5)/n.
    Normal: normalised central-difference gradient (h = 1/n), packed with
    `compress(normals (k,3) float32, shade) -> uint32 (k,)`; a zero word is
    bumped to 1 so the voxel stays non-empty. Returns uint32 (n,n,n) indexed
    [z, y, x] (x fastest in memory, as the .voxel layout).

    Bricks of `brick`^3 voxels whose centre is farther from the surface than
    the brick's half diagonal plus the shell width are skipped: the union of
    exact sphere/torus distances is 1-Lipschitz, so they hold no occupied
    voxel and the result equals a dense evaluation."""
    if compress is None:
        compress = compress_material_np
    prims = sdf_primitives(seed)
    out = np.zeros((n, n, n), dtype=np.uint32)
    thr = 0.5 * math.sqrt(3.0) / n
    h = 1.0 / n
    nb = n // brick
    bc = (np.arange(nb) * brick + brick / 2.0) / n
    Z, Y, X = np.meshgrid(bc, bc, bc, indexing="ij")
    db = sdf_eval(prims, np.stack([X, Y, Z], axis=-1))
    reach = (brick / 2.0) * math.sqrt(3.0) / n + thr + 2.0 / n
    near = np.argwhere(np.abs(db) <= reach)  # (k, 3) as [bz, by, bx]
    if len(near) == 0:
        return out
    off = np.stack(np.meshgrid(np.arange(brick), np.arange(brick), np.arange(brick), indexing="ij"), -1).reshape(-1, 3)
    for s0 in range(0, len(near), 4096):
        nb_ = near[s0:s0 + 4096]
        vox = (nb_[:, None, :] * brick + off[None, :, :]).reshape(-1, 3)  # [z, y, x]
        P = (vox[:, ::-1] + 0.5) / n  # (x, y, z)
        d = sdf_eval(prims, P)
        occ = np.abs(d) <= thr
        if not occ.any():
            continue
        Q = P[occ]
        g = np.stack([sdf_eval(prims, Q + h * e) - sdf_eval(prims, Q - h * e)
                      for e in np.eye(3)], axis=-1)
        g /= np.maximum(np.linalg.norm(g, axis=-1, keepdims=True), 1e-30)
        w = compress(g.astype(np.float32), shade)
        w[w == 0] = 1
        v = vox[occ]
        out[v[:, 0], v[:, 1], v[:, 2]] = w
    return out


def compress_material_np(n, shade):
    """Vectorised compressMaterial (src/Util.hpp:64-84) for float32 normals
    and a scalar or per-normal shade."""
    n = np.asarray(n, dtype=np.float32)
    a = np.abs(n)
    face = np.zeros(len(n), dtype=np.int64)
    dom = a[:, 0].copy()
    m1 = a[:, 1] > dom
    face[m1] = 1
    dom[m1] = a[m1, 1]
    m2 = a[:, 2] > dom
    face[m2] = 2
    dom[m2] = a[m2, 2]
    rows = np.arange(len(n))
    sign = (n[rows, face] < 0).astype(np.uint32)
    with np.errstate(divide="ignore", invalid="ignore"):
        n1 = (n[rows, (face + 1) % 3] / dom).astype(np.float32)
        n2 = (n[rows, (face + 2) % 3] / dom).astype(np.float32)
    f32 = np.float32
    u = ((n1 * f32(0.5) + f32(0.5)) * f32(2047.0)).astype(np.float32)
    v = ((n2 * f32(0.5) + f32(0.5)) * f32(2047.0)).astype(np.float32)
    c = (np.broadcast_to(np.asarray(shade, dtype=f32), (len(n),)) * f32(127.0)).astype(f32)

    def trunc(x):
        bad = ~(np.abs(x) < 2147483648.0)
        r = np.where(bad, 0, x).astype(np.int64)
        r[bad] = -(1 << 31)
        return r

    ui = np.minimum(trunc(u), 0x7FF).astype(np.int64) & 0xFFFFFFFF
    vi = np.minimum(trunc(v), 0x7FF).astype(np.int64) & 0xFFFFFFFF
    ci = np.minimum(trunc(c), 0x7F).astype(np.int64) & 0xFFFFFFFF
    w = (sign.astype(np.int64) << 31) | (face << 29) | ((ui << 18) & 0xFFFFFFFF) | ((vi << 7) & 0xFFFFFFFF) | ci
    return (w & 0xFFFFFFFF).astype(np.uint32)


# ---------------------------------------------------------------------------
# fBm heightfield terrain
# ---------------------------------------------------------------------------
#
# Definition shared bit for bit with the GPU terrain source (float32, every
# operation separately rounded, in this order):
#   lattice value  v(i, k, s) = (mix(i, k, s) >> 8) * 2^-24, mix = uint32 hash below
#   noise(u, w)    bilinear over the cell with smoothstep weights c*c*(3 - 2c)
#   h(x, z)        sum_{o<8} 0.5^o * noise(u*f_o, w*f_o, seed+o) / sum_o 0.5^o,
#                  u = (x+0.5)/n, w = (z+0.5)/n, f_o = 4*2^o
#   height H       n/8 + int(h * (n/2)); voxel (x, y, z) is solid for y <= H(x, z)
#   exposed shell  y in [min(H, Hmin + 1), H], Hmin = least height of the
#                  4-neighbour columns inside the grid (outside counts as solid)
#   normal         normalize(-gx, 1, -gz), g = (hv(+1) - hv(-1)) * 0.5 with
#                  hv = h * (n/2) and neighbour indices clamped to the grid
#   word           compressMaterial(normal, 0.35 + 0.65*h), 0 bumped to 1
TERRAIN_OCTAVES = 8
TERRAIN_BASE_FREQ = 4.0


def _mix32(i, k, s):
    """uint32 lattice hash (wrapping arithmetic)."""
    u32 = np.uint32
    with np.errstate(over="ignore"):
        h = (i.astype(u32) * u32(0x8DA6B343)) ^ (k.astype(u32) * u32(0xD8163841)) ^ (u32(s) * u32(0x9E3779B9))
        h ^= h >> u32(15)
        h *= u32(0x2C1B3C6D)
        h ^= h >> u32(12)
        h *= u32(0x297A2D39)
        h ^= h >> u32(15)
    return h


def _lattice(i, k, s):
    return (_mix32(i, k, s) >> np.uint32(8)).astype(np.float32) * np.float32(1.0 / 16777216.0)


def _smooth(c):
    f32 = np.float32
    return (c * c) * (f32(3.0) - f32(2.0) * c)


def terrain_h(x, z, n, seed):
    """Normalised fBm height h in [0, 1) for integer column coordinates."""
    f32 = np.float32
    u = (np.asarray(x).astype(f32) + f32(0.5)) / f32(n)
    w = (np.asarray(z).astype(f32) + f32(0.5)) / f32(n)
    total = np.zeros(u.shape, f32)
    norm, amp, freq = f32(0.0), f32(1.0), f32(TERRAIN_BASE_FREQ)
    for o in range(TERRAIN_OCTAVES):
        pu, pw = u * freq, w * freq
        fu, fw = np.floor(pu), np.floor(pw)
        cu, cw = _smooth(pu - fu), _smooth(pw - fw)
        iu, iw = fu.astype(np.int64), fw.astype(np.int64)
        s = (seed + o) & 0xFFFFFFFF
        v00, v10 = _lattice(iu, iw, s), _lattice(iu + 1, iw, s)
        v01, v11 = _lattice(iu, iw + 1, s), _lattice(iu + 1, iw + 1, s)
        a0 = v00 + (v10 - v00) * cu
        a1 = v01 + (v11 - v01) * cu
        total = total + amp * (a0 + (a1 - a0) * cw)
        norm = f32(norm + amp)
        amp = f32(amp * f32(0.5))
        freq = f32(freq * f32(2.0))
    return total / norm


def terrain_heights(n, seed, x=None, z=None):
    """Integer heights H and scaled heights hv = h*(n/2) on columns (x, z)
    (default: the whole grid, indexed [z, x])."""
    if x is None:
        z, x = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    h = terrain_h(x, z, n, seed)
    hv = h * np.float32(n // 2)
    return (n // 8 + hv.astype(np.int32)).astype(np.int32), hv, h


def terrain_volume(n, seed, compress=None):
    """Dense uint32 (n,n,n) [z, y, x] exposed-shell volume of the terrain (small n)."""
    if compress is None:
        compress = compress_material_np
    f32 = np.float32
    H, hv, h = terrain_heights(n, seed)
    big = np.iinfo(np.int32).max
    pad = np.pad(H, 1, constant_values=big)
    hmin = np.minimum(np.minimum(pad[1:-1, :-2], pad[1:-1, 2:]), np.minimum(pad[:-2, 1:-1], pad[2:, 1:-1]))
    lo = np.minimum(H, np.where(hmin == big, H, hmin + 1))
    idx = np.arange(n)
    xm, xp = np.maximum(idx - 1, 0), np.minimum(idx + 1, n - 1)
    gx = (hv[:, xp] - hv[:, xm]) * f32(0.5)
    gz = (hv[xp, :] - hv[xm, :]) * f32(0.5)
    a = np.stack([-gx, np.ones_like(gx), -gz], axis=-1).astype(f32)
    inv = f32(1.0) / np.sqrt((a[..., 0] * a[..., 0] + a[..., 1] * a[..., 1]) + a[..., 2] * a[..., 2])
    nrm = a * inv[..., None]
    shade = f32(0.35) + f32(0.65) * h
    words = compress(nrm.reshape(-1, 3), shade.reshape(-1)).reshape(n, n)
    words[words == 0] = 1
    vol = np.zeros((n, n, n), np.uint32)
    y = np.arange(n)[None, :, None]
    occ = (y >= lo[:, None, :]) & (y <= H[:, None, :])
    vol[occ] = np.broadcast_to(words[:, None, :], vol.shape)[occ]
    return vol


def terrain_view(seed):
    """Seeded camera above the terrain and a directional light: orbit pitch
    U[-50,-25] (negative pitch looks down), yaw U[0,360), radius 1; light
    elevation U[20,60] degrees, azimuth U[0,360). Returns (pitch, yaw, light
    float32[3] unit vector toward the light)."""
    rng = np.random.default_rng(seed)
    pitch = float(np.float32(rng.uniform(-50.0, -25.0)))
    yaw = float(np.float32(rng.uniform(0.0, 360.0)))
    el = math.radians(rng.uniform(20.0, 60.0))
    az = math.radians(rng.uniform(0.0, 360.0))
    light = np.array([math.cos(el) * math.cos(az), math.sin(el), math.cos(el) * math.sin(az)], np.float64)
    return pitch, yaw, (light / np.linalg.norm(light)).astype(np.float32)


def terrain_volume_words(n, seed, x, z, compress=None):
    """Leaf words of columns (x, z) (every voxel of a column shares one word)."""
    if compress is None:
        compress = compress_material_np
    f32 = np.float32
    x, z = np.asarray(x), np.asarray(z)
    half = f32(n // 2)
    h = terrain_h(x, z, n, seed)
    hx = [terrain_h(np.minimum(x + 1, n - 1), z, n, seed), terrain_h(np.maximum(x - 1, 0), z, n, seed)]
    hz = [terrain_h(x, np.minimum(z + 1, n - 1), n, seed), terrain_h(x, np.maximum(z - 1, 0), n, seed)]
    gx = (hx[0] * half - hx[1] * half) * f32(0.5)
    gz = (hz[0] * half - hz[1] * half) * f32(0.5)
    a = np.stack([-gx, np.ones_like(gx), -gz], axis=-1).astype(f32)
    inv = f32(1.0) / np.sqrt((a[:, 0] * a[:, 0] + a[:, 1] * a[:, 1]) + a[:, 2] * a[:, 2])
    w = compress((a * inv[:, None]).astype(f32), f32(0.35) + f32(0.65) * h)
    w[w == 0] = 1
    return w


def terrain_shell_count(H):
    """Number of exposed-shell voxels implied by the column heights H [z, x]."""
    big = np.iinfo(np.int32).max
    H = np.asarray(H, np.int32)
    m = np.full(H.shape, big, np.int32)
    m[:, 1:] = np.minimum(m[:, 1:], H[:, :-1])
    m[:, :-1] = np.minimum(m[:, :-1], H[:, 1:])
    m[1:, :] = np.minimum(m[1:, :], H[:-1, :])
    m[:-1, :] = np.minimum(m[:-1, :], H[1:, :])
    lo = np.minimum(H, np.where(m == big, H, m + 1))
    return int((H.astype(np.int64) - lo + 1).sum())


def write_voxel(path, vol):
    """Raw .voxel: int32 w,h,d then uint32 data, x fastest (vol is [z,y,x])."""
    d, h, w = vol.shape
    with open(path, "wb") as fp:
        fp.write(struct.pack("<iii", w, h, d))
        fp.write(np.ascontiguousarray(vol, dtype=np.uint32).tobytes())


# ---------------------------------------------------------------------------
# cameras
# ---------------------------------------------------------------------------


def orbit_angles(seed, count):
    """Seeded orbit cameras: yaw U[0,360), pitch U[-60,60] (degrees)."""
    rng = np.random.default_rng(seed)
    yaw = rng.uniform(0.0, 360.0, count).astype(np.float32)
    pitch = rng.uniform(-60.0, 60.0, count).astype(np.float32)
    return [(float(p), float(y)) for p, y in zip(pitch, yaw)]


# ---------------------------------------------------------------------------
# multi-octree, multi-view workload (batched / served rendering)
# ---------------------------------------------------------------------------


def mixed_octrees(seed=4096, count=64):
    """64 octree specs: resolution 2^round(U[7,11]); even entries are SDF
    scenes, odd ones displaced icospheres whose triangle count scales with
    resolution^2 (frequency 316 at 2048^3 = 1,997,120 triangles)."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        res = 1 << int(np.round(rng.uniform(7.0, 11.0)))
        s = int(rng.integers(0, 1 << 31))
        if i % 2 == 0:
            specs.append({"kind": "sdf", "resolution": res, "seed": s})
        else:
            specs.append({"kind": "mesh", "resolution": res, "seed": s,
                          "frequency": max(1, int(round(316 * res / 2048)))})
    return specs


def mixed_views(seed=8192, count=256, trees=64):
    """256 view specs: pixel count log-uniform in [320*240, 1920*1080], aspect
    4:3 or 16:9, seeded orbit camera, octree id U{0..trees-1}."""
    rng = np.random.default_rng(seed)
    views = []
    for _ in range(count):
        pixels = math.exp(rng.uniform(math.log(320 * 240), math.log(1920 * 1080)))
        aspect = 4.0 / 3.0 if rng.random() < 0.5 else 16.0 / 9.0
        w = max(8, int(round(math.sqrt(pixels * aspect))))
        h = max(8, int(round(w / aspect)))
        views.append({"tree": int(rng.integers(0, trees)), "width": w, "height": h,
                      "pitch": float(np.float32(rng.uniform(-60.0, 60.0))),
                      "yaw": float(np.float32(rng.uniform(0.0, 360.0)))})
    return views


def build_mixed_octree(pkg, spec):
    """Builds one mixed_octrees() entry on the GPU (SDF source or mesh builder
    with 8 threads' partition and a 1 GiB budget)."""
    if spec["kind"] == "sdf":
        return pkg.VoxelOctree.build_from_sdf(sdf_primitives(spec["seed"]), spec["resolution"])
    v, f = icosphere(spec["frequency"], seed=spec["seed"] % 100003, displacement=0.05)
    tab, has_n, counts, idx = mesh_vertex_table(v, f)
    return pkg.VoxelOctree.build_from_mesh(tab, has_n, counts, idx, spec["resolution"], threads=8,
                                           memory_budget=1 << 30)
