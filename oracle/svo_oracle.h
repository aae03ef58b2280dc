/*
 * svo_oracle.h — CPU restatement of tunabrain/sparse-voxel-octrees' hot paths.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the parity oracle: tests/, the
 * smoke check and bench.py's CPU-baseline leg may call it, as the checker; the
 * product (libsvo_hip.so) never links, loads or calls it.
 *
 * Every function restates a reference function in plain C, keeping the
 * reference's IEEE single-precision operation order (the library is built with
 * -ffp-contract=off for baseline x86-64, so no FMA is formed). Citations are
 * path:line relative to the reference repository root.
 *
 * Hit codes used throughout: 0 = miss, 1 = leaf hit (normal word read),
 * 2 = level-of-detail hit (raymarch returned at `maxTC*rayScale >= scaleExp2`;
 * the reference leaves `normal` unwritten there).
 */
#ifndef SVO_ORACLE_H
#define SVO_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_LEAF_NONE UINT64_MAX

/* ---- arithmetic primitives (src/Util.hpp) --------------------------------- */
float oracle_inv_sqrt(float x);                                   /* Util.hpp:47-58 */
uint32_t oracle_compress_material(const float n[3], float shade); /* Util.hpp:64-84 */
void oracle_decompress_material(uint32_t w, float n[3], float *shade); /* :86-100 */
int oracle_tribox(const float c[3], const float h[3], const float tv[9]); /* tribox3.c:132-206 */
/* batch forms: normals n x 3 + shades n -> words; words -> normals n x 3 + shades;
 * records n x 15 {center, half, v0, v1, v2} -> overlap */
void oracle_compress_material_batch(int64_t n, const float *normals, const float *shades, uint32_t *out);
void oracle_decompress_material_batch(int64_t n, const uint32_t *w, float *normals, float *shades);
void oracle_tribox_batch(int64_t n, const float *rec, int32_t *out);

/* ---- raymarch (src/VoxelOctree.cpp:207-346) ------------------------------- */
/* counts (may be NULL) accumulates {descriptor, far, leaf} word reads. Returns
 * the hit code, or -1 if the walk leaves the word array (malformed tree). A ray
 * with a NaN coefficient (NaN component, or an infinite origin on an axis whose
 * direction is infinite) returns 0: the reference never ends it. */
int oracle_raymarch(const uint32_t *tree, uint64_t nwords, const float o[3], const float d[3],
                    float rayScale, uint32_t *normal, float *t, uint64_t *leaf, uint64_t counts[3]);

/* Batch over n rays with `threads` POSIX threads. rayScale may be NULL (then
 * scale0 is used for every ray). Any output pointer may be NULL. On a miss
 * normal = 0, t = 0, leaf = ORACLE_LEAF_NONE; on a LOD hit normal = 0 and
 * leaf = ORACLE_LEAF_NONE. Returns 0, or -1 if any ray hit a malformed tree. */
int oracle_raymarch_batch(const uint32_t *tree, uint64_t nwords, int64_t n, const float *o,
                          const float *d, const float *rayScale, float scale0, uint8_t *hit,
                          uint32_t *normal, float *t, uint64_t *leaf, uint64_t counts[3],
                          int threads);

/* ---- camera and primary rays (src/Main.cpp, src/math/Mat4.cpp) ------------ */
/* Orbit camera as src/Main.cpp:212,244-250: MODEL = rotXYZ(pitch,0,0)*rotXYZ(0,yaw,0),
 * VIEW = translate(0,0,-radius), tform = INV_MODELVIEW (MatrixStack.cpp:65-66).
 * Writes the 16 floats of tform (row-major, before the translation column is
 * cleared) and the ray origin pos = tform*0 + center + 1 (Main.cpp:150-152). */
void oracle_orbit_camera(float pitch, float yaw, float radius, const float center[3],
                         float tform[16], float pos[3]);

/* Primary ray directions for a width x height image with closed-form
 * dx = -1 + x*scale, dy = aspect - y*scale (the per-pixel expression of
 * src/Main.cpp:97-113 without incremental accumulation). dirs: w*h*3. */
void oracle_primary_dirs(const float tform[16], int width, int height, float *dirs);

/* Full viewer frame: src/Main.cpp renderBatch/renderTile/shade (:81-202) over
 * `stripes` horizontal stripes as the viewer's threads split the image
 * (:351-361): per stripe a coarse pass (one LOD ray per 8x8 tile corner,
 * rayScale = 2*8/(planeDist*H)) and, for tiles with a finite corner depth,
 * per-pixel rays from pos + dir*max(minT - 0.03, 0), shaded to RGBA; pixels of
 * skipped tiles stay 0. tform is INV_MODELVIEW with its translation (as
 * oracle_orbit_camera writes it); stride is 1 (3 while the viewer drags).
 * depth (may be NULL) receives the coarse depths, stripe after stripe. */
void oracle_render_frame(const uint32_t *tree, uint64_t nwords, const float tform[16], const float center[3],
                         int width, int height, int stripes, int stride, uint32_t *rgba, float *depth);
/* Number of coarse depth entries oracle_render_frame writes. */
int64_t oracle_frame_depth_count(int width, int height, int stripes);

/* ---- .oct decode (src/VoxelOctree.cpp:57-90) ------------------------------ */
/* Parses an in-memory .oct image. Allocates *words (free with oracle_free). */
int oracle_oct_decode(const uint8_t *file, uint64_t size, float center[3], uint32_t **words,
                      uint64_t *nwords);

/* Structural walk of a word array. kind (nwords bytes, may be NULL):
 * 1 descriptor, 2 far word, 3 leaf, 0 unreferenced. desc_per_level: up to
 * 24 entries. Returns 0, or -1 if malformed. */
int oracle_octree_walk(const uint32_t *tree, uint64_t nwords, uint8_t *kind,
                       uint64_t desc_per_level[24], uint64_t *leaves, uint64_t *far_words,
                       int *depth);

/* ---- construction (src/PlyLoader.cpp, src/VoxelData.*, src/VoxelOctree.cpp) */
typedef struct oracle_mesh oracle_mesh;

/* verts: nv x {x,y,z,nx,ny,nz,r,g,b} as PlyLoader::readVertices stores them
 * (absent normals = 0, absent colours = 255). faces: CSR (face_counts[nf],
 * face_idx flat). has_normals as PlyLoader::_hasNormals. */
oracle_mesh *oracle_mesh_create(const float *verts, int64_t nv, int has_normals,
                                const int32_t *face_counts, const int32_t *face_idx,
                                int64_t nfaces);
void oracle_mesh_free(oracle_mesh *m);
int64_t oracle_mesh_tri_count(const oracle_mesh *m);
/* out: ntri x 33 floats {v1.pos,v1.n,v1.c, v2.., v3.., lower, upper} */
void oracle_mesh_tris(const oracle_mesh *m, float *out);
/* scaled bounds after PlyLoader::rescaleVertices: lower[3], upper[3] */
void oracle_mesh_bounds(const oracle_mesh *m, float lower[3], float upper[3]);

/* Build parameters the reference derives from the host: `threads` feeds
 * findBestBlockPartition (PlyLoader.cpp:418), `mem` is VoxelData's budget. */
typedef struct {
    int data_dims[3];      /* VoxelData::_dataW/H/D */
    int side_length;       /* VoxelData::sideLength() */
    int cache_size;        /* VoxelData::_maxCacheableSize */
    int sub_block[3];      /* PlyLoader::_subBlockW/H/D (mesh source) */
    int grid[3];           /* PlyLoader::_gridW/H/D (mesh source) */
    uint64_t bin_entries;  /* PlyLoader block-list entries (mesh source) */
    uint64_t leaves;       /* leaf words emitted */
    uint64_t hazard_voxels;/* voxels present in the volume but never emitted */
} oracle_build_info;

int oracle_build_mesh(oracle_mesh *m, int resolution, int threads, uint64_t mem,
                      uint32_t **words, uint64_t *nwords, float center[3],
                      oracle_build_info *info);

/* File-source build (VoxelData(path, mem), src/VoxelData.cpp:37-48) with the
 * raw volume already in memory: grid is w*h*d uint32, x fastest. */
int oracle_build_grid(const uint32_t *grid, int w, int h, int d, uint64_t mem, uint32_t **words,
                      uint64_t *nwords, float center[3], oracle_build_info *info);

/* Dense voxelization of the whole mesh as the mesh source sees it: every
 * voxel the reference's processBlock writes, for the given sub-block dims
 * and cache size (no octree). out: dims[0]*dims[1]*dims[2] uint32. */
int oracle_voxelize_mesh(oracle_mesh *m, int resolution, int threads, uint64_t mem,
                         uint32_t *out, int dims[3]);

/* PlyLoader::convertToVolume (src/PlyLoader.cpp:506-534): the .voxel body
 * (dims[0]*dims[1]*dims[2] uint32, malloc'ed into *out) the reference writes
 * after its w, h, d header. *sliceZ = slices per slab. Returns -2 when one
 * slice does not fit the memory budget. */
int oracle_convert_to_volume(oracle_mesh *m, int maxSize, int threads, uint64_t mem, uint32_t **out,
                             int dims[3], int *sliceZ);

void oracle_free(void *p);

#ifdef __cplusplus
}
#endif
#endif
