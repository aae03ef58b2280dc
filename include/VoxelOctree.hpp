/*
 * VoxelOctree.hpp — drop-in C++ front end of libsvo_hip.so.
 *
 * Replaces the reference's VoxelOctree (src/VoxelOctree.hpp:37-58) and the
 * constructor surface of its build inputs PlyLoader (src/PlyLoader.hpp:96-113)
 * and VoxelData (src/VoxelData.hpp:101-144), so that the reference's callers
 * (src/Main.cpp viewer and builder) compile unchanged against it:
 *
 *     PlyLoader loader(in);                                 // Main.cpp:323
 *     loader.convertToVolume(tmp, resolution, dataMemory);  // Main.cpp:317 (GPU voxelization)
 *     VoxelData data(&loader, resolution, dataMemory);      // Main.cpp:324
 *     VoxelOctree tree(&data);                              // Main.cpp:325 (GPU build)
 *     tree.save(out);                                       // Main.cpp:326
 *     VoxelOctree viewer(file);                             // Main.cpp:333
 *     viewer.raymarch(o, d, rayScale, normal, t);           // Main.cpp:118,181
 *
 * Every member forwards to the C ABI in svo.h; nothing is computed here.
 * Build inputs only record their arguments: the PLY is parsed and voxelized
 * on the GPU when the octree is constructed. The ray and centre vector type
 * is the caller's Vec3 (the reference's math/Vec3.hpp by default; override
 * with -DSVO_VEC3_HEADER="..."), which must have float x, y, z members and a
 * (float, float, float) constructor.
 *
 * Errors: the reference reports none (failed loads leave an empty tree,
 * src/VoxelOctree.cpp:57-60). Here a failed call throws svo_error, unless
 * SVO_NO_EXCEPTIONS is defined, in which case ok() turns false and
 * error() holds the message.
 */
#ifndef SVO_VOXELOCTREE_HPP
#define SVO_VOXELOCTREE_HPP

#ifndef SVO_VEC3_HEADER
#define SVO_VEC3_HEADER "math/Vec3.hpp"
#endif
#include SVO_VEC3_HEADER

#include <cstdint>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "svo.h"

class svo_error : public std::runtime_error {
public:
    svo_error(int code, const std::string &msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

namespace svo_shim {
inline bool check(int code, std::string *err) {
    if (code == SVO_OK) return true;
#ifndef SVO_NO_EXCEPTIONS
    throw svo_error(code, svo_last_error());
#else
    if (err) *err = svo_last_error();
    return false;
#endif
}
}  // namespace svo_shim

/* PlyLoader(path), src/PlyLoader.cpp:66-79: remembers the mesh file. The
 * thread count fixes the sub-block partition, as for VoxelData below. */
class PlyLoader {
public:
    explicit PlyLoader(const char *path) : path_(path) {}
    const std::string &path() const { return path_; }
    void setThreads(int threads) { threads_ = threads; }

    /* convertToVolume, src/PlyLoader.cpp:506-534 (builder mode 1,
     * src/Main.cpp:315-320): the dense .voxel file, voxelized on the GPU. */
    void convertToVolume(const char *path, int maxSize, size_t memoryBudget) {
        svo_build_opts o{};
        o.memory_budget = memoryBudget;
        o.threads = threads_;
        svo_shim::check(svo_convert_to_volume(path_.c_str(), maxSize, &o, path), nullptr);
    }

private:
    std::string path_;
    int threads_ = 0;
};

/* VoxelData(PlyLoader*, sideLength, mem), src/VoxelData.cpp:50-56, and
 * VoxelData(path, mem) for raw .voxel files, src/VoxelData.cpp:37-48. The
 * reference's host thread count fixes the mesh sub-block partition
 * (src/PlyLoader.cpp:418); it is taken from this host as there, or set with
 * setThreads() to reproduce another machine's output bytes. */
class VoxelData {
public:
    VoxelData(PlyLoader *loader, int sideLength, size_t mem)
        : mesh_(loader ? loader->path() : std::string()), side_(sideLength), mem_(mem), fromMesh_(true) {}
    VoxelData(const char *path, size_t mem) : voxels_(path), mem_(mem), fromMesh_(false) {}
    void setThreads(int threads) { threads_ = threads; }

    const std::string &meshPath() const { return mesh_; }
    const std::string &voxelPath() const { return voxels_; }
    int resolution() const { return side_; }
    size_t memory() const { return mem_; }
    int threads() const { return threads_; }
    bool fromMesh() const { return fromMesh_; }

private:
    std::string mesh_, voxels_;
    int side_ = 0;
    size_t mem_ = 0;
    int threads_ = 0;
    bool fromMesh_;
};

class VoxelOctree {
public:
    /* VoxelOctree(const char *path), src/VoxelOctree.cpp:57-90 */
    explicit VoxelOctree(const char *path) { svo_shim::check(svo_load(path, &h_), &err_); }

    /* VoxelOctree(VoxelData*), src/VoxelOctree.cpp:125-137: built on the GPU */
    explicit VoxelOctree(VoxelData *voxels) {
        svo_build_opts o{};
        o.memory_budget = voxels->memory();
        o.threads = voxels->threads();
        if (voxels->fromMesh())
            svo_shim::check(svo_build_from_ply(voxels->meshPath().c_str(), voxels->resolution(), &o, &h_), &err_);
        else
            svo_shim::check(svo_build_from_voxels(voxels->voxelPath().c_str(), &o, &h_), &err_);
    }

    ~VoxelOctree() { svo_free(h_); }
    VoxelOctree(const VoxelOctree &) = delete;
    VoxelOctree &operator=(const VoxelOctree &) = delete;

    /* VoxelOctree::save, src/VoxelOctree.cpp:92-123 (byte-identical .oct) */
    void save(const char *path) { svo_shim::check(svo_save(h_, path), &err_); }

    /* VoxelOctree::raymarch, src/VoxelOctree.cpp:207-346. One ray per call is
     * walked on the calling thread by the GPU kernels' traversal source (a
     * device round trip per ray would cost ~30 us); raymarchBatch,
     * renderPrimary and renderFrame run tiles and frames on the GPU. normal is
     * written on a leaf hit, t on any hit, as in the reference. */
    bool raymarch(const Vec3 &o, const Vec3 &d, float rayScale, uint32_t &normal, float &t) {
        const float of[3] = {o.x, o.y, o.z}, df[3] = {d.x, d.y, d.z};
        int hit = 0;
        svo_shim::check(svo_raymarch(h_, of, df, rayScale, &normal, &t, &hit), &err_);
        return hit != 0;
    }

    /* VoxelOctree::center, src/VoxelOctree.hpp:55-57 */
    Vec3 center() const {
        float c[3] = {0, 0, 0};
        svo_center(h_, c);
        return Vec3(c[0], c[1], c[2]);
    }

    /* Many rays per call (host arrays; o, d: n x 3 floats): hit codes
     * 0 miss / 1 leaf / 2 level-of-detail, and the same normal/t. */
    void raymarchBatch(int64_t n, const float *o, const float *d, const float *rayScale, float rayScale0,
                       uint8_t *hit, uint32_t *normal, float *t, uint64_t *leaf = nullptr) {
        svo_shim::check(svo_raymarch_batch(h_, n, o, d, rayScale, rayScale0, hit, normal, t, leaf), &err_);
    }

    /* Primary rays of a whole frame from the viewer's orbit camera. */
    svo_camera orbitCamera(float pitch, float yaw, float radius) const {
        svo_camera c;
        svo_shim::check(svo_orbit_camera(h_, pitch, yaw, radius, &c), const_cast<std::string *>(&err_));
        return c;
    }
    void renderPrimary(const svo_camera &cam, int width, int height, uint8_t *hit, uint32_t *normal, float *t,
                       uint64_t *leaf = nullptr) {
        svo_shim::check(svo_render_primary(h_, &cam, width, height, hit, normal, t, leaf), &err_);
    }

    /* The viewer's frame (src/Main.cpp renderBatch/renderTile/shade, :81-202)
     * as one call: rgba is width*height pixels (0xFF000000 | grey*0x010101,
     * skipped tiles 0); stride 1, or 3 while the camera moves. */
    void renderFrame(const svo_camera &cam, int width, int height, uint32_t *rgba, int stride = 1) {
        svo_shim::check(svo_render_frame(h_, &cam, width, height, 16, stride, rgba, nullptr), &err_);
    }

    svo_octree *handle() const { return h_; }
    bool ok() const { return h_ != nullptr && err_.empty(); }
    const std::string &error() const { return err_; }

private:
    svo_octree *h_ = nullptr;
    std::string err_;
};

#endif /* SVO_VOXELOCTREE_HPP */
