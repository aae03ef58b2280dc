// Test-only command-line driver over the reference's own classes.
//
// This file is part of the oracle: it is compiled together with the
// reference's sources where they lie (see oracle/Makefile, target `ref`) into
// oracle/_ref/svo_ref, and is run only by tests/golden/make_goldens.py in the
// build container to produce golden vectors. Nothing in the product links it.
//
// Commands (all binary files are little-endian, tightly packed):
//   build-ply   <in.ply> <resolution> <threads> <mem_bytes> <out.oct>
//       PlyLoader -> VoxelData(loader, res, mem) -> VoxelOctree -> save
//       (src/Main.cpp:322-327, in-memory builder mode)
//   convert     <in.ply> <resolution> <threads> <mem_bytes> <out.voxel>
//       PlyLoader::convertToVolume (src/PlyLoader.cpp:506-534)
//   build-voxel <in.voxel> <threads> <mem_bytes> <out.oct>
//       VoxelData(path, mem) -> VoxelOctree -> save (src/Main.cpp:315-321)
//   resave      <in.oct> <out.oct>
//       VoxelOctree(path) -> save: exercises the LZ4 1.7.1 stream encoder
//   raymarch    <in.oct> <rays.f32> <out.bin>
//       rays: n x {ox,oy,oz,dx,dy,dz,rayScale} float32
//       out:  n x {uint32 hit, uint32 normal, float32 t}; normal/t are 0 when
//             the reference leaves them unwritten (miss or LOD hit)
//   tribox      <in.f32> <out.i32>
//       in: n x {center[3], half[3], v0[3], v1[3], v2[3]} float32 -> triBoxOverlap
//   material    <in.f32> <out.bin>
//       in: n x {nx,ny,nz,shade} -> compressMaterial, then decompressMaterial
//       out: n x {uint32 word, float nx', ny', nz', shade'}
//   camera      <in.f32> <out.f32>
//       in: n x {pitch, yaw, radius, cx, cy, cz} -> INV_MODELVIEW matrix as
//       built by src/Main.cpp:212,244-250 and src/math/MatrixStack.cpp:65-66
//       out: n x 16 floats
#include "VoxelOctree.hpp"
#include "VoxelData.hpp"
#include "PlyLoader.hpp"
#include "Util.hpp"
#include "math/Mat4.hpp"
#include "thread/ThreadUtils.hpp"
#include "third-party/tribox3.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

template <typename T>
static std::vector<T> readAll(const char *path) {
    std::vector<T> out;
    FILE *fp = fopen(path, "rb");
    if (!fp) { fprintf(stderr, "cannot open %s\n", path); exit(2); }
    fseek(fp, 0, SEEK_END);
    long n = ftell(fp);
    fseek(fp, 0, SEEK_SET);
    out.resize(size_t(n) / sizeof(T));
    if (!out.empty() && fread(out.data(), sizeof(T), out.size(), fp) != out.size()) exit(3);
    fclose(fp);
    return out;
}

template <typename T>
static void writeAll(const char *path, const std::vector<T> &v) {
    FILE *fp = fopen(path, "wb");
    if (!fp) { fprintf(stderr, "cannot write %s\n", path); exit(2); }
    if (!v.empty()) fwrite(v.data(), sizeof(T), v.size(), fp);
    fclose(fp);
}

int main(int argc, char **argv) {
    if (argc < 2) return 1;
    std::string cmd = argv[1];

    if (cmd == "build-ply" && argc == 7) {
        ThreadUtils::startThreads(atoi(argv[4]));
        size_t mem = size_t(strtoull(argv[5], 0, 10));
        std::unique_ptr<PlyLoader> loader(new PlyLoader(argv[2]));
        std::unique_ptr<VoxelData> data(new VoxelData(loader.get(), atoi(argv[3]), mem));
        std::unique_ptr<VoxelOctree> tree(new VoxelOctree(data.get()));
        tree->save(argv[6]);
        return 0;
    }
    if (cmd == "convert" && argc == 7) {
        ThreadUtils::startThreads(atoi(argv[4]));
        size_t mem = size_t(strtoull(argv[5], 0, 10));
        std::unique_ptr<PlyLoader> loader(new PlyLoader(argv[2]));
        loader->convertToVolume(argv[6], atoi(argv[3]), mem);
        return 0;
    }
    if (cmd == "build-voxel" && argc == 6) {
        ThreadUtils::startThreads(atoi(argv[3]));
        size_t mem = size_t(strtoull(argv[4], 0, 10));
        std::unique_ptr<VoxelData> data(new VoxelData(argv[2], mem));
        std::unique_ptr<VoxelOctree> tree(new VoxelOctree(data.get()));
        tree->save(argv[5]);
        return 0;
    }
    if (cmd == "resave" && argc == 4) {
        std::unique_ptr<VoxelOctree> tree(new VoxelOctree(argv[2]));
        tree->save(argv[3]);
        return 0;
    }
    if (cmd == "raymarch" && argc == 5) {
        std::unique_ptr<VoxelOctree> tree(new VoxelOctree(argv[2]));
        std::vector<float> rays = readAll<float>(argv[3]);
        size_t n = rays.size() / 7;
        std::vector<uint32> out(n * 3, 0);
        for (size_t i = 0; i < n; ++i) {
            const float *r = &rays[i * 7];
            uint32 normal = 0;
            float t = 0.0f;
            bool hit = tree->raymarch(Vec3(r[0], r[1], r[2]), Vec3(r[3], r[4], r[5]), r[6], normal, t);
            out[i * 3 + 0] = hit ? 1u : 0u;
            out[i * 3 + 1] = normal;
            memcpy(&out[i * 3 + 2], &t, 4);
        }
        writeAll(argv[4], out);
        return 0;
    }
    if (cmd == "tribox" && argc == 4) {
        std::vector<float> in = readAll<float>(argv[2]);
        size_t n = in.size() / 15;
        std::vector<int> out(n);
        for (size_t i = 0; i < n; ++i) {
            float c[3], h[3], tv[3][3];
            memcpy(c, &in[i * 15 + 0], 12);
            memcpy(h, &in[i * 15 + 3], 12);
            memcpy(tv, &in[i * 15 + 6], 36);
            out[i] = triBoxOverlap(c, h, tv);
        }
        writeAll(argv[3], out);
        return 0;
    }
    if (cmd == "material" && argc == 4) {
        std::vector<float> in = readAll<float>(argv[2]);
        size_t n = in.size() / 4;
        std::vector<uint32> out(n * 5);
        for (size_t i = 0; i < n; ++i) {
            Vec3 nrm(in[i * 4 + 0], in[i * 4 + 1], in[i * 4 + 2]);
            uint32 w = compressMaterial(nrm, in[i * 4 + 3]);
            Vec3 dn;
            float shade;
            decompressMaterial(w, dn, shade);
            out[i * 5 + 0] = w;
            memcpy(&out[i * 5 + 1], &dn.x, 4);
            memcpy(&out[i * 5 + 2], &dn.y, 4);
            memcpy(&out[i * 5 + 3], &dn.z, 4);
            memcpy(&out[i * 5 + 4], &shade, 4);
        }
        writeAll(argv[3], out);
        return 0;
    }
    if (cmd == "camera" && argc == 4) {
        std::vector<float> in = readAll<float>(argv[2]);
        size_t n = in.size() / 6;
        std::vector<float> out(n * 16);
        for (size_t i = 0; i < n; ++i) {
            const float *p = &in[i * 6];
            Mat4 model = Mat4::rotXYZ(Vec3(p[0], 0.0f, 0.0f)) * Mat4::rotXYZ(Vec3(0.0f, p[1], 0.0f));
            Mat4 view = Mat4::translate(Vec3(0.0f, 0.0f, -p[2]));
            Mat4 tform = model.pseudoInvert() * view;
            memcpy(&out[i * 16], tform.a, 64);
        }
        writeAll(argv[3], out);
        return 0;
    }
    fprintf(stderr, "bad command\n");
    return 1;
}
