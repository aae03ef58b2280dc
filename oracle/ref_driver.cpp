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
//       PlyLoader::convertToVolume (src/PlyLoader.cpp:506-534), its loop restated
//       on the reference's members (the function itself dies with SIGFPE)
//   build-voxel <in.voxel> <threads> <mem_bytes> <out.oct>
//       VoxelData(path, mem) -> VoxelOctree -> save (src/Main.cpp:315-321)
//   resave      <in.oct> <out.oct>
//       VoxelOctree(path) -> save: exercises the LZ4 1.7.1 stream encoder
//   raymarch    <in.oct> <rays.f32> <out.bin> [threads]
//       (prints the wall time of the raymarch loop alone)
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
//   frame       <in.oct> <cams.f32> <width> <height> <stride> <out.bin> [workers]
//       cams: n x {pitch, yaw, radius}; per camera the viewer frame of
//       src/Main.cpp:81-202 over 16 stripes (src/Main.cpp:351-361): Main.cpp
//       needs SDL and cannot be compiled here, so this loop restates it on the
//       reference's own VoxelOctree::raymarch, Mat4, Vec3, invSqrt and
//       decompressMaterial.
//       out: per camera width*height uint32 RGBA, then the stripes' coarse
//       depth tables back to back (float32)
#include "VoxelOctree.hpp"
#include "VoxelData.hpp"
#include "PlyLoader.hpp"
#include "Util.hpp"
#include "math/Mat4.hpp"
#include "thread/ThreadUtils.hpp"
#include "third-party/tribox3.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <thread>
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
        // the builder's in-memory mode (src/Main.cpp:322-327), each phase timed:
        // PLY parse, voxelization + octree build, LZ4 save
        ThreadUtils::startThreads(atoi(argv[4]));
        size_t mem = size_t(strtoull(argv[5], 0, 10));
        auto t0 = std::chrono::steady_clock::now();
        std::unique_ptr<PlyLoader> loader(new PlyLoader(argv[2]));
        auto t1 = std::chrono::steady_clock::now();
        std::unique_ptr<VoxelData> data(new VoxelData(loader.get(), atoi(argv[3]), mem));
        std::unique_ptr<VoxelOctree> tree(new VoxelOctree(data.get()));
        auto t2 = std::chrono::steady_clock::now();
        tree->save(argv[6]);
        auto t3 = std::chrono::steady_clock::now();
        auto sec = [](std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
            return std::chrono::duration<double>(b - a).count();
        };
        printf("build_seconds parse %.6f build %.6f save %.6f threads %d\n", sec(t0, t1), sec(t1, t2), sec(t2, t3),
               atoi(argv[4]));
        return 0;
    }
    if (cmd == "convert" && argc == 7) {
        // PlyLoader::convertToVolume (src/PlyLoader.cpp:506-534) restated on the
        // reference's own public members. The function itself never calls
        // isBlockEmpty, so processBlock's progress line divides by a zero
        // _numNonZeroBlocks (PlyLoader.cpp:466) and the process dies with SIGFPE
        // before the first slab is written. Here isBlockEmpty runs once per slab
        // first (it only feeds that counter), and the slab buffer is zeroed
        // (the reference's is uninitialised memory, zero when freshly mapped);
        // the voxels come from the reference's setupBlockProcessing and
        // processBlock, slab after slab into the same buffer, as there.
        ThreadUtils::startThreads(atoi(argv[4]));
        const int maxSize = atoi(argv[3]);
        size_t mem = size_t(strtoull(argv[5], 0, 10));
        PlyLoader loader(argv[2]);
        int w, h, d;
        loader.suggestedDimensions(maxSize, w, h, d);
        size_t sliceCost = loader.blockMemRequirement(w, h, 1);
        int sliceZ = std::min((int)(mem / sliceCost), d);
        if (sliceZ == 0) return 1;
        FILE *fp = fopen(argv[6], "wb");
        if (!fp) return 1;
        std::unique_ptr<uint32[]> data(new uint32[size_t(w) * size_t(h) * size_t(sliceZ)]());
        loader.setupBlockProcessing(maxSize, w, h, sliceZ, w, h, d);
        for (int z = 0; z < d; z += sliceZ) loader.isBlockEmpty(0, 0, z);
        fwrite(&w, 4, 1, fp);
        fwrite(&h, 4, 1, fp);
        fwrite(&d, 4, 1, fp);
        for (int z = 0; z < d; z += sliceZ) {
            loader.processBlock(data.get(), 0, 0, z, w, h, sliceZ);
            fwrite(data.get(), sizeof(uint32), size_t(w) * size_t(h) * size_t(std::min(sliceZ, d - z)), fp);
        }
        fclose(fp);
        loader.teardownBlockProcessing();
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
    if (cmd == "raymarch" && (argc == 5 || argc == 6)) {
        std::unique_ptr<VoxelOctree> tree(new VoxelOctree(argv[2]));
        std::vector<float> rays = readAll<float>(argv[3]);
        size_t n = rays.size() / 7;
        std::vector<uint32> out(n * 3, 0);
        // the viewer calls raymarch from many threads on one tree (src/Main.cpp:342-367)
        const int threads = argc == 6 ? std::max(1, atoi(argv[5])) : 1;
        std::atomic<size_t> next(0);
        auto work = [&]() {
            for (size_t b; (b = next.fetch_add(4096)) < n;)
                for (size_t i = b; i < std::min(n, b + 4096); ++i) {
                    const float *r = &rays[i * 7];
                    uint32 normal = 0;
                    float t = 0.0f;
                    bool hit = tree->raymarch(Vec3(r[0], r[1], r[2]), Vec3(r[3], r[4], r[5]), r[6], normal, t);
                    out[i * 3 + 0] = hit ? 1u : 0u;
                    out[i * 3 + 1] = normal;
                    memcpy(&out[i * 3 + 2], &t, 4);
                }
        };
        const auto t0 = std::chrono::steady_clock::now();
        std::vector<std::thread> pool;
        for (int k = 1; k < threads; ++k) pool.emplace_back(work);
        work();
        for (std::thread &th : pool) th.join();
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        printf("raymarch_seconds %.6f rays %zu threads %d\n", sec, n, threads);
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
    if (cmd == "frame" && (argc == 8 || argc == 9)) {
        VoxelOctree tree(argv[2]);
        std::vector<float> cams = readAll<float>(argv[3]);
        const int W = atoi(argv[4]), H = atoi(argv[5]), stride = atoi(argv[6]);
        const int threads = 16, tile = 8;
        const int workers = argc == 9 ? std::max(1, atoi(argv[8])) : 1;  // the viewer runs one thread per stripe
        double seconds = 0.0;
        const float aspect = H / (float)W;
        std::vector<uint32_t> out;
        for (size_t c = 0; c + 3 <= cams.size(); c += 3) {
            Mat4 model = Mat4::rotXYZ(Vec3(cams[c], 0.0f, 0.0f)) * Mat4::rotXYZ(Vec3(0.0f, cams[c + 1], 0.0f));
            Mat4 view = Mat4::translate(Vec3(0.0f, 0.0f, -cams[c + 2]));
            Mat4 tform = model.pseudoInvert() * view;
            Vec3 pos = tform * Vec3() + tree.center() + Vec3(1.0);
            tform.a14 = tform.a24 = tform.a34 = 0.0f;
            const float scale = 2.0f / W, tileScale = tile * scale;
            const float planeDist = 1.0f / std::tan(float(M_PI) / 6.0f);
            const float zx = planeDist * tform.a13, zy = planeDist * tform.a23, zz = planeDist * tform.a33;
            const float coarseScale = 2.0f * tile / (planeDist * H);
            const Vec3 light = (tform * Vec3(-1.0, 1.0, -1.0)).normalize();
            std::vector<uint32_t> img(size_t(W) * H, 0u);
            std::vector<float> depths;
            auto direction = [&](float dx, float dy) {
                Vec3 dir(dx * tform.a11 + dy * tform.a12 + zx, dx * tform.a21 + dy * tform.a22 + zy,
                         dx * tform.a31 + dy * tform.a32 + zz);
                return dir * invSqrt(dir.x * dir.x + dir.y * dir.y + dir.z * dir.z);
            };
            const int sh = (H - 1) / threads + 1;
            std::vector<std::vector<float>> stripeDepths(threads);
            auto stripe = [&](int i) {
                const int y0 = i * sh, y1 = std::min((i + 1) * sh, H);
                const int tilesX = (W - 1) / tile + 2, tilesY = (y1 - y0 - 1) / tile + 2;
                std::vector<float> &db = stripeDepths[size_t(i)];
                db.assign(size_t(tilesX) * tilesY, 0.0f);
                float dy = aspect - y0 * scale;
                for (int y = 0, idx = 0; y < tilesY; y++, dy -= tileScale) {
                    float dx = -1.0f + 0 * scale;
                    for (int x = 0; x < tilesX; x++, dx += tileScale, idx++) {
                        uint32_t nrm;
                        float t;
                        db[idx] = tree.raymarch(pos, direction(dx, dy), coarseScale, nrm, t) ? t : 1e10f;
                        if (x == 0 || y == 0) continue;
                        const float minT = std::min(std::min(db[idx], db[idx - 1]),
                                                    std::min(db[idx - tilesX], db[idx - tilesX - 1]));
                        if (minT == 1e10f) continue;
                        const int tx0 = (x - 1) * tile, ty0 = (y - 1) * tile + y0;
                        const int tx1 = std::min(tx0 + tile, W), ty1 = std::min(ty0 + tile, y1);
                        const float m = std::max(minT - 0.03f, 0.0f);
                        float py = aspect - ty0 * scale;
                        for (int yy = ty0; yy < ty1; ++yy, py -= scale) {
                            float px = -1.0f + tx0 * scale;
                            for (int xx = tx0; xx < tx1; ++xx, px += scale) {
                                const int cx = xx - ((xx - tx0) % stride), cy = yy - ((yy - ty0) % stride);
                                if (cx != xx || cy != yy) {
                                    img[size_t(yy) * W + xx] = img[size_t(cy) * W + cx];
                                    continue;
                                }
                                const Vec3 dir = direction(px, py);
                                Vec3 col;
                                if (tree.raymarch(pos + dir * m, dir, 0.0f, nrm, t)) {
                                    Vec3 n;
                                    float sc;
                                    decompressMaterial(nrm, n, sc);
                                    const float d = std::max(light.dot(dir.reflect(n)), 0.0f);
                                    col = sc * 0.9f * std::abs(light.dot(n)) + d * d * 0.2f;
                                }
                                const uint32_t g = uint32_t(std::min(col.x, 1.0f) * 255.0);
                                img[size_t(yy) * W + xx] = g | (g << 8) | (g << 16) | 0xFF000000u;
                            }
                        }
                    }
                }
            };
            std::atomic<int> nextStripe(0);
            auto worker = [&]() {
                for (int i; (i = nextStripe.fetch_add(1)) < threads;) stripe(i);
            };
            const auto t0 = std::chrono::steady_clock::now();
            std::vector<std::thread> pool;
            for (int k = 1; k < workers; ++k) pool.emplace_back(worker);
            worker();
            for (std::thread &th : pool) th.join();
            seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            for (const std::vector<float> &db : stripeDepths) depths.insert(depths.end(), db.begin(), db.end());
            out.insert(out.end(), img.begin(), img.end());
            for (float v : depths) {
                uint32_t u;
                memcpy(&u, &v, 4);
                out.push_back(u);
            }
        }
        printf("frame_seconds %.6f frames %zu workers %d\n", seconds, cams.size() / 3, workers);
        writeAll(argv[7], out);
        return 0;
    }
    fprintf(stderr, "bad command\n");
    return 1;
}
