// Caller of include/VoxelOctree.hpp written the way the reference's Main.cpp
// uses VoxelOctree (load, center, save; build from PlyLoader/VoxelData;
// raymarch). Modes:
//   shim_main cpu <in.oct> <out.oct>            load + center + save, error path
//   shim_main rays <in.oct> <rays.f32> <out.bin> per-ray raymarch of n x 7 floats;
//                                              prints the calls per second
//   shim_main build <in.ply> <res> <threads> <mem> <out.oct>
//   shim_main frame <in.oct> <pitch> <yaw> <radius> <w> <h> <stride> <out.rgba>
#include <chrono>
#include <cstdio>
#include <memory>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "VoxelOctree.hpp"

int main(int argc, char **argv) {
    if (argc < 2) return 1;
    const std::string mode = argv[1];
    if (mode == "cpu" && argc == 4) {
        VoxelOctree tree(argv[2]);
        Vec3 c = tree.center();
        std::printf("center %.9g %.9g %.9g\n", c.x, c.y, c.z);
        tree.save(argv[3]);
        try {
            VoxelOctree missing("/nonexistent/file.oct");
            return 2;
        } catch (const svo_error &e) {
            std::printf("error %d %s\n", e.code(), e.what());
        }
        return 0;
    }
    if (mode == "rays" && argc == 5) {
        VoxelOctree tree(argv[2]);
        FILE *fp = std::fopen(argv[3], "rb");
        std::vector<float> rays;
        float buf[7];
        while (std::fread(buf, 4, 7, fp) == 7) rays.insert(rays.end(), buf, buf + 7);
        std::fclose(fp);
        std::vector<uint32_t> out;
        const size_t n = rays.size() / 7;
        // the reference's per-pixel call pattern, timed over at least 2^20 calls
        size_t calls = 0;
        uint32_t sink = 0;
        const auto t0 = std::chrono::steady_clock::now();
        while (calls < (size_t(1) << 20)) {
            for (size_t i = 0; i < n; ++i, ++calls) {
                const float *r = &rays[i * 7];
                uint32_t normal = 0;
                float t = 0.0f;
                sink += tree.raymarch(Vec3(r[0], r[1], r[2]), Vec3(r[3], r[4], r[5]), r[6], normal, t) ? normal : 0u;
            }
        }
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("calls_per_s %.0f calls %zu sink %u\n", double(calls) / sec, calls, sink);
        for (size_t i = 0; i < n; ++i) {
            const float *r = &rays[i * 7];
            uint32_t normal = 0;
            float t = 0.0f;
            bool hit = tree.raymarch(Vec3(r[0], r[1], r[2]), Vec3(r[3], r[4], r[5]), r[6], normal, t);
            uint32_t tb;
            std::memcpy(&tb, &t, 4);
            out.push_back(hit ? 1u : 0u);
            out.push_back(normal);
            out.push_back(tb);
        }
        fp = std::fopen(argv[4], "wb");
        std::fwrite(out.data(), 4, out.size(), fp);
        std::fclose(fp);
        return 0;
    }
    if (mode == "build" && argc == 7) {
        PlyLoader loader(argv[2]);
        VoxelData data(&loader, std::atoi(argv[3]), size_t(std::strtoull(argv[5], nullptr, 10)));
        data.setThreads(std::atoi(argv[4]));
        VoxelOctree tree(&data);
        tree.save(argv[6]);
        return 0;
    }
    if (mode == "convert" && argc == 8) {
        // the builder's on-disk mode (src/Main.cpp:315-320): convertToVolume, then
        // VoxelData(path, mem) -> VoxelOctree -> save
        std::unique_ptr<PlyLoader> loader(new PlyLoader(argv[2]));
        loader->setThreads(std::atoi(argv[4]));
        const size_t mem = size_t(std::strtoull(argv[5], nullptr, 10));
        loader->convertToVolume(argv[6], std::atoi(argv[3]), mem);
        std::unique_ptr<VoxelData> data(new VoxelData(argv[6], mem));
        std::unique_ptr<VoxelOctree> tree(new VoxelOctree(data.get()));
        tree->save(argv[7]);
        return 0;
    }
    if (mode == "frame" && argc == 10) {
        VoxelOctree tree(argv[2]);
        svo_camera cam = tree.orbitCamera(float(std::atof(argv[3])), float(std::atof(argv[4])), float(std::atof(argv[5])));
        const int w = std::atoi(argv[6]), h = std::atoi(argv[7]);
        std::vector<uint32_t> pixels(size_t(w) * h);
        tree.renderFrame(cam, w, h, pixels.data(), std::atoi(argv[8]));
        FILE *fp = std::fopen(argv[9], "wb");
        std::fwrite(pixels.data(), 4, pixels.size(), fp);
        std::fclose(fp);
        return 0;
    }
    return 1;
}
