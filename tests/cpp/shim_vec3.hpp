// Minimal float3 for compiling the VoxelOctree.hpp front end in tests (the
// reference's own math/Vec3.hpp plays this role in its tree).
#pragma once
struct Vec3 {
    float x, y, z;
    Vec3() : x(0), y(0), z(0) {}
    Vec3(float a, float b, float c) : x(a), y(b), z(c) {}
};
