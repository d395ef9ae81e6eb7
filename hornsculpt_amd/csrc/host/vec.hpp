// Small FP64 vector algebra shared by host planning and device kernels.
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#define HS_HD __host__ __device__
#else
#define HS_HD
#endif

namespace hsbem {

struct Vec3 {
    double x{}, y{}, z{};
};

HS_HD inline Vec3 vsub(Vec3 a, Vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
HS_HD inline Vec3 vadd(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
HS_HD inline Vec3 vscale(Vec3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
HS_HD inline double vdot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
HS_HD inline Vec3 vcross(Vec3 a, Vec3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
HS_HD inline double vnorm2(Vec3 a) { return vdot(a, a); }

} // namespace hsbem
