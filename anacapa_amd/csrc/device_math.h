// device_math.h — small fp32 vector helpers and the reference's math conventions
// (include/anacapa/core/Types.h) for the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ahip {

struct V3 {
    float x, y, z;
};
struct V2 {
    float x, y;
};

__device__ __forceinline__ V3 mk(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 ld3(const float* p) { return V3{p[0], p[1], p[2]}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ V3 operator/(V3 a, V3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
__device__ __forceinline__ V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float len(V3 v) { return sqrtf(v.x * v.x + v.y * v.y + v.z * v.z); }
__device__ __forceinline__ V3 normalize(V3 v) { return v * (1.f / len(v)); }
__device__ __forceinline__ V3 safeNormalize(V3 v) {
    float l = len(v);
    return l > 1e-8f ? v * (1.f / l) : V3{0.f, 0.f, 1.f};
}
__device__ __forceinline__ bool isBlack(V3 s) { return s.x == 0.f && s.y == 0.f && s.z == 0.f; }
__device__ __forceinline__ float maxc(V3 v) {
    float m = v.x;
    if (m < v.y) m = v.y;
    if (m < v.z) m = v.z;
    return m;
}
__device__ __forceinline__ bool finite3(V3 v) { return isfinite(v.x) && isfinite(v.y) && isfinite(v.z); }
__device__ __forceinline__ float luminance(V3 s) { return 0.2722287f * s.x + 0.6740818f * s.y + 0.0536895f * s.z; }
__device__ __forceinline__ float stdmax(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float stdmin(float a, float b) { return (b < a) ? b : a; }

// Frisvad / Duff orthonormal basis (Types.h:108-114)
__device__ __forceinline__ void onb(V3 n, V3& t, V3& bt) {
    float sign = copysignf(1.f, n.z);
    float a = -1.f / (sign + n.z);
    float b = n.x * n.y * a;
    t = V3{1.f + sign * n.x * n.x * a, sign * b, -sign * n.x};
    bt = V3{b, sign + n.y * n.y * a, -n.y};
}

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.31830988618379067154f;
constexpr float k2Pi = 6.28318530717958647692f;

}  // namespace ahip
