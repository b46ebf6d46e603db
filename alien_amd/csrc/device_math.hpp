// device_math.hpp — scalar/vector math of the physics step for gfx950.
//
// Every function is written so that one source operation is one IEEE float operation (the
// library is compiled with -ffp-contract=off and correctly rounded division/sqrt), which makes the
// results reproducible bit for bit on any IEEE single-precision machine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace alien {

constexpr float kPi = 3.1415926535897932384626433832795f;  // Math.cuh:9
constexpr float kRadToDeg = 180.0f / kPi;                   // Math.cuh:11
constexpr float kKernelNorm = 3.0f / (2.0f * kPi);          // 3/(2 pi), CellProcessor.cuh:103
constexpr float kNearZero = 0.00001f;                       // Macros.cuh:63
constexpr float kTwoThirds = 2.0f / 3.0f;
constexpr float kHalfPi = 1.57079637050628662109375f;      // nearest float to pi/2

struct f2 {
    float x, y;
};
__device__ __forceinline__ f2 add(f2 a, f2 b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ f2 sub(f2 a, f2 b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ f2 mul(f2 a, float m) { return {a.x * m, a.y * m}; }
__device__ __forceinline__ f2 dvd(f2 a, float m) { return {a.x / m, a.y / m}; }
__device__ __forceinline__ f2 neg(f2 a) { return {-a.x, -a.y}; }
__device__ __forceinline__ f2 to_f2(float2 v) { return {v.x, v.y}; }
__device__ __forceinline__ float2 to_float2(f2 v) { return make_float2(v.x, v.y); }

__device__ __forceinline__ float dot2(f2 a, f2 b) { return a.x * b.x + a.y * b.y; }
__device__ __forceinline__ float lensq(f2 v) { return v.x * v.x + v.y * v.y; }
__device__ __forceinline__ float len2(f2 v) { return __builtin_sqrtf(v.x * v.x + v.y * v.y); }

// unit vector with the (1,0) fallback below 1e-5 (Math.cuh:246-261), as the product with the
// reciprocal length
__device__ __forceinline__ f2 unit(f2 v) {
    float l = __builtin_sqrtf(v.x * v.x + v.y * v.y);
    if (l > kNearZero) {
        const float inv = 1.0f / l;
        return {v.x * inv, v.y * inv};
    }
    return {1.0f, 0.0f};
}

// floor to int exactly like Base.cuh:83-90
__device__ __forceinline__ int floor_int(float v) {
    int r = static_cast<int>(v);
    if (static_cast<float>(r) > v) --r;
    return r;
}
// x mod size in [0, size); one period off either way (the common case) without a division
__device__ __forceinline__ int wrap_int(int x, int size) {
    if (x >= 0 && x < size) return x;
    if (x < 0 && x >= -size) return x + size;
    if (x >= size && x - size < size) return x - size;
    return ((x % size) + size) % size;
}

// a - b on the torus for wrapped coordinates (BaseMap::correctDirection, Map.cuh:35-39, applied to
// a - b).  The period is subtracted from the operand beyond it before the difference is taken: that
// subtraction is exact, so a difference across the seam is as accurate as one inside the world.
__device__ __forceinline__ float torus_delta1(float a, float b, float size) {
    const float d = a - b;
    const float half = size / 2.0f;
    if (d > half) return (a - size) - b;
    if (d < -half) return a - (b - size);
    return d;
}
__device__ __forceinline__ f2 torus_delta(f2 a, f2 b, float W, float H) {
    return {torus_delta1(a.x, b.x, W), torus_delta1(a.y, b.y, H)};
}
// torus_delta for points known not to straddle the seam (|a - b| <= size / 2 per axis) unless
// `seam` is set: then the plain difference is the same value
__device__ __forceinline__ f2 stencil_delta(f2 a, f2 b, float W, float H, bool seam) {
    if (seam) return torus_delta(a, b, W, H);
    return {a.x - b.x, a.y - b.y};
}

// BaseMap::correctPosition(float2&) (Map.cuh:20-26)
__device__ __forceinline__ f2 torus_pos(f2 p, int W, int H) {
    int ix = floor_int(p.x), iy = floor_int(p.y);
    float fx = p.x - static_cast<float>(ix);
    float fy = p.y - static_cast<float>(iy);
    ix = wrap_int(ix, W);
    iy = wrap_int(iy, H);
    return {static_cast<float>(ix) + fx, static_cast<float>(iy) + fy};
}

// arcsine: argument reduction asin(a) = pi/2 - 2 asin(sqrt((1-a)/2)) above 1/2, odd polynomial
// x + x^3 P(x^2) below (coefficients least-squares fitted for this library; |err| < 1.7e-7 rad)
__device__ __forceinline__ float asin_poly(float v) {
    float a = v < 0.0f ? -v : v;
    bool big = a > 0.5f;
    float z = big ? (1.0f - a) * 0.5f : a * a;
    float s = big ? __builtin_sqrtf(z) : a;
    float p = 0.037105727940797806f;
    p = p * z + 0.014683568850159645f;
    p = p * z + 0.0317583866417408f;
    p = p * z + 0.044520482420921326f;
    p = p * z + 0.07500503212213516f;
    p = p * z + 0.16666659712791443f;
    float r = s + (s * z) * p;
    if (big) r = kHalfPi - (r + r);
    return v < 0.0f ? -r : r;
}

// Math::angleOfVector (Math.cuh:140-155): degrees, 0 = (0,-1)
__device__ __forceinline__ float angle_of(f2 v) {
    if (len2(v) < kNearZero) return 0.0f;
    float nvy = -v.y / len2(v);
    nvy = fmaxf(-1.0f, fminf(1.0f, nvy));
    float s = asin_poly(nvy) * kRadToDeg;
    if (v.x >= 0.0f) return 90.0f - s;
    return s + 270.0f;
}
// Math::subtractAngle (Math.cuh:300-310)
__device__ __forceinline__ float sub_angle(float a, float b) {
    float d = a - b;
    if (d > 360.0f) d -= 360.0f;
    if (d < 0.0f) d += 360.0f;
    return d;
}
// Math::isAngleInBetween (Math.cuh:223-236)
__device__ __forceinline__ bool angle_between(float a1, float a2, float c) {
    if (a1 == a2 && a1 != c) return false;
    if (c < a1) {
        c += 360.0f;
        a2 += 360.0f;
    }
    if (a2 < c) a2 += 360.0f;
    return a2 - a1 < 360.0f;
}
// Math::crossing (Math.cuh:363-395)
__device__ __forceinline__ bool segments_cross(f2 s, f2 e, f2 os, f2 oe) {
    if ((s.x == os.x && s.y == os.y) || (s.x == oe.x && s.y == oe.y) || (e.x == os.x && e.y == os.y) ||
        (e.x == oe.x && e.y == oe.y))
        return false;
    f2 v1 = sub(e, s), v2 = sub(oe, os);
    float divisor = v2.x * v1.y - v2.y * v1.x;
    if (fabsf(divisor) < kNearZero) return false;
    float mue = (v1.x * (os.y - s.y) - v1.y * (os.x - s.x)) / divisor;
    if (mue < -kNearZero || mue > 1.0f + kNearZero) return false;
    float lambda;
    if (fabsf(v1.x) > kNearZero)
        lambda = (os.x - s.x + mue * v2.x) / v1.x;
    else if (fabsf(v1.y) > kNearZero)
        lambda = (os.y - s.y + mue * v2.y) / v1.y;
    else
        return false;
    return lambda >= kNearZero && lambda <= 1.0f - kNearZero;
}
__device__ __forceinline__ void rot_cw(f2& v) {  // Math.cuh:157-162
    float t = v.x;
    v.x = -v.y;
    v.y = t;
}
__device__ __forceinline__ void rot_ccw(f2& v) {  // Math.cuh:164-169
    float t = v.x;
    v.x = v.y;
    v.y = -t;
}

// Counter-based random draws (the reference's table with a shared atomic cursor,
// CudaNumberGenerator.cuh:12-104, depends on thread timing).  Philox4x32-10 of Salmon, Moraes,
// Dror and Shaw ("Parallel random numbers: as easy as 1, 2, 3", SC 2011) with counter
// {id lo, id hi, timestep lo, purpose} and key {seed lo, seed hi}; the first output word's upper
// 24 bits give a float in [0, 1).  A draw is a pure function of its arguments.
constexpr uint32_t kStreamForceDecay = 1;  // checkForces, CellProcessor.cuh:367-386
// CellProcessor::radiation (CellProcessor.cuh:725-773): whether the cell radiates this step, the
// fraction of the energy loss drawn, the direction of the velocity perturbation
constexpr uint32_t kStreamRadiation = 2, kStreamRadiationEnergy = 3, kStreamRadiationAngle = 4;
// RadiationProcessor::splitting (RadiationProcessor.cuh:172-204): whether the particle splits, the
// direction of the split
constexpr uint32_t kStreamSplit = 5, kStreamSplitAngle = 6;
__host__ __device__ inline float uniform_draw(unsigned long long seed, unsigned long long timestep,
                                              unsigned long long id, uint32_t purpose) {
    uint32_t c0 = uint32_t(id), c1 = uint32_t(id >> 32), c2 = uint32_t(timestep), c3 = purpose;
    uint32_t k0 = uint32_t(seed), k1 = uint32_t(seed >> 32);
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const unsigned long long p0 = (unsigned long long)0xD2511F53u * c0;
        const unsigned long long p1 = (unsigned long long)0xCD9E8D57u * c2;
        const uint32_t n0 = uint32_t(p1 >> 32) ^ c1 ^ k0, n1 = uint32_t(p1);
        const uint32_t n2 = uint32_t(p0 >> 32) ^ c3 ^ k1, n3 = uint32_t(p0);
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return float(c0 >> 8) * (1.0f / 16777216.0f);
}

// Math::unitVectorOfAngle (Math.cuh:134-138), {sin a, -cos a} for a in degrees, [0, 360) here.
// Written with single-precision operations in a fixed order instead of sinf/cosf, whose last bit
// is not the same on every platform: the angle is reduced to a quarter turn r in [0, 90) degrees,
// x = r pi / 180, and sin x and cos x are their Taylor polynomials to x^13 and x^14 in Horner form
// (truncation below 7e-10 on [0, pi/2]); the quarter turns then swap and negate them.
__host__ __device__ inline float2 unit_vector_of_angle(float deg) {
    int q = int(deg / 90.0f);
    q = q < 0 ? 0 : (q > 3 ? 3 : q);
    const float x = (deg - 90.0f * float(q)) * 0.017453292f;
    const float x2 = x * x;
    const float sn =
        x * (1.0f + x2 * (-0.16666667f +
                          x2 * (0.008333334f +
                                x2 * (-0.0001984127f + x2 * (2.7557319e-06f + x2 * (-2.5052108e-08f + x2 * 1.6059044e-10f))))));
    const float cs =
        1.0f + x2 * (-0.5f + x2 * (0.041666668f +
                                   x2 * (-0.0013888889f +
                                         x2 * (2.4801588e-05f + x2 * (-2.7557319e-07f + x2 * (2.0876757e-09f + x2 * -1.1470746e-11f))))));
    float s = sn, co = cs;  // sin and cos of the angle
    if (q == 1) {
        s = cs;
        co = -sn;
    } else if (q == 2) {
        s = -sn;
        co = -cs;
    } else if (q == 3) {
        s = -cs;
        co = sn;
    }
    return make_float2(s, -co);
}

// SPH cubic spline and derivative (CellProcessor.cuh:92-119)
__device__ __forceinline__ float sph_w(float q) {
    float r;
    if (q < 1.0f)
        r = kTwoThirds - q * q + 0.5f * q * q * q;
    else if (q < 2.0f) {
        r = 2.0f - q;
        r = r * r * r / 6.0f;
    } else
        r = 0.0f;
    r *= kKernelNorm;
    return r;
}
__device__ __forceinline__ float sph_dw(float q) {
    float r;
    if (q < 1.0f)
        r = -2.0f * q + 1.5f * q * q;
    else if (q < 2.0f)
        r = -0.5f * (2.0f - q) * (2.0f - q);
    else
        r = 0.0f;
    r *= kKernelNorm;
    return r;
}

}  // namespace alien
