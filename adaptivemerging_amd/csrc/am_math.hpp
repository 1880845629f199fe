// am_math.hpp - double-precision helpers for the gfx950 kernels.
//
// Every helper evaluates in the operation order of the javax.vecmath call it
// stands for, so that with contraction disabled (-ffp-contract=off, and the
// pragma below) a kernel rounds exactly like the reference's Java code.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

#define AMD_INLINE __device__ __forceinline__

// Java Math.max / Math.min: NaN propagates, -0.0 orders below +0.0.
// Written as selects so divergent lanes do not branch.
AMD_INLINE double jmax(double a, double b) {
    double r = (a >= b) ? a : b;
    r = ((a == 0.0) & (b == 0.0) & (__signbit(a) != 0)) ? b : r;
    return (a != a) ? a : r;
}
AMD_INLINE double jmin(double a, double b) {
    double r = (a <= b) ? a : b;
    r = ((a == 0.0) & (b == 0.0) & (__signbit(b) != 0)) ? b : r;
    return (a != a) ? a : r;
}

struct d3 {
    double x, y, z;
};

AMD_INLINE d3 mk(double x, double y, double z) { return d3{x, y, z}; }
AMD_INLINE d3 ld3(const double* p) { return d3{p[0], p[1], p[2]}; }
AMD_INLINE void st3(double* p, d3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }
AMD_INLINE double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// Vector3d.cross(a, b)
AMD_INLINE d3 cross(d3 a, d3 b) {
    return d3{a.y * b.z - a.z * b.y, b.x * a.z - b.z * a.x, a.x * b.y - a.y * b.x};
}
AMD_INLINE double length(d3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
// Vector3d.normalize: multiply by 1/sqrt(|a|^2)
AMD_INLINE d3 normalized(d3 a) {
    double n = 1.0 / sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    return d3{a.x * n, a.y * n, a.z * n};
}
AMD_INLINE d3 sub(d3 a, d3 b) { return d3{a.x - b.x, a.y - b.y, a.z - b.z}; }
AMD_INLINE d3 add(d3 a, d3 b) { return d3{a.x + b.x, a.y + b.y, a.z + b.z}; }
AMD_INLINE d3 scale(double s, d3 a) { return d3{s * a.x, s * a.y, s * a.z}; }
// Tuple3d.scaleAdd(s, t1, t2) = s*t1 + t2
AMD_INLINE d3 scale_add(double s, d3 t1, d3 t2) { return d3{s * t1.x + t2.x, s * t1.y + t2.y, s * t1.z + t2.z}; }
AMD_INLINE double distance(d3 a, d3 b) {
    double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    return sqrt(dx * dx + dy * dy + dz * dz);
}
AMD_INLINE double distance_sq(d3 a, d3 b) {
    double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    return dx * dx + dy * dy + dz * dz;
}

// Row-major 3x3 (m00 m01 m02 m10 ...).
struct m33 {
    double m[9];
};
AMD_INLINE m33 ldm(const double* p) {
    m33 r;
#pragma unroll
    for (int i = 0; i < 9; i++) r.m[i] = p[i];
    return r;
}
AMD_INLINE void stm(double* p, const m33& a) {
#pragma unroll
    for (int i = 0; i < 9; i++) p[i] = a.m[i];
}
// Matrix3d.transform(v)
AMD_INLINE d3 mul(const m33& M, d3 v) {
    return d3{M.m[0] * v.x + M.m[1] * v.y + M.m[2] * v.z, M.m[3] * v.x + M.m[4] * v.y + M.m[5] * v.z,
              M.m[6] * v.x + M.m[7] * v.y + M.m[8] * v.z};
}
// R^T v as in RigidTransform3D.inverseTransform
AMD_INLINE d3 mul_t(const m33& M, d3 v) {
    return d3{M.m[0] * v.x + M.m[3] * v.y + M.m[6] * v.z, M.m[1] * v.x + M.m[4] * v.y + M.m[7] * v.z,
              M.m[2] * v.x + M.m[5] * v.y + M.m[8] * v.z};
}
AMD_INLINE d3 xform(const m33& R, d3 t, d3 p) {
    d3 q = mul(R, p);
    return d3{q.x + t.x, q.y + t.y, q.z + t.z};
}
AMD_INLINE d3 inv_xform(const m33& R, d3 t, d3 p) { return mul_t(R, sub(p, t)); }
// Matrix3d.mul(A, B)
AMD_INLINE m33 mmul(const m33& A, const m33& B) {
    m33 r;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            r.m[i * 3 + j] = A.m[i * 3 + 0] * B.m[0 * 3 + j] + A.m[i * 3 + 1] * B.m[1 * 3 + j] + A.m[i * 3 + 2] * B.m[2 * 3 + j];
    return r;
}
// Matrix3d.mulTransposeRight(A, B) = A B^T
AMD_INLINE m33 mmul_bt(const m33& A, const m33& B) {
    m33 r;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            r.m[i * 3 + j] = A.m[i * 3 + 0] * B.m[j * 3 + 0] + A.m[i * 3 + 1] * B.m[j * 3 + 1] + A.m[i * 3 + 2] * B.m[j * 3 + 2];
    return r;
}
// RigidTransform3D.computeRM0RT: R M R^T
AMD_INLINE m33 rm0rt(const m33& R, const m33& M) { return mmul_bt(mmul(R, M), R); }
AMD_INLINE d3 col(const m33& M, int j) { return d3{M.m[j], M.m[3 + j], M.m[6 + j]}; }
AMD_INLINE double comp(d3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
AMD_INLINE void set_comp(d3& v, int i, double s) {
    if (i == 0) v.x = s;
    else if (i == 1) v.y = s;
    else v.z = s;
}
