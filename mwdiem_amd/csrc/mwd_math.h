// Small FP64 vector and quaternion helpers for the gfx950 kernels.
//
// Every helper evaluates in a fixed operator order (left to right, as the
// reference's CUDA helpers do, gpu_Verlet.cu:1-83 and
// gpu_fine_contact_detection.cu:1-34 / :605-643) and the library is built with
// -ffp-contract=off, so results round exactly as written and can be compared
// bit for bit with a CPU evaluation of the same expressions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MWD_HD __host__ __device__ __forceinline__

struct d3 {
    double x, y, z;
};
struct q4 {
    double w, x, y, z;
};

MWD_HD d3 mk3(double x, double y, double z) {
    d3 r;
    r.x = x;
    r.y = y;
    r.z = z;
    return r;
}
MWD_HD d3 operator+(d3 a, d3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
MWD_HD d3 operator-(d3 a, d3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
MWD_HD d3 operator*(d3 a, double s) { return mk3(a.x * s, a.y * s, a.z * s); }
MWD_HD d3 operator/(d3 a, double s) { return mk3(a.x / s, a.y / s, a.z / s); }
MWD_HD d3 neg3(d3 a) { return mk3(-a.x, -a.y, -a.z); }
MWD_HD double dot3(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
MWD_HD double nrm2(d3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
MWD_HD d3 cross3(d3 a, d3 b) { return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
MWD_HD bool zero3(d3 a) { return a.x == 0.0 && a.y == 0.0 && a.z == 0.0; }
MWD_HD d3 ld3(const double* p) { return mk3(p[0], p[1], p[2]); }
MWD_HD void st3(double* p, d3 a) {
    p[0] = a.x;
    p[1] = a.y;
    p[2] = a.z;
}

MWD_HD q4 mkq(double w, d3 v) {
    q4 q;
    q.w = w;
    q.x = v.x;
    q.y = v.y;
    q.z = v.z;
    return q;
}
MWD_HD d3 qv(q4 q) { return mk3(q.x, q.y, q.z); }
MWD_HD q4 qconj(q4 q) {
    q4 r;
    r.w = q.w;
    r.x = -q.x;
    r.y = -q.y;
    r.z = -q.z;
    return r;
}
// Hamilton product (w1 w2 - v1.v2, v2 w1 + v1 w2 + v1 x v2).
MWD_HD q4 qmul(q4 a, q4 b) {
    d3 v1 = qv(a), v2 = qv(b);
    d3 s = (v2 * a.w + v1 * b.w) + cross3(v1, v2);
    double f = a.w * b.w - dot3(v1, v2);
    return mkq(f, s);
}
// Vector part of q1 (0,v) q2.
MWD_HD d3 qrot(q4 q1, d3 v, q4 q2) { return qv(qmul(q1, qmul(mkq(0.0, v), q2))); }
MWD_HD q4 ldq(const double* p) {
    q4 q;
    q.w = p[0];
    q.x = p[1];
    q.y = p[2];
    q.z = p[3];
    return q;
}
MWD_HD void stq(double* p, q4 q) {
    p[0] = q.w;
    p[1] = q.x;
    p[2] = q.y;
    p[3] = q.z;
}
// Row-major 3x3 times vector, ((m0 x + m1 y) + m2 z) per row.
MWD_HD d3 mat3v(const double* m, d3 v) {
    return mk3(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z,
               m[6] * v.x + m[7] * v.y + m[8] * v.z);
}

// Bounding radius with a rounding margin: every quantity the support-map
// test computes from these coordinates rounds well inside it.
MWD_HD double inflate_radius(double r2max, d3 com) {
    const double cm = fmax(fabs(com.x), fmax(fabs(com.y), fabs(com.z)));
    return sqrt(r2max) * (1.0 + 1e-9) + 1e-9 * (1.0 + cm);
}
