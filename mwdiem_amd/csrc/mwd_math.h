// Small vector and quaternion helpers for the gfx950 kernels, templated on the
// scalar type: double for the forward path, Dual<N> (a value and N tangents)
// for the tangent path.
//
// Every helper evaluates in a fixed operator order (left to right, as the
// reference's CUDA helpers do, gpu_Verlet.cu:1-83 and
// gpu_fine_contact_detection.cu:1-34 / :605-643) and the library is built with
// -ffp-contract=off, so results round exactly as written and can be compared
// bit for bit with a CPU evaluation of the same expressions. Dual<N> computes
// its value with exactly the double operations, so a tangent run reproduces
// the forward run's values bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MWD_HD __host__ __device__ __forceinline__

// ------------------------------------------------------------------ Dual
template <int N>
struct Dual {
    double v;
    double d[N];
    MWD_HD Dual() {}
    MWD_HD Dual(double x) : v(x) {
#pragma unroll
        for (int k = 0; k < N; ++k) d[k] = 0.0;
    }
};
template <int N> MWD_HD Dual<N> operator+(Dual<N> a, Dual<N> b) {
    Dual<N> r; r.v = a.v + b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
template <int N> MWD_HD Dual<N> operator-(Dual<N> a, Dual<N> b) {
    Dual<N> r; r.v = a.v - b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
template <int N> MWD_HD Dual<N> operator-(Dual<N> a) {
    Dual<N> r; r.v = -a.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
    return r;
}
template <int N> MWD_HD Dual<N> operator*(Dual<N> a, Dual<N> b) {
    Dual<N> r; r.v = a.v * b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
template <int N> MWD_HD Dual<N> operator/(Dual<N> a, Dual<N> b) {
    Dual<N> r; r.v = a.v / b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
    return r;
}
template <int N> MWD_HD Dual<N> operator+(Dual<N> a, double b) { return a + Dual<N>(b); }
template <int N> MWD_HD Dual<N> operator+(double a, Dual<N> b) { return Dual<N>(a) + b; }
template <int N> MWD_HD Dual<N> operator-(Dual<N> a, double b) { return a - Dual<N>(b); }
template <int N> MWD_HD Dual<N> operator-(double a, Dual<N> b) { return Dual<N>(a) - b; }
template <int N> MWD_HD Dual<N> operator*(Dual<N> a, double b) {
    Dual<N> r; r.v = a.v * b;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b;
    return r;
}
template <int N> MWD_HD Dual<N> operator*(double a, Dual<N> b) {
    Dual<N> r; r.v = a * b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a * b.d[k];
    return r;
}
template <int N> MWD_HD Dual<N> operator/(Dual<N> a, double b) {
    Dual<N> r; r.v = a.v / b;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] / b;
    return r;
}
template <int N> MWD_HD Dual<N> operator/(double a, Dual<N> b) { return Dual<N>(a) / b; }
template <int N> MWD_HD bool operator<(Dual<N> a, Dual<N> b) { return a.v < b.v; }
template <int N> MWD_HD bool operator>(Dual<N> a, Dual<N> b) { return a.v > b.v; }
template <int N> MWD_HD bool operator<=(Dual<N> a, Dual<N> b) { return a.v <= b.v; }
template <int N> MWD_HD bool operator>=(Dual<N> a, Dual<N> b) { return a.v >= b.v; }
template <int N> MWD_HD bool operator==(Dual<N> a, Dual<N> b) { return a.v == b.v; }
template <int N> MWD_HD bool operator!=(Dual<N> a, Dual<N> b) { return a.v != b.v; }
template <int N> MWD_HD bool operator<(Dual<N> a, double b) { return a.v < b; }
template <int N> MWD_HD bool operator>(Dual<N> a, double b) { return a.v > b; }
template <int N> MWD_HD bool operator<=(Dual<N> a, double b) { return a.v <= b; }
template <int N> MWD_HD bool operator>=(Dual<N> a, double b) { return a.v >= b; }
template <int N> MWD_HD bool operator==(Dual<N> a, double b) { return a.v == b; }
template <int N> MWD_HD bool operator!=(Dual<N> a, double b) { return a.v != b; }

// At 0 the derivative is taken as 0: every sqrt here is the norm of a vector,
// and a vector that is exactly zero has a zero-norm direction of no preference.
template <int N> MWD_HD Dual<N> dsqrt(Dual<N> a) {
    Dual<N> r; r.v = sqrt(a.v);
    const double s = r.v == 0.0 ? 0.0 : 0.5 / r.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * s;
    return r;
}
template <int N> MWD_HD Dual<N> dsin(Dual<N> a) {
    Dual<N> r; r.v = sin(a.v);
    const double c = cos(a.v);
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * c;
    return r;
}
template <int N> MWD_HD Dual<N> dcos(Dual<N> a) {
    Dual<N> r; r.v = cos(a.v);
    const double s = -sin(a.v);
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * s;
    return r;
}
template <int N> MWD_HD Dual<N> dlog(Dual<N> a) {
    Dual<N> r; r.v = log(a.v);
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] / a.v;
    return r;
}
MWD_HD double dsqrt(double a) { return sqrt(a); }
MWD_HD double dlog(double a) { return log(a); }
MWD_HD double dsin(double a) { return sin(a); }
MWD_HD double dcos(double a) { return cos(a); }
MWD_HD double val(double a) { return a; }
template <int N> MWD_HD double val(Dual<N> a) { return a.v; }

// ----------------------------------------------------------- vectors
template <class T>
struct V3 {
    T x, y, z;
};
template <class T>
struct Q4 {
    T w, x, y, z;
};
typedef V3<double> d3;
typedef Q4<double> q4;

template <class T> MWD_HD V3<T> mk3(T x, T y, T z) {
    V3<T> r;
    r.x = x;
    r.y = y;
    r.z = z;
    return r;
}
template <class T> MWD_HD V3<T> zero3() { return mk3<T>(T(0.0), T(0.0), T(0.0)); }
template <class T> MWD_HD V3<T> operator+(V3<T> a, V3<T> b) { return mk3<T>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class T> MWD_HD V3<T> operator-(V3<T> a, V3<T> b) { return mk3<T>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class T, class S> MWD_HD V3<T> operator*(V3<T> a, S s) { return mk3<T>(a.x * s, a.y * s, a.z * s); }
template <class T, class S> MWD_HD V3<T> operator/(V3<T> a, S s) { return mk3<T>(a.x / s, a.y / s, a.z / s); }
template <class T> MWD_HD V3<T> neg3(V3<T> a) { return mk3<T>(-a.x, -a.y, -a.z); }
template <class T> MWD_HD T dot3(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> MWD_HD T nrm2(V3<T> a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
template <class T> MWD_HD V3<T> cross3(V3<T> a, V3<T> b) {
    return mk3<T>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class T> MWD_HD bool zero3v(V3<T> a) { return a.x == 0.0 && a.y == 0.0 && a.z == 0.0; }
MWD_HD bool zero3(d3 a) { return a.x == 0.0 && a.y == 0.0 && a.z == 0.0; }
MWD_HD d3 ld3(const double* p) { return mk3<double>(p[0], p[1], p[2]); }
MWD_HD void st3(double* p, d3 a) {
    p[0] = a.x;
    p[1] = a.y;
    p[2] = a.z;
}
MWD_HD d3 mk3(double x, double y, double z) { return mk3<double>(x, y, z); }

template <class T> MWD_HD Q4<T> mkq(T w, V3<T> v) {
    Q4<T> q;
    q.w = w;
    q.x = v.x;
    q.y = v.y;
    q.z = v.z;
    return q;
}
template <class T> MWD_HD V3<T> qv(Q4<T> q) { return mk3<T>(q.x, q.y, q.z); }
template <class T> MWD_HD Q4<T> qconj(Q4<T> q) {
    Q4<T> r;
    r.w = q.w;
    r.x = -q.x;
    r.y = -q.y;
    r.z = -q.z;
    return r;
}
// Hamilton product (w1 w2 - v1.v2, v2 w1 + v1 w2 + v1 x v2).
template <class T> MWD_HD Q4<T> qmul(Q4<T> a, Q4<T> b) {
    V3<T> v1 = qv(a), v2 = qv(b);
    V3<T> s = (v2 * a.w + v1 * b.w) + cross3(v1, v2);
    T f = a.w * b.w - dot3(v1, v2);
    return mkq(f, s);
}
// Vector part of q1 (0,v) q2.
template <class T> MWD_HD V3<T> qrot(Q4<T> q1, V3<T> v, Q4<T> q2) {
    return qv(qmul(q1, qmul(mkq(T(0.0), v), q2)));
}
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
// Row-major 3x3 (constant) times vector, ((m0 x + m1 y) + m2 z) per row.
template <class T> MWD_HD V3<T> mat3v(const double* m, V3<T> v) {
    return mk3<T>(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z,
                  m[6] * v.x + m[7] * v.y + m[8] * v.z);
}

// Bounding radius with a rounding margin: every quantity the support-map
// test computes from these coordinates rounds well inside it.
MWD_HD double inflate_radius(double r2max, d3 com) {
    const double cm = fmax(fabs(com.x), fmax(fabs(com.y), fabs(com.z)));
    return sqrt(r2max) * (1.0 + 1e-9) + 1e-9 * (1.0 + cm);
}

// ------------------------------------------------ value/tangent arrays
// A state array and (for the tangent path) its N tangent arrays of the same
// layout. Loads and stores of T = double touch only the values.
template <int N>
struct TArr {
    double* v;
    double* t[N];
};
MWD_HD double ldT(const double* v, double* const* /*t*/, int64_t i, double*) { return v[i]; }
template <int N> MWD_HD Dual<N> ldT(const double* v, double* const* t, int64_t i, Dual<N>*) {
    Dual<N> r;
    r.v = v[i];
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = t[k] ? t[k][i] : 0.0;
    return r;
}
MWD_HD void stT(double* v, double* const* /*t*/, int64_t i, double x) { v[i] = x; }
template <int N> MWD_HD void stT(double* v, double* const* t, int64_t i, Dual<N> x) {
    v[i] = x.v;
#pragma unroll
    for (int k = 0; k < N; ++k)
        if (t[k]) t[k][i] = x.d[k];
}
template <class T, int N> MWD_HD T ld1(const TArr<N>& a, int64_t i) { return ldT(a.v, a.t, i, (T*)nullptr); }
template <class T, int N> MWD_HD V3<T> ldv(const TArr<N>& a, int64_t i) {
    return mk3<T>(ldT(a.v, a.t, 3 * i, (T*)nullptr), ldT(a.v, a.t, 3 * i + 1, (T*)nullptr),
                  ldT(a.v, a.t, 3 * i + 2, (T*)nullptr));
}
template <class T, int N> MWD_HD Q4<T> ldqT(const TArr<N>& a, int64_t i) {
    Q4<T> q;
    q.w = ldT(a.v, a.t, 4 * i, (T*)nullptr);
    q.x = ldT(a.v, a.t, 4 * i + 1, (T*)nullptr);
    q.y = ldT(a.v, a.t, 4 * i + 2, (T*)nullptr);
    q.z = ldT(a.v, a.t, 4 * i + 3, (T*)nullptr);
    return q;
}
template <class T, int N> MWD_HD void stv(const TArr<N>& a, int64_t i, V3<T> x) {
    stT(a.v, a.t, 3 * i, x.x);
    stT(a.v, a.t, 3 * i + 1, x.y);
    stT(a.v, a.t, 3 * i + 2, x.z);
}
template <class T, int N> MWD_HD void stqT(const TArr<N>& a, int64_t i, Q4<T> q) {
    stT(a.v, a.t, 4 * i, q.w);
    stT(a.v, a.t, 4 * i + 1, q.x);
    stT(a.v, a.t, 4 * i + 2, q.y);
    stT(a.v, a.t, 4 * i + 3, q.z);
}
