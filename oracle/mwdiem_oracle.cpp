// TEST INFRASTRUCTURE ONLY -- the oracle.
//
// A CPU, FP64 restatement of the reference's per-timestep DEM path for convex
// polyhedra. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg may load this library, and only as the checker or the CPU baseline; the
// product (mwdiem_amd/) never calls it.
//
// What it restates (file:line of the reference, kgillemot/MWDiEM):
//   cell keys ........... MWDiem.py:813-816
//   wall slab test ...... MWDiem.py:807-811
//   Linked_cell_hash .... LIBRARIES/LIBRARY_Coarse_contact_detection.py:84-155
//                         (CPython dict insertion order and CPython set
//                          iteration order for int members are emulated)
//   Support_funtion ..... LIBRARIES/gpu_fine_contact_detection.cu:55-101
//   Check_linesegment ... gpu_fine_contact_detection.cu:103-124
//   Check_triangle ...... gpu_fine_contact_detection.cu:127-186
//   Check_Tetrahedron ... gpu_fine_contact_detection.cu:188-254
//   CorrectFlat ......... gpu_fine_contact_detection.cu:256-268
//   AddTriangle ......... gpu_fine_contact_detection.cu:270-325
//   EPA ................. gpu_fine_contact_detection.cu:328-603
//   Force ............... gpu_fine_contact_detection.cu:647-871
//   contact_detection ... gpu_fine_contact_detection.cu:874-1322
//   Verlet .............. LIBRARIES/gpu_Verlet.cu:85-404
//   force assembly ...... MWDiem.py:676-677,977-981
//
// Deliberate, documented differences from the reference text:
//   * FP64 everywhere (the reference contact path is FP32, MWDiem.py:43).
//   * EPA face storage is EPA_MAX_FACES (the reference's T/n/dist/valid[50]
//     overflow is undefined behaviour, .cu:359-373); on overflow the expansion
//     stops and status bit 2 is raised.
//   * The per-particle force sum runs in ascending pair order (the reference
//     sums with FP32 atomics in an unspecified order, .cu:1304-1319).
//   * displacement_tang is indexed [4P+k] (the reference's [P+k], .cu:974-976,
//     races between neighbouring pairs; the driver always passes zeros).
//
// Compiled with -ffp-contract=off so every expression rounds as written.

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <unordered_map>

#define ORC_EXPORT extern "C" __attribute__((visibility("default")))

namespace {

const int EPA_MAX_FACES = 128;
const int EPA_MAX_VERTS = 32;

// ---------------------------------------------------------------- flop counter
// CD is a double that counts every floating-point add/sub/mul/div/sqrt it
// executes (FMA is never formed: the library is built with -ffp-contract=off).
static thread_local uint64_t g_flops = 0;  // per thread; orc_count_flops sums them
struct CD {
    double v;
    CD() : v(0) {}
    CD(double x) : v(x) {}
};
inline CD operator+(CD a, CD b) { ++g_flops; return CD(a.v + b.v); }
inline CD operator-(CD a, CD b) { ++g_flops; return CD(a.v - b.v); }
inline CD operator*(CD a, CD b) { ++g_flops; return CD(a.v * b.v); }
inline CD operator/(CD a, CD b) { ++g_flops; return CD(a.v / b.v); }
inline CD operator-(CD a) { return CD(-a.v); }
inline bool operator<(CD a, CD b) { return a.v < b.v; }
inline bool operator>(CD a, CD b) { return a.v > b.v; }
inline bool operator<=(CD a, CD b) { return a.v <= b.v; }
inline bool operator>=(CD a, CD b) { return a.v >= b.v; }
inline bool operator==(CD a, CD b) { return a.v == b.v; }
inline bool operator!=(CD a, CD b) { return a.v != b.v; }
inline CD sqrt_(CD a) { ++g_flops; return CD(std::sqrt(a.v)); }
inline double sqrt_(double a) { return std::sqrt(a); }
inline CD fabs_(CD a) { return CD(std::fabs(a.v)); }
inline double fabs_(double a) { return std::fabs(a); }
inline double val(double a) { return a; }
inline double val(CD a) { return a.v; }

template <typename T> struct V3 { T x, y, z; };
template <typename T> inline V3<T> mk(T x, T y, T z) { V3<T> r; r.x = x; r.y = y; r.z = z; return r; }
template <typename T> inline V3<T> operator+(V3<T> a, V3<T> b) { return mk<T>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <typename T> inline V3<T> operator-(V3<T> a, V3<T> b) { return mk<T>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <typename T> inline V3<T> operator*(V3<T> a, T b) { return mk<T>(a.x * b, a.y * b, a.z * b); }
template <typename T> inline V3<T> operator/(V3<T> a, T b) { return mk<T>(a.x / b, a.y / b, a.z / b); }
template <typename T> inline V3<T> neg(V3<T> a) { return mk<T>(-a.x, -a.y, -a.z); }
template <typename T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T> inline V3<T> cross(V3<T> a, V3<T> b) {
    return mk<T>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <typename T> inline T norm2(V3<T> a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
template <typename T> inline V3<T> load3(const double* p) { return mk<T>(T(p[0]), T(p[1]), T(p[2])); }
template <typename T> inline bool is_zero(V3<T> a) { return a.x == T(0.0) && a.y == T(0.0) && a.z == T(0.0); }

// ------------------------------------------------------------------ contact law
// Pair-independent coefficients of Force (.cu:685-701), evaluated in the
// reference's operator order.
// mode 0: the reference law (Q3, Q4, Q5 as written). mode 1: the physical
// variant -- contact point (cpA+cpB)/2, damping from the effective masses
// (normal 1/m_n = 1/mA + 1/mB, tangential 1/m_t = 1/m_n + rA^2/IA + rB^2/IB,
// walls infinitely heavy), tangential force on the linear momentum too.
struct Law {
    double Kn, mu, ln_en, ln_et, A_en, tc, Gn, Kt_num, inv_mu, dt;
    double cn, ct, ratio;
    int mode;
};

Law make_law(double Kn, double en, double et, double mu, double dt, int mode = 0) {
    Law L;
    L.mode = mode;
    const double pi = 3.141592653589793;
    const double pi2 = pi * pi;
    L.Kn = Kn;
    L.mu = mu;
    L.ln_en = std::log(en);
    L.ln_et = std::log(et);
    L.A_en = pi2 + L.ln_en * L.ln_en;
    L.tc = std::sqrt((mu * L.A_en) / Kn);
    L.Gn = ((-2.0 * mu) / L.tc) * L.ln_en;
    L.Kt_num = ((Kn * (pi2 + L.ln_et * L.ln_et)) / L.A_en) * 1.0;
    L.inv_mu = 1.0 / mu;
    L.dt = dt;
    L.cn = (-2.0 * L.ln_en) / std::sqrt(L.A_en);
    L.ct = (-2.0 * L.ln_et) / std::sqrt(pi2 + L.ln_et * L.ln_et);
    L.ratio = (pi2 + L.ln_et * L.ln_et) / L.A_en;
    return L;
}

// ------------------------------------------------------------------ quaternions
template <typename T> struct Q4 { T w, x, y, z; };
template <typename T> inline Q4<T> qmk(T w, V3<T> v) { Q4<T> q; q.w = w; q.x = v.x; q.y = v.y; q.z = v.z; return q; }
template <typename T> inline V3<T> qvec(Q4<T> q) { return mk<T>(q.x, q.y, q.z); }
template <typename T> inline Q4<T> qinv(Q4<T> q) { Q4<T> r; r.w = q.w; r.x = -q.x; r.y = -q.y; r.z = -q.z; return r; }
// Quat_mult, gpu_Verlet.cu:43-54 / .cu:613-624
template <typename T> inline Q4<T> qmul(Q4<T> a, Q4<T> b) {
    V3<T> v1 = qvec(a), v2 = qvec(b);
    V3<T> s = (v2 * a.w + v1 * b.w) + cross(v1, v2);
    T f = a.w * b.w - dot(v1, v2);
    return qmk(f, s);
}
// Quat_tripl, gpu_Verlet.cu:57-74: vector part of q1 (0,v) q2
template <typename T> inline V3<T> qtripl(Q4<T> q1, V3<T> v, Q4<T> q2) {
    Q4<T> vq = qmk(T(0.0), v);
    return qvec(qmul(q1, qmul(vq, q2)));
}
template <typename T> inline Q4<T> qload(const double* p) { Q4<T> q; q.w = T(p[0]); q.x = T(p[1]); q.y = T(p[2]); q.z = T(p[3]); return q; }

// ------------------------------------------------------------------ support map
template <typename T> struct Body {
    const double* v;  // [n][3]
    int n;
};

// Support_funtion, .cu:55-101: first index of the strict maximum.
template <typename T>
inline V3<T> support(const Body<T>& A, const Body<T>& B, V3<T> d, int* ia, int* ib) {
    T best = T(-1e30);
    int i1 = 0;
    for (int i = 0; i < A.n; ++i) {
        T t = dot(d, load3<T>(A.v + 3 * i));
        if (t > best) { best = t; i1 = i; }
    }
    V3<T> nd = neg(d);
    T best2 = T(-1e30);
    int i2 = 0;
    for (int i = 0; i < B.n; ++i) {
        T t = dot(nd, load3<T>(B.v + 3 * i));
        if (t > best2) { best2 = t; i2 = i; }
    }
    *ia = i1;
    *ib = i2;
    return load3<T>(A.v + 3 * i1) - load3<T>(B.v + 3 * i2);
}

// Check_linesegment, .cu:103-124
template <typename T>
inline void check_segment(V3<T> P1, V3<T> P2, int* coll, V3<T>* d) {
    *coll = -1;
    *d = mk<T>(T(0.0), T(0.0), T(0.0));
    V3<T> V21 = P2 - P1;
    if (dot(V21, P2) < T(0.0)) { *coll = 0; return; }
    V3<T> t = cross(V21, P2);
    if (is_zero(t)) { *coll = 1; return; }
    V3<T> Vt = cross(V21, neg(P2));
    *d = cross(Vt, V21);
}

// Check_triangle, .cu:127-186
template <typename T>
inline void check_triangle(V3<T> P1, V3<T> P2, V3<T> P3, int* coll, V3<T>* d) {
    *coll = -1;
    *d = mk<T>(T(0.0), T(0.0), T(0.0));
    int c2, c3;
    check_segment(P2, P3, &c2, d);
    if (c2 == 1) { *coll = 1; return; }
    check_segment(P1, P3, &c3, d);
    if (c3 == 1) { *coll = 1; return; }
    if (c2 + c3 > -1) { *coll = 0; return; }
    V3<T> V32 = P3 - P2;
    V3<T> V13 = P1 - P3;
    V3<T> n = cross(V32, V13);
    V3<T> n2 = cross(V32, n);
    V3<T> n3 = cross(V13, n);
    if (dot(n2, P2) < T(0.0)) { *d = n2; return; }
    if (dot(n3, P3) < T(0.0)) { *d = n3; return; }
    if (dot(P3, n) == T(0.0)) { *coll = 1; return; }
    if (dot(n, P3) > T(0.0)) *d = neg(n); else *d = n;
}

// Check_Tetrahedron, .cu:188-254
template <typename T>
inline void check_tetra(V3<T> P1, V3<T> P2, V3<T> P3, V3<T> P4, int* coll, V3<T>* d, int* todel) {
    *coll = -1;
    *d = mk<T>(T(0.0), T(0.0), T(0.0));
    *todel = 0;
    int c1, c2, c3;
    check_triangle(P1, P2, P4, &c1, d);
    if (c1 == 1) { *coll = 1; return; }
    check_triangle(P2, P3, P4, &c2, d);
    if (c2 == 1) { *coll = 1; return; }
    check_triangle(P3, P1, P4, &c3, d);
    if (c3 == 1) { *coll = 1; return; }
    if (c1 + c2 + c3 > -2) { *coll = 0; return; }
    V3<T> V13 = P1 - P3, V21 = P2 - P1, V32 = P3 - P2, V41 = P4 - P1;
    V3<T> n2 = cross(V21, V41);
    V3<T> V42 = P4 - P2;
    V3<T> n3 = cross(V32, V42);
    V3<T> V43 = P4 - P3;
    V3<T> n4 = cross(V13, V43);
    if (dot(n2, P4) < T(0.0)) { *d = n2; *todel = 3; }
    else if (dot(n3, P4) < T(0.0)) { *d = n3; *todel = 1; }
    else if (dot(n4, P4) < T(0.0)) { *d = n4; *todel = 2; }
    else *coll = 1;
}

// CorrectFlat, .cu:256-268
template <typename T>
inline int correct_flat(V3<T> S1, V3<T> S2, V3<T> S3, V3<T> S4) {
    T det1 = S1.x * (S2.y * S3.z - S3.y * S2.z) - S2.x * (S1.y * S3.z - S3.y * S1.z) + S3.x * (S1.y * S2.z - S2.y * S1.z);
    if (det1 < T(0.000001) && det1 > T(-0.000001)) {
        T det2 = S1.x * (S2.y * S4.z - S4.y * S2.z) - S2.x * (S1.y * S4.z - S4.y * S1.z) + S4.x * (S1.y * S2.z - S2.y * S1.z);
        if (det2 < T(0.000001) && det2 > T(-0.000001)) return 1;
    }
    return 0;
}

// Physical law: a tetrahedron is flat when six times its volume is at most
// 1e-9 of the cube of its longest edge (squared: v6^2 <= 1e-18 L2^3). Unlike
// CorrectFlat's test, this measures the tetrahedron and not where the origin
// lies (not in the reference).
template <typename T>
inline bool flat_tetra(V3<T> S0, V3<T> S1, V3<T> S2, V3<T> S3) {
    const V3<T> e1 = S1 - S0, e2 = S2 - S0, e3 = S3 - S0;
    const T v6 = dot(cross(e1, e2), e3);
    T L2 = norm2(e1);
    const T c2 = norm2(e2), c3 = norm2(e3), c4 = norm2(S2 - S1), c5 = norm2(S3 - S1), c6 = norm2(S3 - S2);
    if (c2 > L2) L2 = c2;
    if (c3 > L2) L2 = c3;
    if (c4 > L2) L2 = c4;
    if (c5 > L2) L2 = c5;
    if (c6 > L2) L2 = c6;
    return v6 * v6 <= T(1e-18) * ((L2 * L2) * L2);
}

// ------------------------------------------------------------------------ EPA
// Face distances at or above this are EPA's placeholders (1e30 for a face
// whose plane misses the origin's projection, 3e20 for a deleted face), not
// penetration depths.
constexpr double EPA_PLACEHOLDER = 1e20;
template <typename T> struct Poly {
    V3<T> V[EPA_MAX_VERTS];
    int opA[EPA_MAX_VERTS], opB[EPA_MAX_VERTS];
    int Tr[EPA_MAX_FACES][3];
    V3<T> n[EPA_MAX_FACES];
    T dist[EPA_MAX_FACES];
    int valid[EPA_MAX_FACES];
};

// AddTriangle, .cu:270-325. With `inner` (the physical law's EPA) the face
// is turned to face away from that interior point instead of away from the
// origin, and dd keeps its sign (negative: the origin is outside the face).
template <typename T>
inline void add_triangle(Poly<T>& P, int e0, int e1, int e2, int id, const V3<T>* inner = nullptr) {
    P.Tr[id][0] = e0; P.Tr[id][1] = e1; P.Tr[id][2] = e2;
    V3<T> nn = cross(P.V[e1] - P.V[e0], P.V[e2] - P.V[e1]);
    T len = sqrt_(norm2(nn));
    nn = nn / len;
    T dd = dot(nn, P.V[e0]);
    P.valid[id] = 1;
    const bool turn = inner ? dot(nn, P.V[e0] - *inner) < T(0.0) : dd < T(0.0);
    if (turn) {
        int o = P.Tr[id][2]; P.Tr[id][2] = P.Tr[id][1]; P.Tr[id][1] = o;
        dd = dd * T(-1.0);
        nn = nn * T(-1.0);
    }
    const V3<T>& a = P.V[P.Tr[id][0]];
    const V3<T>& b = P.V[P.Tr[id][1]];
    const V3<T>& c = P.V[P.Tr[id][2]];
    bool inside = dot(cross(b - a, nn), neg(a)) < T(0.0) &&
                  dot(cross(c - b, nn), neg(b)) < T(0.0) &&
                  dot(cross(a - c, nn), neg(c)) < T(0.0);
    if (!inside) { P.valid[id] = -2; dd = T(1e30); }
    P.n[id] = nn;
    P.dist[id] = dd;
}

struct EpaOut { double depth; double n[3]; double cpA[3]; double cpB[3]; int status; int faces; int iters; };

// EPA, .cu:328-603. status bit 1: expansion cap reached (.cu:443-447);
// bit 2: face storage exhausted (reference: undefined behaviour).
// robust (the physical law; not in the reference): faces are turned away from
// the starting tetrahedron's centroid, an interior point of every polytope
// EPA grows from it, so that a face whose plane passes through or beside the
// origin (a contact of near-zero depth) keeps its outward normal and its
// signed distance; and the expansion also stops when the support point gains
// at most 1e-5 |d| + 1e-10 R over the closest face (R: the largest distance
// of a starting vertex from the origin).
template <typename T>
void epa(const V3<T> S[4], const int ia[4], const int ib[4], const Body<T>& A, const Body<T>& B,
         T* depth, V3<T>* cpA, V3<T>* cpB, V3<T>* normal, int* status, int* faces_used, int* iters,
         int* face = nullptr, bool robust = false) {
    Poly<T> P;
    T xmin_old = T(1e29);
    T proj = T(2.0) * xmin_old + T(1.0);
    int counter = 0, nv = 4, nt = 0, indmin = 0;
    T xmin_new = T(0.0);
    int st = 0;
    for (int k = 0; k < 4; ++k) { P.V[k] = S[k]; P.opA[k] = ia[k]; P.opB[k] = ib[k]; }
    const V3<T> ctr = (((S[0] + S[1]) + S[2]) + S[3]) * T(0.25);
    const V3<T>* inner = robust ? &ctr : nullptr;
    T R2 = norm2(S[0]);
    for (int k = 1; k < 4; ++k) if (norm2(S[k]) > R2) R2 = norm2(S[k]);
    const T tol_abs = T(1e-10) * sqrt_(R2);
    add_triangle(P, 0, 1, 2, 0, inner);
    add_triangle(P, 1, 2, 3, 1, inner);
    add_triangle(P, 2, 0, 3, 2, inner);
    add_triangle(P, 0, 1, 3, 3, inner);
    nt = 4;
    int it = 0;
    while (xmin_old < proj) {
        ++it;
        indmin = 0;
        T xm = P.dist[0];
        for (int k = 0; k < nt; ++k) if (P.dist[k] < xm) { xm = P.dist[k]; indmin = k; }
        xmin_new = P.dist[indmin];
        xmin_old = xmin_new;
        int a, b;
        V3<T> Suj = support(A, B, P.n[indmin], &a, &b);
        proj = dot(P.n[indmin], Suj);
        P.opA[nv] = a;
        P.opB[nv] = b;
        if (xmin_old > proj || (xmin_old * T(1.00001) > proj && xmin_old * T(0.99999) < proj)) break;
        if (robust && proj - xmin_old <= T(1e-5) * fabs_(xmin_old) + tol_abs) break;
        if (counter > 49) { st |= 1; break; }
        counter = counter + 1;
        P.V[nv] = Suj;
        nv = nv + 1;
        for (int i = 0; i < nt; ++i)
            if ((P.valid[i] == 1 || P.valid[i] == -2) && dot(P.V[nv - 1] - P.V[P.Tr[i][0]], P.n[i]) >= T(0.0))
                P.valid[i] = -1;
        // delpair (.cu:470-493): an edge of a visible face is on the horizon iff
        // no other visible face contains the same undirected edge.
        bool horizon[EPA_MAX_FACES][3];
        for (int i = 0; i < nt; ++i) {
            if (P.valid[i] != -1) continue;
            for (int e = 0; e < 3; ++e) {
                int u = P.Tr[i][e], v = P.Tr[i][(e + 1) % 3];
                int cnt = 0;
                for (int j = 0; j < nt; ++j) {
                    if (P.valid[j] != -1) continue;
                    for (int f = 0; f < 3; ++f) {
                        int x = P.Tr[j][f], y = P.Tr[j][(f + 1) % 3];
                        if ((x == u && y == v) || (x == v && y == u)) ++cnt;
                    }
                }
                horizon[i][e] = cnt < 2;
            }
        }
        bool overflow = false;
        int nt0 = nt;
        for (int i = 0; i < nt0 && !overflow; ++i) {
            if (P.valid[i] != -1) continue;
            for (int e = 0; e < 3; ++e) {
                if (!horizon[i][e]) continue;
                if (nt == EPA_MAX_FACES) { overflow = true; break; }
                add_triangle(P, P.Tr[i][e], P.Tr[i][(e + 1) % 3], nv - 1, nt, inner);
                nt = nt + 1;
            }
            if (overflow) break;
            P.valid[i] = 0;
            P.dist[i] = T(3e20);
        }
        if (overflow) { st |= 2; break; }
        counter = counter + 1;
    }
    *depth = xmin_new;
    int a = P.Tr[indmin][0], b = P.Tr[indmin][1], c = P.Tr[indmin][2];
    V3<T> nrm = P.n[indmin];
    V3<T> Proj = nrm * (*depth);
    V3<T> Aa = P.V[a], Bb = P.V[b], Cc = P.V[c];
    V3<T> v0 = Bb - Aa, v1 = Cc - Aa, v2 = Proj - Aa;
    if (dot(cross(v0, v1), nrm) < T(0.0)) { v0 = Cc - Aa; v1 = Bb - Aa; }
    T v00 = dot(v0, v0), v01 = dot(v0, v1), v11 = dot(v1, v1), v20 = dot(v2, v0), v21 = dot(v2, v1);
    T denom = v00 * v11 - v01 * v01;
    T vv = (v11 * v20 - v01 * v21) / denom;
    T ww = (v00 * v21 - v01 * v20) / denom;
    T uu = T(1.0) - ww - vv;
    *cpA = (load3<T>(A.v + 3 * P.opA[a]) * uu + load3<T>(A.v + 3 * P.opA[b]) * vv) + load3<T>(A.v + 3 * P.opA[c]) * ww;
    *cpB = (load3<T>(B.v + 3 * P.opB[a]) * uu + load3<T>(B.v + 3 * P.opB[b]) * vv) + load3<T>(B.v + 3 * P.opB[c]) * ww;
    *normal = nrm;
    *status = st;
    *faces_used = nt;
    *iters = it;
    if (face) {  // support indices of the final face's vertices (stored order) on A, then on B
        face[0] = P.opA[a]; face[1] = P.opA[b]; face[2] = P.opA[c];
        face[3] = P.opB[a]; face[4] = P.opB[b]; face[5] = P.opB[c];
    }
}

// ----------------------------------------------------------------------- Force
template <typename T> struct PairIn {
    V3<T> comA, comB, vA, vB, wA, wB, dispt;
    Q4<T> qA;
    T isA, isB;
    T imA, imB;  // inverse masses (0 for walls), used by the physical law
    int wall;
};
template <typename T> struct ForceOut { V3<T> Fn, tA, tB, disp; };

// Force, .cu:647-871 (Q1, Q3, Q4, Q5 reproduced in mode 0). o.Fn is the
// force on A's linear momentum: the normal force (mode 0) or normal plus
// tangential (mode 1).
template <typename T>
ForceOut<T> force_law(const PairIn<T>& in, T depth, V3<T> n, V3<T> cpA, V3<T> cpB, const Law& L) {
    ForceOut<T> o;
    const bool phys = L.mode == 1;
    if (phys && !(depth < T(EPA_PLACEHOLDER) && depth > T(0.0))) {
        // EPA ended on a face whose distance is its "outside" or "deleted"
        // placeholder (1e30 / 3e20): no depth was resolved; or the signed
        // depth is not positive (the bodies touch without overlapping): the
        // physical law applies no force for the step (the reference law
        // takes the value)
        o.Fn = o.tA = o.tB = o.disp = mk<T>(T(0.0), T(0.0), T(0.0));
        return o;
    }
    V3<T> cp = phys ? (cpA + cpB) / T(2.0) : (cpA - cpB) / T(2.0);
    V3<T> dA = cp - in.comA, dB = cp - in.comB;
    T dAn = norm2(dA), dBn = norm2(dB);
    T Kt, Gt, Gn;
    if (!phys) {
        T C = (T(L.inv_mu) + dAn / in.isA) + dBn / in.isB;
        Kt = T(L.Kt_num) / C;
        Gt = (T(-2.0) / C) * T(L.ln_et);
        Gn = T(L.Gn);
    } else {
        T imn = in.imA + in.imB;
        T mn = T(1.0) / imn;
        T imt = (imn + dAn * (in.imA == T(0.0) ? T(0.0) : T(1.0) / in.isA)) +
                dBn * (in.imB == T(0.0) ? T(0.0) : T(1.0) / in.isB);
        T mt = T(1.0) / imt;
        Gn = T(L.cn) * sqrt_(mn * T(L.Kn));
        Kt = (T(L.Kn) * T(L.ratio)) * (mt / mn);
        Gt = T(L.ct) * sqrt_(mt * Kt);
    }
    if (dot(in.comB - in.comA, n) < T(0.0)) n = n * T(-1.0);
    T nn = sqrt_(norm2(n));
    n = n / nn;
    V3<T> disp_n = n * depth;
    V3<T> u = ((in.vA + cross(in.wA, cpA - in.comA)) - in.vB) - cross(in.wB, cpB - in.comB);
    V3<T> un = n * dot(u, n);
    V3<T> ut = u - un;
    T utn = sqrt_(norm2(ut));
    T dtn = sqrt_(norm2(in.dispt));
    V3<T> t = mk<T>(T(0.0), T(0.0), T(0.0));
    if (utn != T(0.0)) t = ut / utn;
    if (utn == T(0.0) && dtn != T(0.0)) t = in.dispt / dtn;
    V3<T> Fn = (disp_n * T(-1.0)) * T(L.Kn) - un * Gn;
    T Fnn = sqrt_(norm2(Fn));
    V3<T> Ft = (in.dispt * T(-1.0)) * Kt - ut * Gt;
    T Ftn = sqrt_(norm2(Ft));
    T muF = T(L.mu) * Fnn;
    V3<T> dnew = mk<T>(T(0.0), T(0.0), T(0.0));
    if (Ftn > muF) Ft = ((t * Fnn) * T(-1.0)) * T(L.mu);
    if (Ftn <= muF) {
        V3<T> r = qtripl(in.qA, in.dispt, qinv(in.qA));
        dnew = r + ut * T(L.dt);
    } else if (!phys) {
        // .cu:800-806 as written: the slip displacement makes the next step's
        // spring force point along the sliding (the sign of the mu term)
        dnew = ((((t * Fnn) * T(L.mu)) * T(-1.0)) - ut * Gt) / Kt;
    } else {
        // physical law: the displacement whose spring-dashpot force is the
        // Coulomb limit opposing the sliding, -Kt d - Gt ut = -mu |Fn| t
        dnew = (((t * Fnn) * T(L.mu)) - ut * Gt) / Kt;
    }
    V3<T> Ff = Ft + Fn;
    o.Fn = phys ? Ff : Fn;
    o.tA = cross(cp - in.comA, Ff);
    o.tB = cross(cp - in.comB, Ff * T(-1.0));
    o.disp = dnew;
    return o;
}

// ------------------------------------------------------- contact_detection
struct PairResult {
    int collision;
    int epa_status;
    int faces;
    int iters;
    int face[6];
    int gjk_iters;  // tetrahedron iterations of GJK (.cu:1104-1288)
    int flat;       // CorrectFlat re-supported the fourth vertex (.cu:1119-1128)
    double depth, n[3], cpA[3], cpB[3], Fn[3], tA[3], tB[3], disp[4];
};

// contact_detection, .cu:874-1322, for one pair.
template <typename T>
PairResult contact_one(const Body<T>& A, const Body<T>& B, const PairIn<T>& in, const Law& L) {
    PairResult r;
    std::memset(&r, 0, sizeof(r));
    r.collision = -1;
    V3<T> S[4];
    int ia[4] = {0, 0, 0, 0}, ib[4] = {0, 0, 0, 0};
    V3<T> d1 = in.comB - in.comA;
    S[0] = support(A, B, d1, &ia[0], &ib[0]);
    if (dot(d1, S[0]) < T(0.0)) { r.collision = 0; return r; }
    V3<T> d2 = neg(d1);
    S[1] = support(A, B, d2, &ia[1], &ib[1]);
    if (dot(d2, S[1]) < T(0.0)) { r.collision = 0; return r; }
    int coll;
    V3<T> d3;
    check_segment(S[0], S[1], &coll, &d3);
    if (coll == 0) { r.collision = 0; return r; }
    S[2] = support(A, B, d3, &ia[2], &ib[2]);
    if (dot(d3, S[2]) < T(0.0)) { r.collision = 0; return r; }
    V3<T> d4;
    check_triangle(S[0], S[1], S[2], &coll, &d4);
    if (coll == 0) { r.collision = 0; return r; }
    for (int zz = 0; zz < 50; ++zz) {
        r.gjk_iters = zz + 1;
        S[3] = support(A, B, d4, &ia[3], &ib[3]);
        if (dot(d4, S[3]) < T(0.0)) { r.collision = 0; return r; }
        bool hit = false;
        if (is_zero(S[3])) {
            hit = true;
        } else {
            int todel;
            check_tetra(S[0], S[1], S[2], S[3], &coll, &d4, &todel);
            if (coll == 0) { r.collision = 0; return r; }
            if (coll == 1) hit = true;
            else {
                V3<T> n1, n2, n3;
                int a1, b1, a2, b2, a3, b3;
                if (todel == 1) { n1 = S[1]; a1 = ia[1]; b1 = ib[1]; n2 = S[2]; a2 = ia[2]; b2 = ib[2]; }
                else if (todel == 2) { n1 = S[2]; a1 = ia[2]; b1 = ib[2]; n2 = S[0]; a2 = ia[0]; b2 = ib[0]; }
                else { n1 = S[0]; a1 = ia[0]; b1 = ib[0]; n2 = S[1]; a2 = ia[1]; b2 = ib[1]; }
                n3 = S[3]; a3 = ia[3]; b3 = ib[3];
                S[0] = n1; ia[0] = a1; ib[0] = b1;
                S[1] = n2; ia[1] = a2; ib[1] = b2;
                S[2] = n3; ia[2] = a3; ib[2] = b3;
            }
        }
        if (hit) {
            const bool robust = L.mode == 1;
            if (robust) {
                // physical law: a flat tetrahedron gets its fourth vertex from
                // the normal of the first three, on the side where the
                // Minkowski difference reaches farther (so EPA starts from a
                // solid; not in the reference)
                if (flat_tetra(S[0], S[1], S[2], S[3])) {
                    r.flat = 1;
                    const V3<T> nT = cross(S[1] - S[0], S[2] - S[0]);
                    int ap, bp, am, bm;
                    const V3<T> Sp = support(A, B, nT, &ap, &bp);
                    const V3<T> Sm = support(A, B, neg(nT), &am, &bm);
                    if (dot(nT, Sp - S[0]) >= dot(nT, S[0] - Sm)) { S[3] = Sp; ia[3] = ap; ib[3] = bp; }
                    else { S[3] = Sm; ia[3] = am; ib[3] = bm; }
                }
            } else if (correct_flat(S[0], S[1], S[2], S[3])) {
                r.flat = 1;
                d4 = cross(S[0], S[1]);
                S[3] = support(A, B, d4, &ia[3], &ib[3]);
            }
            r.collision = 1;
            T depth;
            V3<T> cpA, cpB, nrm;
            epa(S, ia, ib, A, B, &depth, &cpA, &cpB, &nrm, &r.epa_status, &r.faces, &r.iters, r.face, robust);
            if (!(val(depth) < EPA_PLACEHOLDER)) r.epa_status |= 4;
            ForceOut<T> f = force_law(in, depth, nrm, cpA, cpB, L);
            r.depth = val(depth);
            r.n[0] = val(nrm.x); r.n[1] = val(nrm.y); r.n[2] = val(nrm.z);
            r.cpA[0] = val(cpA.x); r.cpA[1] = val(cpA.y); r.cpA[2] = val(cpA.z);
            r.cpB[0] = val(cpB.x); r.cpB[1] = val(cpB.y); r.cpB[2] = val(cpB.z);
            r.Fn[0] = val(f.Fn.x); r.Fn[1] = val(f.Fn.y); r.Fn[2] = val(f.Fn.z);
            r.tA[0] = val(f.tA.x); r.tA[1] = val(f.tA.y); r.tA[2] = val(f.tA.z);
            r.tB[0] = val(f.tB.x); r.tB[1] = val(f.tB.y); r.tB[2] = val(f.tB.z);
            r.disp[0] = val(f.disp.x); r.disp[1] = val(f.disp.y); r.disp[2] = val(f.disp.z); r.disp[3] = 1.0;
            return r;
        }
    }
    return r;  // tetrahedron cap exhausted: collision stays -1, no force (.cu:1104,1287)
}

// ----------------------------------------------------------------------- Verlet
struct StateView {
    double *v, *v12, *w12bf, *q12, *q, *qall, *wbf, *w;
    const double *F, *tau;
    double *verts;
    double *com;
    const double *mass, *Iinv, *I;
    double *dxall;
};

// Verlet, gpu_Verlet.cu:85-404, for particle P.
// rot_mode 0: vertices rotated by the absolute orientation q each step, as the
// reference does (gpu_Verlet.cu:279-305, the rotation compounds); 1: by the
// step's increment q q_old^-1 (the physical variant).
void verlet_one(const StateView& s, int64_t P, int vmax, int nrot, double dt, const double* box, int rot_mode) {
    typedef double T;
    V3<T> velocity = load3<T>(s.v + 3 * P);
    V3<T> v12 = load3<T>(s.v12 + 3 * P);
    V3<T> w12bf = load3<T>(s.w12bf + 3 * P);
    V3<T> wbf = load3<T>(s.wbf + 3 * P);
    V3<T> w = load3<T>(s.w + 3 * P);
    V3<T> com = load3<T>(s.com + 3 * P);
    V3<T> F = load3<T>(s.F + 3 * P);
    V3<T> tau = load3<T>(s.tau + 3 * P);
    double m = s.mass[P];
    const double* Ii = s.Iinv + 9 * P;
    const double* I = s.I + 9 * P;
    Q4<T> q = qload<T>(s.q + 4 * P), q12 = qload<T>(s.q12 + 4 * P);
    V3<T> dxall = load3<T>(s.dxall + 3 * P);

    V3<T> v32 = v12 + (F / m) * dt;
    Q4<T> qi = qinv(q);
    V3<T> tbf = qtripl(qi, tau, q);
    V3<T> p1 = mk(I[0] * wbf.x + I[1] * wbf.y + I[2] * wbf.z,
                  I[3] * wbf.x + I[4] * wbf.y + I[5] * wbf.z,
                  I[6] * wbf.x + I[7] * wbf.y + I[8] * wbf.z);
    V3<T> p2 = tbf - cross(wbf, p1);
    V3<T> abf = mk(Ii[0] * p2.x + Ii[1] * p2.y + Ii[2] * p2.z,
                   Ii[3] * p2.x + Ii[4] * p2.y + Ii[5] * p2.z,
                   Ii[6] * p2.x + Ii[7] * p2.y + Ii[8] * p2.z);
    V3<T> w32bf = w12bf + abf * dt;
    T wn = std::sqrt(norm2(w));
    V3<T> t11 = mk(0.0, 0.0, 0.0);
    if (wn != 0) t11 = (w / wn) * std::sin((wn * dt) / 2);
    T t112 = std::cos((wn * dt) / 2);
    Q4<T> q32 = qmul(qmk(t112, t11), q12);
    v12 = v32;
    q12 = q32;
    w12bf = w32bf;
    V3<T> dx = v12 * dt;
    dxall = dxall + dx;
    V3<T> w34bf = w12bf + (abf * dt) / 4.0;
    V3<T> w34 = qtripl(q12, w34bf, qinv(q12));
    T n34 = std::sqrt(norm2(w34));
    V3<T> t1 = mk(0.0, 0.0, 0.0);
    if (n34 != 0) t1 = (w34 / n34) * std::sin((n34 * dt) / 4);
    T t12 = std::cos((n34 * dt) / 4);
    Q4<T> qold = q;
    q = qmul(qmk(t12, t1), q12);
    Q4<T> qall = qmul(q, qold);
    wbf = w12bf + (abf * dt) / 2.0;
    w = qtripl(q, wbf, qinv(q));
    velocity = v12 + (F * dt) / m;
    Q4<T> qr = rot_mode == 1 ? qmul(q, qinv(qold)) : q;
    Q4<T> qinv2 = qinv(qr);
    double* vp = s.verts + (int64_t)vmax * 3 * P;
    for (int i = 0; i < nrot; ++i) {
        V3<T> r = qtripl(qr, load3<T>(vp + 3 * i) - com, qinv2);
        V3<T> nvx = r + com;
        vp[3 * i] = nvx.x; vp[3 * i + 1] = nvx.y; vp[3 * i + 2] = nvx.z;
    }
    com = com + dx;
    for (int i = 0; i < nrot; ++i) {
        vp[3 * i] = vp[3 * i] + dx.x; vp[3 * i + 1] = vp[3 * i + 1] + dx.y; vp[3 * i + 2] = vp[3 * i + 2] + dx.z;
    }
    // box order: x_max, x_min, y_max, y_min, z_max, z_min (gpu_Verlet.cu:338)
    if (com.x > box[0] || com.x < box[1] || com.y > box[2] || com.y < box[3] || com.z > box[4] || com.z < box[5]) {
        velocity = mk(0.0, 0.0, 0.0);
        v12 = mk(0.0, 0.0, 0.0);
        w12bf = mk(0.0, 0.0, 0.0);
        wbf = mk(0.0, 0.0, 0.0);
        w = mk(0.0, 0.0, 0.0);
        V3<T> du = dx * -1.0;
        com = com + du;
        for (int i = 0; i < nrot; ++i) {
            vp[3 * i] = vp[3 * i] + du.x; vp[3 * i + 1] = vp[3 * i + 1] + du.y; vp[3 * i + 2] = vp[3 * i + 2] + du.z;
        }
    }
    auto st3 = [&](double* p, V3<T> a) { p[3 * P] = a.x; p[3 * P + 1] = a.y; p[3 * P + 2] = a.z; };
    auto st4 = [&](double* p, Q4<T> a) { p[4 * P] = a.w; p[4 * P + 1] = a.x; p[4 * P + 2] = a.y; p[4 * P + 3] = a.z; };
    st3(s.v, velocity); st3(s.v12, v12); st3(s.w12bf, w12bf); st3(s.com, com); st3(s.dxall, dxall);
    st3(s.wbf, wbf); st3(s.w, w);
    st4(s.q12, q12); st4(s.q, q); st4(s.qall, qall);
}

// ------------------------------------------------------------ CPython set order
// Iteration order of a CPython (3.7+) set of non-negative ints filled in the
// given order: open addressing with 9 linear probes then perturbation,
// growth to 4*used when fill*5 >= mask*3 (Objects/setobject.c).
void probe_insert(std::vector<int64_t>& tab, size_t mask, int64_t key) {
    size_t h = (size_t)key;
    size_t perturb = h;
    size_t i = h & mask;
    for (;;) {
        if (tab[i] < 0) { tab[i] = key; return; }
        if (i + 9 <= mask) {
            for (size_t j = 1; j <= 9; ++j)
                if (tab[i + j] < 0) { tab[i + j] = key; return; }
        }
        perturb >>= 5;
        i = (i * 5 + 1 + perturb) & mask;
    }
}

std::vector<int64_t> set_order(const std::vector<int64_t>& keys) {
    size_t size = 8;
    std::vector<int64_t> tab(size, -1);
    size_t fill = 0;
    for (int64_t k : keys) {
        probe_insert(tab, size - 1, k);
        ++fill;
        if (fill * 5 >= (size - 1) * 3) {
            size_t minused = fill > 50000 ? fill * 2 : fill * 4;
            size_t ns = 8;
            while (ns <= minused) ns <<= 1;
            std::vector<int64_t> nt(ns, -1);
            for (int64_t e : tab) if (e >= 0) probe_insert(nt, ns - 1, e);
            tab.swap(nt);
            size = ns;
        }
    }
    std::vector<int64_t> out;
    for (int64_t e : tab) if (e >= 0) out.push_back(e);
    return out;
}

}  // namespace

// =================================================================== C ABI
ORC_EXPORT int orc_epa_max_faces() { return EPA_MAX_FACES; }

ORC_EXPORT void orc_set_order(const int64_t* ids, int64_t n, int64_t* out) {
    std::vector<int64_t> k(ids, ids + n);
    std::vector<int64_t> o = set_order(k);
    for (size_t i = 0; i < o.size(); ++i) out[i] = o[i];
}

// Cell keys, MWDiem.py:813-816: truncation toward zero, no origin offset.
ORC_EXPORT void orc_cell_keys(const double* com, int64_t n, double csx, double csy, double csz,
                              int64_t ncx, int64_t ncy, int64_t* keys) {
    for (int64_t i = 0; i < n; ++i) {
        int64_t xc = (int64_t)(com[3 * i] / csx);
        int64_t yc = (int64_t)(com[3 * i + 1] / csy);
        int64_t zc = (int64_t)(com[3 * i + 2] / csz);
        keys[i] = xc + yc * ncx + zc * ncx * ncy;
    }
}

// Wall slab test, MWDiem.py:807-811: walls k in [0,W) outer, particles j in
// [W,n) inner, pair [j,k] when |n_k . (x_j - c_k)| < thr. Returns the count;
// writes at most cap pairs.
ORC_EXPORT int64_t orc_wall_pairs(const double* com, int64_t W, int64_t n, const double* wall_n,
                                  const double* wall_c, double thr, int32_t* pairs, int64_t cap) {
    int64_t cnt = 0;
    for (int64_t k = 0; k < W; ++k) {
        for (int64_t j = W; j < n; ++j) {
            double d0 = com[3 * j] - wall_c[3 * k], d1 = com[3 * j + 1] - wall_c[3 * k + 1], d2 = com[3 * j + 2] - wall_c[3 * k + 2];
            double dd = wall_n[3 * k] * d0 + wall_n[3 * k + 1] * d1 + wall_n[3 * k + 2] * d2;
            if (std::fabs(dd) < thr) {
                if (cnt < cap) { pairs[2 * cnt] = (int32_t)j; pairs[2 * cnt + 1] = (int32_t)k; }
                ++cnt;
            }
        }
    }
    return cnt;
}

// Linked_cell_hash, LIBRARY_Coarse_contact_detection.py:84-155, over ids
// [pmin,pmax) with keys[i]. Returns the count; writes at most cap pairs.
// stencil 0: the reference's 7 forward offsets; 1: the full 13-cell forward
// half shell (dz>0, or dz=0 and dy>0, or dz=dy=0 and dx>0).
ORC_EXPORT int64_t orc_linked_cell_hash(const int64_t* keys, int64_t pmin, int64_t pmax, int64_t ncx,
                                        int64_t ncy, int32_t* pairs, int64_t cap, int stencil) {
    std::unordered_map<int64_t, size_t> where;
    std::vector<int64_t> order;
    std::vector<std::vector<int64_t>> members;
    for (int64_t i = pmin; i < pmax; ++i) {
        auto it = where.find(keys[i]);
        if (it == where.end()) {
            where[keys[i]] = members.size();
            order.push_back(keys[i]);
            members.push_back(std::vector<int64_t>(1, i));
        } else {
            members[it->second].push_back(i);
        }
    }
    for (auto& m : members) m = set_order(m);
    std::vector<int64_t> offs;
    if (stencil == 1) {
        const int h[13][3] = {{1, 0, 0},  {-1, 1, 0}, {0, 1, 0},  {1, 1, 0},  {-1, -1, 1}, {0, -1, 1}, {1, -1, 1},
                              {-1, 0, 1}, {0, 0, 1},  {1, 0, 1},  {-1, 1, 1}, {0, 1, 1},   {1, 1, 1}};
        for (int k = 0; k < 13; ++k) offs.push_back(h[k][0] + h[k][1] * ncx + h[k][2] * ncx * ncy);
    } else {
        offs = {1, ncx, ncx * ncy, ncx * ncy + 1, ncx * ncy + 1 + ncx, ncx * ncy + ncx, ncx + 1};
    }
    int64_t cnt = 0;
    auto emit = [&](int64_t a, int64_t b) {
        if (cnt < cap) { pairs[2 * cnt] = (int32_t)a; pairs[2 * cnt + 1] = (int32_t)b; }
        ++cnt;
    };
    for (size_t c = 0; c < order.size(); ++c) {
        const std::vector<int64_t>& L = members[c];
        for (size_t j = 0; j < L.size(); ++j)
            for (size_t k = j + 1; k < L.size(); ++k) emit(L[j], L[k]);
        for (size_t o = 0; o < offs.size(); ++o) {
            auto it = where.find(order[c] + offs[o]);
            if (it == where.end()) continue;
            const std::vector<int64_t>& M = members[it->second];
            for (size_t j = 0; j < L.size(); ++j)
                for (size_t k = 0; k < M.size(); ++k) emit(L[j], M[k]);
        }
    }
    return cnt;
}

// Law coefficients as the product computes them (for cross-checks).
ORC_EXPORT void orc_law(double Kn, double en, double et, double mu, double dt, double* out13) {
    Law L = make_law(Kn, en, et, mu, dt);
    double v[13] = {L.Kn, L.mu, L.ln_en, L.ln_et, L.A_en, L.tc, L.Gn, L.Kt_num, L.inv_mu, L.dt, L.cn, L.ct, L.ratio};
    for (int i = 0; i < 13; ++i) out13[i] = v[i];
}

// Per-pair fine contact in the reference's own input layout
// (gpu_fine_contact_detection.py:76-119 with FP64 arrays):
//   comA/comB/vA/vB/wA/wB [P,3], vertsA/vertsB [P,vmax,3], nA/nB/idA/idB [P],
//   dispt [P,4], qA/qB [P,4], isA/isB [P].
// Outputs per pair: collision, epa status, depth, normal, cpA, cpB, Fn, tauA,
// tauB, disp [P,4]. Per-particle force/torque [N,3] are accumulated in
// ascending pair order onto their current contents.
ORC_EXPORT void orc_contact_pairs(int64_t P, int vmax,
                                  const double* comA, const double* comB,
                                  const double* vertsA, const double* vertsB,
                                  const double* vA, const double* vB, const double* wA, const double* wB,
                                  const int32_t* nA, const int32_t* nB, const int32_t* idA, const int32_t* idB,
                                  const double* dispt, const double* qA, const double* qB,
                                  const double* isA, const double* isB,
                                  const double* law_p, const double* law_w, int64_t W, double dt,
                                  int32_t* collision, int32_t* epa_status, double* depth, double* normal,
                                  double* cpA, double* cpB, double* Fn, double* tauA, double* tauB,
                                  double* disp_out, double* force, double* torque, int32_t* faces_used) {
    Law Lp = make_law(law_p[0], law_p[1], law_p[2], law_p[3], dt, (int)law_p[4]);
    Law Lw = make_law(law_w[0], law_w[1], law_w[2], law_w[3], dt, (int)law_w[4]);
    (void)qB;
    std::unordered_map<int64_t, std::vector<double>> sums;
    for (int64_t p = 0; p < P; ++p) {
        Body<double> A{vertsA + (int64_t)vmax * 3 * p, nA[p]};
        Body<double> B{vertsB + (int64_t)vmax * 3 * p, nB[p]};
        PairIn<double> in;
        in.comA = load3<double>(comA + 3 * p); in.comB = load3<double>(comB + 3 * p);
        in.vA = load3<double>(vA + 3 * p); in.vB = load3<double>(vB + 3 * p);
        in.wA = load3<double>(wA + 3 * p); in.wB = load3<double>(wB + 3 * p);
        in.dispt = load3<double>(dispt + 4 * p);
        in.qA = qload<double>(qA + 4 * p);
        in.isA = isA[p]; in.isB = isB[p];
        in.imA = 0.0; in.imB = 0.0;  // the per-pair layout carries no masses (reference law only)
        in.wall = (idA[p] < W || idB[p] < W) ? 1 : 0;
        PairResult r = contact_one(A, B, in, in.wall ? Lw : Lp);
        collision[p] = r.collision;
        if (epa_status) epa_status[p] = r.epa_status;
        if (faces_used) faces_used[p] = r.faces;
        if (depth) depth[p] = r.depth;
        for (int k = 0; k < 3; ++k) {
            if (normal) normal[3 * p + k] = r.n[k];
            if (cpA) cpA[3 * p + k] = r.cpA[k];
            if (cpB) cpB[3 * p + k] = r.cpB[k];
            if (Fn) Fn[3 * p + k] = r.Fn[k];
            if (tauA) tauA[3 * p + k] = r.tA[k];
            if (tauB) tauB[3 * p + k] = r.tB[k];
        }
        if (disp_out) for (int k = 0; k < 4; ++k) disp_out[4 * p + k] = r.disp[k];
        if (r.collision == 1 && force) {
            int a = idA[p], b = idB[p];
            std::vector<double>& sa = sums[a];
            std::vector<double>& sb = sums[b];
            if (sa.empty()) sa.assign(6, 0.0);
            if (sb.empty()) sb.assign(6, 0.0);
            for (int k = 0; k < 3; ++k) {
                sa[k] = sa[k] + r.Fn[k];
                sb[k] = sb[k] + (-r.Fn[k]);
                sa[3 + k] = sa[3 + k] + r.tA[k];
                sb[3 + k] = sb[3 + k] + r.tB[k];
            }
        }
    }
    // contact sums (from zero, ascending pair order) are added once to the
    // caller's force/torque rows.
    if (force)
        for (auto& kv : sums)
            for (int k = 0; k < 3; ++k) {
                force[3 * kv.first + k] = force[3 * kv.first + k] + kv.second[k];
                torque[3 * kv.first + k] = torque[3 * kv.first + k] + kv.second[3 + k];
            }
}

// FP64 flop count of the fine-contact stage (GJK + EPA + force law) over a
// pair list indexed into per-particle state (the device-resident layout).
ORC_EXPORT uint64_t orc_count_flops(int64_t P, const int32_t* pa, const int32_t* pb, int vmax,
                                    const double* com, const double* verts, const int32_t* nverts,
                                    const double* vel, const double* angvel, const double* q,
                                    const double* isize, const double* mass, const double* law_p, const double* law_w,
                                    int64_t W, double dt) {
    Law Lp = make_law(law_p[0], law_p[1], law_p[2], law_p[3], dt, (int)law_p[4]);
    Law Lw = make_law(law_w[0], law_w[1], law_w[2], law_w[3], dt, (int)law_w[4]);
    uint64_t total = 0;
#pragma omp parallel reduction(+ : total)
    {
    g_flops = 0;
#pragma omp for schedule(dynamic, 64)
    for (int64_t p = 0; p < P; ++p) {
        int a = pa[p], b = pb[p];
        Body<CD> A{verts + (int64_t)vmax * 3 * a, nverts[a]};
        Body<CD> B{verts + (int64_t)vmax * 3 * b, nverts[b]};
        PairIn<CD> in;
        in.comA = load3<CD>(com + 3 * a); in.comB = load3<CD>(com + 3 * b);
        in.vA = load3<CD>(vel + 3 * a); in.vB = load3<CD>(vel + 3 * b);
        in.wA = load3<CD>(angvel + 3 * a); in.wB = load3<CD>(angvel + 3 * b);
        in.dispt = mk<CD>(CD(0.0), CD(0.0), CD(0.0));
        in.qA = qload<CD>(q + 4 * a);
        in.isA = CD(isize[a]); in.isB = CD(isize[b]);
        in.imA = CD(a < W ? 0.0 : 1.0 / mass[a]); in.imB = CD(b < W ? 0.0 : 1.0 / mass[b]);
        in.wall = (a < W || b < W) ? 1 : 0;
        contact_one(A, B, in, in.wall ? Lw : Lp);
    }
    total += g_flops;
    }
    return total;
}

// Verlet over rows [pmin,pmax) of the reference's arrays (gpu_Verlet.py:53-92).
// nrot: vertex slots rotated per particle (vmax for the reference behaviour).
ORC_EXPORT void orc_verlet(double* v, double* v12, double* w12bf, double* q12, double* q, double* qall,
                           double* wbf, double* w, const double* F, const double* tau, double* verts,
                           double* com, const double* mass, const double* Iinv, const double* I,
                           double* dxall, int vmax, const int32_t* nrot, double dt, const double* box,
                           int64_t pmin, int64_t pmax, int rot_mode) {
    StateView s{v, v12, w12bf, q12, q, qall, wbf, w, F, tau, verts, com, mass, Iinv, I, dxall};
#pragma omp parallel for schedule(static)
    for (int64_t P = pmin; P < pmax; ++P) verlet_one(s, P, vmax, nrot ? nrot[P] : vmax, dt, box, rot_mode);
}

// Per-particle contact sums (ascending pair order) for a pair list indexed into
// per-particle state; writes collision codes and returns the contact count.
ORC_EXPORT int64_t orc_contact_state(int64_t P, const int32_t* pa, const int32_t* pb, int vmax,
                                     const double* com, const double* verts, const int32_t* nverts,
                                     const double* vel, const double* angvel, const double* q,
                                     const double* isize, const double* mass, const double* law_p, const double* law_w,
                                     int64_t W, double dt, int32_t* collision, double* depth,
                                     double* normal, double* cpA, double* cpB, double* force,
                                     double* torque, int32_t* epa_status, const double* dispt,
                                     double* disp_out) {
    Law Lp = make_law(law_p[0], law_p[1], law_p[2], law_p[3], dt, (int)law_p[4]);
    Law Lw = make_law(law_w[0], law_w[1], law_w[2], law_w[3], dt, (int)law_w[4]);
    int64_t K = 0;
    // pairs are independent: each chunk is evaluated in parallel (OpenMP),
    // then its contacts are added to the particle sums one by one in pair
    // order, so the sums do not depend on the thread count
    const int64_t CH = int64_t(1) << 18;
    std::vector<double> fb(9 * (size_t)std::min<int64_t>(P, CH));
    for (int64_t p0 = 0; p0 < P; p0 += CH) {
        const int64_t p1 = std::min<int64_t>(P, p0 + CH);
#pragma omp parallel for schedule(dynamic, 64)
        for (int64_t p = p0; p < p1; ++p) {
            int a = pa[p], b = pb[p];
            Body<double> A{verts + (int64_t)vmax * 3 * a, nverts[a]};
            Body<double> B{verts + (int64_t)vmax * 3 * b, nverts[b]};
            PairIn<double> in;
            in.comA = load3<double>(com + 3 * a); in.comB = load3<double>(com + 3 * b);
            in.vA = load3<double>(vel + 3 * a); in.vB = load3<double>(vel + 3 * b);
            in.wA = load3<double>(angvel + 3 * a); in.wB = load3<double>(angvel + 3 * b);
            // tangential history of the pair ([P,4] per pair; none: reset, Q1)
            in.dispt = dispt ? mk<double>(dispt[4 * p], dispt[4 * p + 1], dispt[4 * p + 2]) : mk<double>(0.0, 0.0, 0.0);
            in.qA = qload<double>(q + 4 * a);
            in.isA = isize[a]; in.isB = isize[b];
            in.imA = a < W ? 0.0 : 1.0 / mass[a]; in.imB = b < W ? 0.0 : 1.0 / mass[b];
            in.wall = (a < W || b < W) ? 1 : 0;
            PairResult r = contact_one(A, B, in, in.wall ? Lw : Lp);
            collision[p] = r.collision;
            if (epa_status) epa_status[p] = r.epa_status;
            if (depth) depth[p] = r.depth;
            for (int k = 0; k < 3; ++k) {
                if (normal) normal[3 * p + k] = r.n[k];
                if (cpA) cpA[3 * p + k] = r.cpA[k];
                if (cpB) cpB[3 * p + k] = r.cpB[k];
            }
            if (disp_out && r.collision == 1)
                for (int k = 0; k < 4; ++k) disp_out[4 * p + k] = r.disp[k];
            double* f = fb.data() + 9 * (p - p0);
            for (int k = 0; k < 3; ++k) { f[k] = r.Fn[k]; f[3 + k] = r.tA[k]; f[6 + k] = r.tB[k]; }
        }
        for (int64_t p = p0; p < p1; ++p) {
            if (collision[p] != 1) continue;
            ++K;
            const int a = pa[p], b = pb[p];
            const double* f = fb.data() + 9 * (p - p0);
            for (int k = 0; k < 3; ++k) {
                force[3 * a + k] = force[3 * a + k] + f[k];
                force[3 * b + k] = force[3 * b + k] + (-f[k]);
                torque[3 * a + k] = torque[3 * a + k] + f[3 + k];
                torque[3 * b + k] = torque[3 * b + k] + f[6 + k];
            }
        }
    }
    return K;
}

// Contact topology of a pair list: collision codes and, for colliding pairs,
// the support indices of EPA's final face (face[6p..6p+2] on A, +3..+5 on B).
// mode_p / mode_w: the law modes of particle pairs and of wall pairs (rows
// below n_walls), which select the GJK/EPA variant as in contact_state.
ORC_EXPORT void orc_contact_faces(int64_t P, const int32_t* pa, const int32_t* pb, int vmax, const double* com,
                                  const double* verts, const int32_t* nverts, int32_t* collision, int32_t* face,
                                  int mode_p, int mode_w, int64_t n_walls) {
    const Law Lp = make_law(1.0, 0.5, 0.5, 0.5, 1.0, mode_p), Lw = make_law(1.0, 0.5, 0.5, 0.5, 1.0, mode_w);
    for (int64_t p = 0; p < P; ++p) {
        int a = pa[p], b = pb[p];
        const Law& L = (a < n_walls || b < n_walls) ? Lw : Lp;
        Body<double> A{verts + (int64_t)vmax * 3 * a, nverts[a]};
        Body<double> B{verts + (int64_t)vmax * 3 * b, nverts[b]};
        PairIn<double> in;
        std::memset(&in, 0, sizeof(in));
        in.comA = load3<double>(com + 3 * a);
        in.comB = load3<double>(com + 3 * b);
        in.qA.w = 1.0;
        in.isA = 1.0; in.isB = 1.0;
        PairResult r = contact_one(A, B, in, L);
        collision[p] = r.collision;
        for (int k = 0; k < 6; ++k) face[6 * p + k] = r.collision == 1 ? r.face[k] : -1;
    }
}

// GJK's tetrahedron iteration count and collision code per pair (test and
// measurement helper).
ORC_EXPORT void orc_gjk_iters(int64_t P, const int32_t* pa, const int32_t* pb, int vmax, const double* com,
                              const double* verts, const int32_t* nverts, int32_t* collision, int32_t* iters) {
    Law L = make_law(1.0, 0.5, 0.5, 0.5, 1.0);
    for (int64_t p = 0; p < P; ++p) {
        int a = pa[p], b = pb[p];
        Body<double> A{verts + (int64_t)vmax * 3 * a, nverts[a]};
        Body<double> B{verts + (int64_t)vmax * 3 * b, nverts[b]};
        PairIn<double> in;
        std::memset(&in, 0, sizeof(in));
        in.comA = load3<double>(com + 3 * a);
        in.comB = load3<double>(com + 3 * b);
        in.qA.w = 1.0;
        in.isA = 1.0; in.isB = 1.0;
        PairResult r = contact_one(A, B, in, L);
        collision[p] = r.collision;
        iters[p] = r.gjk_iters;
    }
}

// GJK/EPA path of each pair of a list (test helper): collision code,
// tetrahedron iterations, whether CorrectFlat fired, EPA status and EPA
// expansions.
ORC_EXPORT void orc_gjk_trace(int64_t P, const int32_t* pa, const int32_t* pb, int vmax, const double* com,
                              const double* verts, const int32_t* nverts, int32_t* collision, int32_t* iters,
                              int32_t* flat, int32_t* epa_status, int32_t* epa_iters, int robust) {
    Law L = make_law(1.0, 0.5, 0.5, 0.5, 1.0, robust ? 1 : 0);
    for (int64_t p = 0; p < P; ++p) {
        int a = pa[p], b = pb[p];
        Body<double> A{verts + (int64_t)vmax * 3 * a, nverts[a]};
        Body<double> B{verts + (int64_t)vmax * 3 * b, nverts[b]};
        PairIn<double> in;
        std::memset(&in, 0, sizeof(in));
        in.comA = load3<double>(com + 3 * a);
        in.comB = load3<double>(com + 3 * b);
        in.qA.w = 1.0;
        in.isA = 1.0; in.isB = 1.0;
        PairResult r = contact_one(A, B, in, L);
        collision[p] = r.collision;
        iters[p] = r.gjk_iters;
        flat[p] = r.flat;
        epa_status[p] = r.epa_status;
        epa_iters[p] = r.iters;
    }
}

// Threads the parallel loops of the oracle use (OpenMP).
ORC_EXPORT int orc_threads() { return omp_get_max_threads(); }

// ------------------------------------------------------------ test helpers
ORC_EXPORT void orc_quat_mul(const double* a, const double* b, double* out) {
    Q4<double> r = qmul(qload<double>(a), qload<double>(b));
    out[0] = r.w; out[1] = r.x; out[2] = r.y; out[3] = r.z;
}

ORC_EXPORT void orc_quat_tripl(const double* q1, const double* v, const double* q2, double* out) {
    V3<double> r = qtripl(qload<double>(q1), load3<double>(v), qload<double>(q2));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

// GJK + EPA of two vertex sets with given centres (contact_detection's GJK
// driver and EPA, no force law; robust: the physical law's variant).
// out: [depth, n(3), cpA(3), cpB(3)].
ORC_EXPORT int orc_gjk_epa(const double* vA, int nA, const double* vB, int nB, const double* comA, const double* comB,
                           double* out, int* status, int robust) {
    Body<double> A{vA, nA};
    Body<double> B{vB, nB};
    PairIn<double> in;
    std::memset(&in, 0, sizeof(in));
    in.comA = load3<double>(comA);
    in.comB = load3<double>(comB);
    in.qA.w = 1.0;
    in.isA = 1.0; in.isB = 1.0;
    Law L = make_law(1.0, 0.5, 0.5, 0.5, 1.0, robust ? 1 : 0);
    PairResult r = contact_one(A, B, in, L);
    out[0] = r.depth;
    for (int k = 0; k < 3; ++k) { out[1 + k] = r.n[k]; out[4 + k] = r.cpA[k]; out[7 + k] = r.cpB[k]; }
    if (status) *status = r.epa_status;
    return r.collision;
}

// Flop count of one GJK/EPA + law evaluation of a pair (reference law).
ORC_EXPORT uint64_t orc_pair_flops(const double* vA, int nA, const double* vB, int nB, const double* comA,
                                   const double* comB) {
    Body<CD> A{vA, nA};
    Body<CD> B{vB, nB};
    PairIn<CD> in;
    in.comA = load3<CD>(comA);
    in.comB = load3<CD>(comB);
    in.vA = in.vB = in.wA = in.wB = in.dispt = mk<CD>(CD(0.0), CD(0.0), CD(0.0));
    in.qA.w = CD(1.0); in.qA.x = in.qA.y = in.qA.z = CD(0.0);
    in.isA = CD(1.0); in.isB = CD(1.0); in.imA = CD(0.0); in.imB = CD(0.0);
    in.wall = 0;
    Law L = make_law(1.0, 0.5, 0.5, 0.5, 1.0);
    g_flops = 0;
    contact_one(A, B, in, L);
    return g_flops;
}
