// Fine contact: GJK, EPA and the normal/tangential contact law (FP64).
//
// Follows the control flow of the reference's contact_detection kernel
// (LIBRARIES/gpu_fine_contact_detection.cu:874-1322) and its helpers
// (:41-871): the same support-map tie-breaking (first strict maximum), the
// same Voronoi tests and simplex reduction, the same EPA face order,
// termination constants (1e-5 relative, 25 expansions) and contact-point
// construction, and the same law (including the reference's quirks: contact
// point (cpA-cpB)/2, mu in the damping time, normal force only on the linear
// momentum). Evaluated in FP64 with a fixed operator order.
//
// GJK runs G lanes per pair (k_gjk); a colliding pair claims a contact slot
// and stores its final simplex. EPA and the law run one contact per wave,
// all 64 lanes on one polytope in LDS, writing per-contact force/torque and
// two reduction records (row of A, row of B) per contact: either in a launch
// of their own over all contacts (k_epa_force), or in the GJK launch, each
// wave on the contacts it found (k_gjk_epa), which saves a launch and a
// grid-wide dependency per step but leaves the EPA work where the contacts
// fell.
#include <string.h>

#include <type_traits>

#include "mwd_internal.h"
#include "mwd_scan.h"

namespace {

constexpr int EPA_FACES = MWD_EPA_MAX_FACES;
// Grid cap of the grid-strided EPA launch of the per-pair entry point (one
// contact per wave): 2,048 blocks of 4 waves, about what the polytopes' LDS
// lets stay resident (7 blocks per CU); more contacts loop. The stepper's
// launch takes its cap from the tuning (a small grid costs less to dispatch
// when contacts are few).
constexpr int64_t EPA_MAX_BLOCKS = 2048;
constexpr int EPA_VERTS = 32;

struct Side {
    const double* v;  // vertices [n][3]
    int n;
};


// Support map (Support_funtion, .cu:55-101): first index of the strict maximum
// of d.a over A and of (-d).b over B; returns a-b.
__device__ __forceinline__ d3 support(const Side& A, const Side& B, d3 d, int* ia, int* ib) {
    double best = -1e30;
    int i1 = 0;
    for (int i = 0; i < A.n; ++i) {
        const double t = dot3(d, ld3(A.v + 3 * i));
        if (t > best) {
            best = t;
            i1 = i;
        }
    }
    const d3 nd = neg3(d);
    double best2 = -1e30;
    int i2 = 0;
    for (int i = 0; i < B.n; ++i) {
        const double t = dot3(nd, ld3(B.v + 3 * i));
        if (t > best2) {
            best2 = t;
            i2 = i;
        }
    }
    *ia = i1;
    *ib = i2;
    return ld3(A.v + 3 * i1) - ld3(B.v + 3 * i2);
}

struct SerialSup {
    static constexpr bool kQuad = false;  // one lane per pair
    __device__ __forceinline__ int quad_lane() const { return 0; }
    Side A, B;
    __device__ __forceinline__ d3 operator()(d3 d, int* ia, int* ib) const { return support(A, B, d, ia, ib); }
};

// The support map evaluated by a group of G lanes that share one pair: lane g
// of the group holds vertices g, g+G, ..., g+(K-1)G of both bodies in
// registers, takes the first strict maximum over its own, and the group
// merges (value, index) pairs keeping the larger value and, on a tie, the
// smaller index -- the serial loop's choice, bit for bit. The winners'
// coordinates come from their owner lanes by shuffle (no memory round trip).
template <int G, int K>
struct GroupSup {
    static constexpr bool kQuad = true;  // groups of 4, 8 or 16 lanes: whole quads
    __device__ __forceinline__ int quad_lane() const { return g & 3; }
    Side A, B;
    d3 va[K], vb[K];
    int g;
    __device__ __forceinline__ void load(const Side& a, const Side& b, int lane_in_group) {
        A = a;
        B = b;
        g = lane_in_group;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int i = g + G * k;
            va[k] = i < A.n ? ld3(A.v + 3 * i) : mk3(0.0, 0.0, 0.0);
            vb[k] = i < B.n ? ld3(B.v + 3 * i) : mk3(0.0, 0.0, 0.0);
        }
    }
    // One exchange of the merge: each lane meets the lane that the DPP
    // permutation CTRL names (within its quad, 8-lane half or 16-lane row) and
    // keeps the better (value, index).
    template <int CTRL>
    __device__ __forceinline__ static void meet(double& best, int& bi) {
        const int2 w = *reinterpret_cast<const int2*>(&best);
        int2 ow;
        ow.x = __builtin_amdgcn_update_dpp(0, w.x, CTRL, 0xF, 0xF, false);
        ow.y = __builtin_amdgcn_update_dpp(0, w.y, CTRL, 0xF, 0xF, false);
        const double ob = *reinterpret_cast<const double*>(&ow);
        const int oi = __builtin_amdgcn_update_dpp(0, bi, CTRL, 0xF, 0xF, false);
        if (ob > best || (ob == best && oi < bi)) {
            best = ob;
            bi = oi;
        }
    }
    // The group's larger value, and on a tie its smaller index, in every lane
    // of the group. The choice does not depend on the order of the exchanges,
    // so DPP permutations (register-to-register, no LDS round trip) pair the
    // lanes: swap neighbours, swap pairs (quad_perm), then mirror the 8-lane
    // halves (row_half_mirror) and the 16-lane rows (row_mirror).
    __device__ __forceinline__ static void merge(double& best, int& bi) {
        static_assert(G == 4 || G == 8 || G == 16, "groups of 4, 8 or 16 lanes");
        meet<0xB1>(best, bi);  // quad_perm [1,0,3,2]
        meet<0x4E>(best, bi);  // quad_perm [2,3,0,1]
        if constexpr (G >= 8) meet<0x141>(best, bi);  // row_half_mirror
        if constexpr (G >= 16) meet<0x140>(best, bi);  // row_mirror
    }
    __device__ __forceinline__ d3 operator()(d3 d, int* ia, int* ib) const {
        constexpr int NONE = 0x7fffffff;
        double best = -1e30, best2 = -1e30;
        int i1 = NONE, i2 = NONE;
        const d3 nd = neg3(d);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int i = g + G * k;
            if (i < A.n) {
                const double t = dot3(d, va[k]);
                if (t > best) {
                    best = t;
                    i1 = i;
                }
            }
            if (i < B.n) {
                const double t = dot3(nd, vb[k]);
                if (t > best2) {
                    best2 = t;
                    i2 = i;
                }
            }
        }
        merge(best, i1);
        merge(best2, i2);
        if (i1 == NONE) i1 = 0;
        if (i2 == NONE) i2 = 0;
        *ia = i1;
        *ib = i2;
        return pick(va, i1) - pick(vb, i2);
    }
    // vertex i of a set held across the group: slot i / G of lane i % G
    __device__ __forceinline__ static d3 pick(const d3* v, int i) {
        const int k = i / G, src = i - k * G;
        d3 mine = v[0];
#pragma unroll
        for (int j = 1; j < K; ++j)
            if (k == j) mine = v[j];
        return mk3(__shfl(mine.x, src, G), __shfl(mine.y, src, G), __shfl(mine.z, src, G));
    }
};

// The simplex tests below are pure functions of their points. Each one
// evaluates all of its sub-tests unconditionally and then picks the outcome
// in the order of the reference's early returns, which gives the same code
// and direction as running them one after another (a sub-test whose outcome
// is not picked has no effect), while the independent sub-tests interleave
// instead of forming one long chain of dependent FP64 operations.
struct SimplexOut {
    int c;  // -1 go on, 0 apart, 1 origin enclosed
    d3 d;   // the direction the reference leaves behind
};

// Check_linesegment, .cu:103-124
__device__ __forceinline__ SimplexOut check_segment(d3 P1, d3 P2) {
    const d3 V21 = P2 - P1;
    const bool apart = dot3(V21, P2) < 0.0;
    const bool on = zero3(cross3(V21, P2));
    const d3 Vt = cross3(V21, neg3(P2));
    const d3 dd = cross3(Vt, V21);
    SimplexOut o;
    o.c = apart ? 0 : (on ? 1 : -1);
    o.d = (apart || on) ? mk3(0.0, 0.0, 0.0) : dd;
    return o;
}

// Check_triangle, .cu:127-186
__device__ __forceinline__ SimplexOut check_triangle(d3 P1, d3 P2, d3 P3) {
    const SimplexOut s2 = check_segment(P2, P3);
    const SimplexOut s3 = check_segment(P1, P3);
    const d3 V32 = P3 - P2;
    const d3 V13 = P1 - P3;
    const d3 n = cross3(V32, V13);
    const d3 n2 = cross3(V32, n);
    const d3 n3 = cross3(V13, n);
    SimplexOut o;
    if (s2.c == 1) {
        o.c = 1;
        o.d = s2.d;
    } else if (s3.c == 1) {
        o.c = 1;
        o.d = s3.d;
    } else if (s2.c + s3.c > -1) {
        o.c = 0;
        o.d = s3.d;
    } else if (dot3(n2, P2) < 0.0) {
        o.c = -1;
        o.d = n2;
    } else if (dot3(n3, P3) < 0.0) {
        o.c = -1;
        o.d = n3;
    } else if (dot3(P3, n) == 0.0) {
        o.c = 1;
        o.d = s3.d;
    } else {
        o.c = -1;
        o.d = (dot3(n, P3) > 0.0) ? neg3(n) : n;
    }
    return o;
}

__device__ __forceinline__ void check_segment(d3 P1, d3 P2, int* coll, d3* d) {
    const SimplexOut o = check_segment(P1, P2);
    *coll = o.c;
    *d = o.d;
}

__device__ __forceinline__ void check_triangle(d3 P1, d3 P2, d3 P3, int* coll, d3* d) {
    const SimplexOut o = check_triangle(P1, P2, P3);
    *coll = o.c;
    *d = o.d;
}

// Check_Tetrahedron, .cu:188-254
__device__ __forceinline__ void check_tetra(d3 P1, d3 P2, d3 P3, d3 P4, int* coll, d3* d, int* todel) {
    const SimplexOut t1 = check_triangle(P1, P2, P4);
    const SimplexOut t2 = check_triangle(P2, P3, P4);
    const SimplexOut t3 = check_triangle(P3, P1, P4);
    const d3 V13 = P1 - P3, V21 = P2 - P1, V32 = P3 - P2, V41 = P4 - P1;
    const d3 n2 = cross3(V21, V41);
    const d3 V42 = P4 - P2;
    const d3 n3 = cross3(V32, V42);
    const d3 V43 = P4 - P3;
    const d3 n4 = cross3(V13, V43);
    *todel = 0;
    if (t1.c == 1) {
        *coll = 1;
        *d = t1.d;
    } else if (t2.c == 1) {
        *coll = 1;
        *d = t2.d;
    } else if (t3.c == 1) {
        *coll = 1;
        *d = t3.d;
    } else if (t1.c + t2.c + t3.c > -2) {
        *coll = 0;
        *d = t3.d;
    } else if (dot3(n2, P4) < 0.0) {
        *coll = -1;
        *d = n2;
        *todel = 3;
    } else if (dot3(n3, P4) < 0.0) {
        *coll = -1;
        *d = n3;
        *todel = 1;
    } else if (dot3(n4, P4) < 0.0) {
        *coll = -1;
        *d = n4;
        *todel = 2;
    } else {
        *coll = 1;
        *d = t3.d;
    }
}

// Check_Tetrahedron by the four lanes of a quad that share one pair (the
// grouped support map's lanes hold the same simplex): lane q evaluates
// triangle q of the three the reference tests -- (P1,P2,P4), (P2,P3,P4),
// (P3,P1,P4); lane 3 repeats the third -- and the quad exchanges the
// outcomes by DPP broadcasts, so each lane runs one triangle test instead of
// three. The outcome is check_tetra's, bit for bit (the same arithmetic on
// the same values, picked in the same order). Every lane of the quad must be
// active.
template <int J>
__device__ __forceinline__ int quad_bcast(int x) {
    return __builtin_amdgcn_update_dpp(0, x, J * 0x55, 0xF, 0xF, false);  // quad_perm [J,J,J,J]
}
template <int J>
__device__ __forceinline__ double quad_bcast(double x) {
    const int2 w = *reinterpret_cast<const int2*>(&x);
    int2 o;
    o.x = quad_bcast<J>(w.x);
    o.y = quad_bcast<J>(w.y);
    return *reinterpret_cast<const double*>(&o);
}
template <int J>
__device__ __forceinline__ SimplexOut quad_bcast(const SimplexOut& t) {
    SimplexOut o;
    o.c = quad_bcast<J>(t.c);
    o.d = mk3(quad_bcast<J>(t.d.x), quad_bcast<J>(t.d.y), quad_bcast<J>(t.d.z));
    return o;
}

__device__ __forceinline__ void check_tetra_quad(d3 P1, d3 P2, d3 P3, d3 P4, int q, int* coll, d3* d, int* todel) {
    const d3 X = q == 0 ? P1 : (q == 1 ? P2 : P3);
    const d3 Y = q == 0 ? P2 : (q == 1 ? P3 : P1);
    const SimplexOut mine = check_triangle(X, Y, P4);
    const SimplexOut t1 = quad_bcast<0>(mine);
    const SimplexOut t2 = quad_bcast<1>(mine);
    const SimplexOut t3 = quad_bcast<2>(mine);
    const d3 V13 = P1 - P3, V21 = P2 - P1, V32 = P3 - P2, V41 = P4 - P1;
    const d3 n2 = cross3(V21, V41);
    const d3 V42 = P4 - P2;
    const d3 n3 = cross3(V32, V42);
    const d3 V43 = P4 - P3;
    const d3 n4 = cross3(V13, V43);
    *todel = 0;
    if (t1.c == 1) {
        *coll = 1;
        *d = t1.d;
    } else if (t2.c == 1) {
        *coll = 1;
        *d = t2.d;
    } else if (t3.c == 1) {
        *coll = 1;
        *d = t3.d;
    } else if (t1.c + t2.c + t3.c > -2) {
        *coll = 0;
        *d = t3.d;
    } else if (dot3(n2, P4) < 0.0) {
        *coll = -1;
        *d = n2;
        *todel = 3;
    } else if (dot3(n3, P4) < 0.0) {
        *coll = -1;
        *d = n3;
        *todel = 1;
    } else if (dot3(n4, P4) < 0.0) {
        *coll = -1;
        *d = n4;
        *todel = 2;
    } else {
        *coll = 1;
        *d = t3.d;
    }
}

// CorrectFlat, .cu:256-268
__device__ __forceinline__ bool correct_flat(d3 S1, d3 S2, d3 S3, d3 S4) {
    const double det1 = S1.x * (S2.y * S3.z - S3.y * S2.z) - S2.x * (S1.y * S3.z - S3.y * S1.z) +
                        S3.x * (S1.y * S2.z - S2.y * S1.z);
    if (det1 < 0.000001 && det1 > -0.000001) {
        const double det2 = S1.x * (S2.y * S4.z - S4.y * S2.z) - S2.x * (S1.y * S4.z - S4.y * S1.z) +
                            S4.x * (S1.y * S2.z - S2.y * S1.z);
        if (det2 < 0.000001 && det2 > -0.000001) return true;
    }
    return false;
}

// The physical law's flatness test of a GJK tetrahedron: six times its
// volume at most 1e-9 of the cube of its longest edge (squared: v6^2 <=
// 1e-18 L2^3). It measures the tetrahedron, not where the origin lies (not
// in the reference).
__device__ __forceinline__ bool flat_tetra(d3 S0, d3 S1, d3 S2, d3 S3) {
    const d3 e1 = S1 - S0, e2 = S2 - S0, e3 = S3 - S0;
    const double v6 = dot3(cross3(e1, e2), e3);
    double L2 = nrm2(e1);
    const double c2 = nrm2(e2), c3 = nrm2(e3), c4 = nrm2(S2 - S1), c5 = nrm2(S3 - S1), c6 = nrm2(S3 - S2);
    if (c2 > L2) L2 = c2;
    if (c3 > L2) L2 = c3;
    if (c4 > L2) L2 = c4;
    if (c5 > L2) L2 = c5;
    if (c6 > L2) L2 = c6;
    return v6 * v6 <= 1e-18 * ((L2 * L2) * L2);
}

// Whether pair (ra, rb) takes the physical law's GJK/EPA variant.
__device__ __forceinline__ bool robust_pair(const BodyView& A, const BodyView& B, int64_t ra, int64_t rb) {
    if (!A.laws2) return false;
    const int wall = (A.is_wall[ra] || B.is_wall[rb]) ? 1 : 0;
    const int scene = A.scene ? A.scene[ra] : 0;
    return A.laws2[2 * scene + wall].mode == MWD_LAW_PHYSICAL;
}

// GJK of one pair (.cu:1009-1288). Returns -1 (tetrahedron cap), 0 or 1; on
// 1 the simplex handed to EPA (after CorrectFlat) is in S/ia/ib. robust (the
// physical law's pairs; not in the reference): instead of CorrectFlat, a flat
// tetrahedron gets its fourth vertex from the normal of the first three, on
// the side where the Minkowski difference reaches farther.
template <class Sup>
__device__ __forceinline__ int gjk_pair(const Sup& support, d3 comA, d3 comB, d3* S, int* ia, int* ib,
                                        bool robust) {
    // the simplex lives in named scalars (registers); the arrays are written
    // once, on a collision
    d3 S0, S1, S2, S3;
    int a0, a1, a2, a3, b0, b1, b2, b3;
    const d3 d1 = comB - comA;
    S0 = support(d1, &a0, &b0);
    if (dot3(d1, S0) < 0.0) return 0;
    const d3 d2 = neg3(d1);
    S1 = support(d2, &a1, &b1);
    if (dot3(d2, S1) < 0.0) return 0;
    int coll;
    d3 d3v;
    check_segment(S0, S1, &coll, &d3v);
    if (coll == 0) return 0;
    S2 = support(d3v, &a2, &b2);
    if (dot3(d3v, S2) < 0.0) return 0;
    d3 d4;
    check_triangle(S0, S1, S2, &coll, &d4);
    if (coll == 0) return 0;
    for (int zz = 0; zz < 50; ++zz) {
        S3 = support(d4, &a3, &b3);
        if (dot3(d4, S3) < 0.0) return 0;
        bool hit = zero3(S3);
        if (!hit) {
            int todel;
            if constexpr (Sup::kQuad)
                check_tetra_quad(S0, S1, S2, S3, support.quad_lane(), &coll, &d4, &todel);
            else
                check_tetra(S0, S1, S2, S3, &coll, &d4, &todel);
            if (coll == 0) return 0;
            if (coll == 1) {
                hit = true;
            } else {
                // keep the two vertices the reference keeps, then the new one
                d3 n1, n2;
                int x1, y1, x2, y2;
                if (todel == 1) {
                    n1 = S1; x1 = a1; y1 = b1;
                    n2 = S2; x2 = a2; y2 = b2;
                } else if (todel == 2) {
                    n1 = S2; x1 = a2; y1 = b2;
                    n2 = S0; x2 = a0; y2 = b0;
                } else {
                    n1 = S0; x1 = a0; y1 = b0;
                    n2 = S1; x2 = a1; y2 = b1;
                }
                S0 = n1; a0 = x1; b0 = y1;
                S1 = n2; a1 = x2; b1 = y2;
                S2 = S3; a2 = a3; b2 = b3;
            }
        }
        if (hit) {
            if (robust) {
                if (flat_tetra(S0, S1, S2, S3)) {
                    const d3 nT = cross3(S1 - S0, S2 - S0);
                    int ap, bp, am, bm;
                    const d3 Sp = support(nT, &ap, &bp);
                    const d3 Sm = support(neg3(nT), &am, &bm);
                    if (dot3(nT, Sp - S0) >= dot3(nT, S0 - Sm)) {
                        S3 = Sp; a3 = ap; b3 = bp;
                    } else {
                        S3 = Sm; a3 = am; b3 = bm;
                    }
                }
            } else if (correct_flat(S0, S1, S2, S3)) {
                d4 = cross3(S0, S1);
                S3 = support(d4, &a3, &b3);
            }
            S[0] = S0; S[1] = S1; S[2] = S2; S[3] = S3;
            ia[0] = a0; ia[1] = a1; ia[2] = a2; ia[3] = a3;
            ib[0] = b0; ib[1] = b1; ib[2] = b2; ib[3] = b3;
            return 1;
        }
    }
    return -1;
}

// Bounding-sphere test over all pairs. Pairs whose spheres are apart are
// decided here (the first support test of .cu:1013 rejects them); the rest
// are appended to surv with one atomic per wave, so GJK then runs on full
// waves of live pairs instead of waves that mostly idle.
__global__ void __launch_bounds__(1024) k_sphere(BodyView A, BodyView B, int64_t n_cap, const int64_t* n_dev,
                                                int32_t* __restrict__ collision, int32_t* __restrict__ surv,
                                                int64_t* n_surv) {
    int64_t n = n_dev ? *n_dev : n_cap;
    if (n > n_cap) n = n_cap;
    const int lane = threadIdx.x & (MWD_WAVE - 1);
    const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (MWD_WAVE - lane));
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    // p0 is uniform across a wave, so every lane of a wave reaches each ballot
    for (int64_t p0 = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x - lane); p0 < n; p0 += stride) {
        const int64_t p = p0 + lane;
        bool keep = false;
        if (p < n) {
            keep = true;
            if (A.radius) {
                const int64_t ra = view_row(A, p), rb = view_row(B, p);
                const d3 ca = ld3(A.com + 3 * ra), cb = ld3(B.com + 3 * rb);
                const double rr = A.radius[ra] + B.radius[rb];
                keep = !(nrm2(cb - ca) > rr * rr);
            }
            if (!keep) collision[p] = 0;
        }
        const uint64_t m = __ballot(keep);
        if (m) {
            int64_t base = 0;
            if (lane == 0) base = (int64_t)atomicAdd((unsigned long long*)n_surv, (unsigned long long)__popcll(m));
            base = __shfl(base, 0, MWD_WAVE);
            const int64_t at = base + __popcll(m & lt);
            if (keep && at < n_cap) surv[at] = (int32_t)p;
        }
    }
}

// Survivors [b, e) of a GJK launch: [*lo, *hi) (lo NULL: from 0), within
// the list's capacity.
__device__ __forceinline__ void surv_range(const int64_t* lo, const int64_t* hi, int64_t n_cap, int64_t* b,
                                           int64_t* e) {
    int64_t x = *hi;
    if (x > n_cap) x = n_cap;
    int64_t y = lo ? *lo : 0;
    if (y > x) y = x;
    *b = y;
    *e = x;
}

// GJK class of a survivor: ceil(max(nA, nB) / G) - 1 vertices-per-lane
// steps, the last class taking the rest.
__device__ __forceinline__ int gjk_class(const BodyView& A, const BodyView& B, int64_t p, int G, int nb) {
    const int32_t na = A.nverts[view_row(A, p)], nb_ = B.nverts[view_row(B, p)];
    const int v = na > nb_ ? na : nb_;
    int k = (v + G - 1) / G - 1;
    if (k < 0) k = 0;
    if (k > nb - 1) k = nb - 1;
    return k;
}

// Survivor counts per GJK class (LDS histogram, one global add per class and
// block). cnt[nb] zero on entry.
__global__ void __launch_bounds__(256) k_bucket_count(BodyView A, BodyView B, const int32_t* __restrict__ surv,
                                                      const int64_t* n_surv, int64_t n_cap, int G, int nb,
                                                      int64_t* cnt) {
    __shared__ int c[MWD_GJK_BUCKETS];
    if (threadIdx.x < MWD_GJK_BUCKETS) c[threadIdx.x] = 0;
    __syncthreads();
    int64_t n = *n_surv;
    if (n > n_cap) n = n_cap;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        atomicAdd(&c[gjk_class(A, B, surv[i], G, nb)], 1);
    __syncthreads();
    if ((int)threadIdx.x < nb && c[threadIdx.x]) atomicAdd((unsigned long long*)&cnt[threadIdx.x], (unsigned long long)c[threadIdx.x]);
}

// Survivors written class after class into out (any order within a class):
// each block ranks its chunk's members per class in LDS and reserves one
// range per class with a global add on cur[nb] (zero on entry). ofs[nb + 1]
// receives the classes' start offsets (block 0).
__global__ void __launch_bounds__(256) k_bucket_scatter(BodyView A, BodyView B, const int32_t* __restrict__ surv,
                                                        const int64_t* n_surv, int64_t n_cap, int G, int nb,
                                                        const int64_t* cnt, int64_t* cur, int64_t* ofs,
                                                        int32_t* __restrict__ out) {
    __shared__ int64_t off[MWD_GJK_BUCKETS + 1];
    __shared__ int64_t base[MWD_GJK_BUCKETS];
    __shared__ int lc[MWD_GJK_BUCKETS];
    if (threadIdx.x == 0) {
        int64_t a = 0;
        for (int k = 0; k < nb; ++k) {
            off[k] = a;
            a += cnt[k];
        }
        off[nb] = a;
        if (blockIdx.x == 0)
            for (int k = 0; k <= nb; ++k) ofs[k] = off[k];
    }
    int64_t n = *n_surv;
    if (n > n_cap) n = n_cap;
    const int tid = threadIdx.x;
    // a uniform trip count over the block (its barriers)
    for (int64_t s0 = (int64_t)blockIdx.x * blockDim.x; s0 < n; s0 += (int64_t)gridDim.x * blockDim.x) {
        if (tid < MWD_GJK_BUCKETS) lc[tid] = 0;
        __syncthreads();
        int k = -1, r = 0;
        int32_t p = 0;
        if (s0 + tid < n) {
            p = surv[s0 + tid];
            k = gjk_class(A, B, p, G, nb);
            r = atomicAdd(&lc[k], 1);
        }
        __syncthreads();
        if (tid < nb) base[tid] = lc[tid] ? off[tid] + (int64_t)atomicAdd((unsigned long long*)&cur[tid], (unsigned long long)lc[tid]) : 0;
        __syncthreads();
        if (k >= 0) out[base[k] + r] = p;
        __syncthreads();
    }
}

// GJK over the survivors (.cu:983-1322, without EPA), G lanes per pair
// (GroupSup); every lane of a group follows the same path on the same
// values, and its first lane writes the results. K = 0 selects one lane per
// pair with the vertices read from memory (bodies of more than 6G vertices).
// GJK of candidate pair p by lane g of its group (K = 0: one lane, the
// vertices read from memory).
template <int G, int K>
__device__ __forceinline__ int gjk_of(const BodyView& A, const BodyView& B, int64_t p, int g, d3* S, int* ia,
                                      int* ib) {
    const int64_t ra = view_row(A, p), rb = view_row(B, p);
    const d3 ca = ld3(A.com + 3 * ra), cb = ld3(B.com + 3 * rb);
    const Side sa{A.verts + (int64_t)A.vmax * 3 * ra, A.nverts[ra]};
    const Side sb{B.verts + (int64_t)B.vmax * 3 * rb, B.nverts[rb]};
    const bool robust = robust_pair(A, B, ra, rb);
    if constexpr (K > 0) {
        GroupSup<G, K> sup;
        sup.load(sa, sb, g);
        return gjk_pair(sup, ca, cb, S, ia, ib, robust);
    } else {
        return gjk_pair(SerialSup{sa, sb}, ca, cb, S, ia, ib, robust);
    }
}

template <int G, int K>
__global__ void __launch_bounds__(256) k_gjk(BodyView A, BodyView B, int64_t n_cap, const int32_t* __restrict__ surv,
                                             const int64_t* lo, const int64_t* n_surv,
                                             int32_t* __restrict__ collision, Simplex* __restrict__ simplex,
                                             int32_t* __restrict__ sidx, int64_t* slots, int64_t slot_cap,
                                             uint32_t* flags, int32_t* __restrict__ contact_pair) {
    int64_t b0, m;
    surv_range(lo, n_surv, n_cap, &b0, &m);
    constexpr int GG = K > 0 ? G : 1;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int g = (int)(t % GG);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x / GG;
    for (int64_t s = b0 + t / GG; s < m; s += stride) {
        const int64_t p = surv[s];
        d3 S[4];
        int ia[4] = {0, 0, 0, 0}, ib[4] = {0, 0, 0, 0};
        const int c = gjk_of<G, K>(A, B, p, g, S, ia, ib);
        if (g != 0) continue;
        collision[p] = c;
        if (c == 1) {
            const int64_t slot = atomicAdd((unsigned long long*)slots, 1ull);
            sidx[p] = (int32_t)slot;
            if (slot >= slot_cap) {
                atomicOr(flags, MWD_F_CONTACT_OVERFLOW);
                continue;
            }
            if (contact_pair) contact_pair[slot] = (int32_t)p;  // contacts in slot order
            Simplex& x = simplex[slot];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                x.s[k][0] = S[k].x;
                x.s[k][1] = S[k].y;
                x.s[k][2] = S[k].z;
                x.ia[k] = ia[k];
                x.ib[k] = ib[k];
            }
        }
    }
}

// Contact compaction: contact c is the c-th pair with collision == 1, in pair
// order; a count beyond contact_cap is flagged.
struct ContactIO {
    const int32_t* collision;
    int32_t* contact_pair;
    int64_t cap;
    uint32_t* flags;
    __device__ __forceinline__ int32_t load(int64_t p) const { return collision[p] == 1 ? 1 : 0; }
    __device__ __forceinline__ void store(int64_t p, int64_t c, int32_t v) const {
        if (v && c < cap) contact_pair[c] = (int32_t)p;
    }
    __device__ __forceinline__ void finish(int64_t total) const {
        if (total > cap) atomicOr(flags, MWD_F_CONTACT_OVERFLOW);
    }
};

// ------------------------------------------------------------------- EPA
struct EpaResult {
    double depth;
    d3 n, cpA, cpB;
    int status;
    // the final face: support indices of its vertices in stored order, and
    // whether AddTriangle swapped its last two vertices
    int fa[3], fb[3], flip;
};

// ------------------------------------------------- EPA, one wave per contact
// EPA (.cu:328-603) run by the 64 lanes of a wave on a polytope held in LDS
// (structure of arrays, 5.7 KB per wave): every loop of the reference over
// faces or vertices is spread over the lanes, and every sequential choice is
// reproduced exactly. The closest face is the first index of the minimum
// (the reference's strict-less scan from face 0); the support point is the
// first strict maximum (support()); faces that see the new point are listed
// in index order; the horizon edges (.cu:470-540) are taken in (visible face,
// edge) order and each new face gets its slot by a prefix count over that
// order, so faces land where the reference's sequential loop puts them.
// status 1: expansion cap; 2: face storage exhausted.
static_assert(EPA_FACES == 2 * MWD_WAVE, "a lane holds faces lane and lane + 64");
struct PolyW {
    double vx[EPA_VERTS], vy[EPA_VERTS], vz[EPA_VERTS];
    double nx[EPA_FACES], ny[EPA_FACES], nz[EPA_FACES];
    double dist[EPA_FACES];
    uint8_t t0[EPA_FACES], t1[EPA_FACES], t2[EPA_FACES];
    int8_t valid[EPA_FACES];
    uint8_t flip[EPA_FACES];
    uint8_t vis[EPA_FACES];
    uint8_t opA[EPA_VERTS], opB[EPA_VERTS];
};

// Orders the LDS accesses of one wave (its lanes run in lockstep).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ d3 pw_v(const PolyW& P, int k) { return mk3(P.vx[k], P.vy[k], P.vz[k]); }
__device__ __forceinline__ d3 pw_n(const PolyW& P, int k) { return mk3(P.nx[k], P.ny[k], P.nz[k]); }

// AddTriangle (.cu:270-325) into slot id. robust (the physical law's EPA):
// the face is turned away from the interior point `inner` instead of away
// from the origin, and dd keeps its sign (negative: the origin is outside).
__device__ __forceinline__ void add_triangle_w(PolyW& P, int e0, int e1, int e2, int id, bool robust, d3 inner) {
    const d3 V0 = pw_v(P, e0), V1 = pw_v(P, e1), V2 = pw_v(P, e2);
    d3 nn = cross3(V1 - V0, V2 - V1);
    const double len = sqrt(nrm2(nn));
    nn = nn / len;
    double dd = dot3(nn, V0);
    uint8_t fl = 0;
    if (robust ? dot3(nn, V0 - inner) < 0.0 : dd < 0.0) {
        const int o = e2;
        e2 = e1;
        e1 = o;
        dd = dd * -1.0;
        nn = nn * -1.0;
        fl = 1;
    }
    P.flip[id] = fl;
    P.t0[id] = (uint8_t)e0;
    P.t1[id] = (uint8_t)e1;
    P.t2[id] = (uint8_t)e2;
    const d3 a = pw_v(P, e0), b = pw_v(P, e1), c = pw_v(P, e2);
    const bool inside = dot3(cross3(b - a, nn), neg3(a)) < 0.0 && dot3(cross3(c - b, nn), neg3(b)) < 0.0 &&
                        dot3(cross3(a - c, nn), neg3(c)) < 0.0;
    P.valid[id] = inside ? 1 : -2;
    if (!inside) dd = 1e30;
    P.nx[id] = nn.x;
    P.ny[id] = nn.y;
    P.nz[id] = nn.z;
    P.dist[id] = dd;
}

// Wave-wide min / max: exchanges within 16-lane rows by DPP permutes
// (register to register), then across rows by shuffles. min and max do not
// depend on the order of the exchanges.
template <int CTRL>
__device__ __forceinline__ double dpp_d(double x) {
    const int2 w = *reinterpret_cast<const int2*>(&x);
    int2 o;
    o.x = __builtin_amdgcn_update_dpp(0, w.x, CTRL, 0xF, 0xF, false);
    o.y = __builtin_amdgcn_update_dpp(0, w.y, CTRL, 0xF, 0xF, false);
    return *reinterpret_cast<const double*>(&o);
}
template <int CTRL>
__device__ __forceinline__ int dpp_i(int x) {
    return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xF, 0xF, false);
}
template <class T, class Pick, class Dpp>
__device__ __forceinline__ T wave_reduce(T x, Pick pick, Dpp) {
    x = pick(x, Dpp::template get<0xB1>(x));   // quad_perm [1,0,3,2]
    x = pick(x, Dpp::template get<0x4E>(x));   // quad_perm [2,3,0,1]
    x = pick(x, Dpp::template get<0x141>(x));  // row_half_mirror
    x = pick(x, Dpp::template get<0x140>(x));  // row_mirror
    x = pick(x, __shfl_xor(x, 16, MWD_WAVE));
    x = pick(x, __shfl_xor(x, 32, MWD_WAVE));
    return x;
}
struct DppD {
    template <int CTRL>
    __device__ static double get(double x) { return dpp_d<CTRL>(x); }
};
struct DppI {
    template <int CTRL>
    __device__ static int get(int x) { return dpp_i<CTRL>(x); }
};
__device__ __forceinline__ double wave_min(double x) {
    return wave_reduce(x, [](double a, double b) { return b < a ? b : a; }, DppD{});
}
__device__ __forceinline__ double wave_max(double x) {
    return wave_reduce(x, [](double a, double b) { return b > a ? b : a; }, DppD{});
}
__device__ __forceinline__ int wave_min_i(int x) {
    return wave_reduce(x, [](int a, int b) { return b < a ? b : a; }, DppI{});
}

// Index of the first strict maximum of d.v over a body (support(): best
// starts at -1e30, index 0 when nothing beats it).
__device__ __forceinline__ int support_index_w(const Side& S, d3 d, int lane) {
    double best = -1e30;
    int bi = 0x7fffffff;
    for (int i = lane; i < S.n; i += MWD_WAVE) {
        const double t = dot3(d, ld3(S.v + 3 * i));
        if (t > best) {
            best = t;
            bi = i;
        }
    }
    const double m = wave_max(best);
    if (!(m > -1e30)) return 0;
    return wave_min_i(best == m ? bi : 0x7fffffff);
}

// First index of the minimum distance over faces [0, nt) (array_min_finder,
// .cu:41-52: a strict-less scan from face 0, so a NaN never wins unless it is
// face 0's and so the starting value).
__device__ __forceinline__ int closest_face_w(const PolyW& P, int nt, int lane) {
    const double inf = __builtin_inf();
    double x0 = inf, x1 = inf;
    if (lane < nt) {
        x0 = P.dist[lane];
        if (x0 != x0) x0 = inf;
    }
    if (lane + MWD_WAVE < nt) {
        x1 = P.dist[lane + MWD_WAVE];
        if (x1 != x1) x1 = inf;
    }
    const double d0 = P.dist[0];
    if (d0 != d0) return 0;
    const double m = wave_min(x0 < x1 ? x0 : x1);
    const uint64_t lo = __ballot(lane < nt && x0 == m);
    if (lo) return __ffsll((long long)lo) - 1;
    const uint64_t hi = __ballot(lane + MWD_WAVE < nt && x1 == m);
    return MWD_WAVE + __ffsll((long long)hi) - 1;
}

// robust (the physical law; not in the reference): faces are turned away from
// the starting tetrahedron's centroid, an interior point of every polytope
// EPA grows from it, so a face whose plane passes through or beside the
// origin (a contact of near-zero depth) keeps its outward normal and signed
// distance; the expansion also stops when the support point gains at most
// 1e-5 |d| + 1e-10 R over the closest face (R: the largest distance of a
// starting vertex from the origin).
__device__ EpaResult epa_wave(const Simplex& sx, const Side& A, const Side& B, PolyW& P, int lane, bool robust) {
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (MWD_WAVE - lane));
    double xmin_old = 1e29;
    double proj = 2.0 * xmin_old + 1.0;
    int counter = 0, nv = 4, nt = 0, indmin = 0;
    double xmin_new = 0.0;
    int st = 0;
    if (lane < 4) {
        P.vx[lane] = sx.s[lane][0];
        P.vy[lane] = sx.s[lane][1];
        P.vz[lane] = sx.s[lane][2];
        P.opA[lane] = (uint8_t)sx.ia[lane];
        P.opB[lane] = (uint8_t)sx.ib[lane];
    }
    const d3 s0 = mk3(sx.s[0][0], sx.s[0][1], sx.s[0][2]), s1 = mk3(sx.s[1][0], sx.s[1][1], sx.s[1][2]);
    const d3 s2 = mk3(sx.s[2][0], sx.s[2][1], sx.s[2][2]), s3 = mk3(sx.s[3][0], sx.s[3][1], sx.s[3][2]);
    const d3 inner = (((s0 + s1) + s2) + s3) * 0.25;
    double R2 = nrm2(s0);
    if (nrm2(s1) > R2) R2 = nrm2(s1);
    if (nrm2(s2) > R2) R2 = nrm2(s2);
    if (nrm2(s3) > R2) R2 = nrm2(s3);
    const double tol_abs = 1e-10 * sqrt(R2);
    wave_sync();
    if (lane < 4) {
        const int e0 = lane == 2 ? 2 : (lane == 1 ? 1 : 0);
        const int e1 = lane == 0 ? 1 : (lane == 1 ? 2 : (lane == 2 ? 0 : 1));
        const int e2 = lane == 0 ? 2 : 3;
        add_triangle_w(P, e0, e1, e2, lane, robust, inner);
    }
    nt = 4;
    wave_sync();
    while (xmin_old < proj) {
        indmin = closest_face_w(P, nt, lane);
        xmin_new = P.dist[indmin];
        xmin_old = xmin_new;
        const d3 nrm = pw_n(P, indmin);
        const int a = support_index_w(A, nrm, lane);
        const int b = support_index_w(B, neg3(nrm), lane);
        const d3 Suj = ld3(A.v + 3 * a) - ld3(B.v + 3 * b);
        proj = dot3(nrm, Suj);
        if (lane == 0) {
            P.opA[nv] = (uint8_t)a;
            P.opB[nv] = (uint8_t)b;
        }
        if (xmin_old > proj || (xmin_old * 1.00001 > proj && xmin_old * 0.99999 < proj)) break;
        if (robust && proj - xmin_old <= 1e-5 * fabs(xmin_old) + tol_abs) break;
        if (counter > 49) {
            st |= 1;
            break;
        }
        counter = counter + 1;
        if (lane == 0) {
            P.vx[nv] = Suj.x;
            P.vy[nv] = Suj.y;
            P.vz[nv] = Suj.z;
        }
        nv = nv + 1;
        wave_sync();
        // faces that see the new point, listed in index order
        int nvis = 0;
#pragma unroll
        for (int h = 0; h < EPA_FACES / MWD_WAVE; ++h) {
            const int i = h * MWD_WAVE + lane;
            bool see = false;
            if (i < nt && (P.valid[i] == 1 || P.valid[i] == -2)) {
                const int t = P.t0[i];
                see = dot3(Suj - pw_v(P, t), pw_n(P, i)) >= 0.0;
            }
            const uint64_t m = __ballot(see);
            if (see) {
                P.valid[i] = -1;
                P.vis[nvis + __popcll(m & lt)] = (uint8_t)i;
            }
            nvis += __popcll(m);
        }
        wave_sync();
        // horizon: an edge of a visible face that no other visible face holds
        // (.cu:470-540), in (visible face, edge) order
        const int nt0 = nt;
        int added = 0;
        for (int base = 0; base < 3 * nvis; base += MWD_WAVE) {
            const int q = base + lane;
            bool isnew = false;
            int u = 0, v = 0;
            if (q < 3 * nvis) {
                const int vi = q / 3, e = q - 3 * vi;
                const int i = P.vis[vi];
                const int x[3] = {P.t0[i], P.t1[i], P.t2[i]};
                u = x[e];
                v = x[e == 2 ? 0 : e + 1];
                int cnt = 0;
                for (int vj = 0; vj < nvis; ++vj) {
                    const int j = P.vis[vj];
                    const int x0 = P.t0[j], x1 = P.t1[j], x2 = P.t2[j];
                    cnt += ((x0 == u && x1 == v) || (x0 == v && x1 == u));
                    cnt += ((x1 == u && x2 == v) || (x1 == v && x2 == u));
                    cnt += ((x2 == u && x0 == v) || (x2 == v && x0 == u));
                }
                isnew = cnt < 2;
            }
            const uint64_t m = __ballot(isnew);
            const int id = nt0 + added + __popcll(m & lt);
            if (isnew && id < EPA_FACES) add_triangle_w(P, u, v, nv - 1, id, robust, inner);
            added += __popcll(m);
        }
        wave_sync();
        if (added > EPA_FACES - nt0) {  // the sequential loop stops at a full face store
            st |= 2;
            break;
        }
        nt = nt0 + added;
        for (int vi = lane; vi < nvis; vi += MWD_WAVE) {
            const int i = P.vis[vi];
            P.valid[i] = 0;
            P.dist[i] = 3e20;
        }
        counter = counter + 1;
        wave_sync();
    }
    wave_sync();
    EpaResult r;
    r.depth = xmin_new;
    const int a = P.t0[indmin], b = P.t1[indmin], c = P.t2[indmin];
    const d3 nrm = pw_n(P, indmin);
    const d3 Proj = nrm * xmin_new;
    const d3 Aa = pw_v(P, a), Bb = pw_v(P, b), Cc = pw_v(P, c);
    d3 v0 = Bb - Aa, v1 = Cc - Aa;
    const d3 v2 = Proj - Aa;
    if (dot3(cross3(v0, v1), nrm) < 0.0) {
        v0 = Cc - Aa;
        v1 = Bb - Aa;
    }
    const double v00 = dot3(v0, v0), v01 = dot3(v0, v1), v11 = dot3(v1, v1), v20 = dot3(v2, v0),
                 v21 = dot3(v2, v1);
    const double denom = v00 * v11 - v01 * v01;
    const double vv = (v11 * v20 - v01 * v21) / denom;
    const double ww = (v00 * v21 - v01 * v20) / denom;
    const double uu = 1.0 - ww - vv;
    r.cpA = (ld3(A.v + 3 * P.opA[a]) * uu + ld3(A.v + 3 * P.opA[b]) * vv) + ld3(A.v + 3 * P.opA[c]) * ww;
    r.cpB = (ld3(B.v + 3 * P.opB[a]) * uu + ld3(B.v + 3 * P.opB[b]) * vv) + ld3(B.v + 3 * P.opB[c]) * ww;
    r.n = nrm;
    r.status = st;
    r.fa[0] = P.opA[a];
    r.fa[1] = P.opA[b];
    r.fa[2] = P.opA[c];
    r.fb[0] = P.opB[a];
    r.fb[1] = P.opB[b];
    r.fb[2] = P.opB[c];
    r.flip = P.flip[indmin];
    wave_sync();  // the polytope is rewritten by the wave's next contact
    return r;
}

// Depth, normal and contact points of EPA's final face, recomputed from the
// face's support pairs with exactly EPA's operations (AddTriangle, .cu:278-291,
// and the contact construction, .cu:548-601): the values are EPA's, and with
// T = Dual they carry the derivatives with respect to both bodies' vertices.
template <class T>
struct FaceGeo {
    T depth;
    V3<T> n, cpA, cpB;
};
template <class T>
__device__ FaceGeo<T> face_geometry(const TArr<MWD_NT>& VA, int64_t ba, const TArr<MWD_NT>& VB, int64_t bb,
                                    const int* fa, const int* fb, int flip) {
    V3<T> a[3], b[3], S[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        a[k] = ldv<T>(VA, ba + fa[k]);
        b[k] = ldv<T>(VB, bb + fb[k]);
        S[k] = a[k] - b[k];
    }
    // creation order: a swap in AddTriangle exchanged the stored vertices 1 and 2
    const int e1 = flip ? 2 : 1, e2 = flip ? 1 : 2;
    V3<T> nn = cross3(S[e1] - S[0], S[e2] - S[e1]);
    const T len = dsqrt(nrm2(nn));
    nn = nn / len;
    T dd = dot3(nn, S[0]);
    if (flip) {
        dd = dd * -1.0;
        nn = nn * -1.0;
    }
    FaceGeo<T> g;
    g.depth = dd;
    g.n = nn;
    const V3<T> Proj = nn * dd;
    V3<T> v0 = S[1] - S[0], v1 = S[2] - S[0];
    const V3<T> v2 = Proj - S[0];
    if (dot3(cross3(v0, v1), nn) < 0.0) {
        v0 = S[2] - S[0];
        v1 = S[1] - S[0];
    }
    const T v00 = dot3(v0, v0), v01 = dot3(v0, v1), v11 = dot3(v1, v1), v20 = dot3(v2, v0), v21 = dot3(v2, v1);
    const T denom = v00 * v11 - v01 * v01;
    const T vv = (v11 * v20 - v01 * v21) / denom;
    const T ww = (v00 * v21 - v01 * v20) / denom;
    const T uu = (1.0 - ww) - vv;
    g.cpA = (a[0] * uu + a[1] * vv) + a[2] * ww;
    g.cpB = (b[0] * uu + b[1] * vv) + b[2] * ww;
    return g;
}

template <class T>
struct LawOut {
    V3<T> Fn, tA, tB, disp;
};

// The law coefficients a contact needs, as values (T = double) or with their
// derivatives with respect to (Kn, en, mu) (T = Dual).
template <class T>
struct LawT {
    T Kn, mu, Gn, Kt_num, inv_mu, ln_et, cn, ct, ratio;
    double dt;
    int mode;
};
__device__ __forceinline__ LawT<double> law_of(const mwd_law& L, const mwd_law_tan*, double*) {
    LawT<double> o;
    o.Kn = L.Kn; o.mu = L.mu; o.Gn = L.Gn; o.Kt_num = L.Kt_num; o.inv_mu = L.inv_mu; o.ln_et = L.ln_et;
    o.cn = L.cn; o.ct = L.ct; o.ratio = L.ratio; o.dt = L.dt; o.mode = L.mode;
    return o;
}
__device__ __forceinline__ Dual<MWD_NT> dual_of(double v, const double* d) {
    Dual<MWD_NT> r;
    r.v = v;
#pragma unroll
    for (int k = 0; k < MWD_NT; ++k) r.d[k] = d ? d[k] : 0.0;
    return r;
}
__device__ __forceinline__ LawT<Dual<MWD_NT>> law_of(const mwd_law& L, const mwd_law_tan* t, Dual<MWD_NT>*) {
    LawT<Dual<MWD_NT>> o;
    o.Kn = dual_of(L.Kn, t ? t->Kn : nullptr);
    o.mu = dual_of(L.mu, t ? t->mu : nullptr);
    o.Gn = dual_of(L.Gn, t ? t->Gn : nullptr);
    o.Kt_num = dual_of(L.Kt_num, t ? t->Kt_num : nullptr);
    o.inv_mu = dual_of(L.inv_mu, t ? t->inv_mu : nullptr);
    o.ln_et = dual_of(L.ln_et, t ? t->ln_et : nullptr);
    o.cn = dual_of(L.cn, t ? t->cn : nullptr);
    o.ct = dual_of(L.ct, t ? t->ct : nullptr);
    o.ratio = dual_of(L.ratio, t ? t->ratio : nullptr);
    o.dt = L.dt;
    o.mode = L.mode;
    return o;
}

// Force, .cu:647-871. Fn of the result is the force on A's linear momentum:
// the normal force (reference law) or normal plus tangential (physical law).
// imA/imB are inverse masses (0 for walls), read by the physical law only.
template <class T>
__device__ LawOut<T> force_law(V3<T> comA, V3<T> comB, V3<T> vA, V3<T> vB, V3<T> wA, V3<T> wB, V3<T> dispt, Q4<T> qA,
                               double isA, double isB, double imA, double imB, T depth, V3<T> n, V3<T> cpA, V3<T> cpB,
                               const LawT<T>& L) {
    LawOut<T> o;
    const bool phys = L.mode == MWD_LAW_PHYSICAL;
    if (phys && !(depth < EPA_PLACEHOLDER && depth > 0.0)) {
        // EPA ended on a face whose distance is its "outside" or "deleted"
        // placeholder (1e30 / 3e20): no depth was resolved; or the signed
        // depth is not positive (the bodies touch without overlapping): the
        // physical law applies no force for the step (the reference law
        // takes the value)
        o.Fn = o.tA = o.tB = o.disp = zero3<T>();
        return o;
    }
    const V3<T> cp = phys ? (cpA + cpB) / 2.0 : (cpA - cpB) / 2.0;
    const V3<T> dA = cp - comA, dB = cp - comB;
    const T dAn = nrm2(dA), dBn = nrm2(dB);
    T Kt, Gt, Gn;
    if (!phys) {
        const T C = (L.inv_mu + dAn / isA) + dBn / isB;
        Kt = L.Kt_num / C;
        Gt = (-2.0 / C) * L.ln_et;
        Gn = L.Gn;
    } else {
        const double imn = imA + imB;
        const double mn = 1.0 / imn;
        const T imt = (T(imn) + dAn * (imA == 0.0 ? 0.0 : 1.0 / isA)) + dBn * (imB == 0.0 ? 0.0 : 1.0 / isB);
        const T mt = 1.0 / imt;
        Gn = L.cn * dsqrt(mn * L.Kn);
        Kt = (L.Kn * L.ratio) * (mt / mn);
        Gt = L.ct * dsqrt(mt * Kt);
    }
    if (dot3(comB - comA, n) < 0.0) n = n * -1.0;
    n = n / dsqrt(nrm2(n));
    const V3<T> disp_n = n * depth;
    const V3<T> u = ((vA + cross3(wA, cpA - comA)) - vB) - cross3(wB, cpB - comB);
    const V3<T> un = n * dot3(u, n);
    const V3<T> ut = u - un;
    const T utn = dsqrt(nrm2(ut));
    const T dtn = dsqrt(nrm2(dispt));
    V3<T> t = zero3<T>();
    if (utn != 0.0) t = ut / utn;
    if (utn == 0.0 && dtn != 0.0) t = dispt / dtn;
    const V3<T> Fn = (disp_n * -1.0) * L.Kn - un * Gn;
    const T Fnn = dsqrt(nrm2(Fn));
    V3<T> Ft = (dispt * -1.0) * Kt - ut * Gt;
    const T Ftn = dsqrt(nrm2(Ft));
    const T muF = L.mu * Fnn;
    V3<T> dnew;
    if (Ftn > muF) Ft = ((t * Fnn) * -1.0) * L.mu;
    if (Ftn <= muF) {
        dnew = qrot(qA, dispt, qconj(qA)) + ut * L.dt;
    } else if (!phys) {
        // .cu:800-806 as written: the slip displacement makes the next step's
        // spring force point along the sliding (the sign of the mu term)
        dnew = ((((t * Fnn) * L.mu) * -1.0) - ut * Gt) / Kt;
    } else {
        // physical law: the displacement whose spring-dashpot force is the
        // Coulomb limit opposing the sliding, -Kt d - Gt ut = -mu |Fn| t
        dnew = (((t * Fnn) * L.mu) - ut * Gt) / Kt;
    }
    const V3<T> Ff = Ft + Fn;
    o.Fn = phys ? Ff : Fn;
    o.tA = cross3(cp - comA, Ff);
    o.tB = cross3(cp - comB, Ff * -1.0);
    o.disp = dnew;
    return o;
}

// Tangential history table: multiplicative hash, linear probing.
__device__ __forceinline__ int64_t hist_key(int64_t a, int64_t b) {
    const int64_t lo = a < b ? a : b, hi = a < b ? b : a;
    return lo + hi * hi;  // MWDiem.py:872-880
}
__device__ __forceinline__ int64_t hist_slot(int64_t key, int64_t mask) {
    return (int64_t)(((uint64_t)key * 0x9E3779B97F4A7C15ull) >> 17) & mask;
}
__device__ __forceinline__ d3 hist_find(const HistView& h, int t, int64_t key) {
    for (int64_t i = hist_slot(key, h.mask), n = 0; n <= h.mask; i = (i + 1) & h.mask, ++n) {
        const int64_t k = h.keys[t][i];
        if (k == key) return ld3(h.val[t] + 4 * i);
        if (k < 0) break;
    }
    return mk3(0.0, 0.0, 0.0);
}
__device__ __forceinline__ void hist_put(const HistView& h, int t, int64_t key, const double* v4) {
    for (int64_t i = hist_slot(key, h.mask), n = 0; n <= h.mask; i = (i + 1) & h.mask, ++n) {
        const unsigned long long prev =
            atomicCAS((unsigned long long*)&h.keys[t][i], ~0ull, (unsigned long long)key);
        if (prev == ~0ull) {  // claimed
            for (int k = 0; k < 4; ++k) h.val[t][4 * i + k] = v4[k];
            const int64_t u = (int64_t)atomicAdd((unsigned long long*)h.count[t], 1ull);
            h.used[t][u] = (int32_t)i;
            return;
        }
        if ((int64_t)prev == key) return;  // the pair listed twice (a one-cell-wide grid): first one kept
    }
}

__device__ __forceinline__ uint32_t push_record(int32_t* rec_cnt, int32_t* rec_bkt, int32_t* rec_ovf, int64_t* n_ovf,
                                               int64_t ovf_cap, int64_t row, int32_t rec) {
    const int32_t k = atomicAdd(&rec_cnt[row], 1);
    if (k < MWD_ROW_CONTACTS) {
        rec_bkt[row * MWD_ROW_CONTACTS + k] = rec;
        return 0;
    }
    const int64_t o = (int64_t)atomicAdd((unsigned long long*)n_ovf, 1ull);
    if (o >= ovf_cap) return MWD_F_ROW_CONTACTS;
    rec_ovf[2 * o] = (int32_t)row;
    rec_ovf[2 * o + 1] = rec;
    return 0;
}

template <class T>
__device__ __forceinline__ V3<T> body3(const double* p, const TArr<MWD_NT>& t, int64_t r) {
    if constexpr (std::is_same<T, double>::value)
        return ld3(p + 3 * r);
    else
        return ldv<T>(t, r);
}
template <class T>
__device__ __forceinline__ Q4<T> bodyq(const double* p, const TArr<MWD_NT>& t, int64_t r) {
    if constexpr (std::is_same<T, double>::value)
        return ldq(p + 4 * r);
    else
        return ldqT<T>(t, r);
}

// What a contact needs besides its pair and simplex: the law tables, the
// inputs of the pair layout, and where the results go.
struct ContactSink {
    const mwd_law* laws2;
    const mwd_law_tan* lawt2;
    const double* dispt;
    const int32_t* idA;
    const int32_t* idB;
    double* depth;
    double* normal;
    double* cpA;
    double* cpB;
    ContactOut co;
    double* disp_out;
    int32_t* epa_status;
    uint32_t* rec_key;
    int32_t* rec_val;
    int32_t* rec_cnt;
    int32_t* rec_bkt;
    int32_t* rec_ovf;
    int64_t* n_ovf;
    int64_t ovf_cap;
};

// Contact c (pair p, GJK simplex sx): EPA by all 64 lanes of the wave, then
// the law and the writes by its first lane.
template <class T>
__device__ __forceinline__ void contact_wave(int64_t c, int64_t p, const Simplex& sx, const BodyView& A,
                                             const BodyView& B, const BodyTan& bt, const ContactSink& k, PolyW& P,
                                             int lane, uint32_t& fl) {
    const mwd_law* laws2 = k.laws2;
    const mwd_law_tan* lawt2 = k.lawt2;
    const double* dispt = k.dispt;
    const int32_t* idA = k.idA;
    const int32_t* idB = k.idB;
    double* depth = k.depth;
    double* normal = k.normal;
    double* cpA = k.cpA;
    double* cpB = k.cpB;
    const ContactOut& co = k.co;
    double* disp_out = k.disp_out;
    int32_t* epa_status = k.epa_status;
    uint32_t* rec_key = k.rec_key;
    int32_t* rec_val = k.rec_val;
    int32_t* rec_cnt = k.rec_cnt;
    int32_t* rec_bkt = k.rec_bkt;
    int32_t* rec_ovf = k.rec_ovf;
    int64_t* n_ovf = k.n_ovf;
    const int64_t ovf_cap = k.ovf_cap;
    const int64_t ra = view_row(A, p), rb = view_row(B, p);
    const Side sa{A.verts + (int64_t)A.vmax * 3 * ra, A.nverts[ra]};
    const Side sb{B.verts + (int64_t)B.vmax * 3 * rb, B.nverts[rb]};
    const EpaResult e = epa_wave(sx, sa, sb, P, lane, robust_pair(A, B, ra, rb));
    if (lane != 0) return;
    const int wall = (A.is_wall[ra] || B.is_wall[rb]) ? 1 : 0;
    const int scene = A.scene ? A.scene[ra] : 0;
    const LawT<T> L = law_of(laws2[2 * scene + wall], lawt2 ? lawt2 + 2 * scene + wall : nullptr, (T*)nullptr);
    V3<T> dt_in = dispt ? mk3<T>(T(dispt[4 * p]), T(dispt[4 * p + 1]), T(dispt[4 * p + 2])) : zero3<T>();
    int hist_t = -1;
    int64_t hkey = 0;
    if constexpr (std::is_same<T, double>::value) {
        if (co.hist.keys[0]) {  // persistent history: last step's displacement of this pair
            hist_t = *co.hist.idx;
            hkey = hist_key(ra, rb);
            dt_in = hist_find(co.hist, hist_t, hkey);
        }
    }
    const double imA = (A.mass && !A.is_wall[ra]) ? 1.0 / A.mass[ra] : 0.0;
    const double imB = (B.mass && !B.is_wall[rb]) ? 1.0 / B.mass[rb] : 0.0;
    T dep;
    V3<T> nrm, ca, cb;
    if constexpr (std::is_same<T, double>::value) {
        dep = e.depth;
        nrm = e.n;
        ca = e.cpA;
        cb = e.cpB;
    } else {
        const FaceGeo<T> g = face_geometry<T>(bt.verts, (int64_t)A.vmax * ra, bt.verts, (int64_t)B.vmax * rb,
                                              e.fa, e.fb, e.flip);
        // A face EPA left as "outside" (or deleted) carries a placeholder
        // distance instead of its plane's: keep EPA's values, without
        // derivatives, whenever the re-evaluation does not reproduce them.
        const bool same = g.depth.v == e.depth && g.n.x.v == e.n.x && g.n.y.v == e.n.y && g.n.z.v == e.n.z &&
                          g.cpA.x.v == e.cpA.x && g.cpA.y.v == e.cpA.y && g.cpA.z.v == e.cpA.z &&
                          g.cpB.x.v == e.cpB.x && g.cpB.y.v == e.cpB.y && g.cpB.z.v == e.cpB.z;
        if (same) {
            dep = g.depth;
            nrm = g.n;
            ca = g.cpA;
            cb = g.cpB;
        } else {
            dep = T(e.depth);
            nrm = mk3<T>(T(e.n.x), T(e.n.y), T(e.n.z));
            ca = mk3<T>(T(e.cpA.x), T(e.cpA.y), T(e.cpA.z));
            cb = mk3<T>(T(e.cpB.x), T(e.cpB.y), T(e.cpB.z));
        }
    }
    const LawOut<T> f = force_law<T>(body3<T>(A.com, bt.com, ra), body3<T>(B.com, bt.com, rb),
                                     body3<T>(A.vel, bt.vel, ra), body3<T>(B.vel, bt.vel, rb),
                                     body3<T>(A.angvel, bt.angvel, ra), body3<T>(B.angvel, bt.angvel, rb), dt_in,
                                     bodyq<T>(A.quat, bt.quat, ra), A.isize[ra], B.isize[rb], imA, imB, dep, nrm,
                                     ca, cb, L);
    if constexpr (std::is_same<T, double>::value) {
        if (co.face) {  // the final face, for the adjoint of this step
            TArr<MWD_NT> va{}, vb{};
            va.v = const_cast<double*>(A.verts);
            vb.v = const_cast<double*>(B.verts);
            const FaceGeo<double> g = face_geometry<double>(va, (int64_t)A.vmax * ra, vb, (int64_t)B.vmax * rb,
                                                            e.fa, e.fb, e.flip);
            const bool same = g.depth == e.depth && g.n.x == e.n.x && g.n.y == e.n.y && g.n.z == e.n.z &&
                              g.cpA.x == e.cpA.x && g.cpA.y == e.cpA.y && g.cpA.z == e.cpA.z &&
                              g.cpB.x == e.cpB.x && g.cpB.y == e.cpB.y && g.cpB.z == e.cpB.z;
            int32_t* f = co.face + 7 * c;
            for (int k = 0; k < 3; ++k) {
                f[k] = e.fa[k];
                f[3 + k] = e.fb[k];
            }
            f[6] = (e.flip ? 1 : 0) | (same ? 0 : 2);
        }
    }
    depth[c] = e.depth;
    st3(normal + 3 * c, e.n);
    st3(cpA + 3 * c, e.cpA);
    st3(cpB + 3 * c, e.cpB);
    stv(co.Fn, c, f.Fn);
    stv(co.tA, c, f.tA);
    stv(co.tB, c, f.tB);
    const int est = e.status | (e.depth < EPA_PLACEHOLDER ? 0 : 4);
    if (epa_status) epa_status[c] = est;
    if (est & 1) fl |= MWD_F_EPA_CAP;
    if (est & 2) fl |= MWD_F_EPA_FACES;
    if (est & 4) fl |= MWD_F_EPA_UNRESOLVED;
    if (disp_out) {
        double* dp = disp_out + 4 * p;
        dp[0] = val(f.disp.x);
        dp[1] = val(f.disp.y);
        dp[2] = val(f.disp.z);
        dp[3] = 1.0;
    }
    if (hist_t >= 0) {  // this step's displacement, for the next step
        const double d4[4] = {val(f.disp.x), val(f.disp.y), val(f.disp.z), 1.0};
        hist_put(co.hist, 1 - hist_t, hkey, d4);
    }
    if (rec_cnt) {
        // buckets of the moving rows (walls do not integrate); a full
        // bucket spills into the shared overflow list as (row, record).
        // Records are keyed by pair index (2p + side), so a row's sum
        // runs in candidate order whatever order the contacts came in
        if (!A.is_wall[ra]) fl |= push_record(rec_cnt, rec_bkt, rec_ovf, n_ovf, ovf_cap, ra, (int32_t)(2 * p));
        if (!B.is_wall[rb]) fl |= push_record(rec_cnt, rec_bkt, rec_ovf, n_ovf, ovf_cap, rb, (int32_t)(2 * p + 1));
    } else {
        rec_key[2 * c] = (uint32_t)(idA ? idA[p] : ra);
        rec_val[2 * c] = (int32_t)(2 * c);
        rec_key[2 * c + 1] = (uint32_t)(idB ? idB[p] : rb);
        rec_val[2 * c + 1] = (int32_t)(2 * c + 1);
    }
}

// EPA + law per contact, one contact per wave: EPA by the whole wave, then
// the law by its first lane. T = double: the forward path (any pair layout).
// T = Dual: the tangent path, where A and B index one state whose tangents bt
// carries; EPA runs on values and the final face is re-evaluated in T.
template <class T>
__global__ void __launch_bounds__(256) k_epa_force(BodyView A, BodyView B, BodyTan bt, ContactSink sink,
                                                   const int32_t* __restrict__ contact_pair,
                                                   const Simplex* __restrict__ simplex,
                                                   const int32_t* __restrict__ sidx, int64_t contact_cap,
                                                   const int64_t* n_contacts, int64_t* n_records, uint32_t* flags) {
    __shared__ PolyW polys[256 / MWD_WAVE];
    int64_t K = *n_contacts;
    if (K > contact_cap) K = contact_cap;
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid == 0 && n_records) *n_records = 2 * K;
    // one contact per wave: EPA by all 64 lanes, then the law by lane 0
    const int lane = threadIdx.x & (MWD_WAVE - 1);
    PolyW& P = polys[threadIdx.x / MWD_WAVE];
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x / MWD_WAVE);
    uint32_t fl = 0;
    for (int64_t c = gid / MWD_WAVE; c < K; c += nw) {
        const int64_t p = contact_pair[c];
        const int32_t slot = sidx[p];
        if (slot < 0 || slot >= contact_cap) {  // simplex storage overflowed (flagged by k_gjk)
            if (sink.co.face && lane == 0) sink.co.face[7 * c + 6] = 4;   // no face: nothing to reverse
            continue;
        }
        contact_wave<T>(c, p, simplex[slot], A, B, bt, sink, P, lane, fl);
    }
    if (fl) atomicOr(flags, fl);
}

// GJK and, for the pairs it finds colliding, EPA and the law in the same
// wave (the plain FP64 step, grouped support): each round the wave runs GJK
// on 64 / G survivors, the group leaders of the colliding ones claim contact
// slots and park their simplices in LDS, and then the whole wave runs EPA
// and the law on each of them in lane order. Results are those of k_gjk
// followed by k_epa_force<double> (records are keyed by pair index, so the
// per-row sums do not depend on which wave ran a contact).
template <int G, int K>
__global__ void __launch_bounds__(256) k_gjk_epa(BodyView A, BodyView B, BodyTan bt, int64_t n_cap,
                                                 const int32_t* __restrict__ surv, const int64_t* lo,
                                                 const int64_t* n_surv, int32_t* __restrict__ collision,
                                                 Simplex* __restrict__ simplex, int32_t* __restrict__ sidx,
                                                 int64_t* slots, int64_t slot_cap, uint32_t* flags,
                                                 int32_t* __restrict__ contact_pair, ContactSink sink) {
    static_assert(K > 0 && MWD_WAVE % G == 0, "grouped support only");
    constexpr int PW = MWD_WAVE / G;  // pairs per wave and round
    __shared__ PolyW polys[256 / MWD_WAVE];
    __shared__ Simplex parked[256 / MWD_WAVE][PW];
    int64_t b0, m;
    surv_range(lo, n_surv, n_cap, &b0, &m);
    const int lane = threadIdx.x & (MWD_WAVE - 1);
    const int w = threadIdx.x / MWD_WAVE;
    const int g = lane % G;
    PolyW& P = polys[w];
    const int64_t stride = (int64_t)gridDim.x * (blockDim.x / G);
    uint32_t fl = 0;
    // s0, the wave's first survivor of a round, is the same on every lane, so
    // the whole wave reaches each ballot and each EPA
    for (int64_t s0 = b0 + ((int64_t)blockIdx.x * blockDim.x + (threadIdx.x - lane)) / G; s0 < m; s0 += stride) {
        const int64_t s = s0 + lane / G;
        int32_t p = 0;
        int64_t slot = -1;
        if (s < m) {
            p = surv[s];
            d3 S[4];
            int ia[4] = {0, 0, 0, 0}, ib[4] = {0, 0, 0, 0};
            const int c = gjk_of<G, K>(A, B, p, g, S, ia, ib);
            if (g == 0) {
                collision[p] = c;
                if (c == 1) {
                    slot = (int64_t)atomicAdd((unsigned long long*)slots, 1ull);
                    sidx[p] = (int32_t)slot;
                    if (slot >= slot_cap) {
                        fl |= MWD_F_CONTACT_OVERFLOW;
                    } else {
                        contact_pair[slot] = p;
                        Simplex& x = simplex[slot];
                        Simplex& y = parked[w][lane / G];
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            x.s[k][0] = y.s[k][0] = S[k].x;
                            x.s[k][1] = y.s[k][1] = S[k].y;
                            x.s[k][2] = y.s[k][2] = S[k].z;
                            x.ia[k] = y.ia[k] = ia[k];
                            x.ib[k] = y.ib[k] = ib[k];
                        }
                    }
                }
            }
        }
        uint64_t todo = __ballot(slot >= 0 && slot < slot_cap);
        if (todo) wave_sync();  // the parked simplices before any lane reads them
        while (todo) {
            const int L = __ffsll((unsigned long long)todo) - 1;
            todo &= todo - 1;
            const int32_t pl = __shfl(p, L, MWD_WAVE);
            const int32_t cl = __shfl((int32_t)slot, L, MWD_WAVE);
            contact_wave<double>(cl, pl, parked[w][L / G], A, B, bt, sink, P, lane, fl);
            wave_sync();  // this contact's polytope before the next one overwrites it
        }
    }
    if (fl) atomicOr(flags, fl);
}

__global__ void k_segments(const uint32_t* __restrict__ key, int64_t cap, const int64_t* n_dev,
                           int32_t* __restrict__ seg_begin, int32_t* __restrict__ seg_end) {
    int64_t n = *n_dev;
    if (n > cap) n = cap;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n; s += stride) {
        const uint32_t k = key[s];
        if (s == 0 || key[s - 1] != k) seg_begin[k] = (int32_t)s;
        if (s == n - 1 || key[s + 1] != k) seg_end[k] = (int32_t)(s + 1);
    }
}

__global__ void k_add_sums(const int32_t* __restrict__ seg_begin, const int32_t* __restrict__ seg_end,
                           const int32_t* __restrict__ rec_val, const double* __restrict__ Fn,
                           const double* __restrict__ tA, const double* __restrict__ tB, double* __restrict__ force,
                           double* __restrict__ torque, int64_t n_rows) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_rows) return;
    const int32_t b = seg_begin[i], e = seg_end[i];
    if (b == e) return;
    d3 F = mk3(0.0, 0.0, 0.0), T = mk3(0.0, 0.0, 0.0);
    for (int32_t s = b; s < e; ++s) {
        const int32_t r = rec_val[s];
        const int64_t c = r >> 1;
        if (r & 1) {
            F = F + neg3(ld3(Fn + 3 * c));
            T = T + ld3(tB + 3 * c);
        } else {
            F = F + ld3(Fn + 3 * c);
            T = T + ld3(tA + 3 * c);
        }
    }
    st3(force + 3 * i, ld3(force + 3 * i) + F);
    st3(torque + 3 * i, ld3(torque + 3 * i) + T);
}

// Per-pair views of per-contact results: non-contact pairs get zeros.
__global__ void k_zero_pairs(int64_t P, double* depth, double* normal, double* cpA, double* cpB, int32_t* epa_status,
                             double* disp) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    if (depth) depth[p] = 0.0;
    for (int k = 0; k < 3; ++k) {
        if (normal) normal[3 * p + k] = 0.0;
        if (cpA) cpA[3 * p + k] = 0.0;
        if (cpB) cpB[3 * p + k] = 0.0;
    }
    if (epa_status) epa_status[p] = 0;
    if (disp)
        for (int k = 0; k < 4; ++k) disp[4 * p + k] = 0.0;
}

__global__ void k_scatter_contacts(const int32_t* __restrict__ cpair, int64_t cap, const int64_t* n_contacts,
                                   const double* __restrict__ dep, const double* __restrict__ nrm,
                                   const double* __restrict__ ca, const double* __restrict__ cb,
                                   const int32_t* __restrict__ est, double* depth, double* normal, double* cpA,
                                   double* cpB, int32_t* epa_status) {
    int64_t K = *n_contacts;
    if (K > cap) K = cap;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= K) return;
    const int64_t p = cpair[c];
    if (depth) depth[p] = dep[c];
    for (int k = 0; k < 3; ++k) {
        if (normal) normal[3 * p + k] = nrm[3 * c + k];
        if (cpA) cpA[3 * p + k] = ca[3 * c + k];
        if (cpB) cpB[3 * p + k] = cb[3 * c + k];
    }
    if (epa_status) epa_status[p] = est[c];
}

unsigned grid_for(int64_t n, int threads, int64_t max_blocks) {
    int64_t b = (n + threads - 1) / threads;
    if (b < 1) b = 1;
    if (b > max_blocks) b = max_blocks;
    return (unsigned)b;
}

}  // namespace

int mwd_launch_sphere(const BodyView& A, const BodyView& B, int64_t n_cap, const int64_t* n_dev, int32_t* collision,
                      int32_t* surv, int64_t* n_surv, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
                      const PairTune& tu) {
    if (n_cap <= 0) return MWD_OK;
    const int tb = tu.sphere_block;
    MWD_LAUNCH_EV(e0, e1, k_sphere, dim3(grid_for(n_cap, tb, (8192 * 256) / tb)), dim3(tb), 0, s, A, B, n_cap, n_dev,
                  collision, surv, n_surv);
    MWD_LAUNCH_CHECK();
    return MWD_OK;
}

int mwd_launch_gjk(const BodyView& A, const BodyView& B, int64_t n_cap, const int64_t* n_dev, int32_t* collision,
                   int32_t* surv, int64_t* n_surv, Simplex* simplex, int32_t* sidx, int64_t* slots, int64_t slot_cap,
                   uint32_t* flags, hipStream_t s) {
    const int rc = mwd_launch_sphere(A, B, n_cap, n_dev, collision, surv, n_surv, s);
    if (rc) return rc;
    return mwd_launch_gjk_surv(A, B, n_cap, collision, surv, n_surv, simplex, sidx, slots, slot_cap, flags, s);
}

namespace {

// k_gjk<G, K> for K = ceil(vmax / G) up to GJK_MAX_K vertices per lane; the
// serial form (K = 0) beyond.
constexpr int GJK_MAX_K = 6;

template <int G, int K>
void launch_gjk_k(int k, dim3 grid, hipStream_t s, hipEvent_t e0, hipEvent_t e1, const BodyView& A,
                  const BodyView& B, int64_t n_cap, const int32_t* surv, const int64_t* lo, const int64_t* n_surv,
                  int32_t* collision, Simplex* simplex, int32_t* sidx, int64_t* slots, int64_t slot_cap,
                  uint32_t* flags, int32_t* contact_pair) {
    if constexpr (K > GJK_MAX_K) {
        MWD_LAUNCH_EV(e0, e1, (k_gjk<G, 0>), grid, dim3(256), 0, s, A, B, n_cap, surv, lo, n_surv, collision,
                      simplex, sidx, slots, slot_cap, flags, contact_pair);
    } else {
        if (k == K)
            MWD_LAUNCH_EV(e0, e1, (k_gjk<G, K>), grid, dim3(256), 0, s, A, B, n_cap, surv, lo, n_surv, collision,
                          simplex, sidx, slots, slot_cap, flags, contact_pair);
        else
            launch_gjk_k<G, K + 1>(k, grid, s, e0, e1, A, B, n_cap, surv, lo, n_surv, collision, simplex, sidx,
                                   slots, slot_cap, flags, contact_pair);
    }
}

template <int G>
void launch_gjk_g(const BodyView& A, const BodyView& B, int64_t n_cap, int32_t* collision, const int32_t* surv,
                  const int64_t* lo, const int64_t* n_surv, Simplex* simplex, int32_t* sidx, int64_t* slots,
                  int64_t slot_cap, uint32_t* flags, hipStream_t s, hipEvent_t e0, hipEvent_t e1, int max_blocks,
                  int32_t* contact_pair, int K_class) {
    const int K = K_class > 0 ? K_class : (A.vmax + G - 1) / G;
    const bool grouped = K <= GJK_MAX_K;
    // the survivors are few next to the candidates: a grid-strided launch
    // with a capped grid
    const dim3 grid(grid_for(n_cap * (grouped ? G : 1), 256, max_blocks));
    launch_gjk_k<G, 1>(grouped ? K : GJK_MAX_K + 1, grid, s, e0, e1, A, B, n_cap, surv, lo, n_surv, collision,
                       simplex, sidx, slots, slot_cap, flags, contact_pair);
}

}  // namespace

int mwd_launch_gjk_surv(const BodyView& A, const BodyView& B, int64_t n_cap, int32_t* collision,
                        const int32_t* surv, const int64_t* n_surv, Simplex* simplex, int32_t* sidx, int64_t* slots,
                        int64_t slot_cap, uint32_t* flags, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
                        const PairTune& tu, int32_t* contact_pair, const int64_t* lo, int K_class) {
    if (n_cap <= 0) return MWD_OK;
    switch (tu.gjk_group) {
        case 4:
            launch_gjk_g<4>(A, B, n_cap, collision, surv, lo, n_surv, simplex, sidx, slots, slot_cap, flags, s, e0,
                            e1, tu.gjk_blocks, contact_pair, K_class);
            break;
        case 16:
            launch_gjk_g<16>(A, B, n_cap, collision, surv, lo, n_surv, simplex, sidx, slots, slot_cap, flags, s, e0,
                             e1, tu.gjk_blocks, contact_pair, K_class);
            break;
        default:
            launch_gjk_g<8>(A, B, n_cap, collision, surv, lo, n_surv, simplex, sidx, slots, slot_cap, flags, s, e0,
                            e1, tu.gjk_blocks, contact_pair, K_class);
    }
    MWD_LAUNCH_CHECK();
    return MWD_OK;
}

namespace {
__global__ void k_hist_prep(HistView h, int64_t cap) {
    const int t = 1 - *h.idx;  // the table this step writes
    int64_t n = *h.count[t];
    if (n > cap) n = cap;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        h.keys[t][h.used[t][i]] = -1;
}
__global__ void k_hist_count0(HistView h) { *h.count[1 - *h.idx] = 0; }
__global__ void k_hist_flip(HistView h) { *h.idx = 1 - *h.idx; }
}  // namespace

int mwd_launch_hist_prep(const HistView& h, int64_t contact_cap, hipStream_t s) {
    hipLaunchKernelGGL(k_hist_prep, dim3(grid_for(contact_cap, 256, 2048)), dim3(256), 0, s, h, contact_cap);
    MWD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_hist_count0, dim3(1), dim3(1), 0, s, h);
    MWD_LAUNCH_CHECK();
    return MWD_OK;
}

int mwd_launch_hist_flip(const HistView& h, hipStream_t s) {
    hipLaunchKernelGGL(k_hist_flip, dim3(1), dim3(1), 0, s, h);
    MWD_LAUNCH_CHECK();
    return MWD_OK;
}

int mwd_launch_compact_contacts(const int32_t* collision, int64_t n_cap, const int64_t* n_dev, int32_t* contact_pair,
                                int64_t contact_cap, int64_t* n_contacts, uint32_t* flags, int64_t* tmp,
                                int64_t tmp_len, hipStream_t s) {
    return mwd_scan_io(ContactIO{collision, contact_pair, contact_cap, flags}, n_cap, n_dev, n_contacts, tmp, tmp_len,
                       s);
}

static TArr<MWD_NT> plain(double* v) {
    TArr<MWD_NT> a;
    a.v = v;
    for (int k = 0; k < MWD_NT; ++k) a.t[k] = nullptr;
    return a;
}

int mwd_launch_epa_force(const BodyView& A, const BodyView& B, const mwd_law* laws2, const int32_t* contact_pair,
                         const Simplex* simplex, const int32_t* sidx, int64_t contact_cap, const int64_t* n_contacts,
                         const double* dispt,
                         const int32_t* idA, const int32_t* idB, double* depth, double* normal, double* cpA,
                         double* cpB, double* Fn, double* tA, double* tB, double* disp_out, int32_t* epa_status,
                         uint32_t* rec_key, int32_t* rec_val, int64_t* n_records, uint32_t* flags, hipStream_t s) {
    if (contact_cap <= 0) return MWD_OK;
    BodyTan bt;
    memset(&bt, 0, sizeof(bt));
    ContactOut co;
    co.Fn = plain(Fn);
    co.tA = plain(tA);
    co.tB = plain(tB);
    const ContactSink sink{laws2,    nullptr,    dispt,   idA,     idB,     depth,   normal,  cpA,     cpB,   co,
                           disp_out, epa_status, rec_key, rec_val, nullptr, nullptr, nullptr, nullptr, 0};
    hipLaunchKernelGGL(k_epa_force<double>, dim3(grid_for(contact_cap * MWD_WAVE, 256, EPA_MAX_BLOCKS)), dim3(256),
                       0, s, A, B, bt, sink, contact_pair, simplex, sidx, contact_cap, n_contacts, n_records, flags);
    MWD_LAUNCH_CHECK();
    return MWD_OK;
}

int mwd_launch_epa_force_rows(const BodyView& A, const BodyView& B, const BodyTan* bt, const mwd_law* laws2,
                              const mwd_law_tan* lawt2, const int32_t* contact_pair, const Simplex* simplex,
                              const int32_t* sidx, int64_t contact_cap, const int64_t* n_contacts, double* depth,
                              double* normal, double* cpA, double* cpB, const ContactOut& co, int32_t* epa_status,
                              int32_t* rec_cnt, int32_t* rec_bkt, int32_t* rec_ovf, int64_t* n_ovf,
                              int64_t ovf_cap, uint32_t* flags, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
                              int64_t max_blocks) {
    if (contact_cap <= 0) return MWD_OK;
    const dim3 grid(grid_for(contact_cap * MWD_WAVE, 256, max_blocks)), blk(256);
    const ContactSink sink{laws2,   bt ? lawt2 : nullptr, nullptr, nullptr, nullptr, depth,   normal,  cpA,   cpB, co,
                           nullptr, epa_status,           nullptr, nullptr, rec_cnt, rec_bkt, rec_ovf, n_ovf, ovf_cap};
    if (bt) {
        MWD_LAUNCH_EV(e0, e1, k_epa_force<Dual<MWD_NT>>, grid, blk, 0, s, A, B, *bt, sink, contact_pair, simplex,
                      sidx, contact_cap, n_contacts, (int64_t*)nullptr, flags);
    } else {
        BodyTan none;
        memset(&none, 0, sizeof(none));
        MWD_LAUNCH_EV(e0, e1, k_epa_force<double>, grid, blk, 0, s, A, B, none, sink, contact_pair, simplex, sidx,
                      contact_cap, n_contacts, (int64_t*)nullptr, flags);
    }
    MWD_LAUNCH_CHECK();
    return MWD_OK;
}

namespace {

template <int G, int K>
void launch_gjk_epa_k(int k, dim3 grid, hipStream_t s, hipEvent_t e0, hipEvent_t e1, const BodyView& A,
                      const BodyView& B, const BodyTan& bt, int64_t n_cap, const int32_t* surv, const int64_t* lo,
                      const int64_t* n_surv, int32_t* collision, Simplex* simplex, int32_t* sidx, int64_t* slots,
                      int64_t slot_cap, uint32_t* flags, int32_t* contact_pair, const ContactSink& sink) {
    if constexpr (K <= 4) {
        if (k == K)
            MWD_LAUNCH_EV(e0, e1, (k_gjk_epa<G, K>), grid, dim3(256), 0, s, A, B, bt, n_cap, surv, lo, n_surv,
                          collision, simplex, sidx, slots, slot_cap, flags, contact_pair, sink);
        else
            launch_gjk_epa_k<G, K + 1>(k, grid, s, e0, e1, A, B, bt, n_cap, surv, lo, n_surv, collision, simplex,
                                       sidx, slots, slot_cap, flags, contact_pair, sink);
    }
}

template <int G>
int launch_gjk_epa_g(const BodyView& A, const BodyView& B, int64_t n_cap, const int32_t* surv, const int64_t* lo,
                     const int64_t* n_surv, int32_t* collision, Simplex* simplex, int32_t* sidx, int64_t* slots,
                     int64_t slot_cap, uint32_t* flags, int32_t* contact_pair, const ContactSink& sink, hipStream_t s,
                     hipEvent_t e0, hipEvent_t e1, int max_blocks, int K_class) {
    // beyond 4 vertices per lane the fused kernel's registers allow one wave
    // per SIMD: the two launches are faster there
    const int Kv = (A.vmax + G - 1) / G;
    if (Kv < 1 || Kv > 4) return 1;
    const int K = K_class > 0 ? K_class : Kv;
    BodyTan none;
    memset(&none, 0, sizeof(none));
    launch_gjk_epa_k<G, 1>(K, dim3(grid_for(n_cap * G, 256, max_blocks)), s, e0, e1, A, B, none, n_cap, surv, lo,
                           n_surv, collision, simplex, sidx, slots, slot_cap, flags, contact_pair, sink);
    return 0;
}

}  // namespace

int mwd_launch_gjk_epa(const BodyView& A, const BodyView& B, int64_t n_cap, int32_t* collision, const int32_t* surv,
                       const int64_t* n_surv, Simplex* simplex, int32_t* sidx, int64_t* slots, int64_t slot_cap,
                       int32_t* contact_pair, const mwd_law* laws2, double* depth, double* normal, double* cpA,
                       double* cpB, const ContactOut& co, int32_t* epa_status, int32_t* rec_cnt, int32_t* rec_bkt,
                       int32_t* rec_ovf, int64_t* n_ovf, int64_t ovf_cap, uint32_t* flags, hipStream_t s,
                       hipEvent_t e0, hipEvent_t e1, const PairTune& tu, const int64_t* lo, int K_class) {
    if (A.vmax != B.vmax || !contact_pair || !rec_cnt) return 1;
    if (n_cap <= 0) return 0;
    const ContactSink sink{laws2,   nullptr, nullptr, nullptr, nullptr, depth,   normal,  cpA,   cpB,    co,
                           nullptr, epa_status, nullptr, nullptr, rec_cnt, rec_bkt, rec_ovf, n_ovf, ovf_cap};
    int r;
    switch (tu.gjk_group) {
        case 4:
            r = launch_gjk_epa_g<4>(A, B, n_cap, surv, lo, n_surv, collision, simplex, sidx, slots, slot_cap, flags,
                                    contact_pair, sink, s, e0, e1, tu.gjk_blocks, K_class);
            break;
        case 16:
            r = launch_gjk_epa_g<16>(A, B, n_cap, surv, lo, n_surv, collision, simplex, sidx, slots, slot_cap,
                                     flags, contact_pair, sink, s, e0, e1, tu.gjk_blocks, K_class);
            break;
        default:
            r = launch_gjk_epa_g<8>(A, B, n_cap, surv, lo, n_surv, collision, simplex, sidx, slots, slot_cap, flags,
                                    contact_pair, sink, s, e0, e1, tu.gjk_blocks, K_class);
    }
    if (r) return r;
    MWD_LAUNCH_CHECK();
    return 0;
}

int mwd_launch_gjk_buckets(const BodyView& A, const BodyView& B, const int32_t* surv, const int64_t* n_surv,
                           int64_t n_cap, int G, int nb, int64_t* bucket, int32_t* out, hipStream_t s) {
    if (nb < 1 || nb > MWD_GJK_BUCKETS) return MWD_E_INVALID;
    int rc;
    // counts and cursors start at zero (a fill kernel: no memset node in a graph)
    if ((rc = mwd_fill_i32(reinterpret_cast<int32_t*>(bucket), 0, 4 * MWD_GJK_BUCKETS, s))) return rc;
    const dim3 grid(grid_for(n_cap, 256, 1024));
    hipLaunchKernelGGL(k_bucket_count, grid, dim3(256), 0, s, A, B, surv, n_surv, n_cap, G, nb, bucket);
    MWD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_bucket_scatter, grid, dim3(256), 0, s, A, B, surv, n_surv, n_cap, G, nb, bucket,
                       bucket + MWD_GJK_BUCKETS, bucket + 2 * MWD_GJK_BUCKETS, out);
    MWD_LAUNCH_CHECK();
    return MWD_OK;
}

int mwd_launch_segments(const uint32_t* rec_key, int64_t rec_cap, const int64_t* n_rec, int32_t* seg_begin,
                        int32_t* seg_end, hipStream_t s) {
    if (rec_cap <= 0) return MWD_OK;
    hipLaunchKernelGGL(k_segments, dim3(grid_for(rec_cap, 256, 8192)), dim3(256), 0, s, rec_key, rec_cap, n_rec,
                       seg_begin, seg_end);
    MWD_LAUNCH_CHECK();
    return MWD_OK;
}

int mwd_launch_add_sums(const int32_t* seg_begin, const int32_t* seg_end, const int32_t* rec_val, const double* Fn,
                        const double* tA, const double* tB, double* force, double* torque, int64_t n_rows,
                        hipStream_t s) {
    if (n_rows <= 0) return MWD_OK;
    hipLaunchKernelGGL(k_add_sums, dim3((unsigned)((n_rows + 255) / 256)), dim3(256), 0, s, seg_begin, seg_end,
                       rec_val, Fn, tA, tB, force, torque, n_rows);
    MWD_LAUNCH_CHECK();
    return MWD_OK;
}

int mwd_launch_zero_pairs(int64_t P, double* depth, double* normal, double* cpA, double* cpB, int32_t* epa_status,
                          double* disp, hipStream_t s) {
    if (P <= 0) return MWD_OK;
    hipLaunchKernelGGL(k_zero_pairs, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, s, P, depth, normal, cpA, cpB,
                       epa_status, disp);
    MWD_LAUNCH_CHECK();
    return MWD_OK;
}

int mwd_launch_scatter_contacts(const int32_t* cpair, int64_t cap, const int64_t* n_contacts, const double* dep,
                                const double* nrm, const double* ca, const double* cb, const int32_t* est, int64_t P,
                                double* depth, double* normal, double* cpA, double* cpB, int32_t* epa_status,
                                double* disp, hipStream_t s) {
    (void)disp;
    (void)P;
    if (cap <= 0 || !(depth || normal || cpA || cpB || epa_status)) return MWD_OK;
    hipLaunchKernelGGL(k_scatter_contacts, dim3((unsigned)((cap + 255) / 256)), dim3(256), 0, s, cpair, cap,
                       n_contacts, dep, nrm, ca, cb, est, depth, normal, cpA, cpB, epa_status);
    MWD_LAUNCH_CHECK();
    return MWD_OK;
}
