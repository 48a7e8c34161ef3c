// Reverse-mode adjoint of one step on the physical law: hand-written
// vector-Jacobian products of the Verlet step (gpu_Verlet.cu:164-403), of the
// physical contact law (the normal spring-dashpot and Coulomb-capped
// tangential force of .cu:647-871 with the midpoint contact point and
// effective-mass damping) and of EPA's final-face geometry (AddTriangle
// .cu:270-325 and the barycentric contact points .cu:567-601). The contact
// topology of the step (which pairs touch, EPA's final face, whether the
// tangential force is capped) is taken from the values and held fixed.
//
// One call reverses step n: given the state s_n and lambda_{n+1} (the
// adjoint of the step's outputs, laid out as the state arrays), it leaves
// lambda_n in the same arrays and adds d(lambda_{n+1} . s_{n+1})/d theta to
// an accumulator, theta being the three law parameters whose coefficient
// derivatives mwd_set_law_tangents gave. Kernels, after the forward contact
// stages have rerun on s_n (or their outputs have come back from the tape,
// below):
//   k_force_sum    per moving row: the step's contact force and torque (its
//                  records in ascending order, as the forward reduction sums
//                  them) plus the external ones;
//   k_verlet_adj   per moving row: the Verlet step recomputed, then reversed
//                  into lambda of the row's own inputs and of its F and T;
//   k_contact_adj  per contact: face geometry and law recomputed, then
//                  reversed from the lambda_F, lambda_T of its two rows; each
//                  side's contribution and the contact's d/d theta stored;
//   k_gather_adj   per moving row: its contacts' contributions added in the
//                  row's record order (deterministic, no atomics), face-vertex
//                  adjoints into the row's vertex slots; buckets emptied;
//                  each contact's d/d theta added to the row holding its A
//                  side (B when A is a wall);
//   k_theta_sum    one block: the rows' d/d theta summed in a fixed order.
// A square root of zero (a body at rest) takes a zero derivative, as the
// dual-number path does; a face whose re-evaluation does not reproduce EPA's
// values (a placeholder face) keeps its geometry without derivative.
#include "mwd_internal.h"

namespace {

__device__ __forceinline__ q4 qadd(q4 a, q4 b) {
    q4 r;
    r.w = a.w + b.w;
    r.x = a.x + b.x;
    r.y = a.y + b.y;
    r.z = a.z + b.z;
    return r;
}
__device__ __forceinline__ q4 qzero() {
    q4 r;
    r.w = r.x = r.y = r.z = 0.0;
    return r;
}
__device__ __forceinline__ d3 z3() { return mk3(0.0, 0.0, 0.0); }

// c = a x b
__device__ __forceinline__ void cross_adj(d3 a, d3 b, d3 lc, d3& la, d3& lb) {
    la = la + cross3(b, lc);
    lb = lb + cross3(lc, a);
}
// c = a b (Hamilton): right multiplication by b has the transpose of right
// multiplication by conj(b), and likewise on the left
__device__ __forceinline__ void qmul_adj(q4 a, q4 b, q4 lc, q4& la, q4& lb) {
    la = qadd(la, qmul(lc, qconj(b)));
    lb = qadd(lb, qmul(qconj(a), lc));
}
// r = vector part of q1 (0, v) q2
__device__ __forceinline__ void qrot_adj(q4 q1, d3 v, q4 q2, d3 lr, q4& lq1, d3& lv, q4& lq2) {
    const q4 vq = mkq(0.0, v);
    const q4 m = qmul(vq, q2);
    const q4 lp = mkq(0.0, lr);
    lq1 = qadd(lq1, qmul(lp, qconj(m)));
    const q4 lm = qmul(qconj(q1), lp);
    lq2 = qadd(lq2, qmul(qconj(vq), lm));
    lv = lv + qv(qmul(lm, qconj(q2)));
}
// r = q (0, v) conj(q): both uses of q collected in lq
__device__ __forceinline__ void qrotc_adj(q4 q, d3 v, d3 lr, q4& lq, d3& lv) {
    q4 l1 = qzero(), l2 = qzero();
    qrot_adj(q, v, qconj(q), lr, l1, lv, l2);
    lq = qadd(lq, qadd(l1, qconj(l2)));
}
// h = (cos(|w| s), w / |w| sin(|w| s)) (zero vector part at w = 0)
__device__ __forceinline__ q4 half_turn(d3 w, double dt, double div) {
    const double wn = sqrt(nrm2(w));
    d3 t = z3();
    if (wn != 0.0) t = (w / wn) * sin((wn * dt) / div);
    return mkq(cos((wn * dt) / div), t);
}
__device__ __forceinline__ void half_turn_adj(d3 w, double dt, double div, q4 lh, d3& lw) {
    const double n = sqrt(nrm2(w));
    if (n == 0.0) return;
    const double s = dt / div;
    const d3 u = w / n;
    const double sn = sin(n * s), cs = cos(n * s);
    const d3 lv = qv(lh);
    const double ul = dot3(u, lv);
    lw = lw + u * (-sn * s * lh.w) + (lv - u * ul) * (sn / n) + u * (s * cs * ul);
}
// M^T v for a row-major 3x3 M
__device__ __forceinline__ d3 mat3tv(const double* m, d3 v) {
    return mk3(m[0] * v.x + m[3] * v.y + m[6] * v.z, m[1] * v.x + m[4] * v.y + m[7] * v.z,
               m[2] * v.x + m[5] * v.y + m[8] * v.z);
}
__device__ __forceinline__ void add3(double* p, d3 a) {
    p[0] += a.x;
    p[1] += a.y;
    p[2] += a.z;
}

// ------------------------------------------------------------ force sums
// Per-row kernels over contact records give a row GL lanes: the lanes load
// the row's records (its bucket, when it did not overflow) and what each
// record brings, in parallel, and park them in LDS at the record's rank
// (the records below it); then sums run over the ranks in ascending order,
// one lane per summed component, adding in exactly the order of a walk over
// the sorted records. A row whose bucket overflowed takes that walk on one
// lane (for_row_records). GROWS rows per 64-lane block.
constexpr int GL = 16;
constexpr int GROWS = 4;
constexpr int GB = GL * GROWS;
constexpr int RC = MWD_ROW_CONTACTS;

// The row's bucket into rec[0, n) and each lane's records' ranks (n <= RC);
// call with every lane of the block (it holds a barrier).
__device__ __forceinline__ void group_bucket(const int32_t* __restrict__ b, int n, bool grouped, int lane,
                                             int32_t* rec) {
    if (grouped)
        for (int j = lane; j < n; j += GL) rec[j] = b[j];
    __syncthreads();
}
__device__ __forceinline__ int rank_in(const int32_t* rec, int n, int32_t r) {
    int k = 0;
    for (int j = 0; j < n; ++j) k += rec[j] < r ? 1 : 0;
    return k;
}

__global__ void __launch_bounds__(GB) k_force_sum(const int32_t* __restrict__ plist, int64_t M,
                                                  const int32_t* __restrict__ rec_cnt,
                                                  const int32_t* __restrict__ rec_bkt,
                                                  const int32_t* __restrict__ rec_ovf, const int64_t* n_ovf,
                                                  int64_t ovf_cap, const int32_t* __restrict__ sidx,
                                                  const double* __restrict__ Fn,
                                                  const double* __restrict__ tA, const double* __restrict__ tB,
                                                  const double* __restrict__ fext, const double* __restrict__ text,
                                                  double* __restrict__ F, double* __restrict__ T) {
    __shared__ int32_t s_rec[GROWS][RC];
    __shared__ double s_val[GROWS][RC][6];  // the record's force, then its torque
    const int g = threadIdx.x / GL, lane = threadIdx.x % GL;
    const int64_t i = (int64_t)blockIdx.x * GROWS + g;
    const bool live = i < M;
    const int64_t P = live ? plist[i] : 0;
    const int n = live ? rec_cnt[P] : 0;
    const bool grouped = n <= RC;
    group_bucket(rec_bkt + P * RC, n, grouped, lane, s_rec[g]);
    if (grouped)
        for (int j = lane; j < n; j += GL) {
            const int32_t r = s_rec[g][j];
            const int k = rank_in(s_rec[g], n, r);
            const int64_t c = sidx[r >> 1];
            const d3 f = ld3(Fn + 3 * c), t = ld3((r & 1) ? tB + 3 * c : tA + 3 * c);
            double* o = s_val[g][k];
            o[0] = (r & 1) ? -f.x : f.x;
            o[1] = (r & 1) ? -f.y : f.y;
            o[2] = (r & 1) ? -f.z : f.z;
            o[3] = t.x;
            o[4] = t.y;
            o[5] = t.z;
        }
    __syncthreads();
    if (!live) return;
    if (grouped) {
        if (lane < 6) {
            double acc = 0.0;
            for (int k = 0; k < n; ++k) acc += s_val[g][k][lane];
            if (lane < 3) F[3 * P + lane] = acc + fext[3 * P + lane];
            else T[3 * P + lane - 3] = acc + text[3 * P + lane - 3];
        }
    } else if (lane == 0) {
        d3 f = z3(), t = z3();
        for_row_records(P, n, rec_bkt, rec_ovf, n_ovf, ovf_cap, [&](int32_t r) {
            const int64_t c = sidx[r >> 1];
            if (r & 1) {
                f = f + neg3(ld3(Fn + 3 * c));
                t = t + ld3(tB + 3 * c);
            } else {
                f = f + ld3(Fn + 3 * c);
                t = t + ld3(tA + 3 * c);
            }
        });
        st3(F + 3 * P, f + ld3(fext + 3 * P));
        st3(T + 3 * P, t + ld3(text + 3 * P));
    }
}

// ---------------------------------------------------------- Verlet adjoint
struct Lam {
    double *com, *v, *v12, *w, *wbf, *w12bf, *q, *q12, *qall, *dxall, *verts;
};

__global__ void __launch_bounds__(256) k_verlet_adj(VState a, const double* __restrict__ Fa,
                                                    const double* __restrict__ Ta, Lam lam,
                                                    double* __restrict__ lamF, double* __restrict__ lamT,
                                                    const int32_t* __restrict__ plist, int64_t M,
                                                    const int32_t* __restrict__ scene_of,
                                                    const SceneDev* __restrict__ scenes,
                                                    const int32_t* __restrict__ nverts, int vmax, double dt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const int64_t P = plist[i];
    const SceneDev& sc = scenes[scene_of[P]];
    const int rot_mode = sc.law_p.mode == MWD_LAW_PHYSICAL ? 1 : 0;
    const double* box = sc.box;
    // ---- forward, as verlet_row<double>
    const d3 v12 = ld3(a.v12.v + 3 * P), w12bf = ld3(a.w12bf.v + 3 * P), wbf = ld3(a.wbf.v + 3 * P);
    const d3 w = ld3(a.w.v + 3 * P), com = ld3(a.com.v + 3 * P);
    const q4 q = ldq(a.q.v + 4 * P), q12 = ldq(a.q12.v + 4 * P);
    const double m = a.mass[P];
    const double* I = a.I + 9 * P;
    const double* Iinv = a.Iinv + 9 * P;
    const d3 F = ld3(Fa + 3 * P), Tq = ld3(Ta + 3 * P);
    const d3 v32 = v12 + (F / m) * dt;
    const q4 qi = qconj(q);
    const d3 tbf = qrot(qi, Tq, q);
    const d3 p1 = mat3v(I, wbf);
    const d3 p2 = tbf - cross3(wbf, p1);
    const d3 abf = mat3v(Iinv, p2);
    const d3 w32bf = w12bf + abf * dt;
    const q4 hA = half_turn(w, dt, 2.0);
    const q4 q32 = qmul(hA, q12);
    const d3 dx = v32 * dt;
    const d3 w34bf = w32bf + (abf * dt) / 4.0;
    const d3 w34 = qrot(q32, w34bf, qconj(q32));
    const q4 hB = half_turn(w34, dt, 4.0);
    const q4 qn = qmul(hB, q32);
    const d3 wbfn = w32bf + (abf * dt) / 2.0;
    const q4 qr = rot_mode == 1 ? qmul(qn, qconj(q)) : qn;
    const d3 c1 = com + dx;
    const bool out = c1.x > box[0] || c1.x < box[1] || c1.y > box[2] || c1.y < box[3] || c1.z > box[4] ||
                     c1.z < box[5];
    // ---- adjoints of the outputs (those the reset zeroes were constants)
    d3 Lv = ld3(lam.v + 3 * P), Lv12 = ld3(lam.v12 + 3 * P), Lw12bf = ld3(lam.w12bf + 3 * P);
    d3 Lwbf = ld3(lam.wbf + 3 * P), Lw = ld3(lam.w + 3 * P);
    const d3 Lcom = ld3(lam.com + 3 * P), Ldxall = ld3(lam.dxall + 3 * P);
    const q4 Lq12 = ldq(lam.q12 + 4 * P), Lq = ldq(lam.q + 4 * P), Lqall = ldq(lam.qall + 4 * P);
    if (out) Lv = Lv12 = Lw12bf = Lwbf = Lw = z3();
    // ---- reverse
    d3 lv32 = z3(), lw32bf = z3(), labf = z3(), lcom = z3(), ldx = z3(), lF = z3(), lT = z3();
    d3 lwbf = z3(), lw = z3();
    q4 lqn = qzero(), lq32 = qzero(), lq = qzero(), lq12 = qzero(), lqr = qzero();
    // vertices: nv = qr (verts - com) conj(qr) + com + dx (dx undone when out)
    // (VCH vertices' loads issued together, ahead of their stores)
    const int nrot = nverts[P];
    double* const plv = lam.verts + 3 * (int64_t)vmax * P;
    const double* const pv = a.verts.v + 3 * (int64_t)vmax * P;
    constexpr int VCH = 8;
    for (int k0 = 0; k0 < nrot; k0 += VCH) {
        d3 L[VCH], r[VCH];
#pragma unroll
        for (int j = 0; j < VCH; ++j)
            if (k0 + j < nrot) {
                L[j] = ld3(plv + 3 * (k0 + j));
                r[j] = ld3(pv + 3 * (k0 + j)) - com;
            }
#pragma unroll
        for (int j = 0; j < VCH; ++j)
            if (k0 + j < nrot) {
                d3 lr = z3();
                qrotc_adj(qr, r[j], L[j], lqr, lr);
                st3(plv + 3 * (k0 + j), lr);  // lambda of the vertex itself
                lcom = lcom + (L[j] - lr);
                if (!out) ldx = ldx + L[j];
            }
    }
    lcom = lcom + Lcom;
    if (!out) ldx = ldx + Lcom;
    if (rot_mode == 1) {
        q4 lqc = qzero();
        qmul_adj(qn, qconj(q), lqr, lqn, lqc);
        lq = qadd(lq, qconj(lqc));
    } else {
        lqn = qadd(lqn, lqr);
    }
    lv32 = lv32 + Lv;
    lF = lF + Lv * (dt / m);
    d3 lwbfn = Lwbf;
    qrotc_adj(qn, wbfn, Lw, lqn, lwbfn);
    lw32bf = lw32bf + lwbfn;
    labf = labf + lwbfn * (dt / 2.0);
    qmul_adj(qn, q, Lqall, lqn, lq);
    lqn = qadd(lqn, Lq);
    q4 lhB = qzero();
    qmul_adj(hB, q32, lqn, lhB, lq32);
    d3 lw34 = z3();
    half_turn_adj(w34, dt, 4.0, lhB, lw34);
    d3 lw34bf = z3();
    qrotc_adj(q32, w34bf, lw34, lq32, lw34bf);
    lw32bf = lw32bf + lw34bf;
    labf = labf + lw34bf * (dt / 4.0);
    const d3 ldxall = Ldxall;
    ldx = ldx + Ldxall;
    lv32 = lv32 + ldx * dt + Lv12;
    lq32 = qadd(lq32, Lq12);
    lw32bf = lw32bf + Lw12bf;
    q4 lhA = qzero();
    qmul_adj(hA, q12, lq32, lhA, lq12);
    half_turn_adj(w, dt, 2.0, lhA, lw);
    const d3 lw12bf = lw32bf;
    labf = labf + lw32bf * dt;
    const d3 lp2 = mat3tv(Iinv, labf);
    d3 lp1 = z3();
    cross_adj(wbf, p1, neg3(lp2), lwbf, lp1);
    lwbf = lwbf + mat3tv(I, lp1);
    q4 lqi = qzero();
    qrot_adj(qi, Tq, q, lp2, lqi, lT, lq);
    lq = qadd(lq, qconj(lqi));
    const d3 lv12 = lv32;
    lF = lF + lv32 * (dt / m);
    // ---- lambda_n of the row (the step reads no v and no qall)
    st3(lam.v + 3 * P, z3());
    st3(lam.v12 + 3 * P, lv12);
    st3(lam.w12bf + 3 * P, lw12bf);
    st3(lam.wbf + 3 * P, lwbf);
    st3(lam.w + 3 * P, lw);
    st3(lam.com + 3 * P, lcom);
    st3(lam.dxall + 3 * P, ldxall);
    stq(lam.q + 4 * P, lq);
    stq(lam.q12 + 4 * P, lq12);
    stq(lam.qall + 4 * P, qzero());
    st3(lamF + 3 * P, lF);
    st3(lamT + 3 * P, lT);
}

// --------------------------------------------------------- contact adjoint
constexpr int CADJ = 18;  // per side: lambda of com, v, w, then of the three face vertices

__global__ void __launch_bounds__(256) k_contact_adj(BodyView A, BodyView B, const mwd_law* __restrict__ laws2,
                                                     const mwd_law_tan* __restrict__ lawt2,
                                                     const int32_t* __restrict__ contact_pair,
                                                     const int64_t* n_contacts, int64_t contact_cap,
                                                     const int32_t* __restrict__ face,
                                                     const double* __restrict__ depth_e,
                                                     const double* __restrict__ normal_e,
                                                     const double* __restrict__ cpA_e,
                                                     const double* __restrict__ cpB_e,
                                                     const double* __restrict__ lamF,
                                                     const double* __restrict__ lamT, double* __restrict__ cadj,
                                                     double* __restrict__ cdth) {
    int64_t K = *n_contacts;
    if (K > contact_cap) K = contact_cap;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < K; c += stride) {
        double* oa = cadj + (2 * c) * CADJ;
        double* ob = cadj + (2 * c + 1) * CADJ;
        for (int k = 0; k < CADJ; ++k) oa[k] = ob[k] = 0.0;
        cdth[3 * c] = cdth[3 * c + 1] = cdth[3 * c + 2] = 0.0;
        // no depth resolved, or a signed depth that is not positive: no force, nothing to reverse
        if (!(depth_e[c] < EPA_PLACEHOLDER && depth_e[c] > 0.0)) continue;
        const int64_t p = contact_pair[c];
        const int64_t ra = view_row(A, p), rb = view_row(B, p);
        const bool wallA = A.is_wall[ra], wallB = B.is_wall[rb];
        const int scene = A.scene ? A.scene[ra] : 0;
        // the bodies' state, loaded with the face's vertices (used by the law below)
        const d3 xA = ld3(A.com + 3 * ra), xB = ld3(B.com + 3 * rb);
        const d3 vA = ld3(A.vel + 3 * ra), vB = ld3(B.vel + 3 * rb);
        const d3 wA = ld3(A.angvel + 3 * ra), wB = ld3(B.angvel + 3 * rb);
        const double imA = (A.mass && !wallA) ? 1.0 / A.mass[ra] : 0.0;
        const double imB = (B.mass && !wallB) ? 1.0 / B.mass[rb] : 0.0;
        const double iA = imA == 0.0 ? 0.0 : 1.0 / A.isize[ra];
        const double iB = imB == 0.0 ? 0.0 : 1.0 / B.isize[rb];
        const int lw = (wallA || wallB) ? 1 : 0;
        const mwd_law& L = laws2[2 * scene + lw];
        const mwd_law_tan& LT = lawt2[2 * scene + lw];
        const int32_t* f = face + 7 * c;
        if (f[6] & 4) continue;  // not resolved in the forward pass
        bool ok = true;
        for (int k = 0; k < 3; ++k)
            ok = ok && f[k] >= 0 && f[k] < A.nverts[ra] && f[3 + k] >= 0 && f[3 + k] < B.nverts[rb];
        if (!ok) continue;  // not reached: the forward pass wrote this contact's face
        const bool geo = (f[6] & 2) == 0;
        const int flip = f[6] & 1;
        // ---- face geometry, forward (as face_geometry<double>)
        d3 av[3], bv[3], S[3];
        for (int k = 0; k < 3; ++k) {
            av[k] = ld3(A.verts + 3 * ((int64_t)A.vmax * ra + f[k]));
            bv[k] = ld3(B.verts + 3 * ((int64_t)B.vmax * rb + f[3 + k]));
            S[k] = av[k] - bv[k];
        }
        const int e1 = flip ? 2 : 1, e2 = flip ? 1 : 2;
        const d3 E1 = S[e1] - S[0], E2 = S[e2] - S[e1];
        const d3 nn0 = cross3(E1, E2);
        const double len = sqrt(nrm2(nn0));
        const d3 nraw = nn0 / len;
        const double ddraw = dot3(nraw, S[0]);
        const double sg = flip ? -1.0 : 1.0;
        double depth = ddraw * sg;
        d3 nface = nraw * sg;
        const d3 Proj = nface * depth;
        bool swapped = false;
        d3 v0 = S[1] - S[0], v1 = S[2] - S[0];
        const d3 v2 = Proj - S[0];
        if (dot3(cross3(v0, v1), nface) < 0.0) {
            v0 = S[2] - S[0];
            v1 = S[1] - S[0];
            swapped = true;
        }
        const double v00 = dot3(v0, v0), v01 = dot3(v0, v1), v11 = dot3(v1, v1), v20 = dot3(v2, v0),
                     v21 = dot3(v2, v1);
        const double den = v00 * v11 - v01 * v01;
        const double vv = (v11 * v20 - v01 * v21) / den;
        const double ww = (v00 * v21 - v01 * v20) / den;
        const double uu = (1.0 - ww) - vv;
        d3 cpA = (av[0] * uu + av[1] * vv) + av[2] * ww;
        d3 cpB = (bv[0] * uu + bv[1] * vv) + bv[2] * ww;
        if (!geo) {  // EPA's values, as constants
            depth = depth_e[c];
            nface = ld3(normal_e + 3 * c);
            cpA = ld3(cpA_e + 3 * c);
            cpB = ld3(cpB_e + 3 * c);
        }
        // ---- law, forward (force_law<double>, physical branch, no tangential history)
        const d3 cp = (cpA + cpB) / 2.0;
        const d3 dA = cp - xA, dB = cp - xB;
        const double imn = imA + imB;
        const double mn = 1.0 / imn;
        const double imt = (imn + nrm2(dA) * iA) + nrm2(dB) * iB;
        const double mt = 1.0 / imt;
        const double gKn = sqrt(mn * L.Kn);
        const double Gn = L.cn * gKn;
        const double Kt = (L.Kn * L.ratio) * (mt / mn);
        const double gKt = sqrt(mt * Kt);
        const double Gt = L.ct * gKt;
        const double sgn = dot3(xB - xA, nface) < 0.0 ? -1.0 : 1.0;
        const d3 n1 = nface * sgn;
        const double nl = sqrt(nrm2(n1));
        const d3 n = n1 / nl;
        const d3 dn = n * depth;
        const d3 rA = cpA - xA, rB = cpB - xB;
        const d3 u = ((vA + cross3(wA, rA)) - vB) - cross3(wB, rB);
        const double s_un = dot3(u, n);
        const d3 un = n * s_un;
        const d3 ut = u - un;
        const double utn = sqrt(nrm2(ut));
        d3 t = z3();
        if (utn != 0.0) t = ut / utn;
        const d3 Fn = (dn * -1.0) * L.Kn - un * Gn;
        const double Fnn = sqrt(nrm2(Fn));
        const d3 Ft0 = ut * -Gt;
        const bool cap = sqrt(nrm2(Ft0)) > L.mu * Fnn;
        const d3 Ft = cap ? ((t * Fnn) * -1.0) * L.mu : Ft0;
        const d3 Ff = Ft + Fn;
        // ---- reverse of the law
        const d3 lFa = wallA ? z3() : ld3(lamF + 3 * ra), lFb = wallB ? z3() : ld3(lamF + 3 * rb);
        const d3 ltA = wallA ? z3() : ld3(lamT + 3 * ra), ltB = wallB ? z3() : ld3(lamT + 3 * rb);
        d3 lFf = lFa - lFb;
        d3 ldA = z3(), ldB = z3(), lnegF = z3();
        cross_adj(dA, Ff, ltA, ldA, lFf);
        cross_adj(dB, neg3(Ff), ltB, ldB, lnegF);
        lFf = lFf - lnegF;
        const d3 lFt = lFf;
        d3 lFn = lFf;
        double lKn = 0.0, lmu = 0.0, lGt = 0.0, lGn = 0.0, lFnn = 0.0;
        d3 lut = z3();
        if (cap) {
            lmu += -Fnn * dot3(t, lFt);
            lFnn += -L.mu * dot3(t, lFt);
            const d3 lt = lFt * (-L.mu * Fnn);
            if (utn != 0.0) lut = lut + (lt - t * dot3(t, lt)) / utn;
        } else {
            lGt += -dot3(ut, lFt);
            lut = lut + lFt * -Gt;
        }
        if (Fnn != 0.0) lFn = lFn + (Fn / Fnn) * lFnn;
        lKn += -dot3(dn, lFn);
        const d3 ldn = lFn * -L.Kn;
        lGn += -dot3(un, lFn);
        d3 lun = lFn * -Gn;
        d3 lu = lut;
        lun = lun - lut;
        d3 ln = lun * s_un;
        const double ls_un = dot3(n, lun);
        lu = lu + n * ls_un;
        ln = ln + u * ls_un;
        d3 lvA = lu, lvB = neg3(lu), lwA = z3(), lwB = z3(), lrA = z3(), lrB = z3();
        cross_adj(wA, rA, lu, lwA, lrA);
        cross_adj(wB, rB, neg3(lu), lwB, lrB);
        d3 lcpA = lrA, lcpB = lrB, lxA = neg3(lrA), lxB = neg3(lrB);
        ln = ln + ldn * depth;
        double ldepth = dot3(n, ldn);
        const d3 ln1 = (ln - n * dot3(n, ln)) / nl;
        d3 lnface = ln1 * sgn;
        // coefficients
        double lmt = 0.0, lKt = 0.0, lratio = 0.0, lcn = 0.0, lct = 0.0;
        if (gKt != 0.0) {
            lct += lGt * gKt;
            lmt += lGt * L.ct * Kt / (2.0 * gKt);
            lKt += lGt * L.ct * mt / (2.0 * gKt);
        }
        lKn += lKt * L.ratio * (mt / mn);
        lratio += lKt * L.Kn * (mt / mn);
        lmt += lKt * (L.Kn * L.ratio) / mn;
        lcn += lGn * gKn;
        if (gKn != 0.0) lKn += lGn * L.cn * mn / (2.0 * gKn);
        const double limt = -lmt * mt * mt;
        ldA = ldA + dA * (2.0 * limt * iA);
        ldB = ldB + dB * (2.0 * limt * iB);
        const d3 lcp = ldA + ldB;
        lxA = lxA - ldA;
        lxB = lxB - ldB;
        lcpA = lcpA + lcp / 2.0;
        lcpB = lcpB + lcp / 2.0;
        // d/d theta_k through the law coefficients (their derivatives given
        // with the law tangents; ct depends on et only, which is held fixed
        // unless the seed says otherwise)
        for (int k = 0; k < MWD_N_TAN; ++k)
            cdth[3 * c + k] = (((lKn * LT.Kn[k] + lmu * LT.mu[k]) + lcn * LT.cn[k]) + lct * LT.ct[k]) +
                              lratio * LT.ratio[k];
        // ---- reverse of the face geometry
        d3 la[3] = {z3(), z3(), z3()}, lb[3] = {z3(), z3(), z3()};
        if (geo) {
            const double wgt[3] = {uu, vv, ww};
            double lw_[3];
            for (int k = 0; k < 3; ++k) {
                la[k] = la[k] + lcpA * wgt[k];
                lb[k] = lb[k] + lcpB * wgt[k];
                lw_[k] = dot3(lcpA, av[k]) + dot3(lcpB, bv[k]);
            }
            const double lvv = lw_[1] - lw_[0], lww = lw_[2] - lw_[0];
            const double lv00 = lvv * (-vv * v11 / den) + lww * ((v21 - ww * v11) / den);
            const double lv11 = lvv * ((v20 - vv * v00) / den) + lww * (-ww * v00 / den);
            const double lv01 = lvv * ((-v21 + 2.0 * v01 * vv) / den) + lww * ((-v20 + 2.0 * v01 * ww) / den);
            const double lv20 = lvv * (v11 / den) + lww * (-v01 / den);
            const double lv21 = lvv * (-v01 / den) + lww * (v00 / den);
            const d3 lv0 = (v0 * (2.0 * lv00) + v1 * lv01) + v2 * lv20;
            const d3 lv1 = (v0 * lv01 + v1 * (2.0 * lv11)) + v2 * lv21;
            const d3 lv2 = v0 * lv20 + v1 * lv21;
            d3 lS[3] = {z3(), z3(), z3()};
            if (swapped) {
                lS[2] = lS[2] + lv0;
                lS[1] = lS[1] + lv1;
            } else {
                lS[1] = lS[1] + lv0;
                lS[2] = lS[2] + lv1;
            }
            lS[0] = lS[0] - (lv0 + lv1) - lv2;
            // Proj = nface * depth
            lnface = lnface + lv2 * depth;
            ldepth += dot3(nface, lv2);
            // flip: depth = sg * ddraw, nface = sg * nraw
            const double lddraw = ldepth * sg;
            d3 lnraw = lnface * sg;
            lnraw = lnraw + S[0] * lddraw;
            lS[0] = lS[0] + nraw * lddraw;
            const d3 lnn0 = (lnraw - nraw * dot3(nraw, lnraw)) / len;
            d3 lE1 = z3(), lE2 = z3();
            cross_adj(E1, E2, lnn0, lE1, lE2);
            lS[e1] = lS[e1] + lE1 - lE2;
            lS[e2] = lS[e2] + lE2;
            lS[0] = lS[0] - lE1;
            for (int k = 0; k < 3; ++k) {
                la[k] = la[k] + lS[k];
                lb[k] = lb[k] - lS[k];
            }
        }
        // ---- per side: com, v, w, the three face vertices
        st3(oa, lxA);
        st3(oa + 3, lvA);
        st3(oa + 6, lwA);
        st3(ob, lxB);
        st3(ob + 3, lvB);
        st3(ob + 6, lwB);
        for (int k = 0; k < 3; ++k) {
            st3(oa + 9 + 3 * k, la[k]);
            st3(ob + 9 + 3 * k, lb[k]);
        }
    }
}

// ------------------------------------------------------------- gathering
// Per moving row (GL lanes, as k_force_sum): its contacts' contributions
// added in the row's record order -- to its com, v and w adjoints (one lane
// per component, from the value there), to its share of d/d theta (the
// contacts it holds the A side of, the B side when A is a wall; from zero),
// and to the adjoints of the face vertices (one lane per vertex the faces
// name, at its first mention, adding every mention in order); the bucket
// emptied. The same additions in the same order as a walk over the sorted
// records, so the same bits.
constexpr int GO = 21;  // per record in LDS: com, v, w, three face vertices, d/d theta
__global__ void __launch_bounds__(GB) k_gather_adj(const int32_t* __restrict__ plist, int64_t M,
                                                   int32_t* __restrict__ rec_cnt,
                                                   const int32_t* __restrict__ rec_bkt,
                                                   const int32_t* __restrict__ rec_ovf, const int64_t* n_ovf,
                                                   int64_t ovf_cap, const int32_t* __restrict__ sidx,
                                                   const double* __restrict__ cadj,
                                                   const int32_t* __restrict__ face, Lam lam, int vmax,
                                                   const int32_t* __restrict__ pairs,
                                                   const uint8_t* __restrict__ is_wall,
                                                   const double* __restrict__ cdth, double* __restrict__ rowth) {
    __shared__ int32_t s_rec[GROWS][RC];
    __shared__ double s_o[GROWS][RC][GO];
    __shared__ int32_t s_f[GROWS][RC][4];  // the side's three face vertices; whether d/d theta counts
    const int g = threadIdx.x / GL, lane = threadIdx.x % GL;
    const int64_t i = (int64_t)blockIdx.x * GROWS + g;
    const bool live = i < M;
    const int64_t P = live ? plist[i] : 0;
    const int n = live ? rec_cnt[P] : 0;
    const bool grouped = n <= RC;
    group_bucket(rec_bkt + P * RC, n, grouped, lane, s_rec[g]);
    if (grouped)
        for (int j = lane; j < n; j += GL) {
            const int32_t r = s_rec[g][j];
            const int k = rank_in(s_rec[g], n, r);
            const int64_t c = sidx[r >> 1];
            const int side = r & 1;
            const double* o = cadj + (2 * c + side) * CADJ;
            for (int t = 0; t < CADJ; ++t) s_o[g][k][t] = o[t];
            for (int t = 0; t < MWD_N_TAN; ++t) s_o[g][k][CADJ + t] = cdth[MWD_N_TAN * c + t];
            const int32_t* f = face + 7 * c + 3 * side;
            for (int t = 0; t < 3; ++t) s_f[g][k][t] = f[t];
            s_f[g][k][3] = side == (is_wall[pairs[2 * (int64_t)(r >> 1)]] ? 1 : 0);
        }
    __syncthreads();
    if (!live) return;
    if (grouped) {
        if (lane < 9) {
            double* dst = (lane < 3 ? lam.com : lane < 6 ? lam.v : lam.w) + 3 * P + lane % 3;
            double acc = *dst;
            for (int k = 0; k < n; ++k) acc += s_o[g][k][lane];
            if (n > 0) *dst = acc;
        } else if (lane < 9 + MWD_N_TAN) {
            const int t = lane - 9;
            double acc = 0.0;
            for (int k = 0; k < n; ++k)
                if (s_f[g][k][3]) acc += s_o[g][k][CADJ + t];
            rowth[MWD_N_TAN * i + t] = acc;
        }
        // face vertices: mention e = 3k + s is vertex s of the face of the
        // record ranked k
        for (int e = lane; e < 3 * n; e += GL) {
            const int v = s_f[g][e / 3][e % 3];
            if (v < 0 || v >= vmax) continue;
            bool first = true;
            for (int e2 = 0; e2 < e && first; ++e2) first = s_f[g][e2 / 3][e2 % 3] != v;
            if (!first) continue;
            double* dst = lam.verts + 3 * ((int64_t)vmax * P + v);
            double a0 = dst[0], a1 = dst[1], a2 = dst[2];
            for (int e2 = e; e2 < 3 * n; ++e2)
                if (s_f[g][e2 / 3][e2 % 3] == v) {
                    const double* o = s_o[g][e2 / 3] + 9 + 3 * (e2 % 3);
                    a0 += o[0];
                    a1 += o[1];
                    a2 += o[2];
                }
            dst[0] = a0;
            dst[1] = a1;
            dst[2] = a2;
        }
        if (lane == 0) rec_cnt[P] = 0;
    } else if (lane == 0) {
        double th[MWD_N_TAN] = {};
        for_row_records(P, n, rec_bkt, rec_ovf, n_ovf, ovf_cap, [&](int32_t r) {
            const int64_t c = sidx[r >> 1];
            const int side = r & 1;
            if (side == (is_wall[pairs[2 * (int64_t)(r >> 1)]] ? 1 : 0))
                for (int k = 0; k < MWD_N_TAN; ++k) th[k] += cdth[MWD_N_TAN * c + k];
            const double* o = cadj + (2 * c + side) * CADJ;
            add3(lam.com + 3 * P, ld3(o));
            add3(lam.v + 3 * P, ld3(o + 3));
            add3(lam.w + 3 * P, ld3(o + 6));
            const int32_t* f = face + 7 * c + 3 * side;
            for (int k = 0; k < 3; ++k)
                if (f[k] >= 0 && f[k] < vmax) add3(lam.verts + 3 * ((int64_t)vmax * P + f[k]), ld3(o + 9 + 3 * k));
        });
        rec_cnt[P] = 0;
        for (int k = 0; k < MWD_N_TAN; ++k) rowth[MWD_N_TAN * i + k] = th[k];
    }
}

// The rows' shares of d/d theta, summed by one block in a fixed order (thread
// t takes the moving rows t, t + TSUM, ..., then a fixed tree) and added to
// the accumulator: the same sum whatever order GJK gave the contact slots.
constexpr int TSUM = 256;
__global__ void __launch_bounds__(TSUM) k_theta_sum(const double* __restrict__ rowth,
                                                    const int32_t* __restrict__ plist, int64_t M,
                                                    double* __restrict__ dtheta) {
    __shared__ double s[MWD_N_TAN][TSUM];
    double acc[MWD_N_TAN] = {};
#pragma unroll 8
    for (int64_t i = threadIdx.x; i < M; i += TSUM)
        for (int k = 0; k < MWD_N_TAN; ++k) acc[k] += rowth[MWD_N_TAN * i + k];
    for (int k = 0; k < MWD_N_TAN; ++k) s[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int h = TSUM / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
            for (int k = 0; k < MWD_N_TAN; ++k) s[k][threadIdx.x] += s[k][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < MWD_N_TAN; ++k) dtheta[k] += s[k][0];
}

// ------------------------------------------------------------ the tape
// The reverse sweep's tape holds, per recorded step, the step's input state
// (the evolving state arrays, the bounding radii and the wall centres) and
// what the adjoint step reads of the contact stages' outputs: the contact
// counts, the per-row records, and per contact its pair, the pair's two rows,
// depth, normal, contact points, force, torques and EPA's final face. A
// recorded step pushes after its contact stages (which leave the state as it
// came in); a taped adjoint step pops before its kernels. A slot is a header of
// the three counts (pairs, contacts, overflow records; int64) and then each
// array at a 256-byte boundary; an array keeps the prefix its count covers
// (all of it for the per-row arrays). The candidate pairs themselves are not
// kept: a pop writes back, for each contact c of pair p, sidx[p] = c and the
// pair's rows, which is all the adjoint step reads of them.
constexpr int TAPE_MAX = 32;
// Blocks finishing a tape launch count themselves in TAPE_SUB counters (by
// block number), each on its own 128-byte line, and the last of each counter
// in one more: same-address atomics serialise, so one counter for every
// block would cost more than the copy.
constexpr int TAPE_SUB = 16;
constexpr int TAPE_SYNC_STRIDE = 32;
static_assert((TAPE_SUB + 1) * TAPE_SYNC_STRIDE <= MWD_TAPE_SYNC_WORDS, "tape block counters");
constexpr int TAPE_HEAD = 64;   // words (uint32) before the first array
// the counters of the pairs, the contacts and the overflow records
__device__ __forceinline__ int tape_ctr(int k) { return k == 0 ? 3 : (k == 1 ? 4 : 11); }

struct TapeArr {
    uint32_t* ws;   // the workspace array
    int64_t off;    // its offset in a slot (words, a multiple of 64)
    int64_t unit;   // words per element
    int64_t cap;    // elements at most
    int cnt;        // -1: all cap elements; else the count (0 pairs, 1 contacts, 2 overflow records)
    int rows;       // 1: per contact, the rows of its pair (ws: the pairs array)
    int vec;        // the workspace array is 16-byte aligned (16-byte moves)
};

struct TapeDesc {
    TapeArr a[TAPE_MAX];
    int n;
    uint32_t* base;
    int64_t slot_words, slots;
    int64_t* ctr;      // workspace counters; [18] steps held
    uint32_t* sync;    // [TAPE_SUB + 1] * TAPE_SYNC_STRIDE: blocks of this launch done
    uint32_t* flags;
    int32_t* sidx;     // per candidate pair: its contact slot
    const int32_t* cpair;  // per contact slot: its pair
    int64_t cpair_off;     // the slot offset of the contacts' pairs
    int64_t pair_cap;
};

// The slot layout of a workspace (base and slot count left to the caller);
// returns the words per slot. st: the state whose arrays the slot holds
// (NULL: the layout only).
int64_t tape_layout(const mwd_workspace* ws, const mwd_state* st, TapeDesc& d) {
    d = TapeDesc{};
    int64_t off = TAPE_HEAD;
    auto add = [&](void* p, int64_t unit, int64_t cap, int cnt, int rows = 0) {
        TapeArr& t = d.a[d.n++];
        t.ws = (uint32_t*)p;
        t.off = off;
        t.unit = unit;
        t.cap = cap;
        t.cnt = cnt;
        t.rows = rows;
        t.vec = ((uintptr_t)p & 15) == 0;
        off += (unit * cap + 63) / 64 * 64;
    };
    const int64_t C = ws->contact_cap, N = ws->n_rows;
    const mwd_state z{};
    const mwd_state& a = st ? *st : z;
    add(a.com, 6, N, -1);
    add(a.v, 6, N, -1);
    add(a.v12, 6, N, -1);
    add(a.w, 6, N, -1);
    add(a.wbf, 6, N, -1);
    add(a.w12bf, 6, N, -1);
    add(a.q, 8, N, -1);
    add(a.q12, 8, N, -1);
    add(a.qall, 8, N, -1);
    add(a.dxall, 6, N, -1);
    add(a.verts, 6 * (int64_t)ws->vmax, N, -1);
    add(a.radius, 2, N, -1);  // (when the state has radii)
    add(const_cast<double*>(a.wall_c), 6, N, -1);
    add(ws->rec_cnt, 1, ws->n_rows, -1);
    add(ws->rec_bkt, MWD_ROW_CONTACTS, ws->n_rows, -1);
    add(ws->rec_ovf, 2, 2 * C, 2);
    d.cpair_off = off;
    add(ws->contact_pair, 1, C, 1);
    add(ws->pairs, 2, C, 1, 1);
    add(ws->adj_face, 7, C, 1);
    add(ws->depth, 2, C, 1);
    add(ws->normal, 6, C, 1);
    add(ws->cpA, 6, C, 1);
    add(ws->cpB, 6, C, 1);
    add(ws->Fn, 6, C, 1);
    add(ws->tA, 6, C, 1);
    add(ws->tB, 6, C, 1);
    d.ctr = ws->counters;
    d.flags = ws->flags;
    d.sidx = ws->sidx;
    d.cpair = ws->contact_pair;
    d.pair_cap = ws->pair_cap;
    return off;
}

// The live part of every array in quads (four 32-bit words; a contact pair's
// two rows count as half a quad), cut into units of TAPE_UQ quads; the
// blocks take the units in turn, each thread TAPE_QPT quads of a unit a block
// apart, all its loads issued before its stores. PUSH copies the workspace
// and state arrays into slot counters[18] and raises it; a pop lowers it and
// copies that slot back (counts included). The last block to finish moves the
// slot count, so every block has read it before it changes. A push onto a
// full tape or a pop from an empty one copies nothing and raises MWD_F_TAPE.
constexpr int TAPE_QPT = 4;
constexpr int TAPE_UQ = 256 * TAPE_QPT;
constexpr int TAPE_BLOCKS = 256;

template <bool PUSH>
__global__ void __launch_bounds__(256) k_tape(TapeDesc d) {
    __shared__ int64_t s_slot;
    __shared__ int64_t s_cnt[3];
    __shared__ int64_t s_u0[TAPE_MAX + 1];  // each array's first unit; [n]: all of them
    __shared__ int64_t s_w[TAPE_MAX];       // its live words (pair rows: contacts)
    __shared__ int64_t s_q[TAPE_MAX];       // its live quads
    __shared__ TapeArr s_a[TAPE_MAX];
#pragma unroll
    for (int a = 0; a < TAPE_MAX; ++a)
        if ((int)threadIdx.x == a && a < d.n) s_a[a] = d.a[a];
    __syncthreads();
    if (threadIdx.x < 64) {  // wave 0: the slot, the counts, and lane a sizes array a
        const int lane = threadIdx.x;
        const int64_t held = *(volatile int64_t*)&d.ctr[18];
        const int64_t slot = PUSH ? held : held - 1;
        const bool ok = slot >= 0 && slot < d.slots;
        int64_t cnt[3] = {0, 0, 0};
        if (ok) {
            const int64_t* head = (const int64_t*)(d.base + slot * d.slot_words);
            for (int k = 0; k < 3; ++k) cnt[k] = PUSH ? d.ctr[tape_ctr(k)] : head[k];
        }
        int64_t w = 0, q = 0;
        if (lane < d.n && ok) {
            const TapeArr& t = s_a[lane];
            int64_t m = t.cap;
            if (t.cnt >= 0) {
                const int64_t c = t.cnt == 0 ? cnt[0] : (t.cnt == 1 ? cnt[1] : cnt[2]);
                m = c < 0 ? 0 : (c < m ? c : m);
            }
            if (!t.ws) m = 0;  // a state without radii
            w = t.rows ? m : m * t.unit;
            q = t.rows ? (w + 1) / 2 : (w + 3) / 4;
        }
        const int64_t u = (q + TAPE_UQ - 1) / TAPE_UQ;
        int64_t incl = u;  // inclusive scan of the units over the lanes
        for (int o = 1; o < 64; o <<= 1) {
            const int64_t v = __shfl_up(incl, o, 64);
            if (lane >= o) incl += v;
        }
        if (lane < d.n) {
            s_w[lane] = w;
            s_q[lane] = q;
            s_u0[lane] = incl - u;
        }
        if (lane == d.n - 1) s_u0[d.n] = incl;
        if (lane == 0) {
            s_slot = ok ? slot : -1;
            for (int k = 0; k < 3; ++k) s_cnt[k] = cnt[k];
        }
    }
    __syncthreads();
    const int64_t slot = s_slot;
    if (slot >= 0) {
        uint32_t* sw = d.base + slot * d.slot_words;
        if (blockIdx.x == 0 && threadIdx.x < 3) {
            int64_t* head = (int64_t*)sw;
            if (PUSH) head[threadIdx.x] = s_cnt[threadIdx.x];
            else d.ctr[tape_ctr(threadIdx.x)] = s_cnt[threadIdx.x];
        }
        const int n = d.n;
        for (int64_t unit = blockIdx.x; unit < s_u0[n]; unit += gridDim.x) {
            int lo = 0, hi = n - 1;  // the array holding this unit: the last with s_u0 <= unit
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_u0[mid] <= unit) lo = mid;
                else hi = mid - 1;
            }
            const TapeArr& t = s_a[lo];
            const int64_t w = s_w[lo], q = s_q[lo];
            const int64_t q0 = (unit - s_u0[lo]) * TAPE_UQ + threadIdx.x;
            if (t.rows) {  // contacts 2lq, 2lq+1: their pair rows, gathered from or scattered to the pair list
                int2* rw = (int2*)(sw + t.off);
                int2* pr = (int2*)t.ws;
                const int32_t* cp = PUSH ? d.cpair : (const int32_t*)(sw + d.cpair_off);
                for (int k = 0; k < TAPE_QPT; ++k)
                    for (int64_t c = 2 * (q0 + k * 256); c < 2 * (q0 + k * 256) + 2 && c < w; ++c) {
                        const int32_t p = cp[c];
                        if (p >= 0 && p < d.pair_cap) {
                            if (PUSH) {
                                rw[c] = pr[p];
                            } else {
                                pr[p] = rw[c];
                                d.sidx[p] = (int32_t)c;
                            }
                        }
                    }
                continue;
            }
            uint32_t* dst = PUSH ? sw + t.off : t.ws;
            const uint32_t* src = PUSH ? t.ws : sw + t.off;
            uint4 val[TAPE_QPT];
#pragma unroll
            for (int k = 0; k < TAPE_QPT; ++k) {
                const int64_t lq = q0 + k * 256;
                val[k] = make_uint4(0, 0, 0, 0);
                if (lq < q && t.vec && 4 * lq + 4 <= w) val[k] = ((const uint4*)src)[lq];
            }
#pragma unroll
            for (int k = 0; k < TAPE_QPT; ++k) {
                const int64_t lq = q0 + k * 256;
                if (lq >= q) continue;
                if (t.vec && 4 * lq + 4 <= w) {  // the slot's arrays start on 256-byte boundaries
                    ((uint4*)dst)[lq] = val[k];
                } else {
                    for (int64_t j = 4 * lq; j < 4 * lq + 4 && j < w; ++j) dst[j] = src[j];
                }
            }
        }
    }
    __syncthreads();
    // no fence: the slot count moves after every block has read it, and no
    // block reads what another copies
    if (threadIdx.x == 0) {
        const unsigned nb = gridDim.x;
        const unsigned sub = blockIdx.x % TAPE_SUB;
        const unsigned in_sub = nb / TAPE_SUB + (sub < nb % TAPE_SUB ? 1u : 0u);
        uint32_t* cs = d.sync + sub * TAPE_SYNC_STRIDE;
        if (atomicAdd(cs, 1u) == in_sub - 1) {
            *cs = 0;
            uint32_t* ct = d.sync + TAPE_SUB * TAPE_SYNC_STRIDE;
            const unsigned subs = nb < (unsigned)TAPE_SUB ? nb : (unsigned)TAPE_SUB;
            if (atomicAdd(ct, 1u) == subs - 1) {
                *ct = 0;
                if (slot >= 0) d.ctr[18] += PUSH ? 1 : -1;
                else atomicOr(d.flags, MWD_F_TAPE);
            }
        }
    }
}

__global__ void k_tape_rewind(int64_t* ctr, uint32_t* sync) {
    if (threadIdx.x == 0) ctr[18] = 0;
    for (int k = threadIdx.x; k <= TAPE_SUB; k += blockDim.x) sync[k * TAPE_SYNC_STRIDE] = 0;
}

}  // namespace

int64_t mwd_tape_words(const mwd_workspace* ws) {
    TapeDesc d;
    return tape_layout(ws, nullptr, d);
}

int mwd_launch_tape(mwd_workspace* ws, const mwd_state* st, bool push) {
    TapeDesc d;
    d.slot_words = tape_layout(ws, st, d);
    d.base = ws->tape;
    d.slots = ws->tape_slots;
    d.sync = ws->tape_sync;
    if (!d.base || !d.sync || d.slot_words != ws->tape_words) {
        mwd_set_error("the tape is not allocated for this workspace (mwd_tape_init)");
        return MWD_E_INVALID;
    }
    // units at full capacity bound the grid (fixed, so it can be captured)
    int64_t units = 0;
    for (int a = 0; a < d.n; ++a)
        units += ((d.a[a].rows ? (d.a[a].cap + 1) / 2 : (d.a[a].cap * d.a[a].unit + 3) / 4) + TAPE_UQ - 1) / TAPE_UQ;
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(units, TAPE_BLOCKS));
    if (push) hipLaunchKernelGGL(k_tape<true>, dim3(grid), dim3(256), 0, ws->stream, d);
    else hipLaunchKernelGGL(k_tape<false>, dim3(grid), dim3(256), 0, ws->stream, d);
    MWD_LAUNCH_CHECK();
    return MWD_OK;
}

int mwd_launch_tape_rewind(mwd_workspace* ws) {
    if (!ws->tape_sync) {
        mwd_set_error("the tape is not allocated for this workspace (mwd_tape_init)");
        return MWD_E_INVALID;
    }
    hipLaunchKernelGGL(k_tape_rewind, dim3(1), dim3(64), 0, ws->stream, ws->counters, ws->tape_sync);
    MWD_LAUNCH_CHECK();
    return MWD_OK;
}

int mwd_launch_step_adjoint(mwd_workspace* ws, const mwd_state* st, const BodyView& A, const BodyView& B,
                            double* const* lam_arr, double* dtheta) {
    hipStream_t s = ws->stream;
    const int64_t M = ws->n_moving;
    int64_t* ctr = ws->counters;
    Lam lam{lam_arr[MWD_TAN_COM], lam_arr[MWD_TAN_V],   lam_arr[MWD_TAN_V12],  lam_arr[MWD_TAN_W],
            lam_arr[MWD_TAN_WBF], lam_arr[MWD_TAN_W12BF], lam_arr[MWD_TAN_Q],  lam_arr[MWD_TAN_Q12],
            lam_arr[MWD_TAN_QALL], lam_arr[MWD_TAN_DXALL], lam_arr[MWD_TAN_VERTS]};
    const VState a = mwd_vstate(st, nullptr);
    if (M > 0) {
        hipLaunchKernelGGL(k_force_sum, dim3((unsigned)((M + GROWS - 1) / GROWS)), dim3(GB), 0, s, ws->plist, M,
                           ws->rec_cnt,
                           ws->rec_bkt, ws->rec_ovf, ctr + 11, 2 * ws->contact_cap, ws->sidx, ws->Fn, ws->tA,
                           ws->tB, st->fext, st->text, ws->adj_F, ws->adj_T);
        MWD_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_verlet_adj, dim3((unsigned)((M + 63) / 64)), dim3(64), 0, s, a, ws->adj_F, ws->adj_T, lam,
                           ws->adj_lamF, ws->adj_lamT, ws->plist, M, ws->scene_of, ws->scenes_dev, st->nverts,
                           (int)st->vmax, ws->dt);
        MWD_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_contact_adj, dim3((unsigned)std::min<int64_t>((ws->contact_cap + 63) / 64, 8192)), dim3(64), 0, s, A, B, ws->laws,
                       ws->law_tans, ws->contact_pair, ctr + 4, ws->contact_cap, ws->adj_face, ws->depth, ws->normal,
                       ws->cpA, ws->cpB, ws->adj_lamF, ws->adj_lamT, ws->adj_c, ws->adj_cth);
    MWD_LAUNCH_CHECK();
    if (M > 0) {
        hipLaunchKernelGGL(k_gather_adj, dim3((unsigned)((M + GROWS - 1) / GROWS)), dim3(GB), 0, s, ws->plist, M,
                           ws->rec_cnt,
                           ws->rec_bkt, ws->rec_ovf, ctr + 11, 2 * ws->contact_cap, ws->sidx, ws->adj_c,
                           ws->adj_face, lam, (int)st->vmax, ws->pairs, ws->is_wall, ws->adj_cth, ws->adj_rowth);
        MWD_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_theta_sum, dim3(1), dim3(TSUM), 0, s, ws->adj_rowth, ws->plist, M, dtheta);
        MWD_LAUNCH_CHECK();
    }
    return MWD_OK;
}
