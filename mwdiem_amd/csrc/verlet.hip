// Quaternion velocity-Verlet integrator, one particle per lane (FP64).
//
// Semantics follow the reference's Verlet kernel, LIBRARIES/gpu_Verlet.cu:164-403:
// half-step velocity, Euler's rigid-body equation in the body frame, the
// q12 / q / q_all updates, world angular velocity, rotation of the stored world
// vertices about the centre of mass, translation, and the out-of-box reset
// (velocities zeroed, last translation undone, rotation kept, dx_all keeps dx).
//
// The stepper variant fuses the force assembly (MWDiem.py:977-981): the
// per-particle contact force and torque are summed over the particle's
// contact records in ascending pair order (its bucket of records, sorted),
// then the external force/torque is added, then the step runs.
//
// Both are templated on the scalar: double for the forward path, Dual<3> for
// the tangent path (derivatives with respect to Kn, en and mu carried along;
// values bit-identical to the forward path).
#include <type_traits>

#include "mwd_internal.h"

namespace {

// What the vertex pass needs of a particle's step: the rotation, the old
// centre, the translation, whether the out-of-box reset undoes it, and the
// new centre (for the bounding radius).
struct VertJob {
    q4 qr;
    d3 com, dx, cnew;
    int32_t nrot, out;
};

// rot_mode 0: vertices rotated by the absolute orientation q (the reference's
// compounding rotation); 1: by the step's increment q q_old^-1. With job
// given (T = double only), the vertex update is left to the caller's
// block-wide vertex pass and described in *job instead.
template <class T>
__device__ __forceinline__ void verlet_row(const VState& a, int64_t P, V3<T> F, V3<T> tau, int vmax, int nrot,
                                           double dt, const double* box, int rot_mode, VertJob* job = nullptr) {
    V3<T> v12 = ldv<T>(a.v12, P);
    V3<T> w12bf = ldv<T>(a.w12bf, P);
    V3<T> wbf = ldv<T>(a.wbf, P);
    V3<T> w = ldv<T>(a.w, P);
    V3<T> com = ldv<T>(a.com, P);
    const double m = a.mass[P];
    Q4<T> q = ldqT<T>(a.q, P);
    Q4<T> q12 = ldqT<T>(a.q12, P);
    V3<T> dxall = ldv<T>(a.dxall, P);

    V3<T> v32 = v12 + (F / m) * dt;
    Q4<T> qi = qconj(q);
    V3<T> tbf = qrot(qi, tau, q);
    V3<T> p1 = mat3v(a.I + 9 * P, wbf);
    V3<T> p2 = tbf - cross3(wbf, p1);
    V3<T> abf = mat3v(a.Iinv + 9 * P, p2);
    V3<T> w32bf = w12bf + abf * dt;
    T wn = dsqrt(nrm2(w));
    V3<T> t11 = zero3<T>();
    if (wn != 0.0) t11 = (w / wn) * dsin((wn * dt) / 2.0);
    T t112 = dcos((wn * dt) / 2.0);
    Q4<T> q32 = qmul(mkq(t112, t11), q12);
    v12 = v32;
    q12 = q32;
    w12bf = w32bf;
    V3<T> dx = v12 * dt;
    dxall = dxall + dx;
    V3<T> w34bf = w12bf + (abf * dt) / 4.0;
    V3<T> w34 = qrot(q12, w34bf, qconj(q12));
    T n34 = dsqrt(nrm2(w34));
    V3<T> t1 = zero3<T>();
    if (n34 != 0.0) t1 = (w34 / n34) * dsin((n34 * dt) / 4.0);
    T t12 = dcos((n34 * dt) / 4.0);
    Q4<T> qold = q;
    q = qmul(mkq(t12, t1), q12);
    Q4<T> qall = qmul(q, qold);
    wbf = w12bf + (abf * dt) / 2.0;
    w = qrot(q, wbf, qconj(q));
    V3<T> vel = v12 + (F * dt) / m;
    const Q4<T> qr = rot_mode == 1 ? qmul(q, qconj(qold)) : q;
    const Q4<T> qc = qconj(qr);
    const V3<T> c1 = com + dx;
    const bool out = c1.x > box[0] || c1.x < box[1] || c1.y > box[2] || c1.y < box[3] || c1.z > box[4] ||
                     c1.z < box[5];
    const V3<T> du = dx * -1.0;
    V3<T> cnew = c1;
    if (out) cnew = cnew + du;
    if constexpr (std::is_same<T, double>::value) {
        if (job) {
            job->qr = qr;
            job->com = com;
            job->dx = dx;
            job->cnew = cnew;
            job->nrot = nrot;
            job->out = out ? 1 : 0;
        }
    }
    if (!job) {
        double r2 = 0.0;
        const int64_t vb = (int64_t)vmax * P;
        for (int i = 0; i < nrot; ++i) {
            V3<T> r = qrot(qr, ldv<T>(a.verts, vb + i) - com, qc);
            V3<T> nv = (r + com) + dx;
            if (out) nv = nv + du;
            stv(a.verts, vb + i, nv);
            r2 = fmax(r2, val(nrm2(nv - cnew)));
        }
        if (a.radius) a.radius[P] = inflate_radius(r2, mk3(val(cnew.x), val(cnew.y), val(cnew.z)));
    }
    com = com + dx;
    if (out) {
        vel = zero3<T>();
        v12 = zero3<T>();
        w12bf = zero3<T>();
        wbf = zero3<T>();
        w = zero3<T>();
        com = com + du;
    }
    stv(a.v, P, vel);
    stv(a.v12, P, v12);
    stv(a.w12bf, P, w12bf);
    stv(a.com, P, com);
    stv(a.dxall, P, dxall);
    stv(a.wbf, P, wbf);
    stv(a.w, P, w);
    stqT(a.q12, P, q12);
    stqT(a.q, P, q);
    stqT(a.qall, P, qall);
}

struct Box6 {
    double b[6];
};

constexpr int VBLOCK = 256;

// Block-wide vertex pass: the block's particles (rows[j], job[j]) have their
// vertex slots updated by consecutive threads, so a wave reads and writes
// consecutive vertices (the per-lane loop strides by a whole particle). Each
// vertex is transformed by exactly the operations of the per-lane loop; the
// bounding radius takes the maximum through an LDS atomic on the bits of the
// non-negative squared distances (NaNs skipped, as fmax skips them).
// New position of one vertex: exactly the operations of the per-lane loop.
__device__ __forceinline__ d3 vertex_move(const VertJob& J, d3 p) {
    const q4 qr = J.qr;
    const d3 r = qrot(qr, p - J.com, qconj(qr));
    d3 nv = (r + J.com) + J.dx;
    if (J.out) nv = nv + J.dx * -1.0;
    return nv;
}

// PACK false: thread idx takes slot (idx / vmax, idx % vmax) of the block's
// nb x vmax slots and skips slots at or above the particle's count. PACK
// true: the block's vertex counts are prefix-summed first (vbase) and thread
// idx takes the idx-th real vertex, so no lane idles on padding slots.
template <int VB, bool PACK>
__device__ __forceinline__ void vertex_pass(double* __restrict__ verts, const int64_t* rows, const VertJob* job,
                                            unsigned long long* r2bits, int* vbase, int nb, int vmax, double* radius) {
    const int tid = threadIdx.x;
    r2bits[tid] = 0ull;
    if constexpr (!PACK) {
        __syncthreads();
        const int total = nb * vmax;
        for (int idx = tid; idx < total; idx += VB) {
            const int j = idx / vmax;
            const int i = idx - j * vmax;
            if (i >= job[j].nrot) continue;
            const d3 nv = vertex_move(job[j], ld3(verts + ((int64_t)vmax * rows[j] + i) * 3));
            st3(verts + ((int64_t)vmax * rows[j] + i) * 3, nv);
            if (radius) {
                const double r2 = nrm2(nv - job[j].cnew);
                if (r2 == r2) atomicMax(&r2bits[j], (unsigned long long)__double_as_longlong(r2));
            }
        }
    } else {
        // block-wide exclusive scan of the vertex counts (wave scans, then
        // the totals of the waves before)
        constexpr int NW = VB / MWD_WAVE;
        __shared__ int wsum[NW];
        const int lane = tid & (MWD_WAVE - 1);
        const int n = tid < nb ? job[tid].nrot : 0;
        int x = n;
#pragma unroll
        for (int o = 1; o < MWD_WAVE; o <<= 1) {
            const int y = __shfl_up(x, o, MWD_WAVE);
            if (lane >= o) x += y;
        }
        if (lane == MWD_WAVE - 1) wsum[tid / MWD_WAVE] = x;
        __syncthreads();
        int before = 0, total = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            if (w < tid / MWD_WAVE) before += wsum[w];
            total += wsum[w];
        }
        vbase[tid] = before + x - n;
        __syncthreads();
        // VU vertices per thread and round: a thread issues the loads of VU
        // vertices before it stores any of them (the compiler would not move
        // a later round's loads above a store to the same array by itself)
        constexpr int VU = 4;
        for (int base = tid; base < total; base += VB * VU) {
            int jj[VU];
            double* pv[VU];
            d3 p[VU];
#pragma unroll
            for (int u = 0; u < VU; ++u) {
                const int idx = base + u * VB;
                // the last particle whose first vertex is at or before idx
                int lo = 0, hi = nb - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (vbase[mid] <= idx)
                        lo = mid;
                    else
                        hi = mid - 1;
                }
                jj[u] = idx < total ? lo : -1;
                pv[u] = verts + ((int64_t)vmax * rows[lo] + (idx - vbase[lo])) * 3;
                if (jj[u] >= 0) p[u] = ld3(pv[u]);
            }
#pragma unroll
            for (int u = 0; u < VU; ++u) {
                const int j = jj[u];
                if (j < 0) continue;
                const d3 nv = vertex_move(job[j], p[u]);
                st3(pv[u], nv);
                if (radius) {
                    const double r2 = nrm2(nv - job[j].cnew);
                    if (r2 == r2) atomicMax(&r2bits[j], (unsigned long long)__double_as_longlong(r2));
                }
            }
        }
    }
    __syncthreads();
    if (radius && tid < nb)
        radius[rows[tid]] = inflate_radius(__longlong_as_double((long long)r2bits[tid]), job[tid].cnew);
}

__global__ void __launch_bounds__(VBLOCK) k_verlet(VState a, const double* __restrict__ F,
                                                   const double* __restrict__ tau, int vmax,
                                                   const int32_t* __restrict__ nrot, double dt, Box6 box,
                                                   int64_t pmin, int64_t pmax) {
    __shared__ VertJob job[VBLOCK];
    __shared__ int64_t rows[VBLOCK];
    __shared__ unsigned long long r2bits[VBLOCK];
    const int64_t first = pmin + (int64_t)blockIdx.x * VBLOCK;
    const int64_t P = first + threadIdx.x;
    const int nb = (int)(pmax - first < VBLOCK ? pmax - first : VBLOCK);
    if (P < pmax) {
        rows[threadIdx.x] = P;
        verlet_row<double>(a, P, ld3(F + 3 * P), ld3(tau + 3 * P), vmax, nrot ? nrot[P] : vmax, dt, box.b, 0,
                           &job[threadIdx.x]);
    }
    vertex_pass<VBLOCK, false>(a.verts.v, rows, job, r2bits, nullptr, nb, vmax, nullptr);
}

// Record r of a particle: side A (r even) or B (r odd) of the contact of
// candidate pair r >> 1.
template <class T>
__device__ __forceinline__ void add_record(const ContactOut& co, int32_t r, V3<T>& F, V3<T>& Tq) {
    const int64_t c = co.sidx[r >> 1];
    if (r & 1) {
        F = F + neg3(ldv<T>(co.Fn, c));
        Tq = Tq + ldv<T>(co.tB, c);
    } else {
        F = F + ldv<T>(co.Fn, c);
        Tq = Tq + ldv<T>(co.tA, c);
    }
}

// The same record summed in single precision (the precision of the
// reference's force arrays and atomics), still in candidate order.
__device__ __forceinline__ void add_record_f32(const ContactOut& co, int32_t r, float* F, float* Tq) {
    const int64_t c = co.sidx[r >> 1];
    const double* f = co.Fn.v + 3 * c;
    const double* t = (r & 1) ? co.tB.v + 3 * c : co.tA.v + 3 * c;
    const float sg = (r & 1) ? -1.0f : 1.0f;
    for (int k = 0; k < 3; ++k) {
        F[k] += sg * (float)f[k];
        Tq[k] += (float)t[k];
    }
}

// Force assembly fused with the integrator for moving particle i. fp32: the
// contact sums are taken in single precision (double path only).
template <class T>
__device__ __forceinline__ void reduce_verlet_row(const VState& a, const int32_t* __restrict__ plist, int64_t i,
                                                  const int32_t* __restrict__ scene_of,
                                                  const SceneDev* __restrict__ scenes,
                                                  int32_t* __restrict__ rec_cnt,
                                                  const int32_t* __restrict__ rec_bkt,
                                                  const int32_t* __restrict__ rec_ovf, const int64_t* n_ovf,
                                                  int64_t ovf_cap, const ContactOut& co,
                                                  const double* __restrict__ fext, const double* __restrict__ text,
                                                  const int32_t* __restrict__ nverts, int vmax, double dt,
                                                  VertJob* job, bool fp32 = false, int64_t row_base = -1) {
    // one scene: its moving rows are consecutive, entry i is row row_base + i
    const int64_t P = row_base >= 0 ? row_base + i : plist[i];
    V3<T> F = zero3<T>(), Tq = zero3<T>();
    // this step's records of the particle, summed in ascending contact order
    // (the order of the candidate pairs); the bucket is emptied for the next step
    const int n = rec_cnt[P];
    if (n > 0) {
        bool done = false;
        if constexpr (std::is_same<T, double>::value) {
            if (fp32) {
                float f[3] = {0.0f, 0.0f, 0.0f}, t[3] = {0.0f, 0.0f, 0.0f};
                for_row_records(P, n, rec_bkt, rec_ovf, n_ovf, ovf_cap,
                                [&](int32_t r) { add_record_f32(co, r, f, t); });
                F = mk3<T>((double)f[0], (double)f[1], (double)f[2]);
                Tq = mk3<T>((double)t[0], (double)t[1], (double)t[2]);
                done = true;
            }
        }
        if (!done)
            for_row_records(P, n, rec_bkt, rec_ovf, n_ovf, ovf_cap, [&](int32_t r) { add_record<T>(co, r, F, Tq); });
        rec_cnt[P] = 0;
    }
    F = F + mk3<T>(T(fext[3 * P]), T(fext[3 * P + 1]), T(fext[3 * P + 2]));
    Tq = Tq + mk3<T>(T(text[3 * P]), T(text[3 * P + 1]), T(text[3 * P + 2]));
    const SceneDev& sc = scenes[row_base >= 0 ? 0 : scene_of[P]];
    verlet_row<T>(a, P, F, Tq, vmax, nverts[P], dt, sc.box, sc.law_p.mode == MWD_LAW_PHYSICAL ? 1 : 0, job);
}

// The double path: force assembly and integrator per lane, then the block's
// vertex pass; VB particles per block. WPE: the minimum waves per SIMD the
// register allocation must allow (1 leaves the compiler's choice).
template <int VB, int WPE, bool PACK>
__global__ void __launch_bounds__(VB) __attribute__((amdgpu_waves_per_eu(WPE)))
k_reduce_verlet_v(VState a, const int32_t* __restrict__ plist, int64_t M_cap, const int64_t* m_live,
                  const int32_t* __restrict__ scene_of,
                  const SceneDev* __restrict__ scenes, int32_t* __restrict__ rec_cnt,
                  const int32_t* __restrict__ rec_bkt, const int32_t* __restrict__ rec_ovf, const int64_t* n_ovf,
                  int64_t ovf_cap, ContactOut co, const double* __restrict__ fext, const double* __restrict__ text,
                  const int32_t* __restrict__ nverts, int vmax, double dt, int fp32, int64_t row_base) {
    __shared__ VertJob job[VB];
    __shared__ int64_t rows[VB];
    __shared__ unsigned long long r2bits[VB];
    __shared__ int vbase[PACK ? VB : 1];
    const int64_t first = (int64_t)blockIdx.x * VB;
    const int64_t M = *m_live < M_cap ? *m_live : M_cap;  // the moving rows active now
    if (first >= M) return;  // the whole block: none of its rows is active
    const int nb = (int)(M - first < VB ? M - first : VB);
    const int64_t i = first + threadIdx.x;
    if (i < M) {
        rows[threadIdx.x] = row_base >= 0 ? row_base + i : plist[i];
        reduce_verlet_row<double>(a, plist, i, scene_of, scenes, rec_cnt, rec_bkt, rec_ovf, n_ovf, ovf_cap, co, fext,
                                  text, nverts, vmax, dt, &job[threadIdx.x], fp32 != 0, row_base);
    }
    vertex_pass<VB, PACK>(a.verts.v, rows, job, r2bits, vbase, nb, vmax, a.radius);
}

// The tangent path: one lane per particle, vertices and their tangents
// updated in the lane's own loop.
__global__ void __launch_bounds__(256) k_reduce_verlet_t(VState a, const int32_t* __restrict__ plist, int64_t M_cap,
                                                         const int64_t* m_live,
                                                         const int32_t* __restrict__ scene_of,
                                                         const SceneDev* __restrict__ scenes,
                                                         int32_t* __restrict__ rec_cnt,
                                                         const int32_t* __restrict__ rec_bkt,
                                                         const int32_t* __restrict__ rec_ovf, const int64_t* n_ovf,
                                                         int64_t ovf_cap, ContactOut co,
                                                         const double* __restrict__ fext,
                                                         const double* __restrict__ text,
                                                         const int32_t* __restrict__ nverts, int vmax, double dt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t M = *m_live < M_cap ? *m_live : M_cap;
    if (i >= M) return;
    reduce_verlet_row<Dual<MWD_NT>>(a, plist, i, scene_of, scenes, rec_cnt, rec_bkt, rec_ovf, n_ovf, ovf_cap, co,
                                    fext, text,
                                    nverts, vmax, dt, nullptr);
}

__global__ void k_radius(const double* __restrict__ com, const double* __restrict__ verts,
                         const int32_t* __restrict__ nverts, int64_t n, int vmax, double* __restrict__ radius) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const d3 c = ld3(com + 3 * r);
    double r2 = 0.0;
    for (int i = 0; i < nverts[r]; ++i) r2 = fmax(r2, nrm2(ld3(verts + ((int64_t)vmax * r + i) * 3) - c));
    radius[r] = inflate_radius(r2, c);
}

TArr<MWD_NT> tarr(double* v, double* const* t, int k) {
    TArr<MWD_NT> a;
    a.v = v;
    for (int j = 0; j < MWD_NT; ++j) a.t[j] = t ? t[MWD_NT * k + j] : nullptr;
    return a;
}

}  // namespace

VState mwd_vstate(const mwd_state* st, double* const* tan) {
    VState a;
    a.com = tarr(st->com, tan, MWD_TAN_COM);
    a.v = tarr(st->v, tan, MWD_TAN_V);
    a.v12 = tarr(st->v12, tan, MWD_TAN_V12);
    a.w = tarr(st->w, tan, MWD_TAN_W);
    a.wbf = tarr(st->wbf, tan, MWD_TAN_WBF);
    a.w12bf = tarr(st->w12bf, tan, MWD_TAN_W12BF);
    a.q = tarr(st->q, tan, MWD_TAN_Q);
    a.q12 = tarr(st->q12, tan, MWD_TAN_Q12);
    a.qall = tarr(st->qall, tan, MWD_TAN_QALL);
    a.dxall = tarr(st->dxall, tan, MWD_TAN_DXALL);
    a.verts = tarr(st->verts, tan, MWD_TAN_VERTS);
    a.mass = st->mass;
    a.Iinv = st->Iinv;
    a.I = st->I;
    a.radius = st->radius;
    return a;
}

int mwd_launch_verlet(double* v, double* v12, double* w12bf, double* q12, double* q, double* qall, double* wbf,
                      double* w, const double* F, const double* tau, double* verts, double* com,
                      const double* mass, const double* Iinv, const double* I, double* dxall, int32_t vmax,
                      const int32_t* nrot, double dt, const double* box, int64_t pmin, int64_t pmax,
                      hipStream_t s) {
    if (pmax <= pmin) return MWD_OK;
    VState a;
    a.v = tarr(v, nullptr, 0);
    a.v12 = tarr(v12, nullptr, 0);
    a.w12bf = tarr(w12bf, nullptr, 0);
    a.q12 = tarr(q12, nullptr, 0);
    a.q = tarr(q, nullptr, 0);
    a.qall = tarr(qall, nullptr, 0);
    a.wbf = tarr(wbf, nullptr, 0);
    a.w = tarr(w, nullptr, 0);
    a.verts = tarr(verts, nullptr, 0);
    a.com = tarr(com, nullptr, 0);
    a.dxall = tarr(dxall, nullptr, 0);
    a.mass = mass;
    a.Iinv = Iinv;
    a.I = I;
    a.radius = nullptr;
    Box6 b;
    for (int k = 0; k < 6; ++k) b.b[k] = box[k];
    const int64_t n = pmax - pmin;
    hipLaunchKernelGGL(k_verlet, dim3((unsigned)((n + VBLOCK - 1) / VBLOCK)), dim3(VBLOCK), 0, s, a, F, tau,
                       (int)vmax, nrot, dt, b, pmin, pmax);
    MWD_LAUNCH_CHECK();
    return MWD_OK;
}

template <int VB, int WPE, bool PACK>
static void launch_reduce_verlet_v(mwd_workspace* ws, const mwd_state* st, const VState& a, const ContactOut& co) {
    const int64_t M = ws->n_moving;
    MWD_LAUNCH_EV(mwd_kev(ws, MWD_KT_VERLET, 0), mwd_kev(ws, MWD_KT_VERLET, 1), (k_reduce_verlet_v<VB, WPE, PACK>),
                  dim3((unsigned)((M + VB - 1) / VB)), dim3(VB), 0, ws->stream, a, ws->plist, M, ws->m_live,
                  ws->scene_of,
                  ws->scenes_dev, ws->rec_cnt, ws->rec_bkt, ws->rec_ovf, ws->counters + 11, 2 * ws->contact_cap, co,
                  st->fext, st->text, st->nverts, (int)st->vmax, ws->dt, ws->reduce_fp32, ws->row_base);
}

template <int VB, int WPE>
static void launch_reduce_verlet_p(mwd_workspace* ws, const mwd_state* st, const VState& a, const ContactOut& co) {
    if (ws->tune.vertex_pack)
        launch_reduce_verlet_v<VB, WPE, true>(ws, st, a, co);
    else
        launch_reduce_verlet_v<VB, WPE, false>(ws, st, a, co);
}

template <int VB>
static void launch_reduce_verlet_b(mwd_workspace* ws, const mwd_state* st, const VState& a, const ContactOut& co) {
    switch (ws->tune.verlet_waves) {
        case 3:
            launch_reduce_verlet_p<VB, 3>(ws, st, a, co);
            break;
        case 4:
            launch_reduce_verlet_p<VB, 4>(ws, st, a, co);
            break;
        default:
            launch_reduce_verlet_p<VB, 1>(ws, st, a, co);
    }
}

int mwd_launch_reduce_verlet(mwd_workspace* ws, mwd_state* st, double* const* tan, const ContactOut& co) {
    if (ws->n_moving <= 0) return MWD_OK;
    const VState a = mwd_vstate(st, tan);
    const int64_t M = ws->n_moving;
    if (tan)
        hipLaunchKernelGGL(k_reduce_verlet_t, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, ws->stream, a,
                           ws->plist, M, ws->m_live, ws->scene_of, ws->scenes_dev, ws->rec_cnt, ws->rec_bkt, ws->rec_ovf,
                           ws->counters + 11, 2 * ws->contact_cap, co, st->fext, st->text, st->nverts, (int)st->vmax,
                           ws->dt);
    else
        switch (ws->tune.verlet_block) {
            case 64:
                launch_reduce_verlet_b<64>(ws, st, a, co);
                break;
            case 128:
                launch_reduce_verlet_b<128>(ws, st, a, co);
                break;
            default:
                launch_reduce_verlet_b<256>(ws, st, a, co);
        }
    MWD_LAUNCH_CHECK();
    return MWD_OK;
}

namespace {
// Walls moved at their prescribed velocity (a kinematic boundary; the
// reference's walls are static): centre, vertices and slab centre translate
// by v dt at the end of the step.
__global__ void k_move_walls(const int32_t* __restrict__ wlist, int64_t nw, double* __restrict__ com,
                             double* __restrict__ verts, const int32_t* __restrict__ nverts, int vmax,
                             const double* __restrict__ v, double* __restrict__ wall_c, double dt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nw) return;
    const int64_t r = wlist[i];
    const d3 d = ld3(v + 3 * r) * dt;
    if (zero3(d)) return;
    st3(com + 3 * r, ld3(com + 3 * r) + d);
    if (wall_c) st3(wall_c + 3 * r, ld3(wall_c + 3 * r) + d);
    for (int k = 0; k < nverts[r]; ++k) {
        double* p = verts + 3 * ((int64_t)vmax * r + k);
        st3(p, ld3(p) + d);
    }
}
}  // namespace

int mwd_launch_move_walls(mwd_workspace* ws, mwd_state* st) {
    if (ws->n_wall_rows <= 0) return MWD_OK;
    hipLaunchKernelGGL(k_move_walls, dim3((unsigned)((ws->n_wall_rows + 63) / 64)), dim3(64), 0, ws->stream,
                       ws->wlist, ws->n_wall_rows, st->com, st->verts, st->nverts, (int)st->vmax, st->v,
                       const_cast<double*>(st->wall_c), ws->dt);
    MWD_LAUNCH_CHECK();
    return MWD_OK;
}

int mwd_launch_radius(const double* com, const double* verts, const int32_t* nverts, int64_t n, int32_t vmax,
                      double* radius, hipStream_t s) {
    if (n <= 0) return MWD_OK;
    hipLaunchKernelGGL(k_radius, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, com, verts, nverts, n, (int)vmax,
                       radius);
    MWD_LAUNCH_CHECK();
    return MWD_OK;
}
