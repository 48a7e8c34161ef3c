// Quaternion velocity-Verlet integrator, one particle per lane (FP64).
//
// Semantics follow the reference's Verlet kernel, LIBRARIES/gpu_Verlet.cu:164-403:
// half-step velocity, Euler's rigid-body equation in the body frame, the
// q12 / q / q_all updates, world angular velocity, rotation of the stored world
// vertices about the centre of mass by q (which compounds, as the reference's
// does), translation, and the out-of-box reset (velocities zeroed, last
// translation undone, rotation kept, dx_all keeps dx).
//
// The stepper variant fuses the force assembly (MWDiem.py:977-981): the
// per-particle contact force and torque are summed over the particle's
// contact records in ascending pair order (the records were stably sorted by
// row), then the external force/torque is added, then the step runs.
#include "mwd_internal.h"

namespace {

struct VerletArrays {
    double *v, *v12, *w12bf, *q12, *q, *qall, *wbf, *w;
    double* verts;
    double* com;
    const double *mass, *Iinv, *I;
    double* dxall;
    double* radius;  // optional: refreshed bounding radius of the moved vertices
};

// rot_mode 0: vertices rotated by the absolute orientation q (the reference's
// compounding rotation); 1: by the step's increment q q_old^-1.
__device__ __forceinline__ void verlet_row(const VerletArrays& a, int64_t P, d3 F, d3 tau, int vmax, int nrot,
                                           double dt, const double* box, int rot_mode) {
    d3 v12 = ld3(a.v12 + 3 * P);
    d3 w12bf = ld3(a.w12bf + 3 * P);
    d3 wbf = ld3(a.wbf + 3 * P);
    d3 w = ld3(a.w + 3 * P);
    d3 com = ld3(a.com + 3 * P);
    const double m = a.mass[P];
    q4 q = ldq(a.q + 4 * P);
    q4 q12 = ldq(a.q12 + 4 * P);
    d3 dxall = ld3(a.dxall + 3 * P);

    d3 v32 = v12 + (F / m) * dt;
    q4 qi = qconj(q);
    d3 tbf = qrot(qi, tau, q);
    d3 p1 = mat3v(a.I + 9 * P, wbf);
    d3 p2 = tbf - cross3(wbf, p1);
    d3 abf = mat3v(a.Iinv + 9 * P, p2);
    d3 w32bf = w12bf + abf * dt;
    double wn = sqrt(nrm2(w));
    d3 t11 = mk3(0.0, 0.0, 0.0);
    if (wn != 0.0) t11 = (w / wn) * sin((wn * dt) / 2.0);
    double t112 = cos((wn * dt) / 2.0);
    q4 q32 = qmul(mkq(t112, t11), q12);
    v12 = v32;
    q12 = q32;
    w12bf = w32bf;
    d3 dx = v12 * dt;
    dxall = dxall + dx;
    d3 w34bf = w12bf + (abf * dt) / 4.0;
    d3 w34 = qrot(q12, w34bf, qconj(q12));
    double n34 = sqrt(nrm2(w34));
    d3 t1 = mk3(0.0, 0.0, 0.0);
    if (n34 != 0.0) t1 = (w34 / n34) * sin((n34 * dt) / 4.0);
    double t12 = cos((n34 * dt) / 4.0);
    q4 qold = q;
    q = qmul(mkq(t12, t1), q12);
    q4 qall = qmul(q, qold);
    wbf = w12bf + (abf * dt) / 2.0;
    w = qrot(q, wbf, qconj(q));
    d3 vel = v12 + (F * dt) / m;
    const q4 qr = rot_mode == 1 ? qmul(q, qconj(qold)) : q;
    const q4 qc = qconj(qr);
    double* vp = a.verts + (int64_t)vmax * 3 * P;
    const bool out = [&] {
        d3 c1 = com + dx;
        return c1.x > box[0] || c1.x < box[1] || c1.y > box[2] || c1.y < box[3] || c1.z > box[4] || c1.z < box[5];
    }();
    const d3 du = dx * -1.0;
    d3 cnew = com + dx;
    if (out) cnew = cnew + du;
    double r2 = 0.0;
    for (int i = 0; i < nrot; ++i) {
        d3 r = qrot(qr, ld3(vp + 3 * i) - com, qc);
        d3 nv = (r + com) + dx;
        if (out) nv = nv + du;
        st3(vp + 3 * i, nv);
        r2 = fmax(r2, nrm2(nv - cnew));
    }
    if (a.radius) a.radius[P] = inflate_radius(r2, cnew);
    com = com + dx;
    if (out) {
        vel = mk3(0.0, 0.0, 0.0);
        v12 = mk3(0.0, 0.0, 0.0);
        w12bf = mk3(0.0, 0.0, 0.0);
        wbf = mk3(0.0, 0.0, 0.0);
        w = mk3(0.0, 0.0, 0.0);
        com = com + du;
    }
    st3(a.v + 3 * P, vel);
    st3(a.v12 + 3 * P, v12);
    st3(a.w12bf + 3 * P, w12bf);
    st3(a.com + 3 * P, com);
    st3(a.dxall + 3 * P, dxall);
    st3(a.wbf + 3 * P, wbf);
    st3(a.w + 3 * P, w);
    stq(a.q12 + 4 * P, q12);
    stq(a.q + 4 * P, q);
    stq(a.qall + 4 * P, qall);
}

struct Box6 {
    double b[6];
};

__global__ void __launch_bounds__(256) k_verlet(VerletArrays a, const double* __restrict__ F,
                                                const double* __restrict__ tau, int vmax,
                                                const int32_t* __restrict__ nrot, double dt, Box6 box, int64_t pmin,
                                                int64_t pmax) {
    const int64_t P = pmin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (P >= pmax) return;
    verlet_row(a, P, ld3(F + 3 * P), ld3(tau + 3 * P), vmax, nrot ? nrot[P] : vmax, dt, box.b, 0);
}

// Force assembly fused with the integrator, over the moving-particle list.
__global__ void __launch_bounds__(256) k_reduce_verlet(VerletArrays a, const int32_t* __restrict__ plist, int64_t M,
                                                       const int32_t* __restrict__ scene_of,
                                                       const SceneDev* __restrict__ scenes,
                                                       const int32_t* __restrict__ seg_begin,
                                                       const int32_t* __restrict__ seg_end,
                                                       const int32_t* __restrict__ rec_val,
                                                       const double* __restrict__ Fn, const double* __restrict__ tA,
                                                       const double* __restrict__ tB,
                                                       const double* __restrict__ fext,
                                                       const double* __restrict__ text,
                                                       const int32_t* __restrict__ nverts, int vmax, double dt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const int64_t P = plist[i];
    d3 F = mk3(0.0, 0.0, 0.0), T = mk3(0.0, 0.0, 0.0);
    const int32_t b = seg_begin[P], e = seg_end[P];
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
    F = F + ld3(fext + 3 * P);
    T = T + ld3(text + 3 * P);
    const SceneDev& sc = scenes[scene_of[P]];
    verlet_row(a, P, F, T, vmax, nverts[P], dt, sc.box, sc.law_p.mode == MWD_LAW_PHYSICAL ? 1 : 0);
}

}  // namespace

int mwd_launch_verlet(double* v, double* v12, double* w12bf, double* q12, double* q, double* qall, double* wbf,
                      double* w, const double* F, const double* tau, double* verts, double* com,
                      const double* mass, const double* Iinv, const double* I, double* dxall, int32_t vmax,
                      const int32_t* nrot, double dt, const double* box, int64_t pmin, int64_t pmax,
                      hipStream_t s) {
    if (pmax <= pmin) return MWD_OK;
    VerletArrays a{v, v12, w12bf, q12, q, qall, wbf, w, verts, com, mass, Iinv, I, dxall, nullptr};
    Box6 b;
    for (int k = 0; k < 6; ++k) b.b[k] = box[k];
    const int64_t n = pmax - pmin;
    hipLaunchKernelGGL(k_verlet, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, F, tau, (int)vmax, nrot, dt,
                       b, pmin, pmax);
    MWD_LAUNCH_CHECK();
    return MWD_OK;
}

int mwd_launch_reduce_verlet(mwd_workspace* ws, mwd_state* st) {
    if (ws->n_moving <= 0) return MWD_OK;
    VerletArrays a{st->v, st->v12, st->w12bf, st->q12, st->q, st->qall, st->wbf, st->w, st->verts, st->com,
                   st->mass, st->Iinv, st->I, st->dxall, st->radius};
    const int64_t M = ws->n_moving;
    hipLaunchKernelGGL(k_reduce_verlet, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, ws->stream, a, ws->plist, M,
                       ws->scene_of, ws->scenes_dev, ws->seg_begin, ws->seg_end, ws->rec_val, ws->Fn, ws->tA,
                       ws->tB, st->fext, st->text, st->nverts, (int)st->vmax, ws->dt);
    MWD_LAUNCH_CHECK();
    return MWD_OK;
}

namespace {
__global__ void k_radius(const double* __restrict__ com, const double* __restrict__ verts,
                         const int32_t* __restrict__ nverts, int64_t n, int vmax, double* __restrict__ radius) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const d3 c = ld3(com + 3 * r);
    double r2 = 0.0;
    for (int i = 0; i < nverts[r]; ++i) r2 = fmax(r2, nrm2(ld3(verts + ((int64_t)vmax * r + i) * 3) - c));
    radius[r] = inflate_radius(r2, c);
}
}  // namespace

int mwd_launch_radius(const double* com, const double* verts, const int32_t* nverts, int64_t n, int32_t vmax,
                      double* radius, hipStream_t s) {
    if (n <= 0) return MWD_OK;
    hipLaunchKernelGGL(k_radius, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, com, verts, nverts, n, (int)vmax,
                       radius);
    MWD_LAUNCH_CHECK();
    return MWD_OK;
}
