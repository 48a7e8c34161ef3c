// The scene generator's hull bodies on the device: for each point set, its
// convex hull, the hull's vertices in input order, their mean, the volume and
// the inertia tensor about the mean -- what Particle_class builds from its
// random points with scipy's ConvexHull and Calc_inertia
// (Particle_class.py:239-303, 125-156), here one wave per hull.
//
// Facets: every triple (i < a < b) of the points is tested against all other
// points; the triple is a facet when all of them lie on one side of its
// plane, oriented so that they lie behind it. Points in general position
// (no four on one supporting plane, within 1e-11 of the plane's scale) are
// assumed, as for random points on an ellipsoid; a hull that breaks it is
// reported (status 1) rather than triangulated. The facets are listed in
// (i, a, b) order by ballots, so every sum below runs in a fixed order.
//
// Vertices: the points that belong to a facet, in ascending input index (the
// order of scipy's hull.vertices in 3-D). Centre: their running sum divided by
// their count (Particle_class.py:221-224). Volume and inertia: the
// tetrahedra (centre, D, E, F) of the outward facets, Dobrovolskis (1996) with
// the volume element D.(E x F)/6, summed over the facets by a fixed tree.
#include "mwd_internal.h"

#define MWD_API extern "C" __attribute__((visibility("default")))

namespace {

constexpr int GEN_PMAX = 64;     // points per hull
constexpr int GEN_FMAX = 128;    // facets per hull (2n - 4 for n points in general position)
constexpr int GEN_WAVES = 4;     // hulls per block

struct GenLds {
    double x[GEN_PMAX], y[GEN_PMAX], z[GEN_PMAX];
    uint8_t f0[GEN_FMAX], f1[GEN_FMAX], f2[GEN_FMAX];
    unsigned long long mask;
    int bad;
};

__device__ __forceinline__ d3 pt(const GenLds& L, int i) { return mk3(L.x[i], L.y[i], L.z[i]); }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int d = MWD_WAVE / 2; d >= 1; d >>= 1) v += __shfl_xor(v, d, MWD_WAVE);
    return v;
}

__global__ void __launch_bounds__(64 * GEN_WAVES) k_hull_bodies(const double* __restrict__ points,
                                                                const int32_t* __restrict__ npts, int64_t n_hulls,
                                                                int32_t pmax, double density,
                                                                double* __restrict__ verts_out,
                                                                int32_t* __restrict__ nverts_out,
                                                                double* __restrict__ com_out,
                                                                double* __restrict__ vol_out,
                                                                double* __restrict__ I_out,
                                                                double* __restrict__ isize_out,
                                                                int32_t* __restrict__ status_out) {
    __shared__ GenLds lds[GEN_WAVES];
    const int lane = threadIdx.x & (MWD_WAVE - 1);
    const int w = threadIdx.x / MWD_WAVE;
    GenLds& L = lds[w];
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (MWD_WAVE - lane));
    const int64_t stride = (int64_t)gridDim.x * GEN_WAVES;
    for (int64_t h = (int64_t)blockIdx.x * GEN_WAVES + w; h < n_hulls; h += stride) {  // uniform per wave
        const int n = npts[h];
        double* vo = verts_out + (int64_t)pmax * 3 * h;
        for (int k = lane; k < 3 * pmax; k += MWD_WAVE) vo[k] = 0.0;
        if (n < 4 || n > pmax || n > GEN_PMAX) {
            if (lane == 0) {
                nverts_out[h] = 0;
                status_out[h] = 2;  // not a point count this kernel takes
            }
            continue;
        }
        const double* p = points + (int64_t)pmax * 3 * h;
        if (lane < n) {
            L.x[lane] = p[3 * lane];
            L.y[lane] = p[3 * lane + 1];
            L.z[lane] = p[3 * lane + 2];
        }
        if (lane == 0) {
            L.mask = 0ull;
            L.bad = 0;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // the scale of the point set: its largest distance from point 0
        double s2 = 0.0;
        if (lane < n) s2 = nrm2(pt(L, lane) - pt(L, 0));
#pragma unroll
        for (int d = MWD_WAVE / 2; d >= 1; d >>= 1) s2 = fmax(s2, __shfl_xor(s2, d, MWD_WAVE));
        const double scale = sqrt(s2);
        // facets: triple (i, a, b), i < a < b, lane-strided over (a, b)
        int nf = 0;
        for (int i = 0; i < n - 2; ++i) {
            const d3 pi = pt(L, i);
            for (int q0 = 0; q0 < n * n; q0 += MWD_WAVE) {
                const int q = q0 + lane;
                const int a = q / n, b = q - (q / n) * n;
                bool facet = false, flip = false;
                if (q < n * n && a > i && b > a) {
                    const d3 nrm = cross3(pt(L, a) - pi, pt(L, b) - pi);
                    const double tol = 1e-11 * sqrt(nrm2(nrm)) * scale;
                    bool pos = false, neg = false, on = false;
                    for (int l = 0; l < n; ++l) {
                        if (l == i || l == a || l == b) continue;
                        const double d = dot3(nrm, pt(L, l) - pi);
                        pos = pos || d > tol;
                        neg = neg || d < -tol;
                        on = on || !(fabs(d) > tol);
                        if (pos && neg) break;
                    }
                    if (!(pos && neg)) {
                        if (on) {
                            L.bad = 1;  // a fourth point on a supporting plane
                        } else {
                            facet = true;
                            flip = pos;  // the others in front of (a - i) x (b - i): turn the facet
                        }
                    }
                }
                const uint64_t m = __ballot(facet);
                if (facet) {
                    const int at = nf + __popcll(m & lt);
                    if (at < GEN_FMAX) {
                        L.f0[at] = (uint8_t)i;
                        L.f1[at] = (uint8_t)(flip ? b : a);
                        L.f2[at] = (uint8_t)(flip ? a : b);
                    }
                    atomicOr(&L.mask, (1ull << i) | (1ull << a) | (1ull << b));
                }
                nf += __popcll(m);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const uint64_t mask = L.mask;
        const int nv = __popcll(mask);
        if (L.bad || nf > GEN_FMAX || nf != 2 * nv - 4) {
            if (lane == 0) {
                nverts_out[h] = 0;
                status_out[h] = 1;  // degenerate (or not a closed triangulated surface)
            }
            continue;
        }
        // the vertices' running sum in ascending index, divided by their count
        d3 c = mk3(0.0, 0.0, 0.0);
        for (uint64_t r = mask; r; r &= r - 1) c = c + pt(L, __ffsll((long long)r) - 1);
        c = c / (double)nv;
        // the centred vertices, in ascending index
        if (lane < n && ((mask >> lane) & 1ull)) {
            const int m = __popcll(mask & lt);
            const d3 v = pt(L, lane) - c;
            vo[3 * m] = v.x;
            vo[3 * m + 1] = v.y;
            vo[3 * m + 2] = v.z;
        }
        // volume and the products of inertia over the outward facets
        double vol = 0.0, pxx = 0.0, pyy = 0.0, pzz = 0.0, pxy = 0.0, pxz = 0.0, pyz = 0.0;
        for (int f = lane; f < nf; f += MWD_WAVE) {
            const d3 D = pt(L, L.f0[f]) - c, E = pt(L, L.f1[f]) - c, F = pt(L, L.f2[f]) - c;
            const double dV = dot3(D, cross3(E, F)) / 6.0;
            const double k = density * dV / 20.0;
            vol += dV;
            pxx += k * (2.0 * D.x * D.x + 2.0 * E.x * E.x + 2.0 * F.x * F.x + D.x * E.x + D.x * E.x + D.x * F.x +
                        D.x * F.x + E.x * F.x + E.x * F.x);
            pyy += k * (2.0 * D.y * D.y + 2.0 * E.y * E.y + 2.0 * F.y * F.y + D.y * E.y + D.y * E.y + D.y * F.y +
                        D.y * F.y + E.y * F.y + E.y * F.y);
            pzz += k * (2.0 * D.z * D.z + 2.0 * E.z * E.z + 2.0 * F.z * F.z + D.z * E.z + D.z * E.z + D.z * F.z +
                        D.z * F.z + E.z * F.z + E.z * F.z);
            pxy += k * (2.0 * D.x * D.y + 2.0 * E.x * E.y + 2.0 * F.x * F.y + D.x * E.y + D.y * E.x + D.x * F.y +
                        D.y * F.x + E.x * F.y + E.y * F.x);
            pxz += k * (2.0 * D.x * D.z + 2.0 * E.x * E.z + 2.0 * F.x * F.z + D.x * E.z + D.z * E.x + D.x * F.z +
                        D.z * F.x + E.x * F.z + E.z * F.x);
            pyz += k * (2.0 * D.y * D.z + 2.0 * E.y * E.z + 2.0 * F.y * F.z + D.y * E.z + D.z * E.y + D.y * F.z +
                        D.z * F.y + E.y * F.z + E.z * F.y);
        }
        vol = wave_sum(vol);
        pxx = wave_sum(pxx);
        pyy = wave_sum(pyy);
        pzz = wave_sum(pzz);
        pxy = wave_sum(pxy);
        pxz = wave_sum(pxz);
        pyz = wave_sum(pyz);
        if (lane == 0) {
            nverts_out[h] = nv;
            status_out[h] = 0;
            com_out[3 * h] = c.x;
            com_out[3 * h + 1] = c.y;
            com_out[3 * h + 2] = c.z;
            vol_out[h] = vol;
            double* I = I_out + 9 * h;
            I[0] = pyy + pzz;
            I[1] = -pxy;
            I[2] = -pxz;
            I[3] = -pxy;
            I[4] = pxx + pzz;
            I[5] = -pyz;
            I[6] = -pxz;
            I[7] = -pyz;
            I[8] = pxx + pyy;
            isize_out[h] = ((pyy + pzz) + (pxx + pzz) + (pxx + pyy)) / 3.0;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();  // the LDS is rewritten by the wave's next hull
    }
}

}  // namespace

MWD_API int mwd_hull_bodies(const double* points, const int32_t* npts, int64_t n_hulls, int32_t pmax, double density,
                            double* verts_out, int32_t* nverts_out, double* com_out, double* vol_out, double* I_out,
                            double* isize_out, int32_t* status_out, void* stream) {
    if (!points || !npts || n_hulls < 0 || pmax < 4 || pmax > GEN_PMAX || !verts_out || !nverts_out || !com_out ||
        !vol_out || !I_out || !isize_out || !status_out) {
        mwd_set_error("mwd_hull_bodies: invalid argument (pmax must be 4..64)");
        return MWD_E_INVALID;
    }
    if (n_hulls == 0) return MWD_OK;
    const int64_t blocks = (n_hulls + GEN_WAVES - 1) / GEN_WAVES;
    hipLaunchKernelGGL(k_hull_bodies, dim3((unsigned)(blocks < 65536 ? blocks : 65536)), dim3(64 * GEN_WAVES), 0,
                       (hipStream_t)stream, points, npts, n_hulls, pmax, density, verts_out, nverts_out, com_out,
                       vol_out, I_out, isize_out, status_out);
    MWD_LAUNCH_CHECK();
    return MWD_OK;
}
