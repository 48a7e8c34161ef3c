// C ABI of libmwdiem_hip.so (declared in include/mwdiem_hip.h).
#include <math.h>
#include <string.h>

#include <stddef.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "mwd_internal.h"

#define MWD_API extern "C" __attribute__((visibility("default")))

static void destroy_events(mwd_workspace* ws);

namespace {
thread_local std::string g_err;

int bits_for(int64_t span) {
    // number of key bits needed for values in [0, span)
    int b = 0;
    uint64_t v = span > 1 ? (uint64_t)(span - 1) : 0;
    while (v) {
        ++b;
        v >>= 1;
    }
    return b;
}


template <typename T>
int dalloc(T** p, int64_t count) {
    *p = nullptr;
    if (count <= 0) count = 1;
    hipError_t e = hipMalloc((void**)p, sizeof(T) * (size_t)count);
    if (e != hipSuccess) {
        mwd_set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
        return MWD_E_HIP;
    }
    return MWD_OK;
}

void free_ws(mwd_workspace* ws) {
    if (!ws) return;
    void* ptrs[] = {ws->scenes_dev, ws->key_seg, ws->key_small, ws->key_head, ws->row_next, ws->plist, ws->wlist, ws->scene_of, ws->is_wall, ws->keys, ws->keys_alt, ws->vals,
                    ws->vals_alt, ws->cell_start, ws->cell_key, ws->cell_self, ws->cell_seg,
                    ws->row_first, ws->row_rank, ws->cell_cnt, ws->cell_off, ws->perm, ws->key_cnt,
                    ws->row_key, ws->sph,
                    ws->pairs, ws->collision, ws->simplex, ws->sidx, ws->surv, ws->surv_cls, ws->bucket,
                    ws->contact_pair,
                    ws->depth, ws->normal, ws->cpA, ws->cpB, ws->Fn, ws->tA, ws->tB, ws->epa_status, ws->rec_cnt,
                    ws->rec_bkt, ws->rec_ovf, ws->hist,
                    ws->hist_off, ws->scan_tmp, ws->scan_tmp_w, ws->wseg_off, ws->word_desc, ws->hit_words, ws->gjk_words, ws->counters, ws->flags, ws->laws,
                    ws->scene_wt_off, ws->ctan, ws->law_tans, ws->adj_F, ws->adj_T, ws->adj_lamF,
                    ws->adj_lamT, ws->adj_face, ws->adj_c, ws->adj_cth, ws->adj_rowth, ws->tape, ws->tape_sync};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    destroy_events(ws);
    for (int t = 0; t < 2; ++t) {
        if (ws->thist.keys[t]) (void)hipFree(ws->thist.keys[t]);
        if (ws->thist.val[t]) (void)hipFree(ws->thist.val[t]);
        if (ws->thist.used[t]) (void)hipFree(ws->thist.used[t]);
    }
    if (ws->thist.idx) (void)hipFree(ws->thist.idx);
    delete[] ws->scenes_host;
    delete ws;
}

BodyView state_view(const mwd_workspace* ws, const mwd_state* st, int side) {
    BodyView b;
    b.com = st->com;
    b.verts = st->verts;
    b.nverts = st->nverts;
    b.vel = st->v;
    b.angvel = st->w;
    b.quat = st->q;
    b.isize = st->isize;
    b.mass = st->mass;
    b.radius = st->radius;
    b.is_wall = ws->is_wall;
    b.scene = ws->scene_of;
    b.rows = ws->pairs + side;
    b.row_stride = 2;
    b.vmax = st->vmax;
    b.laws2 = ws->laws;
    return b;
}

__global__ void k_wall_from_ids(const int32_t* idA, const int32_t* idB, int64_t W, int64_t P, uint8_t* wA,
                                uint8_t* wB) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    wA[p] = idA[p] < W ? 1 : 0;
    wB[p] = idB[p] < W ? 1 : 0;
}

}  // namespace

void mwd_set_error(const std::string& msg) { g_err = msg; }

// Per step slot: MWD_N_STAGES + 1 stage boundaries, then a start and a stop
// event for each timed kernel.
static inline hipEvent_t& ev_at(mwd_workspace* ws, int slot, int k) { return ws->ev[slot * MWD_EV_PER_STEP + k]; }

// Timing is for eager steps: nothing is recorded while the stream is being
// captured into a graph.
static bool timing_active(mwd_workspace* ws) {
    if (!ws || !ws->timing) return false;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(ws->stream, &cs) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return cs == hipStreamCaptureStatusNone;
}

hipEvent_t mwd_kev(mwd_workspace* ws, int kt, int stop) {
    if (!timing_active(ws)) return nullptr;
    ws->kt_slot[ws->ring_slot] |= 1u << kt;
    return ev_at(ws, ws->ring_slot, MWD_N_STAGES + 1 + 2 * kt + (stop ? 1 : 0));
}

static void flush_timing(mwd_workspace* ws) {
    if (!ws->timing || ws->ring_slot == 0) return;
    (void)hipEventSynchronize(ev_at(ws, ws->ring_slot - 1, MWD_N_STAGES));
    for (int sl = 0; sl < ws->ring_slot; ++sl) {
        for (int k = 0; k < MWD_N_STAGES; ++k) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, ev_at(ws, sl, k), ev_at(ws, sl, k + 1)) == hipSuccess)
                ws->stage_ms[k] += ms;
            else
                ws->timing_errors += 1;
        }
        for (int k = 0; k < MWD_N_KTIMED; ++k) {
            if (!(ws->kt_slot[sl] & (1u << k))) continue;
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, ev_at(ws, sl, MWD_N_STAGES + 1 + 2 * k),
                                    ev_at(ws, sl, MWD_N_STAGES + 2 + 2 * k)) == hipSuccess)
                ws->kernel_ms[k] += ms;
            else
                ws->timing_errors += 1;
        }
        ws->kt_slot[sl] = 0;
        ws->stage_calls += 1;
        ws->kernel_calls += 1;
    }
    (void)hipGetLastError();  // a failed query must not surface at the next launch
    ws->ring_slot = 0;
}

void mwd_mark(mwd_workspace* ws, int stage) {
    if (timing_active(ws) && hipEventRecord(ev_at(ws, ws->ring_slot, stage), ws->stream) != hipSuccess) {
        ws->timing_errors += 1;
        (void)hipGetLastError();
    }
}
void mwd_mark_end(mwd_workspace* ws) {
    if (!timing_active(ws)) return;
    if (hipEventRecord(ev_at(ws, ws->ring_slot, MWD_N_STAGES), ws->stream) != hipSuccess) {
        ws->timing_errors += 1;
        (void)hipGetLastError();
    }
    ws->ring_slot += 1;
    if (ws->ring_slot == MWD_TIMING_RING) flush_timing(ws);
}

static void destroy_events(mwd_workspace* ws) {
    if (!ws->ev) return;
    for (int k = 0; k < MWD_TIMING_RING * MWD_EV_PER_STEP; ++k) (void)hipEventDestroy(ws->ev[k]);
    delete[] ws->ev;
    ws->ev = nullptr;
}

MWD_API int mwd_set_timing(mwd_workspace* ws, int enable) {
    if (!ws) return MWD_E_INVALID;
    if (enable && !ws->ev) {
        ws->ev = new hipEvent_t[MWD_TIMING_RING * MWD_EV_PER_STEP];
        for (int k = 0; k < MWD_TIMING_RING * MWD_EV_PER_STEP; ++k) MWD_HIP_CHECK(hipEventCreate(&ws->ev[k]));
    }
    if (!enable) {
        MWD_HIP_CHECK(hipStreamSynchronize(ws->stream));
        destroy_events(ws);
    }
    ws->timing = enable ? 1 : 0;
    ws->ring_slot = 0;
    ws->timing_errors = 0;
    for (int k = 0; k < MWD_TIMING_RING; ++k) ws->kt_slot[k] = 0;
    for (int k = 0; k < MWD_N_STAGES; ++k) ws->stage_ms[k] = 0.0;
    for (int k = 0; k < MWD_N_KTIMED; ++k) ws->kernel_ms[k] = 0.0;
    ws->stage_calls = 0;
    ws->kernel_calls = 0;
    return MWD_OK;
}

MWD_API int mwd_stage_times(mwd_workspace* ws, double* ms, int64_t* steps) {
    if (!ws || !ms) return MWD_E_INVALID;
    MWD_HIP_CHECK(hipStreamSynchronize(ws->stream));
    flush_timing(ws);
    for (int k = 0; k < MWD_N_STAGES; ++k) {
        ms[k] = ws->stage_ms[k];
        ws->stage_ms[k] = 0.0;
    }
    if (steps) *steps = ws->stage_calls;
    ws->stage_calls = 0;
    if (ws->timing_errors) {
        ws->timing_errors = 0;
        mwd_set_error("mwd_stage_times: some stage events could not be recorded or timed");
        return MWD_E_HIP;
    }
    return MWD_OK;
}

MWD_API int mwd_kernel_times(mwd_workspace* ws, double* ms, int64_t* steps) {
    if (!ws || !ms) return MWD_E_INVALID;
    MWD_HIP_CHECK(hipStreamSynchronize(ws->stream));
    flush_timing(ws);
    for (int k = 0; k < MWD_N_KTIMED; ++k) {
        ms[k] = ws->kernel_ms[k];
        ws->kernel_ms[k] = 0.0;
    }
    if (steps) *steps = ws->kernel_calls;
    ws->kernel_calls = 0;
    if (ws->timing_errors) {
        mwd_set_error("mwd_kernel_times: some kernel events could not be recorded or timed");
        return MWD_E_HIP;
    }
    return MWD_OK;
}

MWD_API const char* mwd_last_error(void) { return g_err.c_str(); }
MWD_API int mwd_version(void) { return 1; }

MWD_API int mwd_law_init_mode(mwd_law* out, double Kn, double en, double et, double mu, double dt,
                              int32_t mode) {
    if (!out || !(Kn > 0) || !(en > 0) || !(et > 0) || !(mu > 0) || (mode != MWD_LAW_REF && mode != MWD_LAW_PHYSICAL)) {
        mwd_set_error("mwd_law_init: Kn, en, et and mu must be positive");
        return MWD_E_INVALID;
    }
    // Force, gpu_fine_contact_detection.cu:685-701, in the reference's operator order.
    const double pi = 3.141592653589793;
    const double pi2 = pi * pi;
    out->Kn = Kn;
    out->mu = mu;
    out->ln_en = log(en);
    out->ln_et = log(et);
    out->A_en = pi2 + out->ln_en * out->ln_en;
    out->tc = sqrt((mu * out->A_en) / Kn);
    out->Gn = ((-2.0 * mu) / out->tc) * out->ln_en;
    out->Kt_num = ((Kn * (pi2 + out->ln_et * out->ln_et)) / out->A_en) * 1.0;
    out->inv_mu = 1.0 / mu;
    out->dt = dt;
    out->cn = (-2.0 * out->ln_en) / sqrt(out->A_en);
    out->ct = (-2.0 * out->ln_et) / sqrt(pi2 + out->ln_et * out->ln_et);
    out->ratio = (pi2 + out->ln_et * out->ln_et) / out->A_en;
    out->mode = mode;
    out->pad_ = 0;
    return MWD_OK;
}

MWD_API int mwd_law_init(mwd_law* out, double Kn, double en, double et, double mu, double dt) {
    return mwd_law_init_mode(out, Kn, en, et, mu, dt, MWD_LAW_REF);
}

MWD_API int mwd_workspace_create(mwd_workspace** out, int device, const mwd_scene* scenes, int32_t n_scenes,
                                 int64_t n_rows, int32_t vmax, int64_t pair_cap, int64_t contact_cap, double dt,
                                 void* stream) {
    if (!out || !scenes || n_scenes <= 0 || n_rows <= 0 || vmax <= 0 || vmax > 255 || pair_cap <= 0 ||
        contact_cap <= 0 || n_rows > INT32_MAX || pair_cap > INT32_MAX / 2 || contact_cap > INT32_MAX / 4) {
        mwd_set_error("mwd_workspace_create: invalid argument");
        return MWD_E_INVALID;
    }
    *out = nullptr;
    MWD_HIP_CHECK(hipSetDevice(device));
    mwd_workspace* ws = new mwd_workspace();  // value-initialised: zeros, then the members' own defaults
    ws->device = device;
    ws->stream = (hipStream_t)stream;
    ws->n_scenes = n_scenes;
    ws->n_rows = n_rows;
    ws->vmax = vmax;
    ws->pair_cap = pair_cap;
    ws->contact_cap = contact_cap;
    ws->dt = dt;
    ws->scenes_host = new mwd_scene[n_scenes];
    memcpy(ws->scenes_host, scenes, sizeof(mwd_scene) * n_scenes);

    std::vector<SceneDev> sd(n_scenes);
    std::vector<int32_t> scene_of(n_rows, 0);
    std::vector<uint8_t> is_wall(n_rows, 0);
    std::vector<int32_t> plist, wlist;
    std::vector<int64_t> wt(n_scenes + 1, 0);
    std::vector<mwd_law> laws(2 * n_scenes);
    int64_t kbase = 0, prev_end = 0;
    for (int s = 0; s < n_scenes; ++s) {
        const mwd_scene& sc = scenes[s];
        if (sc.row0 < prev_end || sc.n_walls < 0 || sc.n_active < sc.n_walls || sc.row0 + sc.n_active > n_rows ||
            sc.kspan <= 0 || sc.ncx <= 0 || sc.ncy <= 0) {
            mwd_set_error("mwd_workspace_create: scene " + std::to_string(s) + " is inconsistent");
            free_ws(ws);
            return MWD_E_INVALID;
        }
        const int64_t row_end = (s + 1 < n_scenes) ? scenes[s + 1].row0 : n_rows;
        for (int64_t r = sc.row0; r < row_end; ++r) scene_of[r] = s;
        for (int64_t r = sc.row0; r < sc.row0 + sc.n_walls; ++r) {
            is_wall[r] = 1;
            wlist.push_back((int32_t)r);
        }
        SceneDev& d = sd[s];
        d.row0 = sc.row0;
        d.n_walls = sc.n_walls;
        d.n_active = sc.n_active;
        memcpy(d.box, sc.box, sizeof(d.box));
        memcpy(d.cell, sc.cell, sizeof(d.cell));
        d.ncx = sc.ncx;
        d.ncy = sc.ncy;
        d.kmin = sc.kmin;
        d.kspan = sc.kspan;
        d.kbase = kbase;
        d.wt_off = wt[s];
        d.plist_off = (int64_t)plist.size();
        d.wall_thr = sc.wall_thr;
        d.law_p = sc.law_p;
        d.law_w = sc.law_w;
        laws[2 * s] = sc.law_p;
        laws[2 * s + 1] = sc.law_w;
        {
            const int64_t X = sc.ncx, XY = sc.ncx * sc.ncy;
            if (sc.law_p.mode == MWD_LAW_PHYSICAL) {
                // full forward half shell: (dx,dy,dz) with dz>0, or dz=0 and dy>0, or dz=dy=0 and dx>0
                const int h[13][3] = {{1, 0, 0},  {-1, 1, 0}, {0, 1, 0},  {1, 1, 0},  {-1, -1, 1}, {0, -1, 1}, {1, -1, 1},
                                      {-1, 0, 1}, {0, 0, 1},  {1, 0, 1},  {-1, 1, 1}, {0, 1, 1},   {1, 1, 1}};
                d.n_off = 13;
                for (int k = 0; k < 13; ++k) d.off[k] = h[k][0] + h[k][1] * X + h[k][2] * XY;
            } else {
                // LIBRARY_Coarse_contact_detection.py:101-129
                const int64_t r[7] = {1, X, XY, XY + 1, XY + 1 + X, XY + X, X + 1};
                d.n_off = 7;
                for (int k = 0; k < 7; ++k) d.off[k] = r[k];
                for (int k = 7; k < 13; ++k) d.off[k] = 0;
            }
        }
        kbase += sc.kspan;
        wt[s + 1] = wt[s] + sc.n_walls * (sc.n_active - sc.n_walls);
        for (int64_t r = sc.row0 + sc.n_walls; r < sc.row0 + sc.n_active; ++r) plist.push_back((int32_t)r);
        prev_end = sc.row0 + sc.n_active;
    }
    if (kbase > (int64_t)0xFFFFFFFFll) {
        mwd_set_error("mwd_workspace_create: total key span exceeds 32 bits");
        free_ws(ws);
        return MWD_E_KEYRANGE;
    }
    const int64_t M = (int64_t)plist.size();
    ws->n_moving = M;
    ws->row_base = n_scenes == 1 ? scenes[0].row0 + scenes[0].n_walls : -1;
    // wall-word segments: per scene and wall, the words covering the scene's
    // moving-particle list entries [plist_off, plist_off + np)
    std::vector<int64_t> wseg_off(1, 0);
    std::vector<int32_t> wseg_scene;
    for (int s = 0; s < n_scenes; ++s) {
        sd[s].wseg0 = (int64_t)wseg_scene.size();
        const int64_t np = scenes[s].n_active - scenes[s].n_walls;
        const int64_t nw = np > 0 ? (sd[s].plist_off + np + 63) / 64 - sd[s].plist_off / 64 : 0;
        for (int64_t k = 0; k < scenes[s].n_walls; ++k) {
            wseg_scene.push_back(s);
            wseg_off.push_back(wseg_off.back() + nw);
        }
    }
    ws->n_wseg = (int64_t)wseg_scene.size();
    ws->n_words = wseg_off.back();
    // per word: its wall row, the row of its bit 0, its index among the words
    // of the moving-entry list and the bits that are its scene's entries
    // (first | end << 8; a scene's first and last words share entries with
    // its neighbours)
    std::vector<int4> word_desc((size_t)std::max<int64_t>(ws->n_words, 1), make_int4(0, 0, 0, 0));
    for (int64_t g = 0; g < ws->n_wseg; ++g) {
        const int s = wseg_scene[g];
        const int64_t k = g - sd[s].wseg0;
        const int64_t first = sd[s].plist_off, end = first + scenes[s].n_active - scenes[s].n_walls;
        for (int64_t w = wseg_off[g]; w < wseg_off[g + 1]; ++w) {
            const int64_t wg = (first >> 6) + (w - wseg_off[g]);  // the word in the list's numbering
            const int64_t e0 = wg * 64;
            const int64_t b0 = first > e0 ? first - e0 : 0, b1 = end - e0 < 64 ? end - e0 : 64;
            word_desc[w] = make_int4((int32_t)(scenes[s].row0 + k),
                                     (int32_t)(scenes[s].row0 + scenes[s].n_walls + e0 - first), (int32_t)wg,
                                     (int32_t)(b0 | (b1 << 8)));
        }
    }
    ws->word_walls_ok = 1;
    for (int s = 0; s < n_scenes; ++s)
        if (scenes[s].n_walls > MWD_WORD_WALLS) ws->word_walls_ok = 0;
    ws->n_wall_tests = wt[n_scenes];
    ws->key_span_total = kbase;
    ws->key_bits = bits_for(kbase);
    ws->row_bits = bits_for(n_rows);
    const int64_t max_sort = M;  // the cell-key sort (contact records go to per-row buckets)
    const int64_t tiles = mwd_sort_tiles(max_sort);
    const int64_t hist_len = (int64_t(1) << MWD_SORT_MAX_BITS) * tiles;
    const int64_t max_scan = std::max({M, n_rows, pair_cap, ws->n_wall_tests, hist_len});
    ws->scan_tmp_len = mwd_scan_tmp_len(max_scan);

    int rc = 0;
#define A(ptr, n)                                  \
    if ((rc = dalloc(&ws->ptr, (n))) != MWD_OK) {  \
        free_ws(ws);                               \
        return rc;                                 \
    }
    A(scenes_dev, n_scenes);
    A(plist, M);
    ws->n_wall_rows = (int64_t)wlist.size();
    A(wlist, std::max<int64_t>(ws->n_wall_rows, 1));
    A(scene_of, n_rows);
    A(is_wall, n_rows);
    A(keys, M);
    A(keys_alt, M);
    A(vals, M);
    A(vals_alt, M);
    A(cell_start, M + 1);
    A(cell_key, M);
    A(cell_self, M);
    A(cell_seg, MWD_MAX_OFF * M);
    A(row_first, n_rows);
    A(row_rank, n_rows);
    A(cell_cnt, M);
    A(cell_off, M);
    A(perm, M);
    // key-indexed tables (12 bytes per key) for the cell pass without a sort
    // (the empty lowest-row mark 0x7F7F7F7F must lie above every row)
    ws->cell_table_ok = ws->key_span_total <= (int64_t(1) << 28) && n_rows < 0x7F7F7F7F;
    ws->cell_table = ws->cell_table_ok;
    if (ws->cell_table_ok) {
        A(key_seg, ws->key_span_total);
        A(key_small, 4 * ws->key_span_total);
        A(key_head, ws->key_span_total);
        A(row_next, n_rows);
        A(key_cnt, 2 * ws->key_span_total);  // (count, lowest row) interleaved: one cache line per key
        ws->key_min = ws->key_cnt + 1;
    }
    A(row_key, n_rows);
    A(sph, 4 * n_rows);
    A(pairs, 2 * pair_cap);
    A(collision, pair_cap);
    A(simplex, contact_cap);
    A(sidx, pair_cap);
    A(surv, pair_cap);
    A(surv_cls, pair_cap);
    A(bucket, 4 * MWD_GJK_BUCKETS);
    A(contact_pair, contact_cap);
    A(depth, contact_cap);
    A(normal, 3 * contact_cap);
    A(cpA, 3 * contact_cap);
    A(cpB, 3 * contact_cap);
    A(Fn, 3 * contact_cap);
    A(tA, 3 * contact_cap);
    A(tB, 3 * contact_cap);
    A(epa_status, contact_cap);
    A(rec_cnt, n_rows);
    A(rec_bkt, n_rows * MWD_ROW_CONTACTS);
    A(rec_ovf, 4 * contact_cap);
    A(hist, hist_len);
    A(hist_off, hist_len);
    A(scan_tmp, ws->scan_tmp_len);
    ws->scan_tmp_w_len = mwd_scan_tmp_len(std::max<int64_t>(ws->n_words, 1));
    A(scan_tmp_w, ws->scan_tmp_w_len);
    A(wseg_off, ws->n_wseg + 1);
    A(word_desc, std::max<int64_t>(ws->n_words, 1));
    A(hit_words, std::max<int64_t>(ws->n_words, 1));
    A(gjk_words, std::max<int64_t>(ws->n_words, 1));
    A(counters, 20);
    A(flags, 1);
    A(laws, 2 * n_scenes);
    A(scene_wt_off, n_scenes + 1);
#undef A
    hipStream_t s = ws->stream;
#define H2D(dst, src, n)                                                                                  \
    if (hipMemcpyAsync(ws->dst, (src), sizeof(*ws->dst) * (size_t)(n), hipMemcpyHostToDevice, s) !=     \
        hipSuccess) {                                                                                     \
        mwd_set_error("mwd_workspace_create: upload failed");                                            \
        free_ws(ws);                                                                                      \
        return MWD_E_HIP;                                                                                 \
    }
    H2D(scenes_dev, sd.data(), n_scenes);
    if (M) H2D(plist, plist.data(), M);
    if (ws->n_wall_rows) H2D(wlist, wlist.data(), ws->n_wall_rows);
    H2D(scene_of, scene_of.data(), n_rows);
    H2D(is_wall, is_wall.data(), n_rows);
    H2D(laws, laws.data(), 2 * n_scenes);
    H2D(scene_wt_off, wt.data(), n_scenes + 1);
    H2D(wseg_off, wseg_off.data(), ws->n_wseg + 1);
    if (ws->n_words) H2D(word_desc, word_desc.data(), ws->n_words);
#undef H2D
    const size_t kbytes = sizeof(int4) * (size_t)ws->key_span_total;
    if ((ws->cell_table_ok && (hipMemsetAsync(ws->key_seg, 0xFF, kbytes, s) != hipSuccess ||
                            hipMemsetAsync(ws->key_head, 0xFF, sizeof(int32_t) * (size_t)ws->key_span_total, s) != hipSuccess ||
                            mwd_fill_key_table(ws->key_cnt, ws->key_span_total, s) != MWD_OK)) ||
        hipMemsetAsync(ws->row_key, 0xFF, sizeof(int32_t) * n_rows, s) != hipSuccess ||
        hipMemsetAsync(ws->counters, 0, sizeof(int64_t) * 20, s) != hipSuccess ||
        hipMemsetAsync(ws->rec_cnt, 0, sizeof(int32_t) * n_rows, s) != hipSuccess ||
        hipMemsetAsync(ws->scan_tmp, 0, sizeof(int64_t) * ws->scan_tmp_len, s) != hipSuccess ||
        hipMemsetAsync(ws->scan_tmp_w, 0, sizeof(int64_t) * ws->scan_tmp_w_len, s) != hipSuccess ||
        hipMemsetAsync(ws->hit_words, 0, sizeof(uint64_t) * std::max<int64_t>(ws->n_words, 1), s) != hipSuccess ||
        hipMemsetAsync(ws->gjk_words, 0, sizeof(uint64_t) * std::max<int64_t>(ws->n_words, 1), s) != hipSuccess ||
        hipMemsetAsync(ws->flags, 0, sizeof(uint32_t), s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        mwd_set_error("mwd_workspace_create: initialisation failed");
        free_ws(ws);
        return MWD_E_HIP;
    }
    // every moving row and wall slab test is live until a spawn schedule says otherwise
    ws->m_live = ws->counters + 14;
    ws->wt_live = ws->counters + 15;
    {
        const int64_t live[2] = {M, ws->n_wall_tests};
        if (hipMemcpy(ws->m_live, live, sizeof(live), hipMemcpyHostToDevice) != hipSuccess) {
            mwd_set_error("mwd_workspace_create: initialisation failed");
            free_ws(ws);
            return MWD_E_HIP;
        }
    }
    *out = ws;
    return MWD_OK;
}

MWD_API int mwd_get_tuning(const mwd_workspace* ws, mwd_tuning* out) {
    if (!ws || !out) {
        mwd_set_error("mwd_get_tuning: null argument");
        return MWD_E_INVALID;
    }
    out->cell_pass = ws->cell_table ? MWD_CELLS_TABLE : MWD_CELLS_SORTED;
    out->gjk_group = ws->tune.gjk_group;
    out->gjk_blocks = ws->tune.gjk_blocks;
    out->sphere_block = ws->tune.sphere_block;
    out->verlet_block = ws->tune.verlet_block;
    out->cell_items = ws->tune.cell_items;
    out->verlet_waves = ws->tune.verlet_waves;
    out->vertex_pack = ws->tune.vertex_pack;
    out->epa_blocks = ws->tune.epa_blocks;
    out->epa_fused = ws->tune.epa_fused;
    out->fill_cells = ws->tune.fill_cells;
    out->gjk_buckets = ws->tune.gjk_buckets;
    out->fill_split = ws->tune.fill_split;
    return MWD_OK;
}

MWD_API int mwd_set_tuning(mwd_workspace* ws, const mwd_tuning* t) {
    if (!ws || !t) {
        mwd_set_error("mwd_set_tuning: null argument");
        return MWD_E_INVALID;
    }
    const bool pass_ok = (t->cell_pass == MWD_CELLS_SORTED && !ws->spawn_on) ||
                         (t->cell_pass == MWD_CELLS_TABLE && ws->cell_table_ok);
    const bool group_ok = t->gjk_group == 4 || t->gjk_group == 8 || t->gjk_group == 16;
    const bool blocks_ok = t->gjk_blocks >= 64 && t->gjk_blocks <= 65536;
    const bool sphere_ok = t->sphere_block >= 64 && t->sphere_block <= 1024 && t->sphere_block % 64 == 0;
    const bool verlet_ok = t->verlet_block == 64 || t->verlet_block == 128 || t->verlet_block == 256;
    const bool items_ok = t->cell_items == 0 || t->cell_items == 1 || t->cell_items == 2 || t->cell_items == 4;
    const bool waves_ok = t->verlet_waves == 1 || t->verlet_waves == 3 || t->verlet_waves == 4;
    const bool pack_ok = t->vertex_pack == 0 || t->vertex_pack == 1;
    const bool epa_ok = t->epa_blocks >= 16 && t->epa_blocks <= 65536 && (t->epa_fused == 0 || t->epa_fused == 1);
    const bool fill_ok = t->fill_cells == 4 || t->fill_cells == 16 || t->fill_cells == 64;
    const bool buckets_ok = t->gjk_buckets == 0 || t->gjk_buckets == 1;
    const bool split_ok = t->fill_split == 1 || t->fill_split == 4 || t->fill_split == 16;
    if (!pass_ok || !group_ok || !blocks_ok || !sphere_ok || !verlet_ok || !items_ok || !waves_ok || !pack_ok ||
        !epa_ok || !fill_ok || !buckets_ok || !split_ok) {
        mwd_set_error(!pass_ok ? "mwd_set_tuning: cell_pass unknown, MWD_CELLS_TABLE without key tables, or "
                                 "MWD_CELLS_SORTED with a spawn schedule"
                               : "mwd_set_tuning: gjk_group, gjk_blocks, sphere_block, verlet_block, cell_items, "
                                 "verlet_waves, vertex_pack, epa_blocks, epa_fused, fill_cells, gjk_buckets or "
                                 "fill_split "
                                 "out of range");
        return MWD_E_INVALID;
    }
    ws->cell_table = t->cell_pass == MWD_CELLS_TABLE;
    ws->tune.gjk_group = t->gjk_group;
    ws->tune.gjk_blocks = t->gjk_blocks;
    ws->tune.sphere_block = t->sphere_block;
    ws->tune.verlet_block = t->verlet_block;
    ws->tune.cell_items = t->cell_items;
    ws->tune.verlet_waves = t->verlet_waves;
    ws->tune.vertex_pack = t->vertex_pack;
    ws->tune.epa_blocks = t->epa_blocks;
    ws->tune.epa_fused = t->epa_fused;
    ws->tune.fill_cells = t->fill_cells;
    ws->tune.gjk_buckets = t->gjk_buckets;
    ws->tune.fill_split = t->fill_split;
    return MWD_OK;
}

MWD_API int mwd_workspace_destroy(mwd_workspace* ws) {
    if (!ws) return MWD_OK;
    (void)hipStreamSynchronize(ws->stream);
    free_ws(ws);
    return MWD_OK;
}

MWD_API int mwd_verlet(double* v, double* v12, double* w12bf, double* q12, double* q, double* qall, double* wbf,
                       double* w, const double* F, const double* tau, double* verts, double* com, const double* mass,
                       const double* Iinv, const double* I, double* dxall, int32_t vmax, const int32_t* nrot,
                       double dt, const double* box, int64_t pmin, int64_t pmax, void* stream) {
    if (!v || !v12 || !w12bf || !q12 || !q || !qall || !wbf || !w || !F || !tau || !verts || !com || !mass ||
        !Iinv || !I || !dxall || !box || vmax <= 0 || pmin < 0 || pmax < pmin) {
        mwd_set_error("mwd_verlet: invalid argument");
        return MWD_E_INVALID;
    }
    return mwd_launch_verlet(v, v12, w12bf, q12, q, qall, wbf, w, F, tau, verts, com, mass, Iinv, I, dxall, vmax,
                             nrot, dt, box, pmin, pmax, (hipStream_t)stream);
}

MWD_API int mwd_cell_keys(const double* com, int64_t n, const double* cell, int64_t ncx, int64_t ncy, int64_t* keys,
                          void* stream) {
    if (!com || !cell || !keys || n < 0) {
        mwd_set_error("mwd_cell_keys: invalid argument");
        return MWD_E_INVALID;
    }
    return mwd_launch_cell_keys(com, n, cell, ncx, ncy, keys, (hipStream_t)stream);
}

MWD_API int mwd_candidates(mwd_workspace* ws, const mwd_state* st, const int64_t* keys) {
    if (!ws || !st || (!st->com && !keys)) {
        mwd_set_error("mwd_candidates: invalid argument");
        return MWD_E_INVALID;
    }
    const int rc = mwd_launch_candidates(ws, st, keys);
    if (!rc) ws->fresh_pairs = 1;  // k_finish_pairs zeroed the fine-contact counters
    return rc;
}

MWD_API int mwd_pairs_view(mwd_workspace* ws, int32_t** pairs, int64_t** n_pairs) {
    if (!ws) return MWD_E_INVALID;
    if (pairs) *pairs = ws->pairs;
    if (n_pairs) *n_pairs = ws->counters + 3;
    return MWD_OK;
}

static TArr<MWD_NT> tarr_of(double* v, double* tan_base, int64_t stride, int a) {
    TArr<MWD_NT> t;
    t.v = v;
    for (int k = 0; k < MWD_NT; ++k) t.t[k] = tan_base ? tan_base + (int64_t)(MWD_NT * a + k) * stride : nullptr;
    return t;
}

// Per-contact outputs: values in the workspace arrays, tangents (when
// allocated) in ws->ctan as [array][k][contact_cap * 3].
static ContactOut contact_out(mwd_workspace* ws, bool tangents) {
    ContactOut co;
    double* tb = tangents ? ws->ctan : nullptr;
    const int64_t stride = ws->contact_cap * 3;
    co.Fn = tarr_of(ws->Fn, tb, stride, 0);
    co.tA = tarr_of(ws->tA, tb, stride, 1);
    co.tB = tarr_of(ws->tB, tb, stride, 2);
    if (!tangents && ws->emit_faces) co.face = ws->adj_face;
    co.sidx = ws->sidx;
    if (!tangents && ws->hist_on) co.hist = ws->thist;
    return co;
}

static int fine_contact(mwd_workspace* ws, const mwd_state* st, double* const* tan);

MWD_API int mwd_fine_contact(mwd_workspace* ws, const mwd_state* st) { return fine_contact(ws, st, nullptr); }

static int fine_contact(mwd_workspace* ws, const mwd_state* st, double* const* tan) {
    if (!ws || !st || st->vmax != ws->vmax) {
        mwd_set_error("mwd_fine_contact: invalid argument");
        return MWD_E_INVALID;
    }
    hipStream_t s = ws->stream;
    const BodyView A = state_view(ws, st, 0), B = state_view(ws, st, 1);
    int64_t* ctr = ws->counters;
    mwd_mark(ws, 5);
    int rc = MWD_OK;
    if (!ws->fresh_pairs) {  // a second pass over the same candidates: restart the contact and survivor counters
        if ((rc = mwd_set_i64(ctr + 4, 0, s)) || (rc = mwd_set_i64(ctr + 10, 0, s)) ||
            (rc = mwd_set_i64(ctr + 11, 0, s)))
            return rc;
    }
    // the candidate pass already sphere-tested these pairs (unless this is a
    // second pass over them)
    const bool tested = ws->fresh_pairs && ws->sphere_folded;
    ws->fresh_pairs = 0;
    if (!tested) {
        rc = mwd_launch_sphere(A, B, ws->pair_cap, ctr + 3, ws->collision, ws->surv, ctr + 10, s,
                               mwd_kev(ws, MWD_KT_SPHERE, 0), mwd_kev(ws, MWD_KT_SPHERE, 1), ws->tune);
        if (rc) return rc;
    }
    mwd_mark(ws, 6);
    if (ws->hist_on && !tan && (rc = mwd_launch_hist_prep(ws->thist, ws->contact_cap, s))) return rc;
    if (ws->bkt_dirty) {  // the previous contacts were not consumed by a reduction
        if ((rc = mwd_fill_i32(ws->rec_cnt, 0, ws->n_rows, s)) || (rc = mwd_set_i64(ctr + 11, 0, s))) return rc;
    }
    // a colliding pair claims contact slot c (counter 4) and writes
    // contact_pair[c]: contacts are in no set order; the per-row sums key
    // their records by pair index, so they are summed in candidate order
    // GJK per class of vertex count (gjk_buckets): the survivors sorted by
    // class, then one launch per class over its range of the sorted list,
    // each with the vertices per lane its class needs; results do not depend
    // on the order in which pairs are processed
    const int G = ws->tune.gjk_group;
    const int Kv = (ws->vmax + G - 1) / G;
    const int nb = ws->tune.gjk_buckets ? (Kv < MWD_GJK_BUCKETS ? Kv : MWD_GJK_BUCKETS) : 1;
    const int32_t* surv = ws->surv;
    if (nb > 1) {
        if ((rc = mwd_launch_gjk_buckets(A, B, ws->surv, ctr + 10, ws->pair_cap, G, nb, ws->bucket, ws->surv_cls, s)))
            return rc;
        surv = ws->surv_cls;
    }
    const int64_t* ofs = ws->bucket + 2 * MWD_GJK_BUCKETS;
    hipEvent_t g0 = mwd_kev(ws, MWD_KT_GJK, 0), g1 = mwd_kev(ws, MWD_KT_GJK, 1);
    bool fused = false;
    if (!tan && ws->tune.epa_fused && Kv <= 4) {
        for (int b = 0; b < nb; ++b) {
            rc = mwd_launch_gjk_epa(A, B, ws->pair_cap, ws->collision, surv, nb > 1 ? ofs + b + 1 : ctr + 10,
                                    ws->simplex, ws->sidx, ctr + 4, ws->contact_cap, ws->contact_pair, ws->laws,
                                    ws->depth, ws->normal, ws->cpA, ws->cpB, contact_out(ws, false), ws->epa_status,
                                    ws->rec_cnt, ws->rec_bkt, ws->rec_ovf, ctr + 11, 2 * ws->contact_cap, ws->flags,
                                    s, b == 0 ? g0 : nullptr, b == nb - 1 ? g1 : nullptr, ws->tune,
                                    nb > 1 ? ofs + b : nullptr, nb > 1 ? (b == nb - 1 ? Kv : b + 1) : 0);
            if (rc < 0) return rc;
            if (rc > 0) break;  // not this form (nothing launched: b == 0)
        }
        fused = rc == 0;
    }
    if (!fused) {
        for (int b = 0; b < nb; ++b) {
            rc = mwd_launch_gjk_surv(A, B, ws->pair_cap, ws->collision, surv, nb > 1 ? ofs + b + 1 : ctr + 10,
                                     ws->simplex, ws->sidx, ctr + 4, ws->contact_cap, ws->flags, s,
                                     b == 0 ? g0 : nullptr, b == nb - 1 ? g1 : nullptr, ws->tune, ws->contact_pair,
                                     nb > 1 ? ofs + b : nullptr, nb > 1 ? (b == nb - 1 ? Kv : b + 1) : 0);
            if (rc) return rc;
        }
    }
    mwd_mark(ws, 7);  // (no compaction stage: kept for the stage numbering)
    mwd_mark(ws, 8);
    if (fused) {
        // EPA ran inside the GJK launch: no EPA kernel time for this step
    } else if (tan) {
        const VState vs = mwd_vstate(st, tan);
        BodyTan bt;
        bt.com = vs.com;
        bt.verts = vs.verts;
        bt.vel = vs.v;
        bt.angvel = vs.w;
        bt.quat = vs.q;
        rc = mwd_launch_epa_force_rows(A, B, &bt, ws->laws, ws->law_tans, ws->contact_pair, ws->simplex, ws->sidx,
                                       ws->contact_cap, ctr + 4, ws->depth, ws->normal, ws->cpA, ws->cpB,
                                       contact_out(ws, true), ws->epa_status, ws->rec_cnt, ws->rec_bkt, ws->rec_ovf,
                                       ctr + 11, 2 * ws->contact_cap, ws->flags, s,
                                       mwd_kev(ws, MWD_KT_EPA, 0), mwd_kev(ws, MWD_KT_EPA, 1));
    } else {
        rc = mwd_launch_epa_force_rows(A, B, nullptr, ws->laws, nullptr, ws->contact_pair, ws->simplex, ws->sidx,
                                       ws->contact_cap, ctr + 4, ws->depth, ws->normal, ws->cpA, ws->cpB,
                                       contact_out(ws, false), ws->epa_status, ws->rec_cnt, ws->rec_bkt, ws->rec_ovf,
                                       ctr + 11, 2 * ws->contact_cap, ws->flags, s, mwd_kev(ws, MWD_KT_EPA, 0),
                                       mwd_kev(ws, MWD_KT_EPA, 1), ws->tune.epa_blocks);
    }
    if (rc) return rc;
    if (ws->hist_on && !tan && (rc = mwd_launch_hist_flip(ws->thist, s))) return rc;
    ws->bkt_dirty = 1;
    return MWD_OK;
}

MWD_API int mwd_contacts_view(mwd_workspace* ws, int32_t** collision, int32_t** contact_pair, double** depth,
                              double** normal, double** cpA, double** cpB, double** Fn, double** tauA, double** tauB,
                              int32_t** epa_status, int64_t** n_contacts) {
    if (!ws) return MWD_E_INVALID;
    if (collision) *collision = ws->collision;
    if (contact_pair) *contact_pair = ws->contact_pair;
    if (depth) *depth = ws->depth;
    if (normal) *normal = ws->normal;
    if (cpA) *cpA = ws->cpA;
    if (cpB) *cpB = ws->cpB;
    if (Fn) *Fn = ws->Fn;
    if (tauA) *tauA = ws->tA;
    if (tauB) *tauB = ws->tB;
    if (epa_status) *epa_status = ws->epa_status;
    if (n_contacts) *n_contacts = ws->counters + 4;
    return MWD_OK;
}

MWD_API int mwd_reduce_verlet(mwd_workspace* ws, mwd_state* st) {
    if (!ws || !st) {
        mwd_set_error("mwd_reduce_verlet: invalid argument");
        return MWD_E_INVALID;
    }
    mwd_mark(ws, 9);
    ws->bkt_dirty = 0;  // the reduction empties the buckets of the moving rows
    return mwd_launch_reduce_verlet(ws, st, nullptr, contact_out(ws, false));
}

namespace {
// Spawn schedule (MWDiem.py:685-714 creates particles every 5th step, at the
// start of the step): at the start of every spawn_every-th step, per more of
// the scene's rows become active, up to cap; the live counts that the
// kernels read follow. One thread.
__global__ void k_spawn(SceneDev* sc, int64_t* counters, int64_t every, int64_t per, int64_t cap) {
    if (threadIdx.x != 0) return;
    const int64_t i = counters[16];
    counters[16] = i + 1;
    if (i % every != 0) return;
    int64_t n = sc->n_active;
    if (n >= cap) return;
    n = n + per < cap ? n + per : cap;
    sc->n_active = n;
    counters[14] = n - sc->n_walls;
    counters[15] = sc->n_walls * (n - sc->n_walls);
}

int spawn_begin(mwd_workspace* ws) {
    if (!ws->spawn_on) return MWD_OK;
    hipLaunchKernelGGL(k_spawn, dim3(1), dim3(64), 0, ws->stream, ws->scenes_dev, ws->counters, ws->spawn_every,
                       ws->spawn_per, ws->spawn_cap);
    MWD_LAUNCH_CHECK();
    return MWD_OK;
}
}  // namespace

MWD_API int mwd_set_spawn(mwd_workspace* ws, int64_t n_active, int64_t every, int64_t per) {
    if (!ws || ws->n_scenes != 1 || every < 1 || per < 1) {
        mwd_set_error("mwd_set_spawn: a workspace of one scene, every >= 1 and per >= 1");
        return MWD_E_INVALID;
    }
    const mwd_scene& sc = ws->scenes_host[0];
    const int64_t cap = sc.n_active;  // the scene's rows as created
    if (n_active < sc.n_walls || n_active > cap) {
        mwd_set_error("mwd_set_spawn: n_active outside [n_walls, the scene's rows]");
        return MWD_E_INVALID;
    }
    if (!ws->cell_table_ok) {
        mwd_set_error("mwd_set_spawn: needs the key-table cell pass");
        return MWD_E_INVALID;
    }
    const int64_t live[3] = {n_active - sc.n_walls, sc.n_walls * (n_active - sc.n_walls), 0};
    MWD_HIP_CHECK(hipStreamSynchronize(ws->stream));
    MWD_HIP_CHECK(hipMemcpy(ws->m_live, live, sizeof(live), hipMemcpyHostToDevice));
    MWD_HIP_CHECK(hipMemcpy((char*)ws->scenes_dev + offsetof(SceneDev, n_active), &n_active, sizeof(int64_t),
                            hipMemcpyHostToDevice));
    ws->spawn_on = 1;
    ws->spawn_every = every;
    ws->spawn_per = per;
    ws->spawn_cap = cap;
    ws->cell_table = true;
    return MWD_OK;
}

MWD_API int mwd_step(mwd_workspace* ws, mwd_state* st) {
    int rc = spawn_begin(ws);
    if (rc) return rc;
    rc = mwd_candidates(ws, st, nullptr);
    if (rc) return rc;
    rc = mwd_fine_contact(ws, st);
    if (rc) return rc;
    rc = mwd_reduce_verlet(ws, st);
    if (rc) return rc;
    if (ws->wall_motion && (rc = mwd_launch_move_walls(ws, st))) return rc;
    mwd_mark_end(ws);
    return MWD_OK;
}

MWD_API int mwd_set_tangential_history(mwd_workspace* ws, int32_t mode) {
    if (!ws || (mode != MWD_HISTORY_RESET && mode != MWD_HISTORY_PERSISTENT)) {
        mwd_set_error("mwd_set_tangential_history: null workspace or unknown mode");
        return MWD_E_INVALID;
    }
    if (mode == MWD_HISTORY_PERSISTENT && !ws->thist.keys[0]) {
        MWD_HIP_CHECK(hipSetDevice(ws->device));
        int64_t H = 1024;
        while (H < 2 * ws->contact_cap) H <<= 1;  // load factor at most 1/2
        for (int t = 0; t < 2; ++t) {
            MWD_HIP_CHECK(hipMalloc((void**)&ws->thist.keys[t], sizeof(int64_t) * H));
            MWD_HIP_CHECK(hipMalloc((void**)&ws->thist.val[t], sizeof(double) * 4 * H));
            MWD_HIP_CHECK(hipMalloc((void**)&ws->thist.used[t], sizeof(int32_t) * ws->contact_cap));
            MWD_HIP_CHECK(hipMemsetAsync(ws->thist.keys[t], 0xFF, sizeof(int64_t) * H, ws->stream));
            ws->thist.count[t] = ws->counters + 12 + t;
        }
        MWD_HIP_CHECK(hipMalloc((void**)&ws->thist.idx, sizeof(int32_t)));
        MWD_HIP_CHECK(hipMemsetAsync(ws->thist.idx, 0, sizeof(int32_t), ws->stream));
        MWD_HIP_CHECK(hipMemsetAsync(ws->counters + 12, 0, 2 * sizeof(int64_t), ws->stream));
        MWD_HIP_CHECK(hipStreamSynchronize(ws->stream));
        ws->thist.mask = H - 1;
    }
    ws->hist_on = mode == MWD_HISTORY_PERSISTENT ? 1 : 0;
    return MWD_OK;
}

MWD_API int mwd_set_wall_motion(mwd_workspace* ws, int32_t on) {
    if (!ws) {
        mwd_set_error("mwd_set_wall_motion: null workspace");
        return MWD_E_INVALID;
    }
    ws->wall_motion = on ? 1 : 0;
    return MWD_OK;
}

MWD_API int mwd_set_reduction(mwd_workspace* ws, int32_t mode) {
    if (!ws || (mode != MWD_REDUCE_FP64 && mode != MWD_REDUCE_FP32)) {
        mwd_set_error("mwd_set_reduction: null workspace or unknown mode");
        return MWD_E_INVALID;
    }
    ws->reduce_fp32 = mode == MWD_REDUCE_FP32 ? 1 : 0;
    return MWD_OK;
}

// Law coefficients as functions of (Kn, en, et, mu), in the operator order of
// mwd_law_init_mode; T = Dual gives their derivatives.
template <class T>
static void law_coeffs(T Kn, T en, T et, T mu, T* out /* Kn mu Gn Kt_num inv_mu ln_et cn ct ratio */) {
    const double pi = 3.141592653589793;
    const double pi2 = pi * pi;
    const T ln_en = dlog(en), ln_et = dlog(et);
    const T A_en = pi2 + ln_en * ln_en;
    const T tc = dsqrt((mu * A_en) / Kn);
    out[0] = Kn;
    out[1] = mu;
    out[2] = ((-2.0 * mu) / tc) * ln_en;
    out[3] = ((Kn * (pi2 + ln_et * ln_et)) / A_en) * 1.0;
    out[4] = 1.0 / mu;
    out[5] = ln_et;
    out[6] = (-2.0 * ln_en) / dsqrt(A_en);
    out[7] = (-2.0 * ln_et) / dsqrt(pi2 + ln_et * ln_et);
    out[8] = (pi2 + ln_et * ln_et) / A_en;
}

MWD_API int mwd_law_tangent(mwd_law_tan* out, double Kn, double en, double et, double mu, const double* seed) {
    if (!out || !seed || !(Kn > 0) || !(en > 0) || !(et > 0) || !(mu > 0)) {
        mwd_set_error("mwd_law_tangent: invalid argument");
        return MWD_E_INVALID;
    }
    typedef Dual<MWD_NT> D;
    D in[4];
    const double v[4] = {Kn, en, et, mu};
    for (int i = 0; i < 4; ++i) {
        in[i].v = v[i];
        for (int k = 0; k < MWD_NT; ++k) in[i].d[k] = seed[MWD_NT * i + k];
    }
    D c[9];
    law_coeffs<D>(in[0], in[1], in[2], in[3], c);
    double* dst[9] = {out->Kn, out->mu, out->Gn, out->Kt_num, out->inv_mu, out->ln_et, out->cn, out->ct, out->ratio};
    for (int j = 0; j < 9; ++j)
        for (int k = 0; k < MWD_NT; ++k) dst[j][k] = c[j].d[k];
    return MWD_OK;
}

MWD_API int mwd_set_law_tangents(mwd_workspace* ws, const mwd_law_tan* laws2) {
    if (!ws || !laws2) {
        mwd_set_error("mwd_set_law_tangents: invalid argument");
        return MWD_E_INVALID;
    }
    MWD_HIP_CHECK(hipSetDevice(ws->device));
    if (!ws->law_tans) MWD_HIP_CHECK(hipMalloc((void**)&ws->law_tans, sizeof(mwd_law_tan) * 2 * ws->n_scenes));
    if (!ws->ctan) MWD_HIP_CHECK(hipMalloc((void**)&ws->ctan, sizeof(double) * 3 * MWD_NT * ws->contact_cap * 3));
    MWD_HIP_CHECK(hipMemcpyAsync(ws->law_tans, laws2, sizeof(mwd_law_tan) * 2 * ws->n_scenes, hipMemcpyHostToDevice,
                                 ws->stream));
    MWD_HIP_CHECK(hipStreamSynchronize(ws->stream));
    return MWD_OK;
}

MWD_API int mwd_step_tangent(mwd_workspace* ws, mwd_state* st, double* const* tan) {
    if (!ws || !st || !tan) {
        mwd_set_error("mwd_step_tangent: invalid argument");
        return MWD_E_INVALID;
    }
    if (!ws->law_tans || !ws->ctan) {
        mwd_set_error("mwd_step_tangent: call mwd_set_law_tangents first");
        return MWD_E_INVALID;
    }
    if (ws->hist_on) {
        mwd_set_error("mwd_step_tangent: not with a persistent tangential history");
        return MWD_E_INVALID;
    }
    for (int i = 0; i < MWD_N_TAN_ARRAYS * MWD_NT; ++i)
        if (!tan[i]) {
            mwd_set_error("mwd_step_tangent: tangent array " + std::to_string(i) + " is NULL");
            return MWD_E_INVALID;
        }
    int rc = spawn_begin(ws);
    if (rc) return rc;
    rc = mwd_candidates(ws, st, nullptr);
    if (rc) return rc;
    rc = fine_contact(ws, st, tan);
    if (rc) return rc;
    mwd_mark(ws, 9);
    ws->bkt_dirty = 0;
    rc = mwd_launch_reduce_verlet(ws, st, tan, contact_out(ws, true));
    if (rc) return rc;
    // walls move independently of theta: their tangents stay zero
    if (ws->wall_motion && (rc = mwd_launch_move_walls(ws, st))) return rc;
    mwd_mark_end(ws);
    return MWD_OK;
}

// What the adjoint step and the tape need of a workspace (name: the entry point).
static int adjoint_ok(mwd_workspace* ws, const char* name) {
    if (ws->spawn_on) {
        mwd_set_error(std::string(name) + ": not with a spawn schedule");
        return MWD_E_INVALID;
    }
    if (ws->hist_on) {
        mwd_set_error(std::string(name) + ": not with a persistent tangential history");
        return MWD_E_INVALID;
    }
    for (int s = 0; s < ws->n_scenes; ++s)
        if (ws->scenes_host[s].law_p.mode != MWD_LAW_PHYSICAL || ws->scenes_host[s].law_w.mode != MWD_LAW_PHYSICAL) {
            mwd_set_error(std::string(name) + ": scene " + std::to_string(s) + " is not on the physical law");
            return MWD_E_INVALID;
        }
    return MWD_OK;
}

// The adjoint step's buffers, at the first adjoint call or tape set-up (so
// that later calls can be captured in a graph).
static int alloc_adjoint(mwd_workspace* ws) {
    if (ws->adj_F) return MWD_OK;
    MWD_HIP_CHECK(hipSetDevice(ws->device));
    const size_t r3 = sizeof(double) * 3 * (size_t)ws->n_rows;
    MWD_HIP_CHECK(hipMalloc((void**)&ws->adj_F, r3));
    MWD_HIP_CHECK(hipMalloc((void**)&ws->adj_T, r3));
    MWD_HIP_CHECK(hipMalloc((void**)&ws->adj_lamF, r3));
    MWD_HIP_CHECK(hipMalloc((void**)&ws->adj_lamT, r3));
    MWD_HIP_CHECK(hipMalloc((void**)&ws->adj_face, sizeof(int32_t) * 7 * (size_t)ws->contact_cap));
    MWD_HIP_CHECK(hipMalloc((void**)&ws->adj_c, sizeof(double) * 36 * (size_t)ws->contact_cap));
    MWD_HIP_CHECK(hipMalloc((void**)&ws->adj_cth, sizeof(double) * 3 * (size_t)ws->contact_cap));
    MWD_HIP_CHECK(hipMalloc((void**)&ws->adj_rowth, r3));
    return MWD_OK;
}

static int adjoint_args(mwd_workspace* ws, mwd_state* st, double* const* lam, double* dtheta, const char* name) {
    if (!ws || !st || !lam || !dtheta || st->vmax != ws->vmax) {
        mwd_set_error(std::string(name) + ": invalid argument");
        return MWD_E_INVALID;
    }
    int rc = adjoint_ok(ws, name);
    if (rc) return rc;
    for (int i = 0; i < MWD_N_TAN_ARRAYS; ++i)
        if (!lam[i]) {
            mwd_set_error(std::string(name) + ": adjoint array " + std::to_string(i) + " is NULL");
            return MWD_E_INVALID;
        }
    if (!ws->law_tans) {
        mwd_set_error(std::string(name) +
                      ": call mwd_set_law_tangents first (the parameter derivatives of the laws)");
        return MWD_E_INVALID;
    }
    return alloc_adjoint(ws);
}

MWD_API int mwd_step_adjoint(mwd_workspace* ws, mwd_state* st, double* const* lam, double* dtheta) {
    int rc = adjoint_args(ws, st, lam, dtheta, "mwd_step_adjoint");
    if (rc) return rc;
    rc = mwd_candidates(ws, st, nullptr);
    if (rc) return rc;
    ws->emit_faces = 1;
    rc = fine_contact(ws, st, nullptr);
    ws->emit_faces = 0;
    if (rc) return rc;
    rc = mwd_launch_step_adjoint(ws, st, state_view(ws, st, 0), state_view(ws, st, 1), lam, dtheta);
    ws->bkt_dirty = 0;  // the gather emptied the buckets
    return rc;
}

MWD_API int mwd_tape_init(mwd_workspace* ws, int64_t n_slots, int64_t* slot_bytes) {
    if (!ws || n_slots < 0) {
        mwd_set_error("mwd_tape_init: null workspace or negative slot count");
        return MWD_E_INVALID;
    }
    int rc = adjoint_ok(ws, "mwd_tape_init");
    if (rc || (rc = alloc_adjoint(ws))) return rc;
    const int64_t words = mwd_tape_words(ws);
    if (slot_bytes) *slot_bytes = 4 * words;
    MWD_HIP_CHECK(hipSetDevice(ws->device));
    MWD_HIP_CHECK(hipStreamSynchronize(ws->stream));
    if (ws->tape) MWD_HIP_CHECK(hipFree(ws->tape));
    ws->tape = nullptr;
    ws->tape_slots = ws->tape_words = 0;
    if (n_slots > 0) {
        MWD_HIP_CHECK(hipMalloc((void**)&ws->tape, sizeof(uint32_t) * (size_t)words * (size_t)n_slots));
        ws->tape_slots = n_slots;
        ws->tape_words = words;
    }
    if (!ws->tape_sync) {
        MWD_HIP_CHECK(hipMalloc((void**)&ws->tape_sync, sizeof(uint32_t) * MWD_TAPE_SYNC_WORDS));
        MWD_HIP_CHECK(hipMemsetAsync(ws->tape_sync, 0, sizeof(uint32_t) * MWD_TAPE_SYNC_WORDS, ws->stream));
    }
    if ((rc = mwd_launch_tape_rewind(ws))) return rc;
    MWD_HIP_CHECK(hipStreamSynchronize(ws->stream));
    return MWD_OK;
}

MWD_API int mwd_tape_rewind(mwd_workspace* ws) {
    if (!ws) return MWD_E_INVALID;
    return mwd_launch_tape_rewind(ws);
}

MWD_API int mwd_step_record(mwd_workspace* ws, mwd_state* st) {
    if (!ws || !st || !ws->tape) {
        mwd_set_error("mwd_step_record: invalid argument, or no tape (mwd_tape_init)");
        return MWD_E_INVALID;
    }
    int rc = adjoint_ok(ws, "mwd_step_record");
    if (rc || (rc = mwd_candidates(ws, st, nullptr))) return rc;
    ws->emit_faces = 1;
    rc = fine_contact(ws, st, nullptr);
    ws->emit_faces = 0;
    if (rc || (rc = mwd_launch_tape(ws, st, true)) || (rc = mwd_reduce_verlet(ws, st))) return rc;
    if (ws->wall_motion && (rc = mwd_launch_move_walls(ws, st))) return rc;
    mwd_mark_end(ws);
    return MWD_OK;
}

MWD_API int mwd_step_adjoint_taped(mwd_workspace* ws, mwd_state* st, double* const* lam, double* dtheta) {
    int rc = adjoint_args(ws, st, lam, dtheta, "mwd_step_adjoint_taped");
    if (rc) return rc;
    if (!ws->tape) {
        mwd_set_error("mwd_step_adjoint_taped: no tape (mwd_tape_init)");
        return MWD_E_INVALID;
    }
    if ((rc = mwd_launch_tape(ws, st, false))) return rc;
    rc = mwd_launch_step_adjoint(ws, st, state_view(ws, st, 0), state_view(ws, st, 1), lam, dtheta);
    ws->bkt_dirty = 0;  // the gather emptied the buckets
    return rc;
}

MWD_API int mwd_copy_arrays(int32_t n, void* const* dst, const void* const* src, const int64_t* bytes, void* stream) {
    if (n < 0 || n > MWD_MAX_COPY || (n > 0 && (!dst || !src || !bytes))) {
        mwd_set_error("mwd_copy_arrays: n must be 0.." + std::to_string(MWD_MAX_COPY) + " with arrays given");
        return MWD_E_INVALID;
    }
    int64_t words[MWD_MAX_COPY];
    for (int i = 0; i < n; ++i) {
        if (bytes[i] < 0 || bytes[i] % 8 || (bytes[i] && (!dst[i] || !src[i])) ||
            ((uintptr_t)dst[i] | (uintptr_t)src[i]) % 8) {
            mwd_set_error("mwd_copy_arrays: array " + std::to_string(i) +
                          ": sizes must be multiples of 8 bytes and pointers 8-byte aligned");
            return MWD_E_INVALID;
        }
        words[i] = bytes[i] / 8;
    }
    return mwd_launch_copy_arrays(n, dst, src, words, (hipStream_t)stream);
}

MWD_API int mwd_radius_init(const double* com, const double* verts, const int32_t* nverts, int64_t n, int32_t vmax,
                            double* radius, void* stream) {
    if (!com || !verts || !nverts || !radius || n < 0 || vmax <= 0) {
        mwd_set_error("mwd_radius_init: invalid argument");
        return MWD_E_INVALID;
    }
    return mwd_launch_radius(com, verts, nverts, n, vmax, radius, (hipStream_t)stream);
}

MWD_API int mwd_step_status(mwd_workspace* ws, mwd_status* out) {
    if (!ws || !out) return MWD_E_INVALID;
    int64_t c[11];
    uint32_t f = 0;
    MWD_HIP_CHECK(hipMemcpyAsync(c, ws->counters, sizeof(c), hipMemcpyDeviceToHost, ws->stream));
    MWD_HIP_CHECK(hipMemcpyAsync(&f, ws->flags, sizeof(f), hipMemcpyDeviceToHost, ws->stream));
    MWD_HIP_CHECK(hipStreamSynchronize(ws->stream));
    out->n_cells = c[0];
    out->n_wall_pairs = c[1];
    out->n_pairs = c[7];
    out->n_contacts = c[4];
    out->flags = f;
    out->n_sphere_pass = c[10];
    return MWD_OK;
}

MWD_API int mwd_clear_flags(mwd_workspace* ws) {
    if (!ws) return MWD_E_INVALID;
    MWD_HIP_CHECK(hipMemsetAsync(ws->flags, 0, sizeof(uint32_t), ws->stream));
    return MWD_OK;
}

// ------------------------------------------------------- per-pair fine contact
namespace {

__global__ void k_put_laws(mwd_law* dst, mwd_law a, mwd_law b) {
    if (threadIdx.x == 0) {
        dst[0] = a;
        dst[1] = b;
    }
}

// Temporaries of one per-pair fine-contact call, carved from one device
// buffer: per pair the collision code, contact slot, EPA status, simplex and
// per-contact results; per record (two per pair) the sort keys and values;
// the radix histograms, the scan scratch, counters and the two laws.
struct PairScratch {
    int32_t *coll, *cpair, *est, *rval, *rval2, *sb, *se, *hist, *sidx, *surv;
    uint8_t *wallA, *wallB;
    Simplex* sx;
    int64_t *ctr, *hist_off, *tmp;
    double *dep, *nrm, *ca, *cb, *fn, *ta, *tb;
    uint32_t *rkey, *rkey2, *flags;
    mwd_law* laws;
    int64_t tmp_len;
};

// Lays the temporaries for P pairs and n_particles force rows out from
// `base` (each buffer 256-byte aligned) and returns the bytes they take;
// with base == nullptr it only counts.
size_t carve_pair_scratch(char* base, int64_t P, int64_t n_particles, bool with_force, PairScratch* ps) {
    size_t off = 0;
    auto take = [&](auto** p, int64_t count) {
        using T = typename std::remove_pointer<typename std::remove_pointer<decltype(p)>::type>::type;
        off = (off + 255) & ~size_t(255);
        *p = base ? reinterpret_cast<T*>(base + off) : nullptr;
        off += sizeof(T) * (size_t)std::max<int64_t>(count, 1);
    };
    const int64_t rec = 2 * P;
    const int64_t tiles = mwd_sort_tiles(rec);
    const int64_t hist_len = (int64_t(1) << MWD_SORT_MAX_BITS) * tiles;
    ps->tmp_len = mwd_scan_tmp_len(std::max({P, rec, hist_len, n_particles}));
    take(&ps->coll, P);
    take(&ps->cpair, P);
    take(&ps->est, P);
    take(&ps->rval, rec);
    take(&ps->rval2, rec);
    take(&ps->rkey, rec);
    take(&ps->rkey2, rec);
    take(&ps->wallA, P);
    take(&ps->wallB, P);
    take(&ps->sx, P);
    take(&ps->sidx, P);
    take(&ps->surv, P);
    take(&ps->ctr, 16);
    take(&ps->hist, hist_len);
    take(&ps->hist_off, hist_len);
    take(&ps->tmp, ps->tmp_len);
    take(&ps->dep, P);
    take(&ps->nrm, 3 * P);
    take(&ps->ca, 3 * P);
    take(&ps->cb, 3 * P);
    take(&ps->fn, 3 * P);
    take(&ps->ta, 3 * P);
    take(&ps->tb, 3 * P);
    take(&ps->flags, 1);
    take(&ps->laws, 2);
    ps->sb = ps->se = nullptr;
    if (with_force) {
        take(&ps->sb, n_particles);
        take(&ps->se, n_particles);
    }
    return off;
}

int check_pair_args(const char* fn, int64_t P, int32_t vmax, const mwd_law* law_p, const mwd_law* law_w,
                    int64_t n_particles) {
    if (P < 0 || P > INT32_MAX / 4 || vmax <= 0 || vmax > 255 || !law_p || !law_w || n_particles < 0) {
        mwd_set_error(std::string(fn) + ": invalid argument");
        return MWD_E_INVALID;
    }
    return MWD_OK;
}

int check_pair_arrays(const char* fn, const double* comA, const double* comB, const double* vertsA,
                      const double* vertsB, const double* vA, const double* vB, const double* wA, const double* wB,
                      const int32_t* nA, const int32_t* nB, const int32_t* idA, const int32_t* idB, const double* qA,
                      const double* isA, const double* isB, const mwd_law* law_p, const mwd_law* law_w,
                      const double* force, const double* torque, int64_t n_particles) {
    if (!comA || !comB || !vertsA || !vertsB || !vA || !vB || !wA || !wB || !nA || !nB || !idA || !idB || !qA ||
        !isA || !isB || ((force || torque) && (!force || !torque || n_particles <= 0))) {
        mwd_set_error(std::string(fn) + ": null array");
        return MWD_E_INVALID;
    }
    if (law_p->mode != MWD_LAW_REF || law_w->mode != MWD_LAW_REF) {
        mwd_set_error(std::string(fn) + ": the per-pair layout carries no masses; use MWD_LAW_REF");
        return MWD_E_INVALID;
    }
    return MWD_OK;
}

// GJK, EPA and the law over P pairs in the reference's per-pair layout, the
// per-pair outputs scattered back to pair order and (with force) the
// per-particle contact sums added in ascending pair order. Every temporary
// is in `ps`; nothing is allocated. Stream-ordered (no synchronisation).
int contact_pairs_core(const PairScratch& ps, int64_t P, int32_t vmax, const double* comA, const double* comB,
                       const double* vertsA, const double* vertsB, const double* vA, const double* vB,
                       const double* wA, const double* wB, const int32_t* nA, const int32_t* nB, const int32_t* idA,
                       const int32_t* idB, const double* dispt, const double* qA, const double* isA,
                       const double* isB, const mwd_law* law_p, const mwd_law* law_w, int64_t n_walls,
                       int32_t* collision, int32_t* epa_status, double* depth, double* normal, double* cpA,
                       double* cpB, double* disp, double* force, double* torque, int64_t n_particles,
                       hipStream_t s) {
    int rc;
    const int64_t rec = 2 * P;
    // counters, scan scratch and flags start at zero (fill kernels, no memset nodes)
    if ((rc = mwd_fill_i32(reinterpret_cast<int32_t*>(ps.ctr), 0, 32, s))) return rc;
    if ((rc = mwd_fill_i32(reinterpret_cast<int32_t*>(ps.tmp), 0, 2 * ps.tmp_len, s))) return rc;
    if ((rc = mwd_fill_i32(reinterpret_cast<int32_t*>(ps.flags), 0, 1, s))) return rc;
    hipLaunchKernelGGL(k_put_laws, dim3(1), dim3(64), 0, s, ps.laws, *law_p, *law_w);
    MWD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_wall_from_ids, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, s, idA, idB, n_walls, P,
                       ps.wallA, ps.wallB);
    MWD_LAUNCH_CHECK();
    BodyView A{comA, vertsA, nA, vA, wA, qA, isA, nullptr, nullptr, ps.wallA, nullptr, nullptr, 1, vmax};
    BodyView B{comB, vertsB, nB, vB, wB, qA, isB, nullptr, nullptr, ps.wallB, nullptr, nullptr, 1, vmax};
    int64_t* ctr = ps.ctr;
    if ((rc = mwd_set_i64(ctr + 3, P, s))) return rc;
    if ((rc = mwd_launch_zero_pairs(P, depth, normal, cpA, cpB, epa_status, disp, s))) return rc;
    if ((rc = mwd_launch_gjk(A, B, P, ctr + 3, ps.coll, ps.surv, ctr + 10, ps.sx, ps.sidx, ctr + 8, P, ps.flags, s)))
        return rc;
    if ((rc = mwd_launch_compact_contacts(ps.coll, P, ctr + 3, ps.cpair, P, ctr + 4, ps.flags, ps.tmp, ps.tmp_len,
                                          s)))
        return rc;
    if ((rc = mwd_launch_epa_force(A, B, ps.laws, ps.cpair, ps.sx, ps.sidx, P, ctr + 4, dispt, idA, idB, ps.dep,
                                   ps.nrm, ps.ca, ps.cb, ps.fn, ps.ta, ps.tb, disp, ps.est, ps.rkey, ps.rval,
                                   ctr + 5, ps.flags, s)))
        return rc;
    // per-pair outputs scattered back from contact order
    if (collision) MWD_HIP_CHECK(hipMemcpyAsync(collision, ps.coll, sizeof(int32_t) * P, hipMemcpyDeviceToDevice, s));
    if ((rc = mwd_launch_scatter_contacts(ps.cpair, P, ctr + 4, ps.dep, ps.nrm, ps.ca, ps.cb, ps.est, P, depth,
                                          normal, cpA, cpB, epa_status, disp, s)))
        return rc;
    if (force) {
        uint32_t* ks;
        int32_t* vs;
        if ((rc = mwd_radix_sort(ps.rkey, ps.rval, ps.rkey2, ps.rval2, rec, ctr + 5, bits_for(n_particles), ps.hist,
                                 ps.hist_off, ctr + 9, ps.tmp, ps.tmp_len, s, &ks, &vs)))
            return rc;
        if ((rc = mwd_fill_i32(ps.sb, 0, n_particles, s))) return rc;
        if ((rc = mwd_fill_i32(ps.se, 0, n_particles, s))) return rc;
        if ((rc = mwd_launch_segments(ks, rec, ctr + 5, ps.sb, ps.se, s))) return rc;
        if ((rc = mwd_launch_add_sums(ps.sb, ps.se, vs, ps.fn, ps.ta, ps.tb, force, torque, n_particles, s)))
            return rc;
    }
    return MWD_OK;
}

int finish_pair_call(const char* fn, int rc, hipStream_t s) {
    if (hipStreamSynchronize(s) != hipSuccess && !rc) {
        mwd_set_error(std::string(fn) + ": kernel failure");
        rc = MWD_E_HIP;
    }
    return rc;
}

}  // namespace

MWD_API int mwd_contact_pairs(int64_t P, int32_t vmax, const double* comA, const double* comB, const double* vertsA,
                              const double* vertsB, const double* vA, const double* vB, const double* wA,
                              const double* wB, const int32_t* nA, const int32_t* nB, const int32_t* idA,
                              const int32_t* idB, const double* dispt, const double* qA, const double* isA,
                              const double* isB, const mwd_law* law_p, const mwd_law* law_w, int64_t n_walls,
                              int32_t* collision, int32_t* epa_status, double* depth, double* normal, double* cpA,
                              double* cpB, double* disp, double* force, double* torque, int64_t n_particles,
                              void* stream) {
    const char* fn = "mwd_contact_pairs";
    int rc = check_pair_args(fn, P, vmax, law_p, law_w, n_particles);
    if (rc || P == 0) return rc;
    if ((rc = check_pair_arrays(fn, comA, comB, vertsA, vertsB, vA, vB, wA, wB, nA, nB, idA, idB, qA, isA, isB, law_p,
                                law_w, force, torque, n_particles)))
        return rc;
    hipStream_t s = (hipStream_t)stream;
    PairScratch ps;
    const size_t bytes = carve_pair_scratch(nullptr, P, n_particles, force != nullptr, &ps);
    char* base = nullptr;
    if ((rc = dalloc(&base, (int64_t)bytes))) return rc;
    carve_pair_scratch(base, P, n_particles, force != nullptr, &ps);
    rc = contact_pairs_core(ps, P, vmax, comA, comB, vertsA, vertsB, vA, vB, wA, wB, nA, nB, idA, idB, dispt, qA, isA,
                            isB, law_p, law_w, n_walls, collision, epa_status, depth, normal, cpA, cpB, disp, force,
                            torque, n_particles, s);
    rc = finish_pair_call(fn, rc, s);
    (void)hipFree(base);
    return rc;
}

// ---------------------------------------------- the allocation-free variant
// The temporaries of mwd_contact_pairs_noalloc live in this array (one copy
// per GPU, resident while the library is loaded). Host code never uses the
// name itself as an address: each device's copy is looked up once with
// hipGetSymbolAddress and that pointer is what the kernels receive.
static __device__ __align__(256) unsigned char scratch[MWD_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxDevices] = {};

// Address of the current device's scratch array (looked up on its first use
// on that device).
int scratch_for_current_device(char** out) {
    int dev = -1;
    MWD_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        mwd_set_error("mwd_contact_pairs_noalloc: device index out of range");
        return MWD_E_INVALID;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            mwd_set_error(std::string("mwd_contact_pairs_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
            return MWD_E_HIP;
        }
        g_scratch_ptr[dev] = static_cast<char*>(p);
    }
    *out = g_scratch_ptr[dev];
    return MWD_OK;
}
}  // namespace

MWD_API int mwd_contact_pairs_noalloc(int64_t P, int32_t vmax, const double* comA, const double* comB,
                                      const double* vertsA, const double* vertsB, const double* vA, const double* vB,
                                      const double* wA, const double* wB, const int32_t* nA, const int32_t* nB,
                                      const int32_t* idA, const int32_t* idB, const double* dispt, const double* qA,
                                      const double* isA, const double* isB, const mwd_law* law_p,
                                      const mwd_law* law_w, int64_t n_walls, int32_t* collision, int32_t* epa_status,
                                      double* depth, double* normal, double* cpA, double* cpB, double* disp,
                                      double* force, double* torque, int64_t n_particles, void* stream) {
    const char* fn = "mwd_contact_pairs_noalloc";
    int rc = check_pair_args(fn, P, vmax, law_p, law_w, n_particles);
    if (rc) return rc;
    if (P > MWD_NOALLOC_MAX_PAIRS || n_particles > MWD_NOALLOC_MAX_PARTICLES) {
        mwd_set_error("mwd_contact_pairs_noalloc: more than MWD_NOALLOC_MAX_PAIRS pairs or "
                      "MWD_NOALLOC_MAX_PARTICLES particles");
        return MWD_E_INVALID;
    }
    if (P == 0) return MWD_OK;
    if ((rc = check_pair_arrays(fn, comA, comB, vertsA, vertsB, vA, vB, wA, wB, nA, nB, idA, idB, qA, isA, isB, law_p,
                                law_w, force, torque, n_particles)))
        return rc;
    PairScratch ps;
    if (carve_pair_scratch(nullptr, P, n_particles, force != nullptr, &ps) > sizeof(scratch)) {
        mwd_set_error("mwd_contact_pairs_noalloc: temporaries exceed MWD_NOALLOC_SCRATCH_BYTES");
        return MWD_E_INVALID;
    }
    char* base = nullptr;
    if ((rc = scratch_for_current_device(&base))) return rc;
    carve_pair_scratch(base, P, n_particles, force != nullptr, &ps);
    hipStream_t s = (hipStream_t)stream;
    rc = contact_pairs_core(ps, P, vmax, comA, comB, vertsA, vertsB, vA, vB, wA, wB, nA, nB, idA, idB, dispt, qA, isA,
                            isB, law_p, law_w, n_walls, collision, epa_status, depth, normal, cpA, cpB, disp, force,
                            torque, n_particles, s);
    return finish_pair_call(fn, rc, s);
}
