// Coarse contact detection: cell keys, wall slab test and the linked-cell
// candidate list, rebuilt as sort + scan + emit on the GPU.
//
// Reference semantics reproduced exactly (so the candidate list is identical,
// order included):
//   * keys: trunc(com/cellsize) per axis, key = xc + yc*ncx + zc*ncx*ncy
//     (MWDiem.py:813-816);
//   * wall pairs first: walls k outer, particles j inner, pair [j,k] when
//     |n_k.(x_j-c_k)| < 6*max_particlesize (MWDiem.py:807-811);
//   * then Linked_cell_hash (LIBRARY_Coarse_contact_detection.py:84-155):
//     cells in order of their lowest particle id (dict insertion order), the
//     members of a cell in CPython set iteration order, same-cell pairs j<k,
//     then the 7 forward offsets +1, +ncx, +ncx*ncy, +ncx*ncy+1,
//     +ncx*ncy+1+ncx, +ncx*ncy+ncx, +ncx+1 with A in the lower cell.
//     Scenes on the physical law use the full 13-cell forward half shell
//     instead (the reference's 7 offsets miss 6 neighbour directions).
//
// Pipeline: keys -> stable radix sort of (key, row) -> cell heads (scan) ->
// cell rank by lowest row (scan over rows) -> per-cell set order, neighbour
// lookup (binary search of the sorted cell keys) and pair count -> scan of
// counts in rank order -> emission. Wall tests are a flag/scan/scatter over
// the (scene, wall, particle) test space.
#include "mwd_internal.h"
#include "mwd_scan.h"

namespace {

__device__ __forceinline__ int64_t key_of(d3 c, const double* cell, int64_t ncx, int64_t ncy) {
    const int64_t xc = (int64_t)(c.x / cell[0]);
    const int64_t yc = (int64_t)(c.y / cell[1]);
    const int64_t zc = (int64_t)(c.z / cell[2]);
    return xc + yc * ncx + zc * ncx * ncy;
}

__global__ void k_cell_keys(const double* __restrict__ com, int64_t n, double cx, double cy, double cz, int64_t ncx,
                            int64_t ncy, int64_t* __restrict__ keys) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double cell[3] = {cx, cy, cz};
    keys[i] = key_of(ld3(com + 3 * i), cell, ncx, ncy);
}

// Global sort keys of the moving particles.
__global__ void k_sort_keys(const double* __restrict__ com, const int64_t* __restrict__ keys_in,
                            const int32_t* __restrict__ plist, int64_t M, const int32_t* __restrict__ scene_of,
                            const SceneDev* __restrict__ scenes, uint32_t* __restrict__ keys,
                            int32_t* __restrict__ vals, uint32_t* flags) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const int32_t row = plist[i];
    const SceneDev& sc = scenes[scene_of[row]];
    const int64_t key = keys_in ? keys_in[row] : key_of(ld3(com + 3 * row), sc.cell, sc.ncx, sc.ncy);
    int64_t loc = key - sc.kmin;
    if (loc < 0 || loc >= sc.kspan) {
        atomicOr(flags, MWD_F_KEYRANGE);
        loc = loc < 0 ? 0 : sc.kspan - 1;
    }
    keys[i] = (uint32_t)(sc.kbase + loc);
    vals[i] = row;
}

// CPython set iteration order of non-negative ints inserted in the given
// order (Objects/setobject.c: 9 linear probes, then perturbation; resize to
// the smallest power of two above 4*used once fill*5 >= mask*3).
__device__ void cpy_probe(int32_t* tab, uint32_t mask, int32_t key) {
    uint64_t perturb = (uint64_t)key;
    uint64_t i = (uint64_t)key & mask;
    for (;;) {
        if (tab[i] < 0) {
            tab[i] = key;
            return;
        }
        if (i + 9 <= mask) {
            for (uint64_t j = 1; j <= 9; ++j)
                if (tab[i + j] < 0) {
                    tab[i + j] = key;
                    return;
                }
        }
        perturb >>= 5;
        i = (i * 5 + 1 + perturb) & mask;
    }
}

constexpr int SET_TAB = 128;  // holds up to MWD_MAX_CELL members

__device__ void set_order(const int32_t* rows, int n, int64_t row0, int32_t* out) {
    if (n <= 1) {
        if (n == 1) out[0] = rows[0];
        return;
    }
    if (n <= 4) {
        // 8-slot table, never resized at this size; with mask 7 no linear
        // probe fits, so each collision perturbs. Slots live in registers.
        uint32_t occ = 0;
        int slot[4] = {8, 8, 8, 8};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (k < n) {
                const uint64_t key = (uint64_t)(rows[k] - row0);
                uint64_t perturb = key;
                uint32_t i = (uint32_t)(key & 7u);
                while ((occ >> i) & 1u) {
                    perturb >>= 5;
                    i = (uint32_t)((i * 5u + 1u + perturb) & 7u);
                }
                occ |= 1u << i;
                slot[k] = (int)i;
            }
        }
        int o = 0;
        for (int sl = 0; sl < 8; ++sl) {
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (slot[k] == sl) out[o++] = rows[k];
        }
        return;
    }
    int32_t tab[SET_TAB], tmp[SET_TAB];
    uint32_t size = 8;
    for (int k = 0; k < SET_TAB; ++k) tab[k] = -1;
    uint32_t fill = 0;
    for (int k = 0; k < n; ++k) {
        cpy_probe(tab, size - 1, (int32_t)(rows[k] - row0));
        ++fill;
        if (fill * 5 >= (size - 1) * 3) {
            const uint32_t minused = fill * 4;
            uint32_t ns = 8;
            while (ns <= minused) ns <<= 1;
            for (uint32_t q = 0; q < ns; ++q) tmp[q] = -1;
            for (uint32_t q = 0; q < size; ++q)
                if (tab[q] >= 0) cpy_probe(tmp, ns - 1, tab[q]);
            for (uint32_t q = 0; q < ns; ++q) tab[q] = tmp[q];
            size = ns;
        }
    }
    int o = 0;
    for (uint32_t q = 0; q < size; ++q)
        if (tab[q] >= 0) out[o++] = (int32_t)(tab[q] + row0);
}

__device__ __forceinline__ int64_t find_cell(const uint32_t* cell_key, int64_t C, uint32_t key) {
    int64_t lo = 0, hi = C;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (cell_key[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return (lo < C && cell_key[lo] == key) ? lo : -1;
}

__global__ void k_cell_info(const int32_t* __restrict__ vals, const int32_t* __restrict__ cell_start,
                            const uint32_t* __restrict__ cell_key, const int64_t* __restrict__ n_cells,
                            int64_t M, const int32_t* __restrict__ scene_of, const SceneDev* __restrict__ scenes,
                            const int64_t* __restrict__ row_rank, int32_t* __restrict__ perm,
                            int32_t* __restrict__ cell_nb, int32_t* __restrict__ cell_by_rank,
                            int32_t* __restrict__ cell_cnt, uint32_t* flags, const int32_t* __restrict__ cell_of) {
    const int64_t C = *n_cells;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const int32_t start = cell_start[c], end = cell_start[c + 1];
    const int n = end - start;
    const int32_t row_min = vals[start];
    const SceneDev& sc = scenes[scene_of[row_min]];
    if (n > MWD_MAX_CELL) {
        atomicOr(flags, MWD_F_CELL);
        for (int k = 0; k < n; ++k) perm[start + k] = vals[start + k];
    } else {
        set_order(vals + start, n, sc.row0, perm + start);
    }
    const int64_t lk = (int64_t)cell_key[c] - sc.kbase + sc.kmin;
    int64_t cnt = (int64_t)n * (n - 1) / 2;
    for (int o = 0; o < MWD_MAX_OFF; ++o) {
        int64_t nb = -1;
        if (o < sc.n_off) {
            const int64_t t = lk + sc.off[o] - sc.kmin;
            if (t >= 0 && t < sc.kspan)
                nb = cell_of ? cell_of[sc.kbase + t] : find_cell(cell_key, C, (uint32_t)(sc.kbase + t));
        }
        cell_nb[c * MWD_MAX_OFF + o] = (int32_t)nb;
        if (nb >= 0) cnt += (int64_t)n * (cell_start[nb + 1] - cell_start[nb]);
    }
    const int64_t r = row_rank[row_min];
    cell_by_rank[r] = (int32_t)c;
    cell_cnt[r] = (int32_t)cnt;
}

// Pair emission, one output pair per thread: a block takes 64 cells by rank,
// stages their pair offsets, members and neighbour cells in LDS, and its
// threads walk the block's output range, so consecutive threads write
// consecutive pairs. Pair q of a cell is decoded in the emission order of
// Linked_cell_hash (LIBRARY_Coarse_contact_detection.py:84-118): same-cell
// pairs (j < k, members in set order), then per neighbour offset o every
// member j against every member k of that cell.
constexpr int FILL_CELLS = 64;    // cells per block
constexpr int FILL_THREADS = 256;

__global__ void __launch_bounds__(FILL_THREADS) k_cell_fill(const int32_t* __restrict__ perm,
                                                          const int32_t* __restrict__ cell_start,
                                                          const int32_t* __restrict__ cell_nb,
                                                          const int32_t* __restrict__ cell_by_rank,
                                                          const int64_t* __restrict__ cell_off,
                                                          const int64_t* __restrict__ counters,
                                                          int32_t* __restrict__ pairs, int64_t cap) {
    __shared__ int64_t s_off[FILL_CELLS + 1];
    __shared__ int32_t s_s0[FILL_CELLS], s_n[FILL_CELLS];
    __shared__ int32_t s_t0[FILL_CELLS][MWD_MAX_OFF], s_tn[FILL_CELLS][MWD_MAX_OFF];
    const int64_t C = counters[0];
    const int64_t r0 = (int64_t)blockIdx.x * FILL_CELLS;
    if (r0 >= C) return;  // uniform per block
    const int nc = (int)(C - r0 < FILL_CELLS ? C - r0 : FILL_CELLS);
    const int tid = threadIdx.x;
    const int64_t base = cell_off[r0];
    if (tid < nc) {
        const int32_t c = cell_by_rank[r0 + tid];
        const int32_t s0 = cell_start[c];
        const int32_t n = cell_start[c + 1] - s0;
        s_s0[tid] = s0;
        s_n[tid] = n;
        int64_t cnt = (int64_t)n * (n - 1) / 2;
        for (int o = 0; o < MWD_MAX_OFF; ++o) {
            const int32_t nb = cell_nb[(int64_t)c * MWD_MAX_OFF + o];
            const int32_t t0 = nb >= 0 ? cell_start[nb] : 0;
            const int32_t tn = nb >= 0 ? cell_start[nb + 1] - t0 : 0;
            s_t0[tid][o] = t0;
            s_tn[tid][o] = tn;
            cnt += (int64_t)n * tn;
        }
        s_off[tid + 1] = cell_off[r0 + tid] - base + cnt;  // end of this cell's range
    }
    if (tid == 0) s_off[0] = 0;
    __syncthreads();
    const int64_t total = s_off[nc];
    const int64_t pos0 = counters[1] + base;
    for (int64_t q = tid; q < total; q += FILL_THREADS) {
        // cell i: s_off[i] <= q < s_off[i + 1]
        int lo = 0, hi = nc;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (s_off[mid] <= q) lo = mid;
            else hi = mid;
        }
        const int i = lo;
        const int32_t n = s_n[i], s0 = s_s0[i];
        int64_t t = q - s_off[i];
        const int64_t tri = (int64_t)n * (n - 1) / 2;
        int32_t a, b;
        if (t < tri) {
            int j = 0;
            int64_t row = n - 1;  // pairs of member j with the later members
            while (t >= row) {
                t -= row;
                ++j;
                --row;
            }
            a = perm[s0 + j];
            b = perm[s0 + j + 1 + (int32_t)t];
        } else {
            t -= tri;
            int o = 0;
            for (; o < MWD_MAX_OFF; ++o) {
                const int64_t sz = (int64_t)n * s_tn[i][o];
                if (t < sz) break;
                t -= sz;
            }
            if (o == MWD_MAX_OFF) continue;  // not reached: q lies inside cell i's range
            const int32_t tn = s_tn[i][o];
            const int32_t j = (int32_t)(t / tn), k = (int32_t)(t - (int64_t)j * tn);
            a = perm[s0 + j];
            b = perm[s_t0[i][o] + k];
        }
        const int64_t pos = pos0 + q;
        if (pos < cap) {
            int2 v;
            v.x = a;
            v.y = b;
            reinterpret_cast<int2*>(pairs)[pos] = v;
        }
    }
}

__device__ __forceinline__ int scene_of_test(const int64_t* wt_off, int S, int64_t t) {
    int lo = 0, hi = S;  // largest s with wt_off[s] <= t
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (wt_off[mid] <= t) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Wall slab test (MWDiem.py:807-811) fused with its compaction: test t of a
// scene pairs wall k with particle j (walls outer, particles inner); a pair
// whose centre lies within wall_thr of the wall's plane is written, as
// [j, k], at its offset among the scene-ordered hits.
struct WallIO {
    const double* com;
    const double* wall_n;
    const double* wall_c;
    const SceneDev* scenes;
    int S;
    const int64_t* wt_off;
    int32_t* pairs;
    int64_t cap;
    __device__ __forceinline__ void decode(int64_t t, const SceneDev** sc, int64_t* rj, int64_t* rk) const {
        const int s = scene_of_test(wt_off, S, t);
        *sc = &scenes[s];
        const int64_t np = (*sc)->n_active - (*sc)->n_walls;
        const int64_t tl = t - wt_off[s];
        const int64_t k = tl / np, j = tl - k * np;
        *rk = (*sc)->row0 + k;
        *rj = (*sc)->row0 + (*sc)->n_walls + j;
    }
    __device__ __forceinline__ int32_t load(int64_t t) const {
        const SceneDev* sc;
        int64_t rj, rk;
        decode(t, &sc, &rj, &rk);
        const d3 d = ld3(com + 3 * rj) - ld3(wall_c + 3 * rk);
        const double dd = dot3(ld3(wall_n + 3 * rk), d);
        return fabs(dd) < sc->wall_thr ? 1 : 0;
    }
    __device__ __forceinline__ void store(int64_t t, int64_t pos, int32_t v) const {
        if (!v || pos >= cap) return;
        const SceneDev* sc;
        int64_t rj, rk;
        decode(t, &sc, &rj, &rk);
        pairs[2 * pos] = (int32_t)rj;
        pairs[2 * pos + 1] = (int32_t)rk;
    }
    __device__ __forceinline__ void finish(int64_t) const {}
};

// Cell heads over the sorted keys fused with their numbering: item s starts
// a cell when its key differs from the previous one; the cell's index is
// its offset. Writes the cell's start, key, the first row's flag and the
// key -> cell table; the last entry of cell_start closes the list.
struct HeadIO {
    const uint32_t* keys;
    const int32_t* vals;
    int64_t M;
    int32_t* cell_start;
    uint32_t* cell_key;
    int32_t* row_first;
    int32_t* cell_of;
    __device__ __forceinline__ int32_t load(int64_t s) const {
        return (s == 0 || keys[s] != keys[s - 1]) ? 1 : 0;
    }
    __device__ __forceinline__ void store(int64_t s, int64_t c, int32_t v) const {
        if (!v) return;
        cell_start[c] = (int32_t)s;
        cell_key[c] = keys[s];
        row_first[vals[s]] = 1;
        if (cell_of) cell_of[keys[s]] = (int32_t)c;
    }
    __device__ __forceinline__ void finish(int64_t n_cells) const { cell_start[n_cells] = (int32_t)M; }
};

// Scan of the per-cell pair counts; its total closes the candidate count
// (counters 1 wall pairs + 2 cell pairs -> 3 capped, 7 uncapped) and
// restarts the counters of the fine-contact stage that consumes the pairs.
struct CellCountIO {
    const int32_t* cnt;
    int64_t* off;
    int64_t* counters;
    int64_t cap;
    uint32_t* flags;
    __device__ __forceinline__ int32_t load(int64_t i) const { return cnt[i]; }
    __device__ __forceinline__ void store(int64_t i, int64_t o, int32_t) const { off[i] = o; }
    __device__ __forceinline__ void finish(int64_t total) const {
        counters[6] = counters[0];  // cells of this step, cleared from the key table next step
        const int64_t P = counters[1] + total;
        counters[3] = P < cap ? P : cap;
        counters[7] = P;
        counters[8] = 0;   // simplex slots of the fine-contact stage
        counters[10] = 0;  // its sphere-test survivors
        counters[11] = 0;  // and its bucket overflow records
        if (P > cap) atomicOr(flags, MWD_F_PAIR_OVERFLOW);
    }
};

// Before the cells of this step: clear the row flags and forget the previous
// step's cells in the key -> cell table.
__global__ void k_cells_prep(int32_t* __restrict__ row_first, int64_t n_rows, const uint32_t* __restrict__ cell_key,
                             const int64_t* __restrict__ n_prev, int32_t* __restrict__ cell_of) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_rows) row_first[i] = 0;
    if (cell_of && i < *n_prev) cell_of[cell_key[i]] = -1;
}

__global__ void k_finish_pairs(int64_t* counters, int64_t cap, uint32_t* flags) {
    counters[6] = counters[0];  // cells of this step, cleared from the key table next step
    const int64_t P = counters[1] + counters[2];
    counters[3] = P < cap ? P : cap;
    counters[7] = P;
    counters[8] = 0;   // simplex slots of the fine-contact stage that consumes these pairs
    counters[10] = 0;  // its sphere-test survivors
    counters[11] = 0;  // and its bucket overflow records
    if (P > cap) atomicOr(flags, MWD_F_PAIR_OVERFLOW);
}

__global__ void k_zero_i64(int64_t* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256 > 0 ? (n + 255) / 256 : 1); }

}  // namespace

int mwd_launch_cell_keys(const double* com, int64_t n, const double* cell, int64_t ncx, int64_t ncy, int64_t* keys,
                         hipStream_t s) {
    if (n <= 0) return MWD_OK;
    hipLaunchKernelGGL(k_cell_keys, dim3(blocks_for(n)), dim3(256), 0, s, com, n, cell[0], cell[1], cell[2], ncx, ncy,
                       keys);
    MWD_LAUNCH_CHECK();
    return MWD_OK;
}

int mwd_launch_candidates(mwd_workspace* ws, const mwd_state* st, const int64_t* keys_in) {
    hipStream_t s = ws->stream;
    const int64_t M = ws->n_moving;
    int64_t* ctr = ws->counters;
    int rc;
    // counters 0..3 = n_cells, n_wall_pairs, n_cell_pairs, n_pairs
    mwd_mark(ws, 0);
    const bool walls = ws->n_wall_tests > 0 && st->wall_n && st->wall_c;
    if (!walls || M <= 0) {  // otherwise the scans below write counters 0..3
        hipLaunchKernelGGL(k_zero_i64, dim3(1), dim3(64), 0, s, ctr, 4);
        MWD_LAUNCH_CHECK();
    }
    // -- wall slab pairs: test, count and write in one pass
    if (walls) {
        const WallIO io{st->com, st->wall_n, st->wall_c, ws->scenes_dev, ws->n_scenes, ws->scene_wt_off,
                        ws->pairs, ws->pair_cap};
        rc = mwd_scan_io(io, ws->n_wall_tests, nullptr, ctr + 1, ws->scan_tmp, ws->scan_tmp_len, s);
        if (rc) return rc;
    }
    // -- linked-cell pairs
    if (M > 0) {
        mwd_mark(ws, 1);
        hipLaunchKernelGGL(k_sort_keys, dim3(blocks_for(M)), dim3(256), 0, s, st->com, keys_in, ws->plist, M,
                           ws->scene_of, ws->scenes_dev, ws->keys, ws->vals, ws->flags);
        MWD_LAUNCH_CHECK();
        uint32_t* ks;
        int32_t* vs;
        rc = mwd_radix_sort(ws->keys, ws->vals, ws->keys_alt, ws->vals_alt, M, nullptr, ws->key_bits, ws->hist,
                            ws->hist_off, ctr + 9, ws->scan_tmp, ws->scan_tmp_len, s, &ks, &vs);
        if (rc) return rc;
        ws->keys_sorted = ks;
        ws->vals_sorted = vs;
        mwd_mark(ws, 2);
        const int64_t prep = ws->n_rows > M ? ws->n_rows : M;
        hipLaunchKernelGGL(k_cells_prep, dim3(blocks_for(prep)), dim3(256), 0, s, ws->row_first, ws->n_rows,
                           ws->cell_key, ctr + 6, ws->cell_of);
        MWD_LAUNCH_CHECK();
        rc = mwd_scan_io(HeadIO{ks, vs, M, ws->cell_start, ws->cell_key, ws->row_first, ws->cell_of}, M, nullptr,
                         ctr + 0, ws->scan_tmp, ws->scan_tmp_len, s);
        if (rc) return rc;
        rc = mwd_scan_i32(ws->row_first, ws->row_rank, nullptr, ws->n_rows, nullptr, ws->scan_tmp,
                          ws->scan_tmp_len, s);
        if (rc) return rc;
        hipLaunchKernelGGL(k_cell_info, dim3(blocks_for(M)), dim3(256), 0, s, vs, ws->cell_start,
                           ws->cell_key, ctr + 0, M, ws->scene_of, ws->scenes_dev, ws->row_rank, ws->perm,
                           ws->cell_nb, ws->cell_by_rank, ws->cell_cnt, ws->flags, ws->cell_of);
        MWD_LAUNCH_CHECK();
        rc = mwd_scan_io(CellCountIO{ws->cell_cnt, ws->cell_off, ctr, ws->pair_cap, ws->flags}, M, ctr + 0,
                         ctr + 2, ws->scan_tmp, ws->scan_tmp_len, s);
        if (rc) return rc;
        mwd_mark(ws, 3);
        MWD_LAUNCH_EV(mwd_kev(ws, MWD_KT_EMIT, 0), mwd_kev(ws, MWD_KT_EMIT, 1), k_cell_fill,
                      dim3((unsigned)((M + FILL_CELLS - 1) / FILL_CELLS)), dim3(FILL_THREADS), 0, s, ws->perm,
                      ws->cell_start, ws->cell_nb, ws->cell_by_rank, ws->cell_off, ctr, ws->pairs, ws->pair_cap);
        MWD_LAUNCH_CHECK();
    }
    mwd_mark(ws, 4);
    if (M <= 0) {  // no cell-count scan closed the candidate count
        hipLaunchKernelGGL(k_finish_pairs, dim3(1), dim3(1), 0, s, ctr, ws->pair_cap, ws->flags);
        MWD_LAUNCH_CHECK();
    }
    return MWD_OK;
}
