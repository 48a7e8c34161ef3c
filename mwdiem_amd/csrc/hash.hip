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
// Two ways to the cells, both ending in per-cell tables by rank (cells in
// order of their lowest row) that the emission reads:
//   key table (default while the summed key span is at most 2^28): a
//     counting sort whose single digit is the whole key. Per particle, count
//     the key's members, keep its lowest row and file the row under the key
//     (atomics on key-indexed tables: the first sixteen members side by side,
//     more in a list); one scan over rows numbers the cells by rank and
//     places their member ranges (a packed two-count scan), stamping each
//     key's (first slot, members) with the pass's epoch; per cell (fused
//     into the scan of pair counts), gather and sort the few members, put
//     them in set order, look up the neighbour cells (one stamped entry
//     each) and count pairs. The tables return to their empty state as they
//     are used, so nothing is cleared between steps.
//   sorted keys: stable radix sort of (key, row) -> cell heads (scan) ->
//     cell rank by lowest row (scan over rows) -> per-cell set order,
//     neighbour lookup (binary search of the sorted cell keys) and pair
//     count -> scan of counts in rank order.
// Then emission, one pair per thread. Wall slab tests: on the key-table path
// the key pass tests each particle against its scene's walls and gathers the
// results by ballot into 64-bit words laid out in the reference's test order
// (scene, wall, particle); a scan over words, in the same launch as the scan
// over rows, places the wall pairs. Otherwise one scan over the (scene, wall,
// particle) test space tests, counts and writes them.
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

// Global key of a moving row: its scene's key offset by the scene's base in
// the concatenated key space (clamped, with a flag, outside the scene's span).
__device__ __forceinline__ uint32_t global_key(const double* __restrict__ com, const int64_t* __restrict__ keys_in,
                                               int32_t row, const SceneDev& sc, uint32_t* flags) {
    const int64_t key = keys_in ? keys_in[row] : key_of(ld3(com + 3 * row), sc.cell, sc.ncx, sc.ncy);
    int64_t loc = key - sc.kmin;
    if (loc < 0 || loc >= sc.kspan) {
        atomicOr(flags, MWD_F_KEYRANGE);
        loc = loc < 0 ? 0 : sc.kspan - 1;
    }
    return (uint32_t)(sc.kbase + loc);
}

// Sorted path: global sort keys of the moving particles.
__global__ void k_sort_keys(const double* __restrict__ com, const int64_t* __restrict__ keys_in,
                            const int32_t* __restrict__ plist, int64_t M, const int32_t* __restrict__ scene_of,
                            const SceneDev* __restrict__ scenes, uint32_t* __restrict__ keys,
                            int32_t* __restrict__ vals, uint32_t* flags, const double* __restrict__ radius,
                            double* __restrict__ sph) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const int32_t row = plist[i];
    keys[i] = global_key(com, keys_in, row, scenes[scene_of[row]], flags);
    vals[i] = row;
    if (sph) pack_sphere(sph, com, radius, row);
}

// Wall slab test of particle row rj (centre x) against wall row rk
// (MWDiem.py:807-811): the pair is listed when |dd| < thr, dd = n.(x - c).
// With the survivor list a listed pair also goes to GJK unless the wall's
// slab |n.(v - c)| <= h (h: the half-extent of the wall's vertices along n)
// and the particle's bounding sphere are apart by more than
// 1e-6 (1 + |dd| + h): then every particle vertex lies outside the slab, the
// bodies are separated and GJK would return 0.
__device__ __forceinline__ double wall_dd(d3 x, const double* __restrict__ wall_n, const double* __restrict__ wall_c,
                                          int64_t rk) {
    return dot3(ld3(wall_n + 3 * rk), x - ld3(wall_c + 3 * rk));
}

__device__ __forceinline__ bool wall_apart(double dd, const double* __restrict__ wall_n,
                                           const double* __restrict__ wall_c, const double* __restrict__ radius,
                                           const double* __restrict__ verts, const int32_t* __restrict__ nverts,
                                           int32_t vmax, int64_t rj, int64_t rk) {
    const d3 n = ld3(wall_n + 3 * rk);
    const d3 c = ld3(wall_c + 3 * rk);
    const double* v = verts + (int64_t)vmax * 3 * rk;
    const int nv = nverts[rk];
    double h = 0.0;
    for (int i0 = 0; i0 < nv; i0 += 8) {  // eight vertex loads in flight
        d3 p[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (i0 + u < nv) p[u] = ld3(v + 3 * (i0 + u));
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (i0 + u < nv) h = fmax(h, fabs(dot3(n, p[u] - c)));
    }
    const double a = fabs(dd);
    return a - h - radius[rj] > 1e-6 * (1.0 + a + h);
}

// 0: not listed; 1: listed; (1 << 32) | 1: listed and goes to GJK.
__device__ __forceinline__ int64_t wall_pair_bits(d3 x, const double* __restrict__ wall_n,
                                                  const double* __restrict__ wall_c,
                                                  const double* __restrict__ radius, const double* __restrict__ verts,
                                                  const int32_t* __restrict__ nverts, int32_t vmax, int64_t rj,
                                                  int64_t rk, double thr, bool with_surv) {
    const double dd = wall_dd(x, wall_n, wall_c, rk);
    if (!(fabs(dd) < thr)) return 0;
    if (!with_surv || wall_apart(dd, wall_n, wall_c, radius, verts, nverts, vmax, rj, rk)) return 1;
    return (int64_t(1) << 32) | 1;
}

// Key-table path, one thread per moving particle: its key, the member count
// of that key and the key's lowest row; the key's first sixteen members go to
// its 64-byte entry of key_small (in arrival order), any more to a list
// (key_head, row_next), so that the cell pass finds a cell's members without
// a scatter of rows into slots. With wall words, also the particle's
// slab tests against every wall of its scene, gathered by ballots into the
// 64-bit words of the tests (a wave is 64 consecutive list entries, so its
// ballot for one scene and wall is that segment's word; a scan over the
// words then places the pairs in the reference's order).
constexpr int32_t KEY_MIN_EMPTY = 0x7F7F7F7F;  // above every row: the mark of an empty key
constexpr int KEY_SMALL = 16;  // members of a key kept side by side (64 bytes); more go to a list

struct WallWordsIn {  // what the slab tests read and write (hit_words NULL: no tests here)
    const double* wall_n;
    const double* wall_c;
    const double* radius;
    const double* verts;
    const int32_t* nverts;
    int32_t vmax;
    bool with_surv;
    const int64_t* wseg_off;
    uint64_t* hit_words;
    uint64_t* gjk_words;
};

__global__ void k_key_count(const double* __restrict__ com, const int64_t* __restrict__ keys_in,
                            const int32_t* __restrict__ plist, int64_t M_cap, const int64_t* m_live,
                            const int32_t* __restrict__ scene_of,
                            const SceneDev* __restrict__ scenes, int32_t* __restrict__ row_key,
                            int32_t* __restrict__ key_cnt, int32_t* __restrict__ key_min, uint32_t* flags,
                            const double* __restrict__ radius, double* __restrict__ sph, WallWordsIn ww,
                            int64_t* __restrict__ epoch, int32_t* __restrict__ key_small,
                            int32_t* __restrict__ key_head, int32_t* __restrict__ row_next, int64_t row_base) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *epoch += 1;  // a new pass: key segments of earlier passes go stale
    const bool live = i < (*m_live < M_cap ? *m_live : M_cap);
    if (!ww.hit_words && !live) return;
    int32_t row = 0, s = -1;
    if (live) {
        // one scene: its moving rows are consecutive, entry i is row row_base + i
        row = row_base >= 0 ? (int32_t)(row_base + i) : plist[i];
        s = row_base >= 0 ? 0 : scene_of[row];
        const SceneDev& sc = scenes[s];
        if (sph) pack_sphere(sph, com, radius, row);
        const uint32_t k = global_key(com, keys_in, row, sc, flags);
        row_key[row] = (int32_t)k;
        const int slot = atomicAdd(&key_cnt[2 * k], 1);
        atomicMin(&key_min[2 * k], row);
        // the key's first sixteen members side by side, the others in a list
        if (slot < KEY_SMALL) key_small[KEY_SMALL * (int64_t)k + slot] = row;
        else row_next[row] = atomicExch(&key_head[k], row);
    }
    if (!ww.hit_words) return;
    // the particle's tests against its scene's walls (at most MWD_WORD_WALLS):
    // first every plane distance (independent loads), then the slab-apart
    // test of the listed pairs
    uint32_t hm = 0, gm = 0;
    if (live) {
        const SceneDev& sc = scenes[s];
        const d3 x = ld3(com + 3 * (int64_t)row);
        const int W = (int)sc.n_walls;
#pragma unroll 8
        for (int k = 0; k < W; ++k)
            if (fabs(wall_dd(x, ww.wall_n, ww.wall_c, sc.row0 + k)) < sc.wall_thr) hm |= 1u << k;
        if (ww.with_surv) {
            for (uint32_t t = hm; t; t &= t - 1) {
                const int k = __ffs(t) - 1;
                const double dd = wall_dd(x, ww.wall_n, ww.wall_c, sc.row0 + k);
                if (!wall_apart(dd, ww.wall_n, ww.wall_c, ww.radius, ww.verts, ww.nverts, ww.vmax, row, sc.row0 + k))
                    gm |= 1u << k;
            }
        }
    }
    // into the words, one scene of the wave's live lanes at a time (usually one)
    const int lane = threadIdx.x & (MWD_WAVE - 1);
    const int64_t w = i >> 6;  // the wave's word in the list entries' numbering
    uint64_t todo = __ballot(live);
    while (todo) {
        const int leader = __ffsll((long long)todo) - 1;
        const int sl = __shfl(s, leader, MWD_WAVE);
        const bool in = live && s == sl;
        const SceneDev& sc = scenes[sl];
        const int64_t w0 = sc.plist_off >> 6;
        const int W = (int)sc.n_walls;
        for (int k = 0; k < W; ++k) {
            const uint64_t hw = __ballot(in && ((hm >> k) & 1u));
            const uint64_t gw = __ballot(in && ((gm >> k) & 1u));
            if (lane == leader) {
                const int64_t at = ww.wseg_off[sc.wseg0 + k] + (w - w0);
                ww.hit_words[at] = hw;
                ww.gjk_words[at] = gw;
            }
        }
        todo &= ~__ballot(in);
    }
}

// CPython set iteration order of non-negative ints inserted in the given
// order (Objects/setobject.c: 9 linear probes, then perturbation; resize to
// the smallest power of two above 4*used once fill*5 >= mask*3).
constexpr int SET_TAB = 128;  // holds up to MWD_MAX_CELL members

// The general case (any n up to MWD_MAX_CELL): the same table walk with the
// table's occupancy as 128 bits in registers and its keys in a slot array
// that is read only where a bit is set (nothing to clear), so a cell costs a
// few accesses per member. Table sizes follow the rule above: 8 slots, 32
// once 5 keys fill the first (5*5 >= 7*3), 128 once 19 fill the second
// (19*5 >= 31*3; 128 is the smallest power of two above 4*19); the next
// resize would come at 77 keys, past MWD_MAX_CELL.
__device__ __forceinline__ bool occ_get(uint64_t lo, uint64_t hi, uint32_t i) {
    return (((i < 64) ? lo : hi) >> (i & 63)) & 1ull;
}
__device__ __forceinline__ void occ_set(uint64_t& lo, uint64_t& hi, uint32_t i) {
    if (i < 64) lo |= 1ull << i;
    else hi |= 1ull << (i - 64);
}
__device__ __forceinline__ uint32_t probe_occ(uint64_t key, uint32_t mask, uint64_t& lo, uint64_t& hi) {
    uint64_t perturb = key;
    uint32_t i = (uint32_t)(key & mask);
    for (;;) {
        if (!occ_get(lo, hi, i)) {
            occ_set(lo, hi, i);
            return i;
        }
        if (i + 9 <= mask) {
            for (uint32_t j = 1; j <= 9; ++j)
                if (!occ_get(lo, hi, i + j)) {
                    occ_set(lo, hi, i + j);
                    return i + j;
                }
        }
        perturb >>= 5;
        i = (uint32_t)((i * 5 + 1 + perturb) & mask);
    }
}

// tab: two tables of SET_TAB slots (slot -> key, row - row0, valid where the
// occupancy bit is set), in the caller's memory: a thread's own array, or
// its share of LDS where the kernel has room for it.
__device__ void set_order(const int32_t* rows, int n, int64_t row0, int32_t* out, int32_t (*tab)[SET_TAB]) {
    if (n <= 1) {
        if (n == 1) out[0] = rows[0];
        return;
    }
    int cur = 0;
    uint64_t lo = 0, hi = 0;
    uint32_t size = 8, fill = 0;
    for (int k = 0; k < n; ++k) {
        const int32_t key = (int32_t)(rows[k] - row0);
        tab[cur][probe_occ((uint64_t)key, size - 1, lo, hi)] = key;
        ++fill;
        if (fill * 5 >= (size - 1) * 3) {  // resize: the keys in the order of their slots
            const uint32_t minused = fill * 4;
            uint32_t ns = 8;
            while (ns <= minused) ns <<= 1;
            uint64_t nlo = 0, nhi = 0;
            const int nxt = 1 - cur;
            for (int half = 0; half < 2; ++half) {
                uint64_t bits = half ? hi : lo;
                while (bits) {
                    const uint32_t q = (uint32_t)(__ffsll((long long)bits) - 1) + 64u * half;
                    bits &= bits - 1;
                    const int32_t key2 = tab[cur][q];
                    tab[nxt][probe_occ((uint64_t)key2, ns - 1, nlo, nhi)] = key2;
                }
            }
            lo = nlo;
            hi = nhi;
            cur = nxt;
            size = ns;
        }
    }
    int o = 0;
    for (int half = 0; half < 2; ++half) {
        uint64_t bits = half ? hi : lo;
        while (bits) {
            const uint32_t q = (uint32_t)(__ffsll((long long)bits) - 1) + 64u * half;
            bits &= bits - 1;
            out[o++] = (int32_t)(tab[cur][q] + row0);
        }
    }
}
__device__ void set_order(const int32_t* rows, int n, int64_t row0, int32_t* out) {
    int32_t tab[2][SET_TAB];
    set_order(rows, n, row0, out, tab);
}

// set_order of n <= 16 members given in registers: the same table walks
// with the tables as occupancy bits. Up to four keys the 8-slot table is
// never resized (perturbation only: with mask 7 no linear probe fits); the
// fifth insertion fills it to 5 of 8 (5*5 >= 7*3), so the five keys move,
// in the order of their slots, to a 32-slot table (the smallest power of two
// above 4*5), where the later keys join (no resize below 19 keys) and linear
// probes of 9 apply where they fit.
__device__ __forceinline__ int probe_bits(uint64_t key, uint32_t mask, uint32_t* occ) {
    uint64_t perturb = key;
    uint32_t i = (uint32_t)(key & mask);
    for (;;) {
        if (!((*occ >> i) & 1u)) {
            *occ |= 1u << i;
            return (int)i;
        }
        if (i + 9 <= mask) {
            for (uint32_t j = 1; j <= 9; ++j)
                if (!((*occ >> (i + j)) & 1u)) {
                    *occ |= 1u << (i + j);
                    return (int)(i + j);
                }
        }
        perturb >>= 5;
        i = (uint32_t)((i * 5 + 1 + perturb) & mask);
    }
}

// Sort of the first n of 16 values in registers (compare-exchanges at fixed
// indices; entries at or past n hold INT32_MAX and stay last): odd-even
// transposition over 8 or 16 entries.
template <int N>
__device__ __forceinline__ void sort_net(int32_t (&v)[16]) {
#pragma unroll
    for (int r = 0; r < N; ++r) {
#pragma unroll
        for (int i = r & 1; i + 1 < N; i += 2) {
            const int32_t a = v[i], b = v[i + 1];
            v[i] = a < b ? a : b;
            v[i + 1] = a < b ? b : a;
        }
    }
}

__device__ __forceinline__ void set_order16(const int32_t (&rows)[16], int n, int64_t row0, int32_t* out) {
    int slot[16];
    uint32_t occ = 0;
    const int na = n < 5 ? n : 5;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        slot[k] = 64;
        if (k < na) slot[k] = probe_bits((uint64_t)(rows[k] - row0), 7u, &occ);
    }
    int tsize = 8;
    if (n > 4) {  // the resize to 32 slots: the five keys in the order of their slots, then the rest
        int s32[16];
        uint32_t occ32 = 0;
#pragma unroll
        for (int k = 0; k < 16; ++k) s32[k] = 64;
        for (int sl = 0; sl < 8; ++sl) {
#pragma unroll
            for (int k = 0; k < 5; ++k)
                if (slot[k] == sl) s32[k] = probe_bits((uint64_t)(rows[k] - row0), 31u, &occ32);
        }
#pragma unroll
        for (int k = 5; k < 16; ++k)
            if (k < n) s32[k] = probe_bits((uint64_t)(rows[k] - row0), 31u, &occ32);
#pragma unroll
        for (int k = 0; k < 16; ++k) slot[k] = s32[k];
        tsize = 32;
    }
    int o = 0;
    for (int sl = 0; sl < tsize; ++sl) {
#pragma unroll
        for (int k = 0; k < 16; ++k)
            if (slot[k] == sl) out[o++] = rows[k];
    }
}

// Scene of a global cell key: the last scene whose key range starts at or
// below it (scenes are in key order).
__device__ __forceinline__ int scene_of_key(const SceneDev* __restrict__ scenes, int S, uint32_t key) {
    int lo = 0, hi = S;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (scenes[mid].kbase <= (int64_t)key) lo = mid;
        else hi = mid;
    }
    return lo;
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

// The per-cell work of both paths: members in set order into perm, the
// cell's and its neighbours' (first slot, members) into the rank-ordered
// tables, and the cell's pair count. Cell c holds member slots
// [cell_start[c], cell_start[c+1]); `sorted` says whether they are already
// in ascending row order. seg_of(t) gives (first slot, members) of the cell
// of global key t, or (0, 0).
// Where the key-table path keeps a key's members (k_key_count): the first
// sixteen in key_small, the rest in a list. The cell pass empties both (the
// count returns to zero, the list head to -1) for the next pass.
struct KeyLists {
    const int4* key_small;
    int32_t* key_head;
    const int32_t* row_next;
    int32_t* key_cnt;
};

// LDS a per-cell thread may be given for a crowded cell: the set-order
// tables, and the members staged for their rank sort (key-table path).
struct CellLds {
    int32_t tab[2][SET_TAB];
    int32_t m[MWD_MAX_CELL];
    int32_t srt[MWD_MAX_CELL];
};

template <class SegOf>
__device__ __forceinline__ int32_t cell_work(int64_t c, int64_t r, bool sorted, int32_t* __restrict__ members,
                                             const KeyLists* kl, const int32_t* __restrict__ cell_start,
                                             const uint32_t* __restrict__ cell_key,
                                             const SceneDev* __restrict__ scenes, int S,
                                             int32_t* __restrict__ perm, int2* __restrict__ cell_self,
                                             int2* __restrict__ cell_seg, uint32_t* flags, const SegOf& seg_of,
                                             CellLds* lds = nullptr) {
    const int32_t start = cell_start[c], end = cell_start[c + 1];
    const uint32_t key = cell_key[c];
    const int n = end - start;
    int32_t* m = members + start;
    // the scene from the key (keys are offset per scene), so that the
    // neighbour lookups wait only for the key
    const SceneDev& sc = scenes[scene_of_key(scenes, S, key)];
    const int64_t lk = (int64_t)key - sc.kbase + sc.kmin;
    int32_t v[16];  // members in registers, INT32_MAX past n
    if (kl) {  // key-table path: the members from the key's entry (one 64-byte line) and list
        const int4* e = kl->key_small + 4 * (int64_t)key;
        const int4 z = make_int4(INT32_MAX, INT32_MAX, INT32_MAX, INT32_MAX);
        const int4 e0 = e[0], e1 = n > 4 ? e[1] : z, e2 = n > 8 ? e[2] : z, e3 = n > 12 ? e[3] : z;
        const int32_t w[16] = {e0.x, e0.y, e0.z, e0.w, e1.x, e1.y, e1.z, e1.w,
                               e2.x, e2.y, e2.z, e2.w, e3.x, e3.y, e3.z, e3.w};
#pragma unroll
        for (int k = 0; k < 16; ++k) v[k] = k < n ? w[k] : INT32_MAX;
        kl->key_cnt[2 * (int64_t)key] = 0;
        if (n > KEY_SMALL) {  // the cell's slots (or its LDS) take all of them: the entry's sixteen, then the list
            int32_t* mm = lds && n <= MWD_MAX_CELL ? lds->m : m;
#pragma unroll
            for (int k = 0; k < KEY_SMALL; ++k) mm[k] = v[k];
            int32_t row = kl->key_head[key];
            for (int k = KEY_SMALL; k < n; ++k) {
                mm[k] = row;
                row = kl->row_next[row];
            }
            kl->key_head[key] = -1;
        }
    } else {
#pragma unroll
        for (int k = 0; k < 16; ++k) v[k] = k < n ? m[k] : INT32_MAX;
    }
    int2 seg[MWD_MAX_OFF];
    if constexpr (SegOf::kSearch) {
        // sorted keys: the neighbours' lower-bound searches run in lockstep
        // (the step count depends on C only), so each step's probes go out
        // together instead of one dependent chain per neighbour
        uint32_t kk[MWD_MAX_OFF];
        bool ok[MWD_MAX_OFF];
        int64_t b[MWD_MAX_OFF];
#pragma unroll
        for (int o = 0; o < MWD_MAX_OFF; ++o) {
            seg[o] = make_int2(0, 0);
            ok[o] = false;
            kk[o] = 0;
            b[o] = 0;
            if (o < sc.n_off) {
                const int64_t t = lk + sc.off[o] - sc.kmin;
                if (t >= 0 && t < sc.kspan) {
                    ok[o] = true;
                    kk[o] = (uint32_t)(sc.kbase + t);
                }
            }
        }
        const uint32_t* __restrict__ ck = seg_of.cell_key;
        const int64_t C = seg_of.C;  // at least 1: this cell
        for (int64_t n = C; n > 1;) {  // b + n <= C throughout
            const int64_t half = n >> 1;
#pragma unroll
            for (int o = 0; o < MWD_MAX_OFF; ++o) b[o] = ck[b[o] + half] < kk[o] ? b[o] + half : b[o];
            n -= half;
        }
#pragma unroll
        for (int o = 0; o < MWD_MAX_OFF; ++o) {
            const int64_t lb = b[o] + (ck[b[o]] < kk[o] ? 1 : 0);  // the first key >= kk[o]
            if (ok[o] && lb < C && ck[lb] == kk[o]) {
                const int32_t a0 = seg_of.cell_start[lb];
                seg[o] = make_int2(a0, seg_of.cell_start[lb + 1] - a0);
            }
        }
    } else {
#pragma unroll
        for (int o = 0; o < MWD_MAX_OFF; ++o) {
            seg[o] = make_int2(0, 0);
            if (o < sc.n_off) {
                const int64_t t = lk + sc.off[o] - sc.kmin;
                if (t >= 0 && t < sc.kspan) seg[o] = seg_of((uint32_t)(sc.kbase + t));
            }
        }
    }
    if (n <= 16) {  // members in registers
        if (!sorted) {
            if (n <= 8) sort_net<8>(v);
            else sort_net<16>(v);
        }
        if (n == 1) perm[start] = v[0];
        else if (n > 1) set_order16(v, n, sc.row0, perm + start);
    } else {
        if (n > MWD_MAX_CELL) {
            atomicOr(flags, MWD_F_CELL);
            for (int k = 0; k < n; ++k) perm[start + k] = m[k];
        } else if (sorted) {
            if (lds) set_order(m, n, sc.row0, perm + start, lds->tab);
            else set_order(m, n, sc.row0, perm + start);
        } else if (lds) {  // sorted by rank (the rows are distinct): no chain of dependent moves
            for (int k = 0; k < n; ++k) {
                const int32_t x = lds->m[k];
                int rk = 0;
                for (int j = 0; j < n; ++j) rk += lds->m[j] < x ? 1 : 0;
                lds->srt[rk] = x;
            }
            set_order(lds->srt, n, sc.row0, perm + start, lds->tab);
        } else {
            int32_t srt[MWD_MAX_CELL];
            for (int k = 0; k < n; ++k) {
                const int32_t x = m[k];
                int rk = 0;
                for (int j = 0; j < n; ++j) rk += m[j] < x ? 1 : 0;
                srt[rk] = x;
            }
            set_order(srt, n, sc.row0, perm + start);
        }
    }
    cell_self[r] = make_int2(start, n);
    int64_t cnt = (int64_t)n * (n - 1) / 2;
#pragma unroll
    for (int o = 0; o < MWD_MAX_OFF; ++o) {
        cell_seg[r * MWD_MAX_OFF + o] = seg[o];
        cnt += (int64_t)n * seg[o].y;
    }
    return cnt < INT32_MAX ? (int32_t)cnt : INT32_MAX;  // beyond only with MWD_F_CELL raised
}

struct SegSorted {  // binary search of the sorted cell keys
    static constexpr bool kSearch = true;
    const uint32_t* cell_key;
    const int32_t* cell_start;
    int64_t C;
    __device__ __forceinline__ int2 operator()(uint32_t key) const {
        const int64_t c = find_cell(cell_key, C, key);
        if (c < 0) return make_int2(0, 0);
        const int32_t a = cell_start[c];
        return make_int2(a, cell_start[c + 1] - a);
    }
};

struct SegTable {  // key -> (first slot, members) this pass; other epochs' entries are stale
    static constexpr bool kSearch = false;
    const int4* key_seg;
    int32_t epoch;
    __device__ __forceinline__ int2 operator()(uint32_t key) const {
        const int4 e = key_seg[key];
        return e.x == epoch ? make_int2(e.y, e.z) : make_int2(0, 0);
    }
};

// Sorted path: one thread per cell (cells in key order, members sorted),
// 64 per block, each with its share of LDS for the set-order tables of a
// crowded cell (scratch memory otherwise).
constexpr int CELL_INFO_BLOCK = 64;
__global__ void __launch_bounds__(CELL_INFO_BLOCK) k_cell_info(int32_t* __restrict__ vals, const int32_t* __restrict__ cell_start,
                            const uint32_t* __restrict__ cell_key, const int64_t* __restrict__ n_cells,
                            const SceneDev* __restrict__ scenes, int S,
                            const int64_t* __restrict__ row_rank, int32_t* __restrict__ perm,
                            int2* __restrict__ cell_self, int2* __restrict__ cell_seg,
                            int32_t* __restrict__ cell_cnt, uint32_t* flags) {
    __shared__ CellLds s_lds[CELL_INFO_BLOCK];
    const int64_t C = *n_cells;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const int64_t r = row_rank[vals[cell_start[c]]];
    cell_cnt[r] = cell_work(c, r, true, vals, nullptr, cell_start, cell_key, scenes, S, perm, cell_self, cell_seg,
                            flags, SegSorted{cell_key, cell_start, C}, &s_lds[threadIdx.x]);
}

// Key-table path, the per-cell work as a kernel of its own (one cell per
// thread, 64 per block, each with its LDS for a crowded cell; the pair-count
// scan follows), instead of inside the scan's load.
__global__ void __launch_bounds__(CELL_INFO_BLOCK)
k_cell_info_table(int32_t* __restrict__ members, KeyLists kl, const int32_t* __restrict__ cell_start,
                  const uint32_t* __restrict__ cell_key, const int4* __restrict__ key_seg,
                  const int64_t* __restrict__ counters, const SceneDev* __restrict__ scenes, int S,
                  int32_t* __restrict__ perm, int2* __restrict__ cell_self, int2* __restrict__ cell_seg,
                  int32_t* __restrict__ cell_cnt, uint32_t* flags) {
    __shared__ CellLds s_lds[CELL_INFO_BLOCK];
    const int64_t C = counters[0];
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    cell_cnt[c] = cell_work(c, c, false, members, &kl, cell_start, cell_key, scenes, S, perm, cell_self, cell_seg,
                            flags, SegTable{key_seg, (int32_t)counters[6]}, &s_lds[threadIdx.x]);
}

// Blocks are dispatched to the 8 XCDs in turn (block b to XCD b % 8), and
// each XCD has its own L2. For the first `live` blocks this maps block b to
// a position such that every XCD takes one contiguous run of positions, so
// neighbouring cells (which share particles) meet in one L2. A bijection of
// [0, live).
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t live) {
    constexpr int NX = 8;
    const int64_t q = live / NX, r = live % NX;
    const int64_t x = b % NX, k = b / NX;
    return x * q + (x < r ? x : r) + k;
}

// Pair emission, one output pair per thread: a block takes 64 cells by rank,
// stages their pair offsets, member ranges and neighbour ranges in LDS
// (coalesced loads of the rank-ordered tables), and its threads walk the
// block's output range, so consecutive threads write consecutive pairs. Pair
// q of a cell is decoded in the emission order of Linked_cell_hash
// (LIBRARY_Coarse_contact_detection.py:84-118): same-cell pairs (j < k,
// members in set order), then per neighbour offset o every member j against
// every member k of that cell.
// The cells a block takes (FILL_CELLS, a launch choice: 64, 16 or 4) set the
// balance: with about one hull per cell 64 cells give a block a few hundred
// pairs; a packed bed holds many hulls per cell and thousands of pairs per
// cell, and fewer cells per block spread them over more blocks. With split
// > 1 (a launch choice: 4 or 16) `split` blocks share one group of cells,
// block s taking rounds s, s + split, ... of its pairs: for beds of few,
// crowded cells, which give too few groups to fill the chip.
constexpr int FILL_THREADS = 256;
constexpr int FILL_SURV = 2048;   // survivors gathered in LDS before a global append

template <int FILL_CELLS>
__global__ void __launch_bounds__(FILL_THREADS) k_cell_fill(const int32_t* __restrict__ perm,
                                                          const int2* __restrict__ cell_self,
                                                          const int2* __restrict__ cell_seg,
                                                          const int64_t* __restrict__ cell_off,
                                                          const int64_t* __restrict__ counters,
                                                          int32_t* __restrict__ pairs, int64_t cap,
                                                          const double* __restrict__ sph,
                                                          int32_t* __restrict__ surv, int64_t* n_surv,
                                                          int32_t* __restrict__ collision, int split) {
    __shared__ int64_t s_off[FILL_CELLS + 1];
    __shared__ int2 s_self[FILL_CELLS];
    __shared__ int2 s_seg[FILL_CELLS * MWD_MAX_OFF];
    // per cell, where each part of its pair range starts: [0] the first
    // neighbour pair (after the same-cell pairs), [o + 1] the pairs after
    // neighbour o. Saturated at UINT32_MAX; a cell's pair count (and so every
    // offset decoded below) stays under 2^31.
    __shared__ uint32_t s_cum[FILL_CELLS][MWD_MAX_OFF + 1];
    // sphere-test survivors gathered in LDS and appended to the global list
    // in batches: one global atomic per batch instead of one per wave and
    // round (in crowded steps, with hundreds of thousands of survivors, the
    // appends to one counter serialise); their order there is free (contacts
    // are summed by pair index)
    __shared__ int32_t s_surv[FILL_SURV];
    __shared__ int s_nsurv;
    __shared__ int64_t s_gbase;
    const int64_t C = counters[0];
    const int64_t live = (C + FILL_CELLS - 1) / FILL_CELLS;
    const int64_t groups = gridDim.x / split;  // the grid is `split` runs of `groups` blocks
    const int64_t g = blockIdx.x % groups, part = blockIdx.x / groups;
    if (g >= live) return;  // uniform per block
    const int64_t r0 = xcd_block(g, live) * FILL_CELLS;
    const int nc = (int)(C - r0 < FILL_CELLS ? C - r0 : FILL_CELLS);
    const int tid = threadIdx.x;
    const int64_t base = cell_off[r0];
    for (int k = tid; k < nc * MWD_MAX_OFF; k += FILL_THREADS) s_seg[k] = cell_seg[r0 * MWD_MAX_OFF + k];
    if (tid < nc) {
        s_self[tid] = cell_self[r0 + tid];
        s_off[tid] = cell_off[r0 + tid] - base;
    }
    if (tid == 0) {
        s_off[nc] = (r0 + nc < C ? cell_off[r0 + nc] : counters[2]) - base;  // counters[2]: all cell pairs
        s_nsurv = 0;
    }
    __syncthreads();
    if (tid < nc) {
        const uint64_t n = (uint64_t)s_self[tid].y;
        uint64_t acc = n * (n - (n > 0)) / 2;
        s_cum[tid][0] = acc < UINT32_MAX ? (uint32_t)acc : UINT32_MAX;
        for (int o = 0; o < MWD_MAX_OFF; ++o) {
            acc += n * (uint64_t)s_seg[tid * MWD_MAX_OFF + o].y;
            s_cum[tid][o + 1] = acc < UINT32_MAX ? (uint32_t)acc : UINT32_MAX;
        }
    }
    __syncthreads();
    const int64_t total = s_off[nc];
    const int64_t pos0 = counters[1] + base;
    // a uniform trip count over the block, so that the survivors' ballot
    // sees every lane of a wave
    int rounds = 0;
    const int64_t qstep = (int64_t)FILL_THREADS * split;
    for (int64_t q0 = part * FILL_THREADS; q0 < total; q0 += qstep) {
        const int64_t q = q0 + tid;
        bool have = q < total;
        int32_t a = 0, b = 0;
        if (have) {
            // cell i: s_off[i] <= q < s_off[i + 1]
            int lo = 0, hi = nc;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (s_off[mid] <= q) lo = mid;
                else hi = mid;
            }
            const int i = lo;
            const int32_t s0 = s_self[i].x, n = s_self[i].y;
            uint32_t t = (uint32_t)(q - s_off[i]);
            if (t < s_cum[i][0]) {
                int j = 0;
                uint32_t row = (uint32_t)n - 1u;  // pairs of member j with the later members
                while (t >= row) {
                    t -= row;
                    ++j;
                    --row;
                }
                a = perm[s0 + j];
                b = perm[s0 + j + 1 + (int32_t)t];
            } else {
                // the last neighbour o whose pairs start at or before t (empty
                // neighbours start where the next one does, so they are passed)
                int o = 0;
#pragma unroll
                for (int step = 8; step > 0; step >>= 1)
                    if (o + step < MWD_MAX_OFF && s_cum[i][o + step] <= t) o += step;
                const int2 sg = s_seg[i * MWD_MAX_OFF + o];
                const uint32_t tt = t - s_cum[i][o];
                const uint32_t j = tt / (uint32_t)sg.y, k = tt - j * (uint32_t)sg.y;
                a = perm[s0 + (int32_t)j];
                b = perm[sg.x + (int32_t)k];
            }
        }
        const int64_t pos = pos0 + q;
        have = have && pos < cap;
        if (have) {
            int2 v;
            v.x = a;
            v.y = b;
            reinterpret_cast<int2*>(pairs)[pos] = v;
        }
        if (sph) {  // apart: decided (collision 0); the others go to GJK, which writes theirs
            const bool keep = have && spheres_meet_packed(sph, a, b);
            if (have && !keep) collision[pos] = 0;
            const uint64_t m = __ballot(keep);
            if (m) {
                const int lane = tid & (MWD_WAVE - 1);
                const int leader = __ffsll((long long)m) - 1;
                int at = 0;
                if (lane == leader) at = atomicAdd(&s_nsurv, __popcll(m));
                at = __shfl(at, leader, MWD_WAVE);
                const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (MWD_WAVE - lane));
                if (keep) s_surv[at + __popcll(m & lt)] = (int32_t)pos;
            }
            // append after FILL_SURV / FILL_THREADS rounds (the buffer is then
            // at most full) and after the last; the round count is uniform,
            // so only the appending rounds pay the barriers
            if (++rounds == FILL_SURV / FILL_THREADS || q0 + qstep >= total) {
                rounds = 0;
                __syncthreads();
                const int cnt = s_nsurv;
                if (tid == 0 && cnt > 0)
                    s_gbase = (int64_t)atomicAdd((unsigned long long*)n_surv, (unsigned long long)cnt);
                __syncthreads();
                const int64_t gb = s_gbase;
                for (int k = tid; k < cnt; k += FILL_THREADS)
                    if (gb + k < cap) surv[gb + k] = s_surv[k];
                __syncthreads();  // every lane has read the count and copied before the next round appends
                if (tid == 0) s_nsurv = 0;
                __syncthreads();
            }
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

// Test t of the wall slab test space: scene s's tests are
// [wt_off[s], wt_off[s+1]), wall k outer and particle j inner.
struct WallTests {
    const SceneDev* scenes;
    int S;
    const int64_t* wt_off;
    __device__ __forceinline__ void decode(int64_t t, const SceneDev** sc, int64_t* rj, int64_t* rk) const {
        const int s = scene_of_test(wt_off, S, t);
        *sc = &scenes[s];
        const int64_t np = (*sc)->n_active - (*sc)->n_walls;
        const int64_t tl = t - wt_off[s];
        int64_t k;
        if ((tl >> 32) == 0) k = (uint32_t)tl / (uint32_t)np;  // the common case: a 32-bit division
        else k = tl / np;
        const int64_t j = tl - k * np;
        *rk = (*sc)->row0 + k;
        *rj = (*sc)->row0 + (*sc)->n_walls + j;
    }
};

// What a listed wall pair writes: [j, k] at its offset among the scene-ordered
// hits; with the survivor list, either its slot there (bit 32 of the count:
// it goes to GJK) or collision 0 (decided by the slab).
struct WallOut {
    int32_t* pairs;
    int64_t cap;
    int32_t* surv;
    int32_t* collision;
    int64_t* counters;  // [1] wall pairs, [17] wall survivors (written by finish)
    __device__ __forceinline__ void put(int64_t rj, int64_t rk, int64_t off, int64_t v) const {
        const int64_t pos = off & 0xffffffffll;
        if (!v || pos >= cap) return;
        pairs[2 * pos] = (int32_t)rj;
        pairs[2 * pos + 1] = (int32_t)rk;
        if (surv) {
            if (v >> 32) surv[off >> 32] = (int32_t)pos;
            else collision[pos] = 0;
        }
    }
    __device__ __forceinline__ void finish(int64_t total) const {
        counters[1] = total & 0xffffffffll;
        counters[17] = total >> 32;
    }
};

// Wall slab test (MWDiem.py:807-811) fused with its compaction: test t of a
// scene pairs wall k with particle j (walls outer, particles inner); a pair
// whose centre lies within wall_thr of the wall's plane is written, as
// [j, k], at its offset among the scene-ordered hits. With bounding radii
// the wall pairs go to GJK as survivors, listed in surv from its start,
// except pairs the wall's slab separates from the particle's bounding sphere,
// which get collision 0 here (a wall plate's bounding sphere, half its
// diagonal, meets nearly every particle in its slab, so the sphere test
// decides nothing for them). The count is packed: bit 0 the pair is listed,
// bit 32 it goes to GJK.
struct WallIO {
    const double* com;
    const double* wall_n;
    const double* wall_c;
    WallTests tests;
    WallOut out;
    const double* radius;
    const double* verts;
    const int32_t* nverts;
    int32_t vmax;
    __device__ __forceinline__ int64_t load(int64_t t) const {
        const SceneDev* sc;
        int64_t rj, rk;
        tests.decode(t, &sc, &rj, &rk);
        return wall_pair_bits(ld3(com + 3 * rj), wall_n, wall_c, radius, verts, nverts, vmax, rj, rk, sc->wall_thr,
                              out.surv != nullptr);
    }
    __device__ __forceinline__ void store(int64_t t, int64_t off, int64_t v) const {
        if (!v) return;
        const SceneDev* sc;
        int64_t rj, rk;
        tests.decode(t, &sc, &rj, &rk);
        out.put(rj, rk, off, v);
    }
    __device__ __forceinline__ void finish(int64_t total) const { out.finish(total); }
};

// The wall pairs placed from the key pass's bit words: a scan over words
// (64 tests each, concatenated in the reference's test order) whose count is
// the word's listed pairs | its pairs for GJK << 32, as WallIO's. Bits of
// entries past the scene's live count (rows not yet spawned) are ignored.
struct WallWordIO {
    const uint64_t* hit_words;
    const uint64_t* gjk_words;
    // per word: wall row, row of bit 0, the word's index in the list of
    // moving entries (bit b: entry 64 * index + b), and the bits that are
    // the scene's entries (first | end << 8)
    const int4* word_desc;
    const int64_t* m_live;  // live moving entries (rows not yet spawned lie beyond)
    WallOut out;
    __device__ __forceinline__ uint64_t bits_of(const int4& d, int64_t live) const {
        const int first = d.w & 0xff, end = d.w >> 8;
        uint64_t m = (end >= 64 ? ~0ull : ((1ull << end) - 1)) & (~0ull << first);
        const int64_t e0 = (int64_t)d.z << 6;
        if (live <= e0) m = 0;
        else if (live - e0 < 64) m &= ~0ull >> (64 - (live - e0));
        return m;
    }
    __device__ __forceinline__ int64_t load(int64_t w) const {
        uint64_t h = hit_words[w], g = gjk_words[w];
        const int4 d = word_desc[w];
        const int64_t live = *m_live;
        const uint64_t m = bits_of(d, live);
        h &= m;
        g &= m;
        return ((int64_t)__popcll(g) << 32) | (int64_t)__popcll(h);
    }
    // The set bits of word w in ascending order, each pair at the word's
    // offset plus the set bits below it (one word per thread: a word holds
    // few pairs, at most 64).
    __device__ __forceinline__ void store(int64_t w, int64_t off, int64_t v) const {
        if (!v) return;
        const int4 d = word_desc[w];
        const uint64_t m = bits_of(d, *m_live);
        uint64_t h = hit_words[w] & m;
        const uint64_t g = gjk_words[w] & m;
        int64_t o = off;
        while (h) {
            const int b = __ffsll((long long)h) - 1;
            h &= h - 1;
            const int64_t gb = (int64_t)((g >> b) & 1u) << 32;
            out.put((int64_t)d.y + b, d.x, o, gb | 1);
            o += gb + 1;
        }
    }
    __device__ __forceinline__ void finish(int64_t total) const { out.finish(total); }
};

// Sorted path: cell heads over the sorted keys fused with their numbering:
// item s starts a cell when its key differs from the previous one; the
// cell's index is its offset. Writes the cell's start, key and the first
// row's flag; the last entry of cell_start closes the list.
struct HeadIO {
    const uint32_t* keys;
    const int32_t* vals;
    int64_t M;
    int32_t* cell_start;
    uint32_t* cell_key;
    int32_t* row_first;
    __device__ __forceinline__ int32_t load(int64_t s) const {
        return (s == 0 || keys[s] != keys[s - 1]) ? 1 : 0;
    }
    __device__ __forceinline__ void store(int64_t s, int64_t c, int32_t v) const {
        if (!v) return;
        cell_start[c] = (int32_t)s;
        cell_key[c] = keys[s];
        row_first[vals[s]] = 1;
    }
    __device__ __forceinline__ void finish(int64_t n_cells) const { cell_start[n_cells] = (int32_t)M; }
};

// Key-table path: one scan over rows numbers the cells by the rank of their
// lowest row and places their member ranges. Row r counts (1 cell, n
// members of its key) when it is its key's lowest row, else nothing; the two
// counts travel packed as cells << 32 | members, so the exclusive offset of
// such a row gives both its cell's rank and its first member slot. The store
// also resets the key's lowest row for the next step.
struct CellTableIO {
    const int32_t* row_key;
    const int32_t* key_cnt;
    int32_t* key_min;
    int32_t* cell_start;
    uint32_t* cell_key;
    int4* key_seg;
    int64_t* n_cells;  // counters + 0; counters[6] holds the pass's epoch
    __device__ __forceinline__ int64_t load(int64_t r) const {
        const int32_t k = row_key[r];
        if (k < 0) return 0;
        const int2 cm = *reinterpret_cast<const int2*>(key_cnt + 2 * (int64_t)k);  // (members, lowest row)
        if (cm.y != (int32_t)r) return 0;
        return (int64_t(1) << 32) | (int64_t)cm.x;
    }
    __device__ __forceinline__ void store(int64_t r, int64_t off, int64_t v) const {
        if (!v) return;
        const int32_t k = row_key[r];
        const int64_t c = off >> 32;
        const int32_t first = (int32_t)(off & 0xffffffffll);
        cell_start[c] = first;
        cell_key[c] = (uint32_t)k;
        key_seg[k] = make_int4((int32_t)n_cells[6], first, (int32_t)(v & 0xffffffffll), 0);
        key_min[2 * (int64_t)k] = KEY_MIN_EMPTY;
    }
    __device__ __forceinline__ void finish(int64_t total) const {
        *n_cells = total >> 32;
        cell_start[total >> 32] = (int32_t)(total & 0xffffffffll);
    }
};

// Scan of the per-cell pair counts (cells by rank); its total closes the
// candidate count (counters 1 wall pairs + 2 cell pairs -> 3 capped, 7
// uncapped) and restarts the counters of the fine-contact stage that
// consumes the pairs.
__device__ __forceinline__ void close_pairs(int64_t* counters, int64_t cell_pairs, int64_t cap, uint32_t* flags,
                                            bool folded) {
    const int64_t P = counters[1] + cell_pairs;
    // survivors of the sphere test: the wall pairs listed for GJK so far
    // when the emission appends the rest, else none yet
    counters[10] = folded ? (counters[17] < cap ? counters[17] : cap) : 0;
    counters[3] = P < cap ? P : cap;
    counters[7] = P;
    counters[4] = 0;   // contacts (slots claimed by GJK) of the fine-contact stage
    counters[8] = 0;   // simplex slots of the per-pair fine-contact entry point
    counters[11] = 0;  // and its bucket overflow records
    if (P > cap) atomicOr(flags, MWD_F_PAIR_OVERFLOW);
}

struct CellCountIO {  // sorted path: counts written by k_cell_info
    const int32_t* cnt;
    int64_t* off;
    int64_t* counters;
    int64_t cap;
    uint32_t* flags;
    bool folded;
    __device__ __forceinline__ int32_t load(int64_t i) const { return cnt[i]; }
    __device__ __forceinline__ void store(int64_t i, int64_t o, int32_t) const { off[i] = o; }
    __device__ __forceinline__ void finish(int64_t total) const { close_pairs(counters, total, cap, flags, folded); }
};

struct CellInfoIO {  // key-table path: the per-cell work runs as the scan's load
    int32_t* members;
    KeyLists kl;
    const int32_t* cell_start;
    const uint32_t* cell_key;
    const int4* key_seg;
    int S;
    const SceneDev* scenes;
    int32_t* perm;
    int2* cell_self;
    int2* cell_seg;
    int64_t* off;
    int64_t* counters;
    int64_t cap;
    uint32_t* flags;
    bool folded;
    __device__ __forceinline__ int32_t load(int64_t c) const {
        return cell_work(c, c, false, members, &kl, cell_start, cell_key, scenes, S, perm, cell_self, cell_seg,
                         flags, SegTable{key_seg, (int32_t)counters[6]});
    }
    __device__ __forceinline__ void store(int64_t i, int64_t o, int32_t) const { off[i] = o; }
    __device__ __forceinline__ void finish(int64_t total) const { close_pairs(counters, total, cap, flags, folded); }
};

// Sorted path, before the cells of this step: clear the row flags.
__global__ void k_cells_prep(int32_t* __restrict__ row_first, int64_t n_rows) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_rows) row_first[i] = 0;
}

__global__ void k_finish_pairs(int64_t* counters, int64_t cap, uint32_t* flags, int folded) {
    close_pairs(counters, counters[2], cap, flags, folded != 0);
}

__global__ void k_zero_i64(int64_t* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256 > 0 ? (n + 255) / 256 : 1); }

__global__ void k_fill_key_table(int2* key_cm, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) key_cm[i] = make_int2(0, KEY_MIN_EMPTY);
}

}  // namespace

int mwd_fill_key_table(int32_t* key_cm, int64_t n, hipStream_t s) {
    if (n <= 0) return MWD_OK;
    hipLaunchKernelGGL(k_fill_key_table, dim3(blocks_for(n)), dim3(256), 0, s, reinterpret_cast<int2*>(key_cm), n);
    MWD_LAUNCH_CHECK();
    return MWD_OK;
}

int mwd_launch_cell_keys(const double* com, int64_t n, const double* cell, int64_t ncx, int64_t ncy, int64_t* keys,
                         hipStream_t s) {
    if (n <= 0) return MWD_OK;
    hipLaunchKernelGGL(k_cell_keys, dim3(blocks_for(n)), dim3(256), 0, s, com, n, cell[0], cell[1], cell[2], ncx, ncy,
                       keys);
    MWD_LAUNCH_CHECK();
    return MWD_OK;
}

namespace {
// The cell-table scan over rows and the wall words' scan in one launch (rows:
// 4 items per thread up to a million, else 16; words: one per thread, as a
// word's pairs are written one after another).
int scan_cells_and_walls(const CellTableIO& ct, int64_t n_rows, int64_t* tmp, int64_t tmp_len, const WallWordIO& wio,
                         int64_t n_words, int64_t* tmp_w, int64_t tmp_w_len, hipStream_t s) {
    if (n_rows <= SCAN_SHORT_MAX)
        return mwd_scan_io2<CellTableIO, SCAN_ITEMS_SHORT, WallWordIO, 1>(
            ct, n_rows, nullptr, tmp, tmp_len, wio, n_words, nullptr, tmp_w, tmp_w_len, s);
    return mwd_scan_io2<CellTableIO, SCAN_ITEMS_LONG, WallWordIO, 1>(
        ct, n_rows, nullptr, tmp, tmp_len, wio, n_words, nullptr, tmp_w, tmp_w_len, s);
}
}  // namespace

int mwd_launch_candidates(mwd_workspace* ws, const mwd_state* st, const int64_t* keys_in) {
    hipStream_t s = ws->stream;
    const int64_t M = ws->n_moving;
    int64_t* ctr = ws->counters;
    int rc;
    // counters 0..3 = n_cells, n_wall_pairs, n_cell_pairs, n_pairs
    mwd_mark(ws, 0);
    const bool walls = ws->n_wall_tests > 0 && st->wall_n && st->wall_c;
    if (!walls || M <= 0) {  // otherwise the scans below write counters 0..3 and 17
        hipLaunchKernelGGL(k_zero_i64, dim3(1), dim3(64), 0, s, ctr, 4);
        MWD_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_zero_i64, dim3(1), dim3(64), 0, s, ctr + 17, 1);
        MWD_LAUNCH_CHECK();
    }
    // with bounding radii the survivors of the sphere test are listed as the
    // pairs are written: the wall pairs (all of them) at their own indices,
    // the cell pairs whose spheres meet appended after them by the emission
    ws->sphere_folded = st->radius != nullptr;
    const WallTests wtests{ws->scenes_dev, ws->n_scenes, ws->scene_wt_off};
    const WallOut wout{ws->pairs, ws->pair_cap, st->radius ? ws->surv : nullptr, ws->collision, ctr};
    // -- wall slab pairs and cell keys. On the key-table path with every
    // scene walled (at most MWD_MASK_WALLS walls each), the key pass also
    // tests each particle against its scene's walls and leaves the results as
    // bits, and the scan that places the wall pairs runs in the launch of the
    // scan over rows. Otherwise the wall scan tests as it goes, first. Either
    // way the step is one chain of launches on one stream (a parallel branch
    // for the wall scan costs more in cross-queue waits inside a replayed
    // graph than it overlaps).
    const bool words = walls && M > 0 && ws->cell_table && ws->n_words > 0 && ws->word_walls_ok;
    if (walls && !words) {
        const WallIO io{st->com, st->wall_n, st->wall_c, wtests, wout, st->radius, st->verts, st->nverts, st->vmax};
        rc = mwd_scan_io(io, ws->n_wall_tests, ws->wt_live, nullptr, ws->scan_tmp, ws->scan_tmp_len, s);
        if (rc) return rc;
    }
    // -- linked-cell pairs
    if (M > 0) {
        mwd_mark(ws, 1);
        if (ws->cell_table) {
            const WallWordsIn ww{st->wall_n,     st->wall_c,   st->radius,
                                 st->verts,      st->nverts,   st->vmax,
                                 st->radius != nullptr,        ws->wseg_off,
                                 words ? ws->hit_words : nullptr, ws->gjk_words};
            hipLaunchKernelGGL(k_key_count, dim3(blocks_for(M)), dim3(256), 0, s, st->com, keys_in, ws->plist, M,
                               ws->m_live, ws->scene_of, ws->scenes_dev, ws->row_key, ws->key_cnt, ws->key_min,
                               ws->flags, st->radius, st->radius ? ws->sph : nullptr, ww, ctr + 6,
                               reinterpret_cast<int32_t*>(ws->key_small), ws->key_head, ws->row_next,
                               ws->row_base);
            MWD_LAUNCH_CHECK();
            mwd_mark(ws, 2);
            const CellTableIO ct{ws->row_key, ws->key_cnt, ws->key_min, ws->cell_start, ws->cell_key,
                                 ws->key_seg, ctr + 0};
            if (words) {
                const WallWordIO wio{ws->hit_words, ws->gjk_words, ws->word_desc, ws->m_live, wout};
                rc = scan_cells_and_walls(ct, ws->n_rows, ws->scan_tmp, ws->scan_tmp_len, wio, ws->n_words,
                                          ws->scan_tmp_w, ws->scan_tmp_w_len, s);
            } else {
                rc = mwd_scan_io(ct, ws->n_rows, nullptr, nullptr, ws->scan_tmp, ws->scan_tmp_len, s);
            }
            if (rc) return rc;
            const KeyLists kl{ws->key_small, ws->key_head, ws->row_next, ws->key_cnt};
            const CellInfoIO io{ws->vals,      kl,             ws->cell_start, ws->cell_key,  ws->key_seg, ws->n_scenes,
                                ws->scenes_dev, ws->perm,      ws->cell_self, ws->cell_seg, ws->cell_off,
                                ctr,           ws->pair_cap,   ws->flags,     ws->sphere_folded != 0};
            switch (ws->tune.cell_items) {
                case 0:  // the per-cell work as its own kernel, then the pair-count scan
                    hipLaunchKernelGGL(k_cell_info_table, dim3((unsigned)((M + CELL_INFO_BLOCK - 1) / CELL_INFO_BLOCK)),
                                       dim3(CELL_INFO_BLOCK), 0, s, ws->vals, kl,
                                       ws->cell_start, ws->cell_key, ws->key_seg, ctr, ws->scenes_dev,
                                       ws->n_scenes, ws->perm, ws->cell_self, ws->cell_seg, ws->cell_cnt,
                                       ws->flags);
                    MWD_LAUNCH_CHECK();
                    rc = mwd_scan_io(CellCountIO{ws->cell_cnt, ws->cell_off, ctr, ws->pair_cap, ws->flags,
                                                 ws->sphere_folded != 0},
                                     M, ctr + 0, ctr + 2, ws->scan_tmp, ws->scan_tmp_len, s);
                    break;
                case 2:
                    rc = mwd_scan_io<CellInfoIO, 2>(io, M, ctr + 0, ctr + 2, ws->scan_tmp, ws->scan_tmp_len, s);
                    break;
                case 4:
                    rc = mwd_scan_io<CellInfoIO, 4>(io, M, ctr + 0, ctr + 2, ws->scan_tmp, ws->scan_tmp_len, s);
                    break;
                default:  // one cell per thread inside the scan
                    rc = mwd_scan_io<CellInfoIO, 1>(io, M, ctr + 0, ctr + 2, ws->scan_tmp, ws->scan_tmp_len, s);
            }
            if (rc) return rc;
        } else {
            hipLaunchKernelGGL(k_sort_keys, dim3(blocks_for(M)), dim3(256), 0, s, st->com, keys_in, ws->plist, M,
                               ws->scene_of, ws->scenes_dev, ws->keys, ws->vals, ws->flags, st->radius,
                               st->radius ? ws->sph : nullptr);
            MWD_LAUNCH_CHECK();
            uint32_t* ks;
            int32_t* vs;
            rc = mwd_radix_sort(ws->keys, ws->vals, ws->keys_alt, ws->vals_alt, M, nullptr, ws->key_bits, ws->hist,
                                ws->hist_off, ctr + 9, ws->scan_tmp, ws->scan_tmp_len, s, &ks, &vs);
            if (rc) return rc;
            mwd_mark(ws, 2);
            hipLaunchKernelGGL(k_cells_prep, dim3(blocks_for(ws->n_rows)), dim3(256), 0, s, ws->row_first,
                               ws->n_rows);
            MWD_LAUNCH_CHECK();
            rc = mwd_scan_io(HeadIO{ks, vs, M, ws->cell_start, ws->cell_key, ws->row_first}, M, nullptr, ctr + 0,
                             ws->scan_tmp, ws->scan_tmp_len, s);
            if (rc) return rc;
            rc = mwd_scan_i32(ws->row_first, ws->row_rank, nullptr, ws->n_rows, nullptr, ws->scan_tmp,
                              ws->scan_tmp_len, s);
            if (rc) return rc;
            hipLaunchKernelGGL(k_cell_info, dim3((unsigned)((M + CELL_INFO_BLOCK - 1) / CELL_INFO_BLOCK)),
                               dim3(CELL_INFO_BLOCK), 0, s, vs, ws->cell_start,
                               ws->cell_key, ctr + 0, ws->scenes_dev, ws->n_scenes, ws->row_rank, ws->perm,
                               ws->cell_self, ws->cell_seg, ws->cell_cnt, ws->flags);
            MWD_LAUNCH_CHECK();
            rc = mwd_scan_io(CellCountIO{ws->cell_cnt, ws->cell_off, ctr, ws->pair_cap, ws->flags,
                                         ws->sphere_folded != 0},
                             M, ctr + 0,
                             ctr + 2, ws->scan_tmp, ws->scan_tmp_len, s);
            if (rc) return rc;
        }
        mwd_mark(ws, 3);
        const hipEvent_t e0 = mwd_kev(ws, MWD_KT_EMIT, 0), e1 = mwd_kev(ws, MWD_KT_EMIT, 1);
        double* sph = st->radius ? ws->sph : nullptr;
        const int split = ws->tune.fill_split;
#define MWD_FILL(CELLS)                                                                                      \
    MWD_LAUNCH_EV(e0, e1, k_cell_fill<CELLS>, dim3((unsigned)(((M + (CELLS)-1) / (CELLS)) * split)),             \
                  dim3(FILL_THREADS), 0, s, ws->perm, ws->cell_self, ws->cell_seg, ws->cell_off, ctr, ws->pairs,  \
                  ws->pair_cap, sph, ws->surv, ctr + 10, ws->collision, split)
        switch (ws->tune.fill_cells) {
            case 4:
                MWD_FILL(4);
                break;
            case 16:
                MWD_FILL(16);
                break;
            default:
                MWD_FILL(64);
        }
#undef MWD_FILL
        MWD_LAUNCH_CHECK();
    }
    mwd_mark(ws, 4);
    if (M <= 0) {  // no cell-count scan closed the candidate count
        hipLaunchKernelGGL(k_finish_pairs, dim3(1), dim3(1), 0, s, ctr, ws->pair_cap, ws->flags,
                           ws->sphere_folded);
        MWD_LAUNCH_CHECK();
    }
    return MWD_OK;
}
