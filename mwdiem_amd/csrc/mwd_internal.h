// Internal declarations shared by the translation units of libmwdiem_hip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include <hip/hip_ext.h>

#include "../../include/mwdiem_hip.h"
#include "mwd_math.h"

#define MWD_WAVE 64
#define MWD_TIMING_RING 512
#define MWD_EV_PER_STEP (MWD_N_STAGES + 1 + 2 * MWD_N_KTIMED)
// timed kernels
#define MWD_KT_EMIT 0
#define MWD_KT_SPHERE 1
#define MWD_KT_GJK 2
#define MWD_KT_EPA 3
#define MWD_KT_VERLET 4

// ------------------------------------------------------------------ errors
void mwd_set_error(const std::string& msg);
#define MWD_HIP_CHECK(expr)                                                                 \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) {                                                             \
            mwd_set_error(std::string(#expr) + ": " + hipGetErrorString(e_));               \
            return MWD_E_HIP;                                                               \
        }                                                                                   \
    } while (0)
#define MWD_LAUNCH_CHECK()                                                                  \
    do {                                                                                    \
        hipError_t e_ = hipGetLastError();                                                  \
        if (e_ != hipSuccess) {                                                             \
            mwd_set_error(std::string("kernel launch: ") + hipGetErrorString(e_));          \
            return MWD_E_HIP;                                                               \
        }                                                                                   \
    } while (0)

// ------------------------------------------------------------ device scene
// Device copy of one scene plus the derived offsets the kernels need.
struct SceneDev {
    int64_t row0, n_walls, n_active;
    double box[6];
    double cell[3];
    int64_t ncx, ncy, kmin, kspan;
    int64_t kbase;      // first global (sorted) key of the scene
    int64_t wt_off;     // first wall-slab test index of the scene
    int64_t plist_off;  // first entry of the scene in the moving-particle list
    int64_t wseg0;      // first wall-word segment of the scene (segment wseg0 + k: wall k)
    double wall_thr;
    mwd_law law_p, law_w;
    int32_t n_off;      // 7 (reference stencil) or 13 (full half shell)
    int64_t off[13];    // key offsets of the neighbour cells, in emission order
};

#define MWD_MAX_OFF 13
#define MWD_WORD_WALLS 32  // walls per scene whose slab tests the key pass can gather into words

// Final GJK simplex of a colliding pair (input of EPA).
struct Simplex {
    double s[4][3];
    int32_t ia[4], ib[4];
};

// A body view: where the kernels find a pair's A (or B) side. Row r of the
// view is rows[p] (or p when rows is NULL) for pair p.
struct BodyView {
    const double* com;
    const double* verts;
    const int32_t* nverts;
    const double* vel;
    const double* angvel;
    const double* quat;
    const double* isize;
    const double* mass;         // per row of the view (physical law only; may be NULL)
    const double* radius;       // per row bounding radius (NULL: no sphere test)
    const uint8_t* is_wall;     // per row of the view
    const int32_t* scene;       // per row of the view (NULL: scene 0)
    const int32_t* rows;        // NULL: identity
    int32_t row_stride;         // pairs[p*row_stride] (2 for the interleaved pair list)
    int32_t vmax;
    const mwd_law* laws2;       // per scene (particle law, wall law); NULL: every pair takes the
                                // reference's GJK/EPA (the physical law's pairs take its variant)
};

// A state array and its MWD_N_TAN tangents (tangent pointers NULL in the
// forward path).
#define MWD_NT MWD_N_TAN
struct VState {
    TArr<MWD_NT> com, v, v12, w, wbf, w12bf, q, q12, qall, dxall, verts;
    const double *mass, *Iinv, *I;
    double* radius;
};
// Persistent tangential history (mwd_set_tangential_history): two
// open-addressing tables of (pair key min(a,b) + max(a,b)^2 -> displacement
// [4]), -1 for an empty key. *idx names the table the previous step wrote,
// which this step reads; this step writes the other one, listing the slots it
// fills (used, count) so that the table can be emptied slot by slot later.
struct HistView {
    int64_t* keys[2] = {nullptr, nullptr};
    double* val[2] = {nullptr, nullptr};
    int32_t* used[2] = {nullptr, nullptr};
    int64_t* count[2] = {nullptr, nullptr};
    int32_t* idx = nullptr;
    int64_t mask = 0;
};

// Per-contact force on A, torques on A and B (with tangents in the tangent path).
struct ContactOut {
    TArr<MWD_NT> Fn, tA, tB;
    // optional (forward path): EPA's final face per contact, 7 ints: support
    // indices on A (3) and B (3), then bit 0 = AddTriangle swapped its last
    // two vertices, bit 1 = re-evaluating the face does not give EPA's values,
    // bit 2 = the contact was not resolved (no face)
    int32_t* face = nullptr;
    // pair -> contact (the GJK slot); bucket records are keyed by pair index
    const int32_t* sidx = nullptr;
    // forward path: tangential history read from and written to these tables
    HistView hist;
};
// Tangents of the body arrays the contact stage reads.
struct BodyTan {
    TArr<MWD_NT> com, verts, vel, angvel, quat;
};

// Launch shapes (results do not depend on them).
struct PairTune {
    int gjk_group = 8;       // lanes sharing one pair's support map: 4, 8 or 16
    int gjk_blocks = 1024;   // grid cap of the grid-strided GJK launch
    int sphere_block = 256;  // threads per block of the sphere test
    int verlet_block = 256;  // particles per block of the fused reduction + integrator: 64, 128 or 256
    int cell_items = 1;      // key-table cell pass: cells per thread of the fused cell scan (1, 2, 4), 0 = split
    int verlet_waves = 1;    // minimum waves per SIMD asked of the reduction + integrator's registers: 1, 3 or 4
    int vertex_pack = 0;     // vertex pass over real vertices only (1) or over every vmax slot (0)
    int epa_blocks = 2048;   // grid cap of the EPA launch (blocks of 4 waves, one contact per wave)
    int epa_fused = 1;       // EPA and the law inside the GJK launch (1) or as their own launch (0)
    int fill_cells = 64;     // cells per block of the pair emission: 64, 16 or 4
    int fill_split = 1;      // blocks per emission block's cells, splitting its rounds: 1, 4 or 16
    int gjk_buckets = 0;     // GJK launched per class of vertex count (survivors sorted first): 1, or one launch (0)
};

// Face distances at or above this are EPA's placeholders (1e30 for a face
// whose plane misses the origin's projection, 3e20 for a deleted face), not
// penetration depths.
constexpr double EPA_PLACEHOLDER = 1e20;

// Row of pair p on one side of a view.
__device__ __forceinline__ int64_t view_row(const BodyView& b, int64_t p) {
    return b.rows ? (int64_t)b.rows[p * b.row_stride] : p;
}

// Bounding spheres of rows a and b meet (the test that decides most
// candidate pairs before GJK; its 'apart' answer is the reference's first
// support test's, .cu:1013).
__device__ __forceinline__ bool spheres_meet(const double* com, const double* radius, int64_t a, int64_t b) {
    const d3 d = ld3(com + 3 * b) - ld3(com + 3 * a);
    const double rr = radius[a] + radius[b];
    return !(nrm2(d) > rr * rr);
}

// Row r's centre and bounding radius packed as 4 doubles (two 16-byte
// halves), so that the sphere test loads each row with two vector loads.
__device__ __forceinline__ void pack_sphere(double* __restrict__ sph, const double* __restrict__ com,
                                            const double* __restrict__ radius, int64_t r) {
    double2* p = reinterpret_cast<double2*>(sph + 4 * r);
    p[0] = make_double2(com[3 * r], com[3 * r + 1]);
    p[1] = make_double2(com[3 * r + 2], radius[r]);
}

// spheres_meet on the packed rows: the same operations on the same values.
__device__ __forceinline__ bool spheres_meet_packed(const double* __restrict__ sph, int64_t a, int64_t b) {
    const double2* pa = reinterpret_cast<const double2*>(sph + 4 * a);
    const double2* pb = reinterpret_cast<const double2*>(sph + 4 * b);
    const double2 a0 = pa[0], a1 = pa[1], b0 = pb[0], b1 = pb[1];
    const d3 d = mk3(b0.x - a0.x, b0.y - a0.y, b1.x - a1.x);
    const double rr = a1.y + b1.y;
    return !(nrm2(d) > rr * rr);
}

// The contact records of row P in ascending record order (2p + side for the
// contact of candidate pair p, so in the order of the candidate pairs,
// whatever order the contacts were found in): its bucket of at most
// MWD_ROW_CONTACTS records sorted in registers, or, past a full bucket, the
// bucket plus the row's entries of the overflow list by repeated selection.
// fn(record) is called once per record.
template <class Fn>
__device__ __forceinline__ void for_row_records(int64_t P, int n, const int32_t* __restrict__ rec_bkt,
                                                const int32_t* __restrict__ rec_ovf, const int64_t* n_ovf,
                                                int64_t ovf_cap, Fn&& fn) {
    if (n <= 0) return;
    const int32_t* b = rec_bkt + P * MWD_ROW_CONTACTS;
    if (n <= MWD_ROW_CONTACTS) {
        int32_t r[MWD_ROW_CONTACTS];
        for (int k = 0; k < n; ++k) {  // insertion sort; records are distinct
            const int32_t x = b[k];
            int j = k;
            while (j > 0 && r[j - 1] > x) {
                r[j] = r[j - 1];
                --j;
            }
            r[j] = x;
        }
        for (int k = 0; k < n; ++k) fn(r[k]);
    } else {
        int64_t no = *n_ovf;
        if (no > ovf_cap) no = ovf_cap;
        int32_t prev = -1;
        for (int t = 0; t < n; ++t) {
            int32_t best = 0x7fffffff;
            for (int k = 0; k < MWD_ROW_CONTACTS; ++k)
                if (b[k] > prev && b[k] < best) best = b[k];
            for (int64_t o = 0; o < no; ++o)
                if (rec_ovf[2 * o] == (int32_t)P && rec_ovf[2 * o + 1] > prev && rec_ovf[2 * o + 1] < best)
                    best = rec_ovf[2 * o + 1];
            if (best == 0x7fffffff) break;  // lost to a full overflow list (flagged)
            fn(best);
            prev = best;
        }
    }
}

// --------------------------------------------------------------- workspace
struct mwd_workspace {
    int device;
    hipStream_t stream;
    int32_t n_scenes;
    int64_t n_rows, pair_cap, contact_cap;
    int32_t vmax;
    double dt;
    mwd_scene* scenes_host;
    SceneDev* scenes_dev;
    int64_t n_moving;          // M: moving particles over all scenes
    int64_t row_base;          // one scene: its first moving row (moving-list entry i is row row_base + i); else -1
    int64_t n_wall_tests;      // sum over scenes of W_s * (n_active_s - W_s)
    int64_t key_span_total;    // sum of kspan
    int key_bits, row_bits;    // significant bits of the sort keys
    // per-row tables
    int32_t* plist;            // [M] moving rows, row order
    int32_t* wlist;            // [n_wall_rows] wall rows, row order
    int64_t n_wall_rows;
    int wall_motion;           // walls translate by v dt each step (mwd_set_wall_motion)
    int reduce_fp32;           // per-particle contact sums in single precision (mwd_set_reduction)
    HistView thist;            // persistent tangential history tables (allocated when first enabled)
    int hist_on;
    int32_t* scene_of;         // [n_rows]
    uint8_t* is_wall;          // [n_rows]
    // coarse stage
    uint32_t *keys, *keys_alt; // [M]
    int32_t *vals, *vals_alt;  // [M]
    int32_t* cell_start;       // [M+1] first member slot of each cell (cell index order)
    uint32_t* cell_key;        // [M] key of each cell
    int2* cell_self;           // [M] by rank: (first member slot, members) of the cell
    int2* cell_seg;            // [M*MWD_MAX_OFF] by rank: (first slot, members) of each neighbour cell, (0,0) if none
    int32_t* row_first;        // [n_rows] (sorted path) 1 where a row is the lowest of its cell
    int64_t* row_rank;         // [n_rows] (sorted path) exclusive scan of row_first
    int32_t* cell_cnt;         // [M] pair count by rank
    int64_t* cell_off;         // [M]
    int32_t* perm;             // [M] members of each cell in set order
    // key-table path (cell_table = 1): cells numbered by rank directly, no sort
    int cell_table;            // key-table path in use (MWD_CELLS_TABLE)
    int sphere_folded;         // host: the last candidate pass also did the bounding-sphere test
    int cell_table_ok;         // its tables exist
    PairTune tune;
    int4* key_seg;             // [key_span_total] per key this pass: (epoch, first member slot, members, 0);
                               // an entry whose epoch is not the pass's (counters[6]) has no cell
    int4* key_small;           // [4 * key_span_total] per key its first sixteen member rows (in arrival order)
    int32_t* key_head;         // [key_span_total] per key the list of its further members (-1: none)
    int32_t* row_next;         // [n_rows] next row of a key's list
    int32_t* key_cnt;          // [2 * key_span_total] per key k: members at 2k (zero between steps) and
    int32_t* key_min;          // lowest row at 2k + 1 (KEY_MIN_EMPTY between steps); key_min = key_cnt + 1
    int32_t* row_key;          // [n_rows] key of each moving row this step, -1 for walls
    double* sph;               // [n_rows*4] (centre, bounding radius) of the moving rows, packed by the key
                               // pass when radii are kept, for the sphere test inside the emission
    // candidate pairs
    int32_t* pairs;            // [pair_cap*2]
    int32_t* collision;        // [pair_cap]
    Simplex* simplex;          // [contact_cap], slot claimed by GJK per colliding pair
    int32_t* sidx;             // [pair_cap] simplex slot of a colliding pair
    int32_t* surv;             // [pair_cap] pairs whose bounding spheres meet (any order)
    int32_t* surv_cls;         // [pair_cap] the same sorted by GJK class (gjk_buckets)
    int64_t* bucket;           // [4 * MWD_GJK_BUCKETS] class counts, cursors and offsets
    int fresh_pairs;           // host: candidates not yet consumed by a fine-contact stage
    // contacts
    int32_t* contact_pair;     // [contact_cap]
    double *depth, *normal, *cpA, *cpB, *Fn, *tA, *tB;
    int32_t* epa_status;       // [contact_cap]
    double* ctan;              // [3 arrays * MWD_NT * contact_cap * 3] tangents of Fn, tA, tB (tangent step)
    mwd_law_tan* law_tans;     // [2*n_scenes] (tangent step, adjoint step)
    // adjoint step (allocated at its first call)
    double *adj_F, *adj_T;         // [n_rows*3] the step's force and torque per row
    double *adj_lamF, *adj_lamT;   // [n_rows*3] their adjoints
    int32_t* adj_face;             // [7*contact_cap] EPA's final face per contact
    double* adj_c;                 // [2*18*contact_cap] per contact side: adjoint contributions
    double* adj_cth;               // [3*contact_cap] per contact: d/d theta
    double* adj_rowth;             // [3*n_rows] per moving-list entry: its row's contacts' share of d/d theta
    int emit_faces;                // host: the next fine-contact pass writes adj_face
    // the reverse sweep's tape (mwd_tape_init): per recorded step, the
    // contact stages' outputs the adjoint step reads; counters[18] slots held
    uint32_t* tape;                // [tape_slots * tape_words]
    uint32_t* tape_sync;           // [MWD_TAPE_SYNC_WORDS] block counters of the tape launches (zero between them)
    int64_t tape_slots, tape_words;
    // per moving row, the records (2c for side A of contact c, 2c+1 for side
    // B) of this step's contacts, in arrival order; the reduction sorts them
    int32_t* rec_cnt;          // [n_rows], zero between steps
    int32_t* rec_bkt;          // [n_rows * MWD_ROW_CONTACTS]
    int32_t* rec_ovf;          // [2 * 2*contact_cap] (row, record) of records beyond a full bucket; count in counters[11]
    int bkt_dirty;             // host: buckets filled and not yet consumed
    // sort / scan scratch
    int32_t* hist;             // [256 * max_tiles]
    int64_t* hist_off;         // [256 * max_tiles]
    int64_t* scan_tmp;         // block sums
    int64_t scan_tmp_len;
    // wall slab tests as bit words, written by the key pass: segment g (one
    // per scene and wall, scene-major) holds the words of its scene's
    // moving-particle list entries, 64 per word (word w: entries 64w ..
    // 64w+63; a scene's first and last words are shared with its
    // neighbours' ranges and its bits there are its own). Concatenated, the
    // words run in the reference's test order, so a scan over words places
    // the pairs. The words' scan has its own scratch: it runs in the launch
    // of the scan over rows.
    int64_t n_wseg;            // segments (sum over scenes of their walls)
    int64_t n_words;           // words over all segments
    int64_t* wseg_off;         // [n_wseg + 1] first word of each segment
    int4* word_desc;           // [n_words] per word: scene, wall row, row of bit 0, entry of bit 0 in the scene
    int word_walls_ok;         // no scene has more than MWD_WORD_WALLS walls
    uint64_t* hit_words;       // [n_words] bit set: the pair is listed
    uint64_t* gjk_words;       // [n_words] bit set: the listed pair goes to GJK
    int64_t* scan_tmp_w;
    int64_t scan_tmp_w_len;
    // counters (device): 0 n_cells, 1 n_wall_pairs, 2 n_cell_pairs, 3 n_pairs
    // (capped), 4 n_contacts, 5 n_records, 6 epoch of the key-table cell pass,
    // 7 n_pairs before capping, 8 simplex slots claimed this step, 9 radix
    // histogram entries (scratch), 10 sphere-test survivors, 11 bucket
    // overflow records, 12-13 slots filled in the tangential history tables,
    // 14 moving rows active now (m_live), 15 wall slab tests now (wt_live),
    // 16 steps taken since the spawn schedule was set, 17 wall pairs listed
    // for GJK (the rest are separated by the wall's slab), 18 steps held on
    // the reverse sweep's tape
    int64_t* counters;
    int64_t* m_live;           // counters + 14: every kernel over the moving rows reads its count here
    int64_t* wt_live;          // counters + 15
    // spawning (one scene): rows [row0 + n_active, spawn_cap) are activated
    // spawn_per at a time every spawn_every steps, at the start of a step
    int spawn_on;
    int64_t spawn_every, spawn_per, spawn_cap;
    uint32_t* flags;
    mwd_law* laws;             // [2*n_scenes] (particle law, wall law)
    int64_t* scene_wt_off;     // [n_scenes+1] wall-test prefix (device)
    // optional per-stage timing (HIP events on the workspace stream)
    int timing;
    hipEvent_t* ev;            // [MWD_TIMING_RING * MWD_EV_PER_STEP]
    int ring_slot;             // steps recorded and not yet accumulated
    int64_t timing_errors;     // event records or elapsed-time queries that failed
    double kernel_ms[MWD_N_KTIMED];
    int64_t kernel_calls;
    uint32_t kt_slot[MWD_TIMING_RING];  // per ring slot: the timed kernels launched with events in that step
    double stage_ms[MWD_N_STAGES];
    int64_t stage_calls;
};

// stage marks (events recorded before each stage when timing is on)
void mwd_mark(mwd_workspace* ws, int stage);
void mwd_mark_end(mwd_workspace* ws);
// Start/stop events of a timed kernel in the current step (NULL when timing
// is off): such kernels are launched with hipExtLaunchKernelGGL, which
// stamps the events with the kernel's own start and end.
hipEvent_t mwd_kev(mwd_workspace* ws, int kt, int stop);

// Launch with start/stop events when e0 is set, else a plain launch.
#define MWD_LAUNCH_EV(e0, e1, kernel, grid, block, shmem, stream, ...)                                   \
    do {                                                                                                 \
        if (e0 || e1)                                                                                    \
            hipExtLaunchKernelGGL(kernel, grid, block, shmem, stream, e0, e1, 0u, __VA_ARGS__);          \
        else                                                                                             \
            hipLaunchKernelGGL(kernel, grid, block, shmem, stream, __VA_ARGS__);                         \
    } while (0)

// ----------------------------------------------------------- primitives
// Exclusive scan of int32 values into int64 offsets. n is n_dev[0] when
// n_dev is given (n_cap bounds it), else n_cap. total (device) may be NULL.
int mwd_scan_i32(const int32_t* in, int64_t* out, int64_t* total, int64_t n_cap, const int64_t* n_dev,
                 int64_t* tmp, int64_t tmp_len, hipStream_t s);
int64_t mwd_scan_tmp_len(int64_t n_cap);
// Stable LSD radix sort of (key, val) pairs on the low key_bits key bits, in
// ceil(key_bits / MWD_SORT_MAX_BITS) passes, ping-ponging between (keys,
// vals) and the alt buffers; *keys_out/*vals_out name the buffers holding the
// result (the alt ones when the pass count is odd). hist/hist_off hold
// 2^MWD_SORT_MAX_BITS * mwd_sort_tiles(n_cap) entries; hcount is one device
// int64 of scratch.
#define MWD_SORT_MAX_BITS 11
int mwd_radix_sort(uint32_t* keys, int32_t* vals, uint32_t* keys_alt, int32_t* vals_alt, int64_t n_cap,
                   const int64_t* n_dev, int key_bits, int32_t* hist, int64_t* hist_off, int64_t* hcount,
                   int64_t* tmp, int64_t tmp_len, hipStream_t s, uint32_t** keys_out, int32_t** vals_out);
int64_t mwd_sort_tiles(int64_t n_cap);
int mwd_fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t s);
// The empty key table: (0 members, no lowest row) for each of n keys.
int mwd_fill_key_table(int32_t* key_cm, int64_t n, hipStream_t s);
int mwd_fill2_i32(int32_t* p, int32_t* q, int32_t v, int64_t n, hipStream_t s);
int mwd_set_i64(int64_t* p, int64_t v, hipStream_t s);

// ------------------------------------------------------------ stages
int mwd_launch_verlet(double* v, double* v12, double* w12bf, double* q12, double* q, double* qall, double* wbf,
                      double* w, const double* F, const double* tau, double* verts, double* com,
                      const double* mass, const double* Iinv, const double* I, double* dxall, int32_t vmax,
                      const int32_t* nrot, double dt, const double* box, int64_t pmin, int64_t pmax,
                      hipStream_t s);
VState mwd_vstate(const mwd_state* st, double* const* tan);
// Force assembly + Verlet over the moving rows; tan != NULL runs the tangent path.
int mwd_launch_reduce_verlet(mwd_workspace* ws, mwd_state* st, double* const* tan, const ContactOut& co);
int mwd_launch_move_walls(mwd_workspace* ws, mwd_state* st);
// Tangential history: empty the table this step writes (start of the fine
// contact stage); make it the one read next step (end of the step).
int mwd_launch_hist_prep(const HistView& h, int64_t contact_cap, hipStream_t s);
int mwd_launch_hist_flip(const HistView& h, hipStream_t s);
int mwd_launch_radius(const double* com, const double* verts, const int32_t* nverts, int64_t n, int32_t vmax,
                      double* radius, hipStream_t s);
int mwd_launch_candidates(mwd_workspace* ws, const mwd_state* st, const int64_t* keys);
// n arrays of 8-byte words: dst[i][0, words[i]) <- src[i][0, words[i]).
int mwd_launch_copy_arrays(int n, void* const* dst, const void* const* src, const int64_t* words, hipStream_t s);
// Adjoint of one step after its contact stages have run on the step's input
// state (adj_face written): lam in/out, dtheta (device, 3) accumulated.
int mwd_launch_step_adjoint(mwd_workspace* ws, const mwd_state* st, const BodyView& A, const BodyView& B,
                            double* const* lam, double* dtheta);
// The tape: words per slot for this workspace; push the input state and the
// contact stages' outputs of the step running on st (push = true), or pop
// the last recorded step's back into st and the workspace arrays; empty it.
#define MWD_TAPE_SYNC_WORDS (17 * 32)
int64_t mwd_tape_words(const mwd_workspace* ws);
int mwd_launch_tape(mwd_workspace* ws, const mwd_state* st, bool push);
int mwd_launch_tape_rewind(mwd_workspace* ws);
int mwd_launch_cell_keys(const double* com, int64_t n, const double* cell, int64_t ncx, int64_t ncy,
                         int64_t* keys, hipStream_t s);
// GJK over pairs [0, n) (n from n_dev), in two launches: a bounding-sphere
// test writes collision = 0 for pairs whose spheres are apart and appends
// the others to surv (counter *n_surv, zero on entry); GJK then runs over the
// survivors only, writing their collision codes. A colliding pair claims a
// simplex slot (counter *slots, zero on entry) and stores it in sidx.
int mwd_launch_sphere(const BodyView& A, const BodyView& B, int64_t n_cap, const int64_t* n_dev, int32_t* collision,
                      int32_t* surv, int64_t* n_surv, hipStream_t s, hipEvent_t e0 = nullptr,
                      hipEvent_t e1 = nullptr, const PairTune& tu = PairTune());
// contact_pair (may be NULL): slot -> pair index, written with the slot.
// lo (may be NULL: 0): the survivors [*lo, *n_surv) only; K_class > 0: the
// launch for survivors of at most K_class * gjk_group vertices per body.
int mwd_launch_gjk_surv(const BodyView& A, const BodyView& B, int64_t n_cap, int32_t* collision,
                        const int32_t* surv, const int64_t* n_surv, Simplex* simplex, int32_t* sidx, int64_t* slots,
                        int64_t slot_cap, uint32_t* flags, hipStream_t s, hipEvent_t e0 = nullptr,
                        hipEvent_t e1 = nullptr, const PairTune& tu = PairTune(), int32_t* contact_pair = nullptr,
                        const int64_t* lo = nullptr, int K_class = 0);
// Survivors sorted by GJK class k = ceil(max(nA, nB) / G) - 1 (the last of
// the nb classes takes the rest) into out, class after class. bucket: int64
// scratch of 4 * MWD_GJK_BUCKETS (counts, cursors, then the nb + 1 class
// start offsets at bucket + 2 * MWD_GJK_BUCKETS).
constexpr int MWD_GJK_BUCKETS = 8;
int mwd_launch_gjk_buckets(const BodyView& A, const BodyView& B, const int32_t* surv, const int64_t* n_surv,
                           int64_t n_cap, int G, int nb, int64_t* bucket, int32_t* out, hipStream_t s);
// Both of the above.
int mwd_launch_gjk(const BodyView& A, const BodyView& B, int64_t n_cap, const int64_t* n_dev,
                   int32_t* collision, int32_t* surv, int64_t* n_surv, Simplex* simplex,
                   int32_t* sidx, int64_t* slots, int64_t slot_cap, uint32_t* flags, hipStream_t s);
// Contact compaction (one fused scan): contact_pair[c] = pair index of the
// c-th collision == 1.
int mwd_launch_compact_contacts(const int32_t* collision, int64_t n_cap, const int64_t* n_dev, int32_t* contact_pair,
                                int64_t contact_cap, int64_t* n_contacts, uint32_t* flags, int64_t* tmp,
                                int64_t tmp_len, hipStream_t s);
// EPA + force law over contacts [0, K). Reduction records are keyed by
// idA[p]/idB[p] when given, else by the view rows.
int mwd_launch_epa_force(const BodyView& A, const BodyView& B, const mwd_law* laws2, const int32_t* contact_pair,
                         const Simplex* simplex, const int32_t* sidx, int64_t contact_cap, const int64_t* n_contacts, const double* dispt,
                         const int32_t* idA, const int32_t* idB, double* depth, double* normal, double* cpA,
                         double* cpB, double* Fn, double* tA, double* tB, double* disp_out, int32_t* epa_status,
                         uint32_t* rec_key, int32_t* rec_val, int64_t* n_records, uint32_t* flags, hipStream_t s);
// The stepper's form over one state: records go to per-row buckets (moving
// rows only) instead of a list; bt (tangent path) carries the state's
// tangents, lawt2 the law derivatives per scene like laws2, and co receives
// values (and tangents) of Fn, tA, tB.
// GJK, EPA and the law in one launch, for the plain FP64 step with bucketed
// records: the results of mwd_launch_gjk_surv followed by
// mwd_launch_epa_force_rows. Returns 1, launching nothing, when the bodies
// have more than 4 * gjk_group vertices or the arguments do not fit this
// form; a negative code on a launch error.
int mwd_launch_gjk_epa(const BodyView& A, const BodyView& B, int64_t n_cap, int32_t* collision, const int32_t* surv,
                       const int64_t* n_surv, Simplex* simplex, int32_t* sidx, int64_t* slots, int64_t slot_cap,
                       int32_t* contact_pair, const mwd_law* laws2, double* depth, double* normal, double* cpA,
                       double* cpB, const ContactOut& co, int32_t* epa_status, int32_t* rec_cnt, int32_t* rec_bkt,
                       int32_t* rec_ovf, int64_t* n_ovf, int64_t ovf_cap, uint32_t* flags, hipStream_t s,
                       hipEvent_t e0, hipEvent_t e1, const PairTune& tu, const int64_t* lo = nullptr,
                       int K_class = 0);
int mwd_launch_epa_force_rows(const BodyView& A, const BodyView& B, const BodyTan* bt, const mwd_law* laws2,
                              const mwd_law_tan* lawt2, const int32_t* contact_pair, const Simplex* simplex,
                              const int32_t* sidx, int64_t contact_cap, const int64_t* n_contacts, double* depth,
                              double* normal, double* cpA, double* cpB, const ContactOut& co, int32_t* epa_status,
                              int32_t* rec_cnt, int32_t* rec_bkt, int32_t* rec_ovf, int64_t* n_ovf,
                              int64_t ovf_cap, uint32_t* flags, hipStream_t s, hipEvent_t e0 = nullptr,
                              hipEvent_t e1 = nullptr, int64_t max_blocks = 2048);
// Zero the per-pair outputs, then copy per-contact results to their pairs.
int mwd_launch_zero_pairs(int64_t P, double* depth, double* normal, double* cpA, double* cpB, int32_t* epa_status,
                          double* disp, hipStream_t s);
int mwd_launch_scatter_contacts(const int32_t* cpair, int64_t cap, const int64_t* n_contacts, const double* dep,
                                const double* nrm, const double* ca, const double* cb, const int32_t* est, int64_t P,
                                double* depth, double* normal, double* cpA, double* cpB, int32_t* epa_status,
                                double* disp, hipStream_t s);
// Segment bounds of sorted records per row.
int mwd_launch_segments(const uint32_t* rec_key, int64_t rec_cap, const int64_t* n_rec, int32_t* seg_begin,
                        int32_t* seg_end, hipStream_t s);
// Add per-row contact sums onto force/torque rows (drop-in contact path).
int mwd_launch_add_sums(const int32_t* seg_begin, const int32_t* seg_end, const int32_t* rec_val,
                        const double* Fn, const double* tA, const double* tB, double* force, double* torque,
                        int64_t n_rows, hipStream_t s);
