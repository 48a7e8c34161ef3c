/*
 * mwdiem_hip.h -- C ABI of libmwdiem_hip.so, the MI355X (gfx950) implementation
 * of MWDiEM's per-timestep DEM loop for convex polyhedra.
 *
 * Conventions
 *   - Every entry point returns int: MWD_OK (0) or a negative MWD_E_* code.
 *     mwd_last_error() returns a short text for the last failure on the
 *     calling thread.
 *   - Array arguments are DEVICE pointers (e.g. torch.Tensor.data_ptr() of a
 *     tensor on the GPU), C-contiguous, in the reference's own layouts: [N,3]
 *     rows for vectors, [N,4] for quaternions (w,x,y,z), [N,3,3] for inertia,
 *     [N,Vmax,3] for vertices. Floating point is FP64 throughout.
 *   - `stream` is a hipStream_t (NULL = the default stream). Calls are
 *     stream-ordered and asynchronous unless stated otherwise.
 *   - Units follow the reference: mm, kg, s.
 *
 * Reference interfaces replaced (kgillemot/MWDiEM):
 *   mwd_verlet ............ __global__ Verlet, LIBRARIES/gpu_Verlet.cu:85-431,
 *                           launched by gpu_Verlett, LIBRARIES/gpu_Verlet.py:18-50
 *   mwd_contact_pairs ..... __global__ contact_detection,
 *                           LIBRARIES/gpu_fine_contact_detection.cu:874-1322,
 *                           launched by gpu_contact_detection,
 *                           LIBRARIES/gpu_fine_contact_detection.py:18-73
 *   mwd_cell_keys ......... cell keys, MWDiem.py:813-816
 *   mwd_candidates ........ wall slab test MWDiem.py:807-811 followed by
 *                           Linked_cell_hash,
 *                           LIBRARIES/LIBRARY_Coarse_contact_detection.py:84-155
 *   mwd_step .............. one iteration of the time loop, MWDiem.py:805-1001
 *                           (candidates, gather, contact_detection, force
 *                           assembly :977-981, Verlet), device-resident
 */
#ifndef MWDIEM_HIP_H
#define MWDIEM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MWD_OK 0
#define MWD_E_INVALID (-1)   /* bad argument (null pointer, negative size ...) */
#define MWD_E_HIP (-2)       /* a HIP runtime call failed */
#define MWD_E_CAPACITY (-3)  /* a workspace buffer is too small; see mwd_step_status() */
#define MWD_E_KEYRANGE (-4)  /* a cell key fell outside the scene's key range */
#define MWD_E_CELL (-5)      /* a cell holds more particles than the set-order emulation supports */

/* Status bits of mwd_step_status().flags */
#define MWD_F_PAIR_OVERFLOW 1u     /* more candidate pairs than pair_cap */
#define MWD_F_CONTACT_OVERFLOW 2u  /* more contacts than contact_cap */
#define MWD_F_KEYRANGE 4u          /* cell key outside [kmin, kmin+kspan) (clamped) */
#define MWD_F_CELL 8u              /* cell occupancy above MWD_MAX_CELL */
#define MWD_F_EPA_CAP 16u          /* some EPA hit its 25-expansion cap (.cu:443-447) */
#define MWD_F_EPA_FACES 32u        /* some EPA ran out of face storage */
#define MWD_F_ROW_CONTACTS 64u     /* the stepper's contact-record overflow list is full */
#define MWD_F_EPA_UNRESOLVED 128u  /* some EPA ended on a placeholder face (epa_status bit 4);
                                      the physical law applies no force for such a contact */
#define MWD_F_TAPE 256u            /* mwd_step_record found the tape full, or
                                      mwd_step_adjoint_taped found it empty (nothing copied) */

#define MWD_EPA_MAX_FACES 128
#define MWD_MAX_CELL 76            /* largest cell occupancy whose set order is emulated */
#define MWD_ROW_CONTACTS 32        /* contact records kept per moving particle and step (stepper);
                                      further ones go to an overflow list of 2*contact_cap entries */

/* Contact-law coefficients of Force (.cu:685-701), filled by mwd_law_init in
 * the reference's operator order.
 * mode MWD_LAW_REF: the reference law as written (contact point (cpA-cpB)/2,
 *   mu in place of a mass in the damping time, only the normal force on the
 *   linear momentum).
 * mode MWD_LAW_PHYSICAL: contact point (cpA+cpB)/2, normal damping from the
 *   effective mass 1/m_n = 1/mA + 1/mB, tangential stiffness and damping from
 *   1/m_t = 1/m_n + rA^2/IA + rB^2/IB (walls infinitely heavy), normal plus
 *   tangential force on the linear momentum. Needs per-row masses (stepper). */
#define MWD_LAW_REF 0
#define MWD_LAW_PHYSICAL 1
typedef struct mwd_law {
    double Kn, mu, ln_en, ln_et, A_en, tc, Gn, Kt_num, inv_mu, dt;
    double cn, ct, ratio;   /* -2 ln(en)/sqrt(pi^2+ln^2 en), same for et, (pi^2+ln^2 et)/(pi^2+ln^2 en) */
    int32_t mode, pad_;
} mwd_law;

/* One scene (one independent bed with its walls, grid and law). Rows
 * [row0, row0+n_walls) are the walls, rows [row0+n_walls, row0+n_active) the
 * particles that move; ids inside a scene are row-row0 (the reference's
 * particle ids). Host struct; the library copies it. */
typedef struct mwd_scene {
    int64_t row0, n_walls, n_active;
    double box[6];          /* syssize x_max, x_min, y_max, y_min, z_max, z_min */
    double cell[3];         /* cellsize_x, _y, _z (MWDiem.py:578-580) */
    int64_t ncx, ncy;       /* numcells_x, numcells_y */
    int64_t kmin, kspan;    /* every key of the scene lies in [kmin, kmin+kspan) */
    double wall_thr;        /* 6*max_particlesize (MWDiem.py:810) */
    mwd_law law_p, law_w;   /* particle-particle and wall laws */
} mwd_scene;

/* Device-resident particle state (all device pointers, rows of all scenes). */
typedef struct mwd_state {
    int64_t n_rows;
    int32_t vmax;
    double *com, *v, *v12, *w, *wbf, *w12bf;   /* [n,3] */
    double *q, *q12, *qall;                    /* [n,4] */
    double *dxall;                             /* [n,3] */
    double *verts;                             /* [n,vmax,3] world vertices */
    const int32_t *nverts;                     /* [n] */
    const double *mass, *I, *Iinv, *isize;     /* [n], [n,3,3], [n,3,3], [n] */
    const double *fext, *text;                 /* [n,3] external force/torque */
    const double *wall_n, *wall_c;             /* [n,3] wall normal and centre (rows of walls) */
    double *radius;                            /* [n] bounding radius about com (optional, see below) */
} mwd_state;
/* radius: when given, radius[r] must bound max |vertex - com| of row r with a
 * margin (mwd_radius_init writes such values; mwd_step keeps them current for
 * moving rows). GJK then decides pairs whose bounding spheres are apart without
 * reading vertices: for them the reference's first support test
 * dot(d1, S1) < 0 (.cu:1013) holds, so the result (collision 0) is the same. */

typedef struct mwd_workspace mwd_workspace;

/* Status of the last mwd_step / mwd_candidates on a workspace (synchronises
 * the workspace's stream). */
typedef struct mwd_status {
    int64_t n_pairs;      /* candidate pairs (wall pairs first) */
    int64_t n_wall_pairs;
    int64_t n_contacts;   /* pairs with collision == 1 */
    int64_t n_cells;
    uint32_t flags;       /* MWD_F_* */
    int64_t n_sphere_pass;  /* pairs whose bounding spheres meet (the pairs GJK ran on) */
} mwd_status;

const char* mwd_last_error(void);
int mwd_version(void);

/* Coefficients of the normal/tangential law for (Kn, en, et, mu) and dt
 * (mode MWD_LAW_REF); mwd_law_init_mode selects the mode. */
int mwd_law_init(mwd_law* out, double Kn, double en, double et, double mu, double dt);
int mwd_law_init_mode(mwd_law* out, double Kn, double en, double et, double mu, double dt, int32_t mode);

/* Workspace owning every scratch buffer of the step: sort buffers, cell
 * tables, the candidate pair list (pair_cap pairs), contacts (contact_cap).
 * Scenes are copied in. Calls on distinct workspaces are independent.
 * Limits: vmax <= 255, n_rows < 2^31. */
int mwd_workspace_create(mwd_workspace** out, int device, const mwd_scene* scenes, int32_t n_scenes,
                         int64_t n_rows, int32_t vmax, int64_t pair_cap, int64_t contact_cap, double dt,
                         void* stream);
int mwd_workspace_destroy(mwd_workspace* ws);

/* Moving walls (no counterpart in the reference, whose walls are static;
 * for the sheared-box configuration): with `on`, each mwd_step (and tangent
 * step) ends by translating every wall row r by v[r]*dt -- its centre, its
 * nverts vertices and its slab centre wall_c (which is then written). A wall's
 * velocity v[r] also enters the contact law as the body's velocity. */
int mwd_set_wall_motion(mwd_workspace* ws, int32_t on);

/* Spawning (MWDiem.py:685-714: the driver creates particles at the start of
 * every 5th step and widens the active range [num_particles_wall,
 * num_particles_temp) that the hash and the Verlet step cover). For a
 * workspace of one scene: from the next step on only rows [row0, row0 +
 * n_active) take part; at the start of every `every`-th step (the first
 * included) `per` more of the scene's rows become active, in row order, up
 * to the rows the scene was created with. The rows waiting to spawn hold
 * their particles as they will enter (the reference writes them at spawn
 * time); nothing reads or moves them before. The counts live on the device,
 * so spawning steps are captured in graphs like any others. Key-table cell
 * pass only; not with mwd_step_adjoint. */
int mwd_set_spawn(mwd_workspace* ws, int64_t n_active, int64_t every, int64_t per);

/* Tangential history (a5, MWDiem.py:649,872-880,953-966). MWD_HISTORY_RESET
 * (default): the reference as it runs -- the history is dropped every step
 * (Q1), so every contact starts from a zero tangential displacement.
 * MWD_HISTORY_PERSISTENT: the history the reference intends -- each
 * contact's new displacement (.cu:770-811) is kept, keyed min(a,b)+max(a,b)^2
 * as the driver keys it, and read back next step if the pair is still in
 * contact; pairs that part lose it. Held in two open-addressing tables on the
 * device (2x contact_cap slots each, allocated on first use). One
 * fine-contact pass per step; not with the tangent or adjoint step. */
#define MWD_HISTORY_RESET 0
#define MWD_HISTORY_PERSISTENT 1
int mwd_set_tangential_history(mwd_workspace* ws, int32_t mode);

/* Precision of the per-particle contact sums (a15-a16): MWD_REDUCE_FP64
 * (default) or MWD_REDUCE_FP32, the precision of the reference's force and
 * torque arrays (gpu_fine_contact_detection.cu:1304-1319, FP32 atomics), still
 * summed in candidate order. Contact forces themselves stay FP64. */
#define MWD_REDUCE_FP64 0
#define MWD_REDUCE_FP32 1
int mwd_set_reduction(mwd_workspace* ws, int32_t mode);

/* Launch choices of a workspace. No counterpart in the reference (its
 * kernels have fixed launches, gpu_fine_contact_detection.py:64-73); every
 * choice gives the same results, bit for bit, and the autotuner
 * (mwdiem_amd/tune.py) picks among them by timing.
 *   cell_pass    MWD_CELLS_TABLE: cells from key-indexed count/lowest-row
 *                tables, no sort (available while the summed key span is at
 *                most 2^28; the default then); MWD_CELLS_SORTED: radix sort
 *                of (key, row).
 *   gjk_group    lanes sharing one pair's support map: 4, 8 (default) or 16,
 *                i.e. 16, 8 or 4 pairs per wave.
 *   gjk_blocks   grid cap of the grid-strided GJK launch, 64..65536 (1024).
 *   sphere_block threads per block of the bounding-sphere test, a multiple
 *                of 64 in 64..1024 (256).
 *   verlet_block particles per block of the fused force reduction and
 *                integrator: 64, 128 or 256 (256).
 *   cell_items   key-table cell pass: cells per thread of the scan that
 *                also does the per-cell work (1, 2 or 4; 1), or 0 for a
 *                separate per-cell kernel followed by the pair-count scan.
 *   verlet_waves minimum waves per SIMD that the register allocation of the
 *                fused reduction + integrator must allow: 1 (the compiler's
 *                choice; default), 3 or 4 (fewer registers, possibly spills).
 *   vertex_pack  1: the vertex pass walks the block's real vertices only
 *                (prefix sum of the vertex counts); 0 (default): every vmax
 *                slot, skipping the padding.
 *   epa_blocks   grid cap of the grid-strided EPA launch (4 contacts per
 *                block at a time), 16..65536 (2048).
 *   epa_fused    1 (default): the step runs EPA and the law inside the GJK
 *                launch, each wave on the colliding pairs it found (bodies
 *                of at most 4 * gjk_group vertices, forward steps); 0: a
 *                separate EPA launch over all contacts.
 *   fill_cells   cells per block of the pair emission: 64 (default), 16 or
 *                4 (fewer for beds packed many hulls to a cell).
 *   fill_split   blocks sharing one emission block's cells, each taking every
 *                fill_split-th round of their pairs: 1 (default), 4 or 16
 *                (more blocks for beds of few, crowded cells).
 *   gjk_buckets  1: the sphere-test survivors are first sorted by
 *                ceil(max(nA, nB) / gjk_group) and GJK runs one launch per
 *                class, each with the vertices-per-lane its class needs
 *                (uneven hull sizes); 0 (default): one launch sized for the
 *                largest hull.
 * mwd_set_tuning applies to later steps (graphs captured before keep their
 * launches) and returns MWD_E_INVALID for a choice out of range. */
#define MWD_CELLS_TABLE 0
#define MWD_CELLS_SORTED 1
typedef struct mwd_tuning {
    int32_t cell_pass;
    int32_t gjk_group;
    int32_t gjk_blocks;
    int32_t sphere_block;
    int32_t verlet_block;
    int32_t cell_items;
    int32_t verlet_waves;
    int32_t vertex_pack;
    int32_t epa_blocks;
    int32_t epa_fused;
    int32_t fill_cells;
    int32_t gjk_buckets;
    int32_t fill_split;
} mwd_tuning;
int mwd_get_tuning(const mwd_workspace* ws, mwd_tuning* out);
int mwd_set_tuning(mwd_workspace* ws, const mwd_tuning* t);

/* Verlet over rows [pmin, pmax) of the reference's arrays
 * (gpu_Verlet.cu:85-431). nrot: per-row count of vertex slots to rotate, or
 * NULL to rotate all vmax slots as the reference does. box as in mwd_scene. */
int mwd_verlet(double* v, double* v12, double* w12bf, double* q12, double* q, double* qall, double* wbf,
               double* w, const double* F, const double* tau, double* verts, double* com, const double* mass,
               const double* Iinv, const double* I, double* dxall, int32_t vmax, const int32_t* nrot, double dt,
               const double* box, int64_t pmin, int64_t pmax, void* stream);

/* Fine contact over P pairs in the reference's per-pair layout
 * (gpu_fine_contact_detection.cu:874-895, FP64):
 *   comA/comB/vA/vB/wA/wB [P,3], vertsA/vertsB [P,vmax,3], nA/nB/idA/idB [P],
 *   dispt [P,4], qA [P,4], isA/isB [P].
 * Outputs (each may be NULL): collision [P] (-1 undecided, 0 apart, 1
 * contact), epa_status [P], depth [P], normal/cpA/cpB [P,3], disp [P,4].
 * Every idA/idB must lie in [0, n_particles); vmax <= 255.
 * force/torque [n_particles,3] receive, per particle, the contact sum taken in
 * ascending pair order (Fn on A, -Fn on B, tauA on A, tauB on B), added once
 * to their current contents. Synchronises `stream` before returning. */
int mwd_contact_pairs(int64_t P, int32_t vmax, const double* comA, const double* comB, const double* vertsA,
                      const double* vertsB, const double* vA, const double* vB, const double* wA, const double* wB,
                      const int32_t* nA, const int32_t* nB, const int32_t* idA, const int32_t* idB,
                      const double* dispt, const double* qA, const double* isA, const double* isB,
                      const mwd_law* law_p, const mwd_law* law_w, int64_t n_walls, int32_t* collision,
                      int32_t* epa_status, double* depth, double* normal, double* cpA, double* cpB, double* disp,
                      double* force, double* torque, int64_t n_particles, void* stream);

/* mwd_contact_pairs without allocation. No counterpart in the reference: it
 * is for callers that cannot allocate device memory (tight serving loops,
 * harnesses that forbid hipMalloc). Same arguments, results and
 * synchronisation as mwd_contact_pairs; the call's temporaries live in a
 * fixed device array of MWD_NOALLOC_SCRATCH_BYTES on each GPU the library
 * runs on (held until the process ends). Limits: P <= MWD_NOALLOC_MAX_PAIRS
 * and n_particles <= MWD_NOALLOC_MAX_PARTICLES, else MWD_E_INVALID (nothing
 * launched). Every call on one GPU uses that GPU's one array: calls that
 * overlap on one GPU (other threads, other streams) race. */
#define MWD_NOALLOC_MAX_PAIRS 65536
#define MWD_NOALLOC_MAX_PARTICLES 262144
#define MWD_NOALLOC_SCRATCH_BYTES (32u << 20)
int mwd_contact_pairs_noalloc(int64_t P, int32_t vmax, const double* comA, const double* comB, const double* vertsA,
                              const double* vertsB, const double* vA, const double* vB, const double* wA,
                              const double* wB, const int32_t* nA, const int32_t* nB, const int32_t* idA,
                              const int32_t* idB, const double* dispt, const double* qA, const double* isA,
                              const double* isB, const mwd_law* law_p, const mwd_law* law_w, int64_t n_walls,
                              int32_t* collision, int32_t* epa_status, double* depth, double* normal, double* cpA,
                              double* cpB, double* disp, double* force, double* torque, int64_t n_particles,
                              void* stream);

/* Cell keys (MWDiem.py:813-816) for rows [0,n): trunc(com/cell) combined as
 * xc + yc*ncx + zc*ncx*ncy, int64. */
int mwd_cell_keys(const double* com, int64_t n, const double* cell, int64_t ncx, int64_t ncy, int64_t* keys,
                  void* stream);

/* Candidate pairs of every scene: wall slab pairs [j,k] (walls k outer,
 * particles j inner) then Linked_cell_hash pairs (cells in order of their
 * lowest id, members in CPython set order, the 7 forward offsets). If keys is
 * NULL they are computed from state->com. Pairs (global rows) go to the
 * workspace; mwd_pairs_view exposes them. */
int mwd_candidates(mwd_workspace* ws, const mwd_state* state, const int64_t* keys);

/* Device pointers of the workspace's candidate list [pair_cap,2] int32 and its
 * count (int64, device). */
int mwd_pairs_view(mwd_workspace* ws, int32_t** pairs, int64_t** n_pairs);

/* Fine contact over the workspace's candidate list, indexed into the state.
 * Per-pair outputs stay in the workspace; mwd_contacts_view exposes them. */
int mwd_fine_contact(mwd_workspace* ws, const mwd_state* state);

/* collision [pair_cap] int32; per contact c < n_contacts: pair index, depth,
 * normal/cpA/cpB [.,3], Fn/tauA/tauB [.,3], epa status. */
int mwd_contacts_view(mwd_workspace* ws, int32_t** collision, int32_t** contact_pair, double** depth,
                      double** normal, double** cpA, double** cpB, double** Fn, double** tauA, double** tauB,
                      int32_t** epa_status, int64_t** n_contacts);

/* Force assembly and Verlet (MWDiem.py:977-1001): per moving particle the
 * contact sum (ascending pair order) plus fext/text, then the integrator.
 * Consumes the contacts of the preceding mwd_fine_contact. */
int mwd_reduce_verlet(mwd_workspace* ws, mwd_state* state);

/* One full step: mwd_candidates + mwd_fine_contact + mwd_reduce_verlet.
 * Stream-ordered, no host synchronisation, capturable in a hipGraph. */
int mwd_step(mwd_workspace* ws, mwd_state* state);

/* Per-stage timing. When enabled, HIP events are recorded on the workspace
 * stream between the stages of mwd_step; mwd_stage_times synchronises and
 * returns the accumulated milliseconds of the MWD_N_STAGES stages and the
 * number of timed steps, then resets the accumulators. The stages: wall slab
 * test (where it does not ride on the key pass); cell keys (with the wall slab
 * tests on the key-table path) and their sort; cell tables; pair emission (one kernel);
 * pair count; bounding-sphere test (one kernel); GJK (one kernel); contact
 * compaction; EPA and contact law (one kernel); force assembly and Verlet
 * (one kernel). */
#define MWD_N_STAGES 10
int mwd_set_timing(mwd_workspace* ws, int enable);
int mwd_stage_times(mwd_workspace* ws, double* ms, int64_t* steps);

/* Kernel-only milliseconds, accumulated like mwd_stage_times over the timed
 * steps, of the MWD_N_KTIMED single-kernel stages (pair emission,
 * bounding-sphere test, GJK, EPA and law, force assembly and Verlet): with
 * timing on they are launched with start/stop events that the runtime stamps
 * at the kernel's own start and end, so launch gaps are excluded. When EPA
 * and the law run inside the GJK launch (epa_fused), the GJK time holds them
 * and the EPA time of that step is 0. Call before mwd_stage_times (which
 * resets both). */
#define MWD_N_KTIMED 5
int mwd_kernel_times(mwd_workspace* ws, double* ms, int64_t* steps);

/* radius[r] = max_k |verts[r,k] - com[r]| * (1 + 1e-9) + 1e-9 * (1 + max|com[r]|)
 * for rows [0, n). */
int mwd_radius_init(const double* com, const double* verts, const int32_t* nverts, int64_t n, int32_t vmax,
                    double* radius, void* stream);

/* The scene generator's hull bodies (replaces the hull construction of
 * Particle_class.Spherical_convex_polygon_maker_3D / Random_convex_polygon_
 * maker_3D, Particle_class.py:184-303, and Calc_inertia, :125-156): for each
 * of n_hulls point sets (points [n_hulls, pmax, 3], npts points each,
 * 4 <= npts <= pmax <= 64), its convex hull's vertices in input order centred
 * on their mean (verts_out [n_hulls, pmax, 3], zero-padded; nverts_out), the
 * mean (com_out [n_hulls, 3]), the volume, the inertia tensor about the mean
 * at `density` (I_out [n_hulls, 3, 3], volume element D.(E x F)/6) and its
 * size trace/3. status_out: 0, 1 (four points on a supporting plane: not
 * triangulated, nverts 0) or 2 (npts out of range). Stream-ordered. */
int mwd_hull_bodies(const double* points, const int32_t* npts, int64_t n_hulls, int32_t pmax, double density,
                    double* verts_out, int32_t* nverts_out, double* com_out, double* vol_out, double* I_out,
                    double* isize_out, int32_t* status_out, void* stream);

/* dst[i] <- src[i] for i < n (n <= MWD_MAX_COPY arrays, bytes[i] a multiple
 * of 8, 8-byte aligned device pointers) in one kernel launch on `stream`.
 * No counterpart in the reference: the reverse sweep uses it to save and
 * restore a step's input state (a dozen arrays) in one launch instead of one
 * per array. */
#define MWD_MAX_COPY 16
int mwd_copy_arrays(int32_t n, void* const* dst, const void* const* src, const int64_t* bytes, void* stream);

/* Counts and flags of the last step (synchronises the workspace stream). */
int mwd_step_status(mwd_workspace* ws, mwd_status* out);
int mwd_clear_flags(mwd_workspace* ws);

/* ---- Tangent (forward-mode) step: derivatives of the state with respect to
 * three law parameters theta_k (normally Kn, en, mu). The reference has no
 * differentiable form; these entry points are additions.
 *
 * d(coefficient)/d(theta_k) of a law, k = 0..MWD_N_TAN-1. */
#define MWD_N_TAN 3
typedef struct mwd_law_tan {
    double Kn[MWD_N_TAN], mu[MWD_N_TAN], Gn[MWD_N_TAN], Kt_num[MWD_N_TAN], inv_mu[MWD_N_TAN];
    double ln_et[MWD_N_TAN], cn[MWD_N_TAN], ct[MWD_N_TAN], ratio[MWD_N_TAN];
} mwd_law_tan;

/* Law coefficient derivatives for (Kn, en, et, mu), given seed[4][MWD_N_TAN]
 * = d(Kn, en, et, mu)/d(theta_k) (row-major). */
int mwd_law_tangent(mwd_law_tan* out, double Kn, double en, double et, double mu, const double* seed);

/* Install the law derivatives of every scene, laws2[2*s] (particle law) and
 * laws2[2*s+1] (wall law), and allocate the per-contact tangent buffers. */
int mwd_set_law_tangents(mwd_workspace* ws, const mwd_law_tan* laws2);

/* The state arrays that carry tangents, in the order of `tan` below. */
#define MWD_TAN_COM 0
#define MWD_TAN_V 1
#define MWD_TAN_V12 2
#define MWD_TAN_W 3
#define MWD_TAN_WBF 4
#define MWD_TAN_W12BF 5
#define MWD_TAN_Q 6
#define MWD_TAN_Q12 7
#define MWD_TAN_QALL 8
#define MWD_TAN_DXALL 9
#define MWD_TAN_VERTS 10
#define MWD_N_TAN_ARRAYS 11

/* One step of mwd_step that also advances tangents: tan[MWD_N_TAN * a + k]
 * is the device array d(state array a)/d(theta_k), laid out like the state
 * array. The values it produces are bit-identical to mwd_step's. Contact
 * topology (which pairs touch, and EPA's final face) is held fixed within the
 * step; depth, normal and contact points are differentiated through that
 * face's support vertices. Requires mwd_set_law_tangents. */
int mwd_step_tangent(mwd_workspace* ws, mwd_state* state, double* const* tan);

/* Reverse-mode adjoint of one step (no counterpart in the reference, which
 * has no gradients; the differentiable form's reverse sweep). state holds
 * the step's input state s_n; lam[MWD_TAN_*] (MWD_N_TAN_ARRAYS device
 * arrays laid out as the state arrays) hold on entry the adjoint of the
 * step's output s_{n+1} and on return that of s_n. dtheta (device, 3
 * doubles) accumulates d(lam . s_{n+1})/d(theta_k) for the parameters whose
 * law-coefficient derivatives mwd_set_law_tangents gave. The step's contact
 * stages rerun on s_n; topology (touching pairs, EPA's final face, capped
 * friction) is held fixed at what the values give. Scenes on the physical
 * law only (MWD_E_INVALID otherwise). Walls receive no adjoint. */
int mwd_step_adjoint(mwd_workspace* ws, mwd_state* state, double* const* lam, double* dtheta);

/* The reverse sweep's tape (no counterpart in the reference). A sweep with
 * checkpoints recomputes each segment's steps before reversing them; with the
 * tape, each recomputed step keeps its input state (the evolving arrays,
 * radii when the state has them, wall centres) and what the adjoint step
 * reads of its contact stages (the contact counts, the per-particle contact
 * records, and per contact its pair, the pair's rows, depth, normal, contact
 * points, force, torques and EPA's final face), so that the adjoint step puts
 * the state back and takes the contacts back instead of rerunning the contact
 * stages. The slot count lives on the device, so a whole segment of recorded
 * steps, and of taped adjoint steps, can be captured in one graph each.
 *
 * mwd_tape_init: room for n_slots steps (one slot: a fixed size from the
 * workspace's rows, vmax and capacities, returned in *slot_bytes when not
 * NULL), allocated on the device with the adjoint step's buffers; n_slots = 0
 * frees it. The tape starts empty. Same conditions as mwd_step_adjoint. */
int mwd_tape_init(mwd_workspace* ws, int64_t n_slots, int64_t* slot_bytes);
/* Empty the tape (on the device, so it can be captured in a graph). */
int mwd_tape_rewind(mwd_workspace* ws);
/* mwd_step, its values bit for bit, also pushing the step's input state and
 * contact-stage outputs onto the tape (MWD_F_TAPE when it is full). */
int mwd_step_record(mwd_workspace* ws, mwd_state* state);
/* mwd_step_adjoint of the step last recorded and not yet taken back: pops
 * its input state into `state` and its contact-stage outputs into the
 * workspace (MWD_F_TAPE when the tape is empty) instead of rerunning the
 * contact stages. The same results as mwd_step_adjoint on that input state,
 * bit for bit. */
int mwd_step_adjoint_taped(mwd_workspace* ws, mwd_state* state, double* const* lam, double* dtheta);

#ifdef __cplusplus
}
#endif
#endif
