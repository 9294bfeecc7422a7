/* avbd.h — C ABI of the MI355X Solver6DOF substep library (libavbd_hip.so).
 *
 * Plain C: pointers, sizes and status codes only. Every entry point returns
 * 0 on success and a negative AVBD_E_* code on failure; avbd_last_error()
 * returns a message for the most recent failure on the calling thread.
 *
 * The Python host (avbd_amd/solver_6dof.py) mirrors the reference
 * Solver6DOF API (src/avbd3d/solver_6dof.py:177-2052) and binds these
 * symbols through ctypes; the reference itself launches Warp kernels
 * (src/avbd3d/kernels_6dof.py) through wp.launch at the call sites cited on
 * each declaration.
 *
 * Threading: one solver handle is driven by one host thread at a time; all
 * device work of a handle is issued on the handle's own HIP stream.
 */
#ifndef AVBD_H_
#define AVBD_H_

#ifdef __cplusplus
extern "C" {
#endif

#define AVBD_OK 0
#define AVBD_E_ARG (-1)        /* invalid argument / shape */
#define AVBD_E_HIP (-2)        /* HIP runtime error */
#define AVBD_E_STATE (-3)      /* call out of order (e.g. step before flush) */
#define AVBD_E_NOMEM (-4)      /* device allocation failed */

/* Coloring modes (solver_6dof.py:597-624). */
#define AVBD_COLOR_JACOBI 0
#define AVBD_COLOR_JONES_PLASSMANN 1

/* Scalars that the reference bakes into each launch
 * (solver_6dof.py:180-201 constructor kwargs, _run_iter_loop :1514-1522). */
typedef struct avbd_params {
  float dt;                    /* substep dt (step() divides by substeps) */
  int iterations;
  float gravity[3];
  float alpha, beta, gamma;
  int post_stabilize;
  float max_linear_speed;      /* +inf disables the cap */
  float max_angular_speed;     /* +inf disables the cap */
  float friction_static_mult;
  int coloring_mode;           /* AVBD_COLOR_* */
  int self_collide;            /* enable_self_collision (solver_6dof.py:570) */
  float self_friction;         /* default_friction for box-box contacts */
  int fixed_capacity;          /* unsafe_fixed_capacity semantics */
  float margin;                /* contact margin, 0.005 (solver_6dof.py:1789) */
} avbd_params;

/* Host-side scene handed over at _flush (solver_6dof.py:904-1174).
 * Body arrays have n_bodies entries (vec3 = 3 floats, quat = 4 floats xyzw,
 * mat33 = 9 floats row-major). Row arrays hold the n_static_rows static rows
 * built by add_floor_contact_box / add_pin_corner. */
typedef struct avbd_scene {
  int n_bodies;
  const float* x;
  const float* q;
  const float* v;
  const float* w;
  const float* prev_v;
  const float* prev_w;
  const float* mass;
  const float* inertia_local;  /* 9 per body */
  const float* half_extents;
  const float* friction;
  const float* priority;       /* default_rng(0xA5BD).uniform -> f32 */

  int n_static_rows;
  const int* type;
  const int* body_a;
  const int* body_b;
  const int* sibling;
  const int* partner;
  const float* anchor;         /* 3 per row */
  const float* off_a;          /* 3 per row */
  const float* off_b;          /* 3 per row */
  const float* stiffness;
  const float* fmin;
  const float* fmax;
  const float* fracture;
  const float* friction_d;
  const float* friction_s;
  const float* lambda;         /* initial row state */
  const float* penalty;
  const int* active;
  const int* was_static;

  /* Capacities sized by the host with the reference formulas
   * (solver_6dof.py:937-943, 1153, 1158, 1811). */
  int max_pairs;
  int n_dyn_capacity;          /* rows */
  int hash_cap;
  int pool_max;
} avbd_scene;

/* What one _step_one did; the host turns these into the reference's
 * warnings and attribute updates (solver_6dof.py:1296-1356). */
typedef struct avbd_report {
  int recolored;               /* 1 if the coloring was rebuilt */
  int n_active_colors;         /* achieved colors, clamped to [1, 32] */
  int n_active_rows;           /* raw row reservation counter */
  int row_capacity;            /* capacity before any regrow this substep */
  int row_overflow;            /* 1 if raw > capacity (pool regrown) */
  int new_row_capacity;
  int pair_count;              /* broadphase pair counter */
  int pair_capacity;
  int pair_grew;               /* pair buffers regrown this substep */
  int pool_count;
  int n_contacts;              /* dynamic contact groups emitted */
  int color_conflicts;         /* conflict rows seen by the recolor gate */
  float broadphase_ms;         /* host wall time of contact generation */
} avbd_report;

const char* avbd_last_error(void);
int avbd_device_count(int* out);

/* ---- lifecycle ------------------------------------------------------ */
typedef struct avbd_solver avbd_solver;
/* Solver6DOF.__init__ (solver_6dof.py:180-442). */
int avbd_create(int device, avbd_solver** out);
int avbd_destroy(avbd_solver* s);
/* Solver6DOF._flush (solver_6dof.py:904-1174): upload bodies and static rows,
 * size every pool. Replaces the previous scene of the handle. */
int avbd_flush(avbd_solver* s, const avbd_scene* scene);

/* ---- the substep ---------------------------------------------------- */
/* Solver6DOF._step_one (solver_6dof.py:1232-1464): reset counters, contact
 * generation (:1277-1309), recolor gate (:1333-1356), body->constraint CSR
 * (:1366-1385), predict + prelude (:1388-1417), iteration loop
 * (_run_iter_loop :1507-1679) and warm-start cache refresh (:1450-1464).
 * color_dirty: the host's _color_dirty flag. */
int avbd_substep(avbd_solver* s, const avbd_params* p, int color_dirty, avbd_report* out);
/* The substep run to a residual instead of a fixed iteration count; no
 * counterpart in the reference (the residual measurement of the learned
 * warm-start study). After each dual sweep the residual
 * r = max over active rows of |lambda_new - lambda| / k (metres; lambda and
 * k before the update) is read back; the substep stops iterating at the
 * first sweep with r <= tol, or at p->iterations. The result equals
 * avbd_substep with iterations = *iterations_used. */
int avbd_substep_to_residual(avbd_solver* s, const avbd_params* p, int color_dirty, float tol,
                             int* iterations_used, float* residual, avbd_report* out);

/* Solver6DOF.step (solver_6dof.py:1178-1230): n_substeps substeps of
 * p->dt each. With fixed_capacity set and AVBD_OPT_GRAPH on, every substep
 * after the first one of a dirty coloring keeps its counts and decisions on
 * the device (launches sized by capacity, the recolor gated on the device)
 * and runs as a captured HIP graph, replayed while the scene layout and the
 * parameters stay the same; the counters are read back once, at the end of
 * the frame (the reference's once-per-frame overflow check, :1199-1230).
 * Results are bit-identical to n_substeps avbd_substep calls. *out reports
 * the frame: counts of its last substep, recolored = any substep recolored. */
int avbd_step_frame(avbd_solver* s, const avbd_params* p, int n_substeps, int color_dirty,
                    avbd_report* out);
/* Device -> host synchronisations of the solver's counters since creation
 * (readbacks inside substeps and at frame ends; reads requested through the
 * read entry points are not counted) and substep graph launches. No
 * counterpart in the reference (its tests count .numpy() calls,
 * tests/test_robustness.py:335-395). */
int avbd_sync_count(avbd_solver* s, long long* syncs, long long* graph_launches);

/* ---- one substep, stage by stage -------------------------------------------
 * The stages of avbd_substep as separate calls, in the reference's launch
 * order (solver_6dof.py:1269-1464): contact generation (:1277-1309),
 * coloring (:1333-1356), body->constraint lists (:1360-1366), prelude
 * (predict_inertial_6dof + substep_prelude_6dof + cache_alpha_C0_6dof,
 * :1368-1394), then sweeps 0 .. iterations + post_stabilize - 1 (each: the
 * primal sweep over every color, :1397-1424, the dual update after sweeps
 * < iterations, :1425-1432, and the finalize after the last dual, :1434-1440),
 * then the end (contact cache refresh, :1442-1464). Called in that order with
 * the same params they give the bits of one avbd_substep; out of order they
 * return AVBD_E_STATE and the substep must start again from the contacts.
 * The coloring stage adds its fields to the report of the contacts stage. */
int avbd_stage_contacts(avbd_solver* s, const avbd_params* p, avbd_report* out);
int avbd_stage_coloring(avbd_solver* s, const avbd_params* p, int color_dirty, avbd_report* out);
int avbd_stage_csr(avbd_solver* s);
int avbd_stage_prelude(avbd_solver* s, const avbd_params* p);
int avbd_stage_sweep(avbd_solver* s, const avbd_params* p, int sweep);
int avbd_stage_end(avbd_solver* s, const avbd_params* p);

/* ---- streams and device memory ---------------------------------------------
 * Every launch and copy of the solver goes to one HIP stream: its own
 * (created with the solver) or one the caller sets (a hipStream_t, NULL to go
 * back to the solver's own). Setting waits for the work queued on the
 * previous stream. The caller keeps a stream it set alive while the solver
 * uses it. The reference's arrays live on its Warp device
 * (solver_6dof.py:904-1174) and its launches go to Warp's current stream. */
int avbd_get_stream(avbd_solver* s, void** stream);
int avbd_set_stream(avbd_solver* s, void* stream);

/* Device addresses of the live state (solver_6dof.py:213-225: x, q, v, omega,
 * c_lambda, c_penalty, c_active), valid until the next substep or flush
 * (the pose buffer alternates between substeps). Element strides in 4-byte
 * words:
 *   body i:   x[i*body_stride + 0..2], q[i*body_stride + 0..3] (xyzw),
 *             v[i*body_stride + 0..2], w[i*body_stride + 0..2];
 *   group g:  lambda[g*group_stride + r], penalty[g*group_stride + r] for its
 *             rows r < nrows; state_bits[g*group_stride]: bit r = row r
 *             active, bit 3 + r = row r static (sticking);
 *             header[g*header_stride + 0..3] = {kind | (nrows - 1) << 2 |
 *             ..., body_a, body_b, first row}: row j of the reference's flat
 *             row arrays is (g, r) with j = first row + r; kind 0 floor,
 *             1 box-box, 2 pin.
 * Groups: the static rows' groups first, then this substep's contacts. */
typedef struct avbd_views {
  int n_bodies;
  float* x;
  float* q;
  float* v;
  float* w;
  int body_stride;
  int n_groups;
  float* lambda;
  float* penalty;
  int* state_bits;
  int group_stride;
  int* header;
  int header_stride;
} avbd_views;
int avbd_device_views(avbd_solver* s, avbd_views* out);

/* ---- readback / write-back ------------------------------------------ */
/* positions()/orientations()/velocities()/angular_velocities()
 * (solver_6dof.py:1920-1938); any pointer may be NULL. */
int avbd_read_bodies(avbd_solver* s, float* x, float* q, float* v, float* w,
                     float* prev_v, float* prev_w);
/* set_position/orientation/velocity/angular_velocity (solver_6dof.py:878-900):
 * field 0 = x, 1 = q, 2 = v, 3 = w, 4 = prev_v, 5 = prev_w; writes n entries
 * starting at body `first`, in place. */
int avbd_write_bodies(avbd_solver* s, int field, int first, int n, const float* values);
/* lambdas()/active() and the c_* row arrays (solver_6dof.py:1940-1948):
 * per-row values for rows [0, n_rows); rows past the live count read as 0.
 * Any pointer may be NULL. */
int avbd_read_rows(avbd_solver* s, int n_rows, float* lambda, float* penalty, int* active,
                   int* was_static, int* type, int* body_a, int* body_b, float* alpha_c0);
/* Write static-row state (lambda, penalty, was_static, active) for rows
 * [0, n) -- used to load a snapshot. */
int avbd_write_static_rows(avbd_solver* s, int n, const float* lambda, const float* penalty,
                           const int* active, const int* was_static);
/* body_color (solver_6dof.py:1011) */
int avbd_read_colors(avbd_solver* s, int* body_color);
int avbd_write_colors(avbd_solver* s, const int* body_color, int n_active_colors);
/* n_active_rows / _bp_pair_count / pool_count device counters. */
int avbd_read_counters(avbd_solver* s, int* n_active_rows, int* pair_count, int* pool_count);
/* count_color_conflicts (solver_6dof.py:846-862 -> kernels_6dof.py:3313-3336). */
int avbd_count_color_conflicts(avbd_solver* s, int* out);
/* read_state_batched packing (kernels_6dof.py:2154-2177, 2540-2555):
 * bodies -> 10 floats [x, q, w]; rows [0, n_rows) -> 3 floats
 * [lambda, active, was_static*16 + type]. */
int avbd_pack_state(avbd_solver* s, float* bodies10, int n_rows, float* rows3);

/* Warm-start cache contents (hash_keys / hash_state, solver_6dof.py:1150-1156):
 * export writes up to `cap` occupied entries (key + 8 floats each) and sets
 * *n; import clears the table and inserts n entries. */
int avbd_export_hash(avbd_solver* s, int cap, int* keys, float* state8, int* n);
int avbd_import_hash(avbd_solver* s, int n, const int* keys, const float* state8);
/* Capacities currently allocated: max_pairs, row capacity, pool_max, hash_cap. */
int avbd_read_capacities(avbd_solver* s, int* max_pairs, int* row_cap, int* pool_max, int* hash_cap);

/* ---- per-stage views for tests ---------------------------------------- */
/* Broadphase pairs of the last substep in canonical (a asc, b asc) order
 * (bvh_broadphase_pairs, kernels_6dof.py:1946-1978); copies min(count, cap). */
int avbd_read_pairs(avbd_solver* s, int cap, int* a, int* b, int* overlap, int* sat_idx,
                    float* n_hat, float* depth, int* count);
/* Contact manifold of the last substep (obb_contact_manifold_6dof,
 * kernels_6dof.py:2221-2537): per pair count/ref_is_a/n/t/b and 4 offsets. */
int avbd_read_manifold(avbd_solver* s, int cap, int* count, int* ref_is_a, float* n_hat,
                       float* t_hat, float* b_hat, float* off_ref, float* off_inc);

/* ---- options -------------------------------------------------------- */
/* Kernel selection; no counterpart in the reference (its primal_group_size /
 * primal_shuffle / primal_fused switches, solver_6dof.py:196-198, choose
 * between its CUDA kernels). AVBD_OPT_PRIMAL_VARIANT: 0 = one lane per body,
 * 1, 2, 3, 4 = one wavefront per 1, 2, 4, 8 bodies, 5 = one wavefront per
 * body for colors under 4096 bodies and per two bodies otherwise, 6 = every
 * color of a sweep in one launch, each two-body tile starting as soon as its
 * lower-color neighbours have moved (default), 7, 8, 9, 10 = the same with
 * one, four, eight, three bodies per tile. All variants give identical results; few
 * contacts per body favour more bodies per tile (Solver6DOF.autotune_primal
 * picks by timing). */
#define AVBD_OPT_PRIMAL_VARIANT 0
/* AVBD_OPT_FUSE_DUAL: 1 = with the dataflow sweep (variants 6-10) and a
 * coloring without same-color neighbours, each contact group's dual update
 * runs inside the sweep, in the tile of its higher-color body; 2 = with the
 * dataflow sweep, the dual runs in trailing workgroups of the sweep's launch,
 * each group once both of its bodies have moved (any coloring); 0 (default)
 * = a separate dual launch after every sweep. All give identical results;
 * the residual mode always uses the separate launch. */
#define AVBD_OPT_FUSE_DUAL 1
/* AVBD_OPT_LEARNED_SCOPE (learned warm start, mode 2): 0 (default) = every
 * box-box and floor group, 1 = only groups whose plain start is empty
 * (normal lambda 0: new contacts, floor corners not pressed last substep);
 * groups with a cached lambda keep the plain start. */
#define AVBD_OPT_LEARNED_SCOPE 2
/* AVBD_OPT_FLOW_GIVE_UP_TEST (tests only; no counterpart in the reference):
 * value N > 0 makes every wait of the dataflow sweep poll an array that is
 * never written and give up after N polls, as a wait that does not end
 * would at its cap; 0 (default) restores normal waits. A flush resets it.
 * The give-up sets the solver's error flag: the next sync (substep
 * readback, avbd_read_bodies, avbd_read_rows) returns AVBD_E_STATE, and the
 * flag stays set until the next avbd_flush. */
#define AVBD_OPT_FLOW_GIVE_UP_TEST 3
/* AVBD_OPT_GRAPH: 1 = avbd_step_frame runs fixed-capacity substeps as a
 * captured HIP graph (see avbd_step_frame); 0 (default) = direct launches
 * with the per-substep readbacks. No counterpart in the reference's API (its
 * CUDA path captures the iteration loop, solver_6dof.py:1419-1447). */
#define AVBD_OPT_GRAPH 4
/* AVBD_OPT_BROAD_LANES: lanes per body in the broad phase's tree traversal
 * (the reference's wp.bvh_query_aabb loop, kernels_6dof.py:1946-1978, walks
 * one thread per body): 1 = one lane per body; 8, 16, 64 = a group of that
 * many lanes sharing a stack, each lane testing one node per round; 0
 * (default) = 64 up to 16,384 bodies, 16 up to 65,536, 8 up to 262,144, 1
 * above. All give identical pair lists. */
#define AVBD_OPT_BROAD_LANES 5
int avbd_set_option(avbd_solver* s, int option, int value);

/* ---- profiling ------------------------------------------------------- */
/* When enabled, the primal, dual and prelude launches are bracketed by HIP
 * events on the solver stream; avbd_profile_read returns accumulated
 * milliseconds, launch counts and algorithmic bytes per kernel class
 * (0 = primal sweep, 1 = dual sweep, 2 = substep prelude, 3 = contacts,
 * 4 = coloring, 5 = csr, 6 = finalize/hash). */
int avbd_profile_enable(avbd_solver* s, int on);
int avbd_profile_read(avbd_solver* s, int n_classes, double* ms, long long* launches,
                      double* bytes);

/* ---- batched scenes --------------------------------------------------
 * No counterpart in the reference (one Solver6DOF per scene there). After a
 * flush of the concatenation of independent scenes (bodies of scene k are
 * [first_body[k], first_body[k+1]); static rows refer to their own bodies),
 * this splits it: the broad phase pairs only bodies of one scene (its tree
 * is built over each scene's boxes moved to its own cell of a grid of pitch
 * `pitch` metres; the pitch only affects speed), and the recolor gate
 * recolors only the scenes with a color conflict. Each scene then steps
 * exactly as it would alone. A flush undoes the split. */
int avbd_set_scenes(avbd_solver* s, int n_scenes, const int* first_body, float pitch);

/* ---- learned warm start ----------------------------------------------
 * No counterpart in the reference: its warm start is the persistent-lambda
 * cache alone (kernels_6dof.py:2695-2912), which stays the fallback. A
 * 24 -> 64 -> 64 -> 4 MLP (SiLU, bf16 weights on MFMA, fp32 accumulation)
 * predicts, per box-box contact group, corrections to the plain warm start
 * of lambda (n, t, b) and k_n; it changes only the initial dual state of a
 * substep, never the iterations after it.
 * Parameters (avbd_mlp_num_params() floats): W1[24][64], b1[64], W2[64][64],
 * b2[64], W3[64][4], b3[4], x_mean[24], x_std[24], y_mean[4], y_std[4].
 * Sample layout: 24 features, 4 targets and one output mask per contact
 * group (bit o set = output o is used; tangents only with friction rows). */
int avbd_mlp_num_params(void);
/* mode 0 = off (plain warm start), 1 = log samples (features at emission,
 * targets at the end of each substep), 2 = apply the network (needs
 * avbd_ml_set_params first). Modes 1 and 2 also track contact ages. */
int avbd_ml_set_mode(avbd_solver* s, int mode);
int avbd_ml_set_params(avbd_solver* s, const float* params, int n);
/* The last substep's logged samples (mode 1): copies min(cap, *n) of them;
 * *n = how many there are. feat [*n][24], targ [*n][4], mask [*n]. */
int avbd_ml_read_samples(avbd_solver* s, int cap, float* feat, float* targ, int* mask, int* n);
/* Standalone network calls on host arrays (device `device`):
 * forward: y [n][4] = de-normalised outputs for x [n][24];
 * grad: gradient of the masked loss (normalised targets, summed over used
 *   outputs, averaged over samples; squared error, or the Huber loss with
 *   threshold `huber` when huber > 0) w.r.t. the first
 *   avbd_mlp_num_params() - 56 parameters (the weights; the normalisation
 *   is fixed), and the loss;
 * train: `steps` Adam steps (lr, betas 0.9 / 0.999) on minibatches of
 *   `batch` samples drawn from seeded shuffles; params updated in place,
 *   loss_hist [steps] = each minibatch's loss (may be NULL). */
int avbd_mlp_forward(int device, int n, const float* x, const float* params, float* y);
int avbd_mlp_grad(int device, int n, const float* x, const float* y, const int* mask,
                  const float* params, float huber, float* grad, float* loss);
int avbd_mlp_train(int device, int n, const float* x, const float* y, const int* mask,
                   float* params, int steps, int batch, float lr, float huber, unsigned seed,
                   float* loss_hist);
/* avbd_mlp_grad without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * that must not call hipMalloc, such as tight serving loops and harnesses
 * that forbid allocation). Same arguments and results as avbd_mlp_grad, for
 * n <= 4096 samples (more returns AVBD_E_ARG). The samples, parameters,
 * gradient, loss and per-workgroup gradient rows are staged in one static
 * device array of 2,638,080 bytes (2.52 MiB),
 * which every GPU the library runs on holds until the process ends; every
 * call on a GPU shares that GPU's array, so calls that overlap on one GPU
 * race. Runs on the null stream. */
int avbd_mlp_grad_noalloc(int device, int n, const float* x, const float* y, const int* mask,
                          const float* params, float huber, float* grad, float* loss);

/* ---- 3-DOF particle solver ---------------------------------------------
 * Solver.step (src/avbd3d/solver.py:1124-1400): particles with pin (x/y/z),
 * distance, floor (with x/z friction rows), plane and tet-volume
 * constraints, and the body-body contact rows the host's broad phase
 * emits (sphere, sphere-box, box-box, each with two friction rows). The host (avbd_amd/solver.py, after
 * Solver._flush :964-1122) supplies the rows in the reference's per-row
 * layout, the body coloring (coloring.py:41-55) and the body -> row and
 * body -> tet lists (solver.py:922-962). Vectors are 3 floats. */
typedef struct avbd_p_params {
  float dt;
  int iterations;
  float gravity[3];
  float alpha, beta, gamma;
  int post_stabilize;
  float max_speed;             /* <= 0 or +inf: no cap */
} avbd_p_params;

typedef struct avbd_p_scene {
  int n_bodies;
  const float* x;
  const float* v;
  const float* prev_v;
  const float* mass;
  const int* color;            /* per body, < n_colors */
  int n_colors;
  int n_rows;
  const int* type;             /* 0-2 pin x/y/z, 3 distance, 4 floor, 5 sphere, 6 tangent,
                                  7 sphere-box, 8 box-box, 10 plane */
  const int* body_a;
  const int* body_b;
  const float* anchor;         /* 3 per row */
  const float* rest;
  const float* stiffness;
  const float* fmin;
  const float* fmax;
  const float* fracture;
  const int* sibling;
  const float* friction;
  const float* lambda;
  const float* penalty;
  const int* active;
  const int* con_start;        /* n_bodies + 1 */
  const int* con_idx;
  int n_tets;
  const int* tet_v;            /* 4 per tet */
  const float* tet_rest_vol;
  const float* tet_stiffness;
  const float* tet_fracture;
  const float* tet_lambda;
  const float* tet_penalty;
  const int* tet_active;
  const int* tet_start;        /* n_bodies + 1 */
  const int* tet_idx;          /* (tet << 2) | vertex slot */
  const float* off_a;          /* 3 per row: body_a's contact offset (5, 7, 8), or NULL */
  const float* off_b;          /* 3 per row: body_b's contact offset (5, 8), or NULL */
} avbd_p_scene;

typedef struct avbd_p_solver avbd_p_solver;
const char* avbd_p_last_error(void);
int avbd_p_create(int device, avbd_p_solver** out);
int avbd_p_destroy(avbd_p_solver* s);
/* Solver._flush: replaces the device state with the scene's. */
int avbd_p_flush(avbd_p_solver* s, const avbd_p_scene* scene);
/* One Solver.step: predict, warm start, alpha C0, iterations of per-color
 * primal launches and the dual, velocity finalize. */
int avbd_p_step(avbd_p_solver* s, const avbd_p_params* p);
/* positions()/velocities()/lambdas()/active(); any pointer may be NULL. */
int avbd_p_read(avbd_p_solver* s, float* x, float* v, float* prev_v, float* lambda,
                float* penalty, int* active, float* tet_lambda, float* tet_penalty,
                int* tet_active);
/* set_position / set_velocity (solver.py:899-920): field 0 = x, 1 = v. */
int avbd_p_write_bodies(avbd_p_solver* s, int field, int first, int n, const float* values);
/* set_plane_contact_bulk (solver.py:868-897): rows [first, first+n); anchor
 * (3 per row) and rest may be NULL. */
int avbd_p_write_rows(avbd_p_solver* s, int first, int n, const float* anchor, const float* rest);
/* No counterpart in the reference (Warp launches each kernel from Python).
 * AVBD_P_OPT_GRAPH: 1 (default) captures a step once as a HIP graph and
 * replays it while the flushed buffers and the parameters stay the same;
 * 0 launches every kernel directly.
 * AVBD_P_OPT_PRIMAL_VARIANT: 0 one lane per body; 1, 2, 3: 16, 32, 64
 * lanes per body evaluating its rows and tets in parallel, one launch per
 * color; 4 (default), 5: 32, 16 lanes per body, one dataflow launch per
 * sweep (a body waits for its lower-color neighbours' per-sweep stamps).
 * Results are bit-identical for every setting. */
#define AVBD_P_OPT_GRAPH 0
#define AVBD_P_OPT_PRIMAL_VARIANT 1
int avbd_p_set_option(avbd_p_solver* s, int option, int value);
/* No counterpart in the reference: per-launch timing with HIP events, by
 * class (0 primal, 1 dual and tet dual, 2 the rest, untimed). Enabling
 * resets the totals and launches directly (no graph) while on. */
int avbd_p_profile_enable(avbd_p_solver* s, int on);
/* ms, launches, algorithmic bytes per class (3 entries each; NULL skips). */
int avbd_p_profile_read(avbd_p_solver* s, double* ms, long long* launches, double* bytes);

#ifdef __cplusplus
}
#endif
#endif /* AVBD_H_ */
