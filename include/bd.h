/*
 * bd.h -- C ABI of the MI355X implicit cloth solve (libbd.so).
 *
 * One projected-Newton substep of a woven yarn-crossing lattice: inertial prediction, grip
 * springs, assembly of the clamp-projected Hessian as a radius-2 block stencil, Galerkin
 * multigrid hierarchy, PCG preconditioned by one symmetric V-cycle with block-tridiagonal
 * yarn-line smoothers, line search over {1, 1/2, 1/4, 1/8}, commit, then the yarn damage field
 * (cohesive damage, pull-out, span rupture) and crossing slip.
 *
 * The entry points replace the WebGPU `GpuSolver` of the reference
 * (src/simulation/gpu/solver.ts:194-838); each declaration cites the member it stands for.
 * Plain C types only: the host owns every input array for the duration of a call (the library
 * copies what it keeps to device memory, as queue.writeBuffer does) and readbacks copy into
 * caller memory. Work is enqueued on one HIP stream per solver handle; calls that return data
 * synchronise that stream. A handle is not thread-safe.
 *
 * Every call returns BD_OK (0) or a negative BD_ERR_* code; bd_last_error() gives the message.
 */
#ifndef BD_H
#define BD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BD_OK 0
#define BD_ERR_INVALID (-1)   /* bad argument or inconsistent scene */
#define BD_ERR_HIP (-2)       /* HIP runtime failure (message has the HIP error string) */
#define BD_ERR_LIMIT (-3)     /* size beyond a documented limit (line length, bundle, nodes) */
#define BD_ERR_NOMEM (-4)     /* device allocation failed */
#define BD_ERR_STATE (-5)     /* call order (e.g. a V-cycle before any assembly) */

/* Limits (the reference's own are MAX_LINE 384, 2^20 contact nodes, 3 layers). */
#define BD_MAX_LINE 1024      /* longest yarn line solved in one workgroup's LDS */
#define BD_MAX_LEVELS 12
#define BD_MAX_NEWTON 8
#define BD_MAX_GRIPS 16

/* Self-contact (gpu/selfcontact.ts:7-33 SELF_CONTACT, selfcontact.wgsl:22-30): K candidate faces
 * per node, a spatial hash of 2^17 cells of 12 mm, reach 3 mm, contact distance 1 mm, cubic
 * penalty k = 500 N/m, friction 0.45 regularised below 2 mm/s, bound share 0.45 above a 0.3 mm
 * hard core. Floor contact (floor.ts FLOOR_CONTACT): bulge 0.3 mm, k = 1300 N/m, friction 0.35. */
#define BD_CONTACT_K 4
#define BD_CONTACT_TABLE_BITS 17
#define BD_CONTACT_NONE 0xFFFFFFFFu
/* reciprocal-bound keys are 64-bit (share code << 32 | node + 1); the reference's 32-bit key holds node + 1
 * in 20 bits (gpu/selfcontact.ts:84). The limit here is the lattice index packing (12 bits per i, j). */
#define BD_CONTACT_MAX_NODES (1 << 24)

/* Cross-layer contact of a scene of 2-3 plies (gpu/contact.ts:8-18 CONTACT, contact.wgsl): always on
 * for a scene of several layers, in place of self-contact (gpu/solver.ts:322, 364). K nearest nodes
 * of the other layers per node, a spatial hash of 2^20 (cell, layer) slots of 1 mm cells, quadratic
 * penalty k = 3000 N/m below the pair's contact distance (half the summed layer thicknesses),
 * friction 0.45 regularised below 2 mm/s, bound share 0.45. In that mode BD_BUF_PAIRS holds K N
 * uint32 candidate nodes, BD_BUF_CELLS 2 x 2^20 uint32 and BD_BUF_ITEMS N uint32 node ids. */
#define BD_PLY_TABLE_BITS 20
#define BD_MAX_CONTACT_LAYERS 3  /* SimParams holds 3 layer materials (common.wgsl:24-33) */

/* Yarn family law and constants (src/simulation/material.ts:18-54). */
typedef struct bd_yarn {
  double linear_density;    /* kg/m */
  double ea_low, ea_high;   /* axial tangent stiffness while decrimping / after, N */
  double crimp_strain;
  double ea_compress;       /* compressive stiffness, N */
  double axial_viscosity;   /* N s */
  double bend_rigidity;     /* N m^2 */
  double bend_viscosity;    /* s */
  double strength;          /* intact peak tension, N */
  double softening_opening; /* m */
  int32_t yarns_per_span;   /* parallel yarns per lattice span (1 unbundled) */
  int32_t pad_;
} bd_yarn;

/* Mechanical (bundled) material of a layer (material.ts:56-85, 198-226). */
typedef struct bd_material {
  double pitch_u, pitch_v;          /* rest lengths of u / v spans, m */
  bd_yarn yarns[2];                 /* [0] u family (picks), [1] v family (warps) */
  double shear_stiffness;           /* diagonal pair stiffness, N/m */
  double crossing_friction;
  double crossing_normal0;          /* N, per bundled crossing */
  double crossing_wrap;
  double air_drag;                  /* N s / m^3 */
  double thickness;                 /* apparent thickness for contact (material.ts:76-77), m */
} bd_material;

/* One lattice layer (scene.ts:73-89) and its yarn damage field layout (yarnField.ts:47-66). */
typedef struct bd_layer {
  int32_t nx, ny, node_offset;
  int32_t bundle_u, bundle_v;       /* warps per column, picks per row */
  int32_t field_offset, field_k, field_m;  /* crossings: K = nx*bu warps, M = ny*bv picks */
  double weave_pitch_u, weave_pitch_v;    /* individual-yarn pitches */
  bd_material material;
} bd_layer;

/*
 * Compiled scene (scene.ts:91-125, CompiledScene). Node n = offset + j*nx + i owns span 2n
 * (to i+1, u family) and 2n+1 (to j+1, v family); tails are [u-, u+, v-, v+] per node; the
 * damage field holds per crossing c and family f entry 2c+f.
 */
typedef struct bd_scene {
  int32_t node_count;
  int32_t layer_count;
  const bd_layer* layers;
  const double* position;          /* 3N */
  const double* mass;              /* N */
  const uint8_t* pinned;           /* N */
  const double* rest_length;       /* 2N */
  const double* softening;         /* 2N */
  const uint16_t* seg_flags;       /* 2N: 1 exists, 2 broken, 4 initial defect, 8 slipped */
  const double* intact;            /* 2N */
  const uint32_t* cut_mask;        /* 2N */
  const double* fail_site;         /* 2N */
  const double* crossing_normal0;  /* N */
  const double* tails;             /* 4N */
  int32_t crossing_count;          /* C */
  int32_t pad_;
  const float* field_strength;     /* 2C, peak tension per yarn segment, N */
  const uint8_t* field_damage;     /* 2C, 255 cut, 254 pulled out */
  const uint8_t* field_area;       /* 2C, cross-section code */
} bd_scene;

/* Solver parameters (gpu/solver.ts:26-72 GpuSolverParams; newtonMultigrid.ts:20-36 MGConfig). */
typedef struct bd_params {
  double h;                    /* timestep, s */
  double gravity[3];
  int32_t newton_iterations;   /* always performed */
  int32_t newton_max;          /* gated extra iterations while max |update| > newton_tolerance */
  double newton_tolerance;     /* m */
  int32_t pcg_iterations;      /* fixed-count mode (pcg_rel_tol == 0) */
  int32_t pcg_max_iterations;  /* cap in tolerance mode */
  double pcg_rel_tol;          /* > 0: iterate until ||b - A d|| / ||b|| <= tol (FP64 norms) */
  int32_t coarse_sweeps;       /* line sweeps on the coarsest level */
  int32_t max_coarsest_nodes;  /* coarsen while a level has more nodes */
  int32_t lag_coarse;          /* reuse the Galerkin levels of Newton iteration 0 */
  int32_t line_search;         /* choose alpha in {1, 1/2, 1/4, 1/8} by the potential */
  int32_t history;             /* damage field + crossing slip after commit */
  int32_t precision;           /* BD_PRECISION_* */
  int32_t self_contact;        /* node-face self-contact (GpuSolverParams.selfContact, gpu/solver.ts:49-50) */
  int32_t coarse_solve;        /* BD_COARSE_*: how the V-cycle solves its coarsest level */
  int32_t smooth_sweeps;       /* pre- and post-smoothing sweeps per level of the V-cycle (the reference: 1) */
  int32_t net_level;           /* level of the learned coarse-space network (0: the first with <= 70,000 nodes) */
} bd_params;

#define BD_COARSE_SWEEPS 0     /* coarse_sweeps colour-ordered line sweeps (the reference's V-cycle) */
#define BD_COARSE_EXACT 1      /* exact FP64 solve: banded Cholesky per hierarchy build, dense inverse applied */
#define BD_COARSE_EXACT_LAGGED 2 /* as BD_COARSE_EXACT, but factored once per substep, at its first Newton
                                    iteration: the later iterations' V-cycles (their Galerkin levels rebuilt)
                                    solve the coarsest level with that factorization, an exact solve of the
                                    substep's first coarsest operator. Outside a substep every build factors.
                                    Saves two factorizations in three, but a later iteration whose coarsest
                                    operator moved far from the first's (new contacts) can converge slowly */

#define BD_PRECISION_FP32 0    /* FP32 operators and smoothers, FP64 dot products */
#define BD_PRECISION_FP16 1    /* the V-cycle (smoother gathers, its residuals) reads each Galerkin coarse level's
                                  operator as FP16 scaled by a power of two per level; the fine operator, PCG
                                  and its matvec, the FP64 residual replacement, line search and the exact
                                  coarsest solve keep FP32 / FP64 */

/* Quadratic grip springs (solver.ts:25-37 ReferenceGrip, interact.wgsl:159-181 gripTerms). */
typedef struct bd_grips {
  int32_t grip_count;           /* <= BD_MAX_GRIPS */
  int32_t entry_count;
  const int32_t* node;          /* per entry */
  const int32_t* grip;          /* per entry: owning grip */
  const double* weight;         /* per entry */
  const double* offset;         /* 3 per entry: rest offset from the goal */
  const double* goal;           /* 3 per grip */
  const double* stiffness;      /* per grip, N/m per unit weight */
} bd_grips;

/* Per-substep statistics (gpu/solver.ts:218-225 counters). */
typedef struct bd_stats {
  int32_t newton_iterations;              /* performed in the last substep */
  int32_t pcg_iterations[BD_MAX_NEWTON];  /* per Newton iteration */
  double rel_residual[BD_MAX_NEWTON];     /* final ||b - A d|| / ||b|| (FP64) */
  double alpha[BD_MAX_NEWTON];            /* line-search step */
  double max_update[BD_MAX_NEWTON];       /* largest nodal |update|, m */
  int32_t rejections;                     /* line searches that returned alpha = 0 */
  int32_t ruptures;                       /* spans separated by the damage field (total) */
  int32_t pulled_out;                     /* spans released by crossing slip (total) */
  int32_t level_count;
  int32_t level_nodes[BD_MAX_LEVELS];
  int32_t poll_fallbacks;                 /* PCG convergence checks that took the synchronizing path (total) */
  int32_t learned_fallbacks;              /* learned solves that fell back to plain MG-PCG (total) */
  double warm_relres[BD_MAX_NEWTON];      /* ||b - A d0|| / ||b|| of each Newton solve's start (1: zero start) */
  double max_speed;                       /* largest committed node speed since the last bd_get_stats, m/s
                                             (step.wgsl:145 counters[2]); reading clears it */
} bd_stats;

/* Buffers for bd_read / bd_write. `level` selects the multigrid level where it applies. */
typedef enum bd_buffer {
  BD_BUF_X = 0,        /* positions, 3N doubles (device FP32 widened) */
  BD_BUF_V,            /* velocities, 3N doubles */
  BD_BUF_DX,           /* current step increment, 3N doubles */
  BD_BUF_PRED,         /* inertial prediction, 3N doubles */
  BD_BUF_DELTA,        /* Newton direction of the last PCG, 3N doubles */
  BD_BUF_RHS,          /* Newton rhs -grad, 3N doubles */
  BD_BUF_LEVEL_X,      /* level solution vector, 3 N_l doubles */
  BD_BUF_LEVEL_B,      /* level rhs vector, 3 N_l doubles */
  BD_BUF_LEVEL_R,      /* level residual vector, 3 N_l doubles */
  BD_BUF_STENCIL,      /* level operator expanded to N_l x 25 x 9 doubles (row-major 3x3) */
  BD_BUF_FIXED,        /* level fixed flags, N_l bytes */
  BD_BUF_PROLONG,      /* prolongation weights of level l -> l+1, 4 N_l doubles */
  BD_BUF_REST,         /* span rest lengths, 2N doubles */
  BD_BUF_INTACT,       /* 2N doubles */
  BD_BUF_CUTMASK,      /* 2N uint32 */
  BD_BUF_SEGFLAGS,     /* 2N uint16 */
  BD_BUF_FAILSITE,     /* 2N doubles */
  BD_BUF_TAILS,        /* 4N doubles */
  BD_BUF_FIELD,        /* 2C uint8 damage codes */
  BD_BUF_LAST_UPDATE,  /* N doubles, norm of each node's last Newton update */
  BD_BUF_ENERGY,       /* 9 doubles: energy changes [4], magnitudes [4] and the chosen step length of the last line search */
  BD_BUF_EXT,          /* external terms, 9N floats per node: gradient xyz, Hessian xx xy xz yy yz zz */
  BD_BUF_CMETA,        /* contact bounds, 8N floats per node: own (normal xyz, share), reciprocal (dir xyz, share) */
  BD_BUF_PAIRS,        /* (K + 2) N uint32 per node: K candidate faces (NONE = empty), bound key, row mark */
  BD_BUF_CELLS,        /* 2 x 2^17 uint32: cell start offsets, then fill counts */
  BD_BUF_ITEMS,        /* 2N records x 8 floats: centroid now, radius, centroid at start, face (uint32 bits) */
  BD_BUF_STATE,        /* N uint8 node states: 0 cloth, 1 detached, 2 retired (state.wgsl); writable */
  BD_BUF_RIGIDK,       /* 12N floats per node (detached nodes): external block, rotation stiffness (xx xy xz yy yz zz) */
  BD_BUF_PIECES,       /* 48 doubles per piece: centroid xyz, valid, E^-1 (36, row-major), rigid step t xyz, 0, w xyz, 0 */
  BD_BUF_CURVE,        /* 7N floats per node: its piece's rotation w xyz, flag, arm r xyz (line search) */
  BD_BUF_COUNT_
} bd_buffer;

typedef struct bd_solver bd_solver;

/* Message of the last failing call on `s` (or of bd_create when s is NULL). */
const char* bd_last_error(const bd_solver* s);

/* new GpuSolver(device, scene, params) (gpu/solver.ts:275). Uploads the scene and sizes every
 * level; fails with BD_ERR_LIMIT beyond BD_MAX_LINE or bundles > 32. */
int bd_create(const bd_scene* scene, const bd_params* params, bd_solver** out);
int bd_destroy(bd_solver* s);

/* setState(position, velocity?) (gpu/solver.ts:831-836): 3N doubles each; v may be NULL (zero). */
int bd_set_state(bd_solver* s, const double* x, const double* v);
/* setSegments(spans, tails) (gpu/solver.ts:815-818); any pointer may be NULL (unchanged). */
int bd_set_segments(bd_solver* s, const double* rest_length, const double* intact, const uint32_t* cut_mask,
                    const uint16_t* seg_flags, const double* fail_site, const double* tails);
/* setField(damage) (gpu/solver.ts:821-824): 2C codes. */
int bd_set_field(bd_solver* s, const uint8_t* damage);
/* Grip springs for the following substeps (gripTerms inputs, interact.wgsl:159-181); NULL clears. */
int bd_set_grips(bd_solver* s, const bd_grips* grips);
/* Move the goals of the current grips (3 doubles per grip), the per-substep handle update of
 * Interaction.stepHandles / gripLead (interaction.ts, interact.wgsl:96-157); copies 32 B per grip. */
int bd_set_grip_goals(bd_solver* s, int grip_count, const double* goals);
/* setTimestep (gpu/solver.ts:796-799), setHistoryEnabled (:766-768), setNewtonTolerance (:771-774). */
int bd_set_timestep(bd_solver* s, double h);
int bd_set_history(bd_solver* s, int on);
int bd_set_newton_tolerance(bd_solver* s, double tol);
/* PCG schedule: fixed `iterations` when rel_tol == 0, else to rel_tol with at most max_iterations. */
int bd_set_pcg(bd_solver* s, int iterations, double rel_tol, int max_iterations);

/* encodeSubstep(encoder, minIterations, maxIterations) (gpu/solver.ts:737-760). */
int bd_substep(bd_solver* s, int min_iterations, int max_iterations);

/* Where a Newton system's PCG runs. BD_PCG_GRID (the default): kernels across the whole chip, the
 * host checking convergence per iteration. BD_PCG_WORKGROUP: the whole solve (every iteration's
 * matvec, updates, FP64 true residuals and V-cycle) in one workgroup of one launch, for systems small
 * enough that the grid path is bound by launch latency (a 64 x 64 sheet, the patches of a batch):
 * the same preconditioned CG, with its FP64 sums in another order. It needs one layer, FP32
 * operators, lines of at most 512 nodes, no detached pieces and the learned path off
 * (BD_ERR_STATE with the reason otherwise). No counterpart in the reference. */
#define BD_PCG_GRID 0
#define BD_PCG_WORKGROUP 1
int bd_set_pcg_mode(bd_solver* s, int mode);

/* A substep recorded once and replayed: the reference records a whole substep into one compute pass
 * with its optional Newton iterations gated on the device (gpu/solver.ts:737-760, gate.wgsl:12-23).
 * bd_graph_capture records one bd_substep (the solver's Newton iterations, no optional ones) into a
 * HIP graph; each Newton solve in tolerance mode then runs a fixed budget of pcg_iterations, those
 * past convergence returning at once on the device's done flag instead of the host's per-iteration
 * check (in fixed-count mode the count is the solver's). A solve that needs more than the budget ends
 * unconverged, visible in bd_get_stats (rel_residual above the tolerance). bd_graph_substep replays it
 * on the solver's stream, reading the grip goals and state that the calls before it left on the
 * device. Refused (BD_ERR_STATE) with profiling on, detached pieces, the learned path or sample
 * recording, or newton_max above newton_iterations. A setter that changes what a substep launches
 * (timestep, history, PCG schedule, floor, grips, smoother variants, pieces, learned mode) makes
 * bd_graph_substep return BD_ERR_STATE until the next capture. */
int bd_graph_capture(bd_solver* s, int pcg_iterations);
int bd_graph_substep(bd_solver* s);
int bd_graph_release(bd_solver* s);

/* Batched independent solves (uneven patches, each its own Newton / MG-PCG system with its own
 * scalars and convergence), advanced together by bd_batch_substep on the batch's stream; the batch
 * waits for each solver's queued work (grip goals, uploads) and each solver's later calls wait for
 * it. Two forms, chosen at creation:
 *  - every solver in BD_PCG_WORKGROUP mode (and without optional Newton iterations): the substeps run
 *    in lockstep and each Newton iteration's PCG solves of all solvers are ONE launch, a workgroup
 *    per system, each iterating to its own tolerance;
 *  - otherwise every solver needs bd_graph_capture first: the captured substeps are the branches of
 *    one graph, one launch per frame.
 * Changing a solver's configuration (as above) or capturing it again makes bd_batch_substep return
 * BD_ERR_STATE. No counterpart in the reference, which runs one scene per solver. */
typedef struct bd_batch bd_batch;
int bd_batch_create(bd_solver* const* solvers, int count, bd_batch** out);
/* A third form, for solvers in any configuration: bd_batch_substep runs each solver's substep on the
 * solver's own stream, from `threads` host threads (0: min(count, 8, the host's cores)) that take the
 * solvers largest first, and returns when every substep has been issued (and, for an eager one, its
 * convergence checks have ended). A solver with a captured substep replays it (bd_graph_substep); the
 * others run bd_substep with no extra Newton bounds. The systems' dependent launches then overlap on
 * the chip instead of queueing behind one another. Results are bit-identical to calling
 * bd_graph_substep or bd_substep on each solver in turn. The solvers must not be used from other
 * threads meanwhile; the first failing solver's error is reported (the others still run). */
int bd_batch_create_concurrent(bd_solver* const* solvers, int count, int threads, bd_batch** out);
int bd_batch_substep(bd_batch* b);
int bd_batch_sync(bd_batch* b);
int bd_batch_destroy(bd_batch* b);
const char* bd_batch_last_error(const bd_batch* b);

/* Stage entry points (gpu/solver.ts:574-728 encodePredict, encodeAssemble, encodeVCycle,
 * encodePCG, encodeLineSearch, encodeSmooth, encodeFinalize, encodeDamage). */
int bd_predict(bd_solver* s);
int bd_assemble(bd_solver* s, int coarse);
/* level-0 V-cycle: LEVEL_X(0) = V(LEVEL_B(0)) from a zero start */
int bd_vcycle(bd_solver* s);
/* one colour-ordered line sweep of `level` on LEVEL_X/LEVEL_B of that level */
int bd_smooth(bd_solver* s, int level, int reverse);
/* LEVEL_R(level) = LEVEL_B - A LEVEL_X (fixed rows 0) */
int bd_residual(bd_solver* s, int level);
/* DELTA = PCG solve of the assembled system with RHS */
int bd_pcg(bd_solver* s);
int bd_line_search(bd_solver* s);
int bd_newton_apply(bd_solver* s, int iteration);
int bd_finalize(bd_solver* s);
int bd_damage(bd_solver* s);
int bd_slip(bd_solver* s);

/* SelfContact.encodeDetect (gpu/selfcontact.ts:139-150): hash the intact faces at the start and
 * predicted positions, K nearest faces per node, directional and reciprocal bounds (CMETA). */
int bd_contact_detect(bd_solver* s);
/* SelfContact.encodeTerms (gpu/selfcontact.ts:153-157): penalty and friction terms at the current
 * iterate into EXT (node and face corners), trust-region tightening of the node's own bound. */
int bd_contact_terms(bd_solver* s);
/* projectStart (step.wgsl:85-95): the predicted step made feasible for the bounds and the floor. */
int bd_project_start(bd_solver* s);
/* setFloor(z | null) (gpu/solver.ts:796-801): floor plane at height z when `enabled`, else none. */
int bd_set_floor(bd_solver* s, int enabled, double z);

/* Smoother kernel variants of a level's line passes (bd_autotune, bd_set_smoother). */
#define BD_SMOOTHER_DEFAULT -1 /* direct for lines longer than 257 nodes, else staged */
#define BD_SMOOTHER_DIRECT 0   /* one workgroup per line gathers its rows into registers and solves */
#define BD_SMOOTHER_STAGED4 4  /* rows gathered into LDS, 4 threads per row */
#define BD_SMOOTHER_STAGED8 8  /* rows gathered into LDS, 8 threads per row */
#define BD_SMOOTHER_SPLIT 16   /* a chip-wide row kernel, then one workgroup per line solves (levels
                                  with lines longer than 257 nodes) */
#define BD_SMOOTHER_SPLIT4 17  /* the split variant whose solve holds four rows per thread: two odd-even
                                  levels in registers before the parallel cyclic reduction */
#define BD_SMOOTHER_FACTORED 18 /* the split variant whose solve reuses each line's elimination, factored
                                   once per operator (four rows per thread, then cyclic reduction): a
                                   pass moves only r and x through it */

/* Line colours per family of one level's smoother (the reference uses 3: gpu/solver.ts:93-95; its CPU
 * solver also has 2, newtonMultigrid.ts:25-26). 3: lines solved together are three apart and share no
 * stencil coupling (the default). 2 (the reference CPU solver's option): lines two apart share a colour,
 * their couplings are majorised onto the diagonal and each reads the other's lagged iterate; only the
 * split kernels run it (levels with lines of more than 257 nodes, else BD_ERR_INVALID), and a level's
 * other kernel choices fall back to BD_SMOOTHER_FACTORED. 1 (no counterpart in the reference): all
 * lines of a family are solved together, each against the lagged iterate of the others (line Jacobi
 * across lines; Gauss-Seidel order only between the two families), in a third of the launches, for
 * coarse levels whose passes are bound by launch latency. The V-cycle stays symmetric (the
 * post-smoothing sweep is the pre-smoothing one reversed). BD_ERR_INVALID for other counts;
 * BD_ERR_STATE in BD_PCG_WORKGROUP mode for anything but 3. */
int bd_set_line_colors(bd_solver* s, int level, int colors);

/* Autotuner (no counterpart in the reference): for every level's v-line and u-line passes of the
 * current hierarchy, the smoother kernel variant (BD_SMOOTHER_*) whose V-cycles run fastest, chosen
 * greedily level by level with each variant timed inside whole V-cycles (zero b and x; the level's
 * candidates in turn over `reps` rounds, the fastest run of each counting). chosen receives 2 entries per level (u, v); may be NULL. Results of a pass differ
 * between variants only by FP32 summation order, and for BD_SMOOTHER_SPLIT4 by its elimination order
 * (the same shifted-pivot line solve, rows reduced in another order). */
int bd_autotune(bd_solver* s, int reps, int32_t* chosen);

/* Pins the smoother kernel variant of one level's u-line (family 0) or v-line (family 1) passes
 * (no counterpart in the reference: for tests and callers that tune offline). BD_ERR_INVALID for
 * a variant the level cannot run (its LDS does not fit, or no rows buffer for BD_SMOOTHER_SPLIT);
 * BD_ERR_STATE before the first assembly. A new hierarchy keeps the pinned variants. */
int bd_set_smoother(bd_solver* s, int level, int family, int variant);

/* Grips and picking (src/simulation/gpu/interaction.ts, interact.wgsl): up to BD_MAX_GRIPS_HELD grips
 * hold disjoint material patches of at most BD_GRIP_SLOT_ENTRIES entries, each a quadratic spring from
 * the node to its target (the handle plus its grab-time offset, gathered by the pinch) whose handle is
 * clamped once per substep to within max_lead of the held patch's centroid (bounded total force). The
 * terms enter the Newton system through the external terms (added to contact's). */
#define BD_MAX_GRIPS_HELD 4
#define BD_GRIP_SLOT_ENTRIES 512
#define BD_PICK_MAX_CANDIDATES 128
#define BD_PICK_MAX_WINDOW_RADIUS 7
/* Interaction.attachPatch (interaction.ts:225-262): grip `slot` holds nodes[0..count) (weights, pad
 * sides -1 / +1 or 0 for a radial pinch) with stiffness per unit weight, largest handle lead, pad
 * pitch and slip distance (0: the hold does not slip); offsets and the handle's origin are taken on
 * the GPU from node `picked` now; targets are in the frame of `anchor`. */
int bd_grip_attach(bd_solver* s, int slot, int32_t picked, const int32_t* nodes, const float* weights, const float* sides,
                   int count, double stiffness, double max_lead, double pad_pitch, double slip, const double* anchor);
int bd_grip_lead(bd_solver* s);  /* Interaction.encodeSubstep: the handles' lead clamp (bd_substep runs it) */
int bd_grip_terms(bd_solver* s); /* Interaction.encode: the grip terms into ext (each Newton iteration) */
int bd_grip_target(bd_solver* s, int slot, const double* target); /* setTarget */
int bd_grip_pinch(bd_solver* s, int slot, double scale);          /* setPinch: spread kept by the pinch */
int bd_grip_hold(bd_solver* s, int slot, double fraction);        /* setHold: fraction of the stiffness */
int bd_grip_release(bd_solver* s, int slot);                      /* release */
/* The result of a pick (interaction.ts PickResult): node (-1: none), kind (1 surface hit, 2 proximity),
 * the picked point, and the connectivity window around the node: per lattice offset in [-R, R]^2,
 * row-major from (-R, -R), bit 0 / 1 = its own span along u / v survives, bit 2 = not cloth, bit 3 =
 * outside the lattice. */
typedef struct bd_pick {
  int32_t node;
  int32_t kind;
  double position[3];
  int32_t window_radius;
  uint32_t window[(2 * BD_PICK_MAX_WINDOW_RADIUS + 1) * (2 * BD_PICK_MAX_WINDOW_RADIUS + 1)];
} bd_pick;
/* Interaction.pickAt (interaction.ts:114-144): the nearest surface (intact face, or strand over an
 * opening within strand_radius) along the ray eye + t dir; over none, the nearest crossing within
 * radius_px of the pointer that no intact face hides (view_proj column-major, drawing-buffer pixels). */
int bd_pick_ray(bd_solver* s, const float* view_proj, double px, double py, double width, double height,
                double radius_px, double strand_radius, int window_radius, const double* eye, const double* dir,
                bd_pick* out);
/* Interaction.nearest (interaction.ts:150-166): the crossing with material nearest to p within radius. */
int bd_pick_near(bd_solver* s, const double* p, double radius, int window_radius, bd_pick* out);

/* Detached pieces (src/simulation/gpu/topology.ts, rigid_frame/apply/spin.wgsl). A piece is a set of
 * detached nodes connected through active spans; its six rigid modes are carried by the Newton solve
 * (PCG start from the piece's exact rigid response, spin start of the substep, line search along the
 * piece's rotation). At most BD_MAX_PIECES pieces are carried; more go without the correction. */
#define BD_MAX_PIECES 4096
#define BD_NO_PIECE 0xFFFFFFFFu
/* GpuSolver.setPieces (gpu/solver.ts:664-676): piece_of[n] = piece of node n or BD_NO_PIECE; piece p
 * holds nodes[start[p] .. start[p + 1]); count = 0 removes every piece. */
int bd_set_pieces(bd_solver* s, const uint32_t* piece_of, const uint32_t* start, int count, const uint32_t* nodes);
/* Topology.request (topology.ts:98-160), synchronous: the active span graph is flooded from the
 * mounted nodes and from each anchor (grip patch cores); cloth nodes it does not reach become
 * detached, and the detached nodes are labelled into pieces and handed to bd_set_pieces. *freed
 * receives the number of newly detached nodes (may be NULL). */
int bd_update_topology(bd_solver* s, const int32_t* anchors, int anchor_count, int32_t* freed);
/* The host search alone (no GPU): bits per node (bit 0 / 1 active span of family u / v owned, bit 2
 * not cloth, bit 3 retired), mounted flags and anchors in; freed flags and the pieces out
 * (piece_start holds count + 1 <= N + 1 entries). Returns the number of freed nodes, or < 0. */
int bd_topology_search(int N, int layer_count, const int32_t* nx, const int32_t* ny, const int32_t* offset,
                       const uint8_t* bits, const uint8_t* mounted, const int32_t* anchors, int anchor_count,
                       uint8_t* freed, uint32_t* piece_of, uint32_t* piece_start, uint32_t* piece_nodes,
                       int32_t* piece_count);

/* Per-kernel timing (the reference's timestamp-query profiling, tools/gpu/scripts/profile.ts):
 * while enabled, every launch on the solver's stream is bracketed by HIP events. bd_profile_read
 * synchronises, sums per kernel class since the last read and clears. `bytes` is the algorithmic
 * HBM traffic of the launches (minimal unique bytes by the layout's byte model). */
typedef struct bd_kernel_profile {
  char name[32];
  int32_t launches;
  int32_t pad_;
  double total_ms;
  double bytes;
} bd_kernel_profile;
int bd_profile_enable(bd_solver* s, int on);
int bd_profile_read(bd_solver* s, bd_kernel_profile* out, int max_entries, int* count);

/* Learned coarse-space correction (no counterpart in the reference). A message-passing network on
 * one coarse level of the Galerkin hierarchy (bd_params.net_level) maps that level's operator and
 * the restricted Newton right-hand side to BD_NET_FIELDS displacement fields Z; prolonged to the fine
 * level they give the Newton solve an A-optimal warm start in span(P Z) and an additive coarse
 * correction P Z (Z^T A_l Z)^-1 Z^T P^T r inside flexible PCG. Weights are BF16 on the matrix cores
 * with FP32 accumulation and FP32 master copies. The solve falls back to plain MG-PCG (correction
 * off, zero start, direction restarted) when the fields give a start that is not finite, and when
 * its residual fails to drop below its best for 16 iterations. Training samples come from plain solves: the net
 * level's operator and, as targets,
 * the restricted right-hand side with the Newton direction's energy and the restricted residual left
 * after BD_NET_SLOW_ITER iterations with the remaining error's energy. */
#define BD_NET_FIELDS 8
#define BD_NET_PARAMS 37216
#define BD_NET_SLOW_ITER 4
/* Record up to `capacity` training samples from the following plain Newton solves (0: stop and free). */
int bd_learn_record(bd_solver* s, int capacity);
/* Adam on the GPU over the recorded samples (`batch` samples per step) for the targets in the mask
 * `targets` (1: the Newton directions, the warm start's job; 2: the slow-mode errors, the deflation's
 * job; 3: both); losses[e] receives epoch e's mean loss (the fraction of the targets' energy that
 * span(P Z) misses), may be NULL. */
int bd_learn_train(bd_solver* s, int epochs, double lr, int batch, int targets, double* losses);
/* How Newton solves use the learned fields (BD_LEARNED_*). */
#define BD_LEARNED_OFF 0      /* plain MG-PCG */
#define BD_LEARNED_WARM 1     /* the A-optimal warm start in span(P Z) (G from the net level's Galerkin operator) */
#define BD_LEARNED_DEFLATE 2  /* the warm start, and every preconditioned residual deflated (W = A P Z per system) */
int bd_set_learned(bd_solver* s, int mode);
/* FP32 master weights (count = BD_NET_PARAMS). */
int bd_learn_get_weights(bd_solver* s, float* out, int count);
int bd_learn_set_weights(bd_solver* s, const float* w, int count);
/* Loss and weight gradient of recorded sample `sample` at the current weights (gradient checks). */
int bd_learn_loss_grad(bd_solver* s, int sample, double* loss, float* grad, int count);
/* A recorded sample: targets (2 x 3 planes of the net level's padded size), energies (2), operator
 * (78 planes) and fixed flags; any pointer may be NULL. */
int bd_learn_read_sample(bd_solver* s, int sample, float* b, double* energy, float* op, uint8_t* fixed);
/* The net level, its node count and padded size, and the number of recorded samples. */
int bd_learn_info(bd_solver* s, int* level, int* nodes, int* padded, int* samples);
/* The last prediction's fields, BD_NET_FIELDS x 3 planes of the net level's padded size. */
int bd_learn_read_fields(bd_solver* s, float* z);

/* y = A x with the assembled operator of `level` (0 fine, k > 0 Galerkin level k; fixed rows 0) and
 * x^T A x accumulated in FP64, for callers that run their own Krylov loops on the solver's operators
 * (the reference's callers encode passes over GpuSolver.levels[k].op, gpu/solver.ts:574-728).
 * x and y: device pointers to 3 planes of the level's padded node count (BD_BUF_LEVEL_X layout);
 * xax: a device pointer to one double. Enqueued on `stream` (a hipStream_t; NULL: the solver's
 * stream) and waited for. Allocates a small reduction workspace for the call. */
int bd_apply(bd_solver* s, int level, const float* x, float* y, double* xax, void* stream);
/* bd_apply without any allocation (no counterpart in the reference): for callers that cannot
 * allocate, such as code under HIP graph capture or a tight serving loop. Its reduction scratch is a
 * fixed device array of (256 + 2) doubles per device, set up by bd_create and shared by every call on
 * that device, so calls that overlap on one device race. The scratch does not grow with the level
 * (the reduction grid is capped), so every level fits; results equal bd_apply's bit for bit. */
int bd_apply_noalloc(bd_solver* s, int level, const float* x, float* y, double* xax, void* stream);

int bd_sync(bd_solver* s);
int bd_read(bd_solver* s, bd_buffer buf, int level, void* dst, size_t bytes);
int bd_write(bd_solver* s, bd_buffer buf, int level, const void* src, size_t bytes);
size_t bd_buffer_bytes(const bd_solver* s, bd_buffer buf, int level);
int bd_get_stats(bd_solver* s, bd_stats* out);

#ifdef __cplusplus
}
#endif

#endif /* BD_H */
