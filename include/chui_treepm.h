/*
 * chui_treepm.h — C ABI of the MI355X TreePM gravity library (libchui_treepm.so).
 *
 * Drop-in boundary for CHUI's periodic TreePM force evaluation. Every entry
 * point takes plain pointers and sizes, returns a chui_status and never lets a
 * C++ exception cross the ABI. The message of the last failure on the calling
 * thread is available from chui_treepm_last_error().
 *
 * Interfaces replaced (reference file:line):
 *   chui_treepm_create / chui_treepm_destroy
 *       gravity::TreePmCoordinator::TreePmCoordinator(PmGridShape)
 *       include/cosmosim/gravity/tree_pm_coupling.hpp:545-551
 *   chui_treepm_solve / chui_treepm_solve_device
 *       TreePmCoordinator::solveActiveSetWithPmCadence(...)
 *       include/cosmosim/gravity/tree_pm_coupling.hpp:563-573
 *       (solveActiveSet, :575-584, is the same call with refresh = 1)
 *   chui_treepm_solve_adaptive_device / chui_treepm_default_adaptive_options
 *       the same call under TreePmAcceptancePolicy::kAdaptiveRelative with the
 *       accumulator's previous_acceleration_magnitude_code lane and the group
 *       walk options (tree_pm_coupling.hpp:127-165, 243-276)
 *   chui_treepm_options_*  TreePmOptions / PmSolveOptions / TreeGravityOptions /
 *       TreePmSplitPolicy (tree_pm_coupling.hpp:243-332, pm_solver.hpp:223-254,
 *       tree_gravity.hpp:30-43, tree_pm_split_kernel.hpp:20-27) and the
 *       [numerics] .param.txt keys mapped in src/workflows/gravity_runtime.cpp:486-590
 *   chui_treepm_kick_device / chui_treepm_drift_device
 *       velocity kick src/workflows/gravity_runtime.cpp:3045-3058 and drift +
 *       canonical wrap src/workflows/reference_runtime_composition.cpp:59-108
 *   chui_treepm_kdk_factors
 *       CosmologicalTimeline::prepareStep src/core/time_integration.cpp:407-464
 *   chui_treepm_split_diagnostics
 *       computeTreePmDiagnostics src/gravity/tree_pm_coupling.cpp:6276-6312
 *   chui_treepm_tree_* (read-back of the last built tree)
 *       TreeGravitySolver::nodes()/ordering() include/cosmosim/gravity/tree_gravity.hpp:93-129,186-190
 *   chui_treepm_morton_device
 *       buildMortonOrderingInPlace src/gravity/tree_ordering.cpp:128-173
 *   chui_treepm_unwrap_device
 *       unwrapPeriodicAxis src/gravity/tree_pm_coupling.cpp:141-218
 *   chui_treepm_ewald_device (FP64 periodic Ewald sum on sample targets)
 *       tests/support/periodic_ewald_reference.cpp:202-400 (test-only in the reference)
 */
#ifndef CHUI_TREEPM_H
#define CHUI_TREEPM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CHUI_TREEPM_ABI_VERSION 1

/* Status codes. The C++ adapter maps them back onto the reference's exception
 * classes (tree_pm_coupling.cpp / tree_index.hpp:21-28). */
typedef enum chui_status {
  CHUI_OK = 0,
  CHUI_ERR_INVALID_ARGUMENT = 1, /* std::invalid_argument */
  CHUI_ERR_RUNTIME = 2,          /* std::runtime_error */
  CHUI_ERR_OVERFLOW = 3,         /* std::overflow_error (32-bit index / depth limits) */
  CHUI_ERR_LOGIC = 4,            /* std::logic_error (internal invariant) */
  CHUI_ERR_DEVICE = 5,           /* HIP runtime failure */
  CHUI_ERR_OUT_OF_MEMORY = 6,    /* device allocation failed */
  CHUI_ERR_UNSUPPORTED = 7       /* option accepted by the reference, not built here yet */
} chui_status;

typedef enum chui_precision {
  CHUI_PRECISION_FP64 = 0,  /* FP64 everywhere (walk, FFT, deposit, interpolation) */
  CHUI_PRECISION_MIXED = 1  /* 32-bit fixed-point positions, FP32 walk/FFT/interp with
                               compensated accumulation; FP64 drift/kick/reductions */
} chui_precision;

typedef enum chui_assignment {
  CHUI_ASSIGN_CIC = 0,
  CHUI_ASSIGN_TSC = 1
} chui_assignment;

typedef enum chui_opening_criterion {
  CHUI_OPEN_GEOMETRIC = 0,     /* kBarnesHutGeometric */
  CHUI_OPEN_COM_DISTANCE = 1,  /* kBarnesHutComDistance (reference default) */
  CHUI_OPEN_RELATIVE = 2       /* kRelativeForceError */
} chui_opening_criterion;

typedef enum chui_acceptance_policy {
  CHUI_ACCEPT_STRICT = 0,   /* kStrictReference (reference default) */
  CHUI_ACCEPT_ADAPTIVE = 1  /* kAdaptiveRelative */
} chui_acceptance_policy;

/* How the strict walks map targets onto 64-wide waves (this library; no
 * reference counterpart: the reference walks targets one by one). The forces
 * are the same pairs and multipoles either way, summed in another order.
 * Walks with counters or (FP64) the nearest-pair lane, and the adaptive
 * walks, always run one target per lane. */
typedef enum chui_target_mapping {
  CHUI_TARGETS_AUTO = 0,      /* per wave when at most a fraction min(0.25, f0 (2^24/n)^0.75)
                                 of the n sources is active (f0 = 0.025 mixed, 0.03 FP64),
                                 else per lane */
  CHUI_TARGETS_PER_LANE = 1,  /* 64 visit-ordered targets per wave sharing one walk */
  CHUI_TARGETS_PER_WAVE = 2   /* one target per wave: lanes split its nodes and pairs */
} chui_target_mapping;

/* Flat options: the fields of TreePmOptions that change a force evaluation.
 * Defaults (chui_treepm_default_options) equal the reference's struct defaults. */
typedef struct chui_treepm_options {
  uint32_t struct_size;              /* = sizeof(chui_treepm_options) */
  /* PmSolveOptions */
  double box_size[3];                /* comoving axis lengths (Lx, Ly, Lz) */
  double scale_factor;
  double gravitational_constant;     /* G in code units */
  int32_t assignment_scheme;         /* chui_assignment */
  int32_t window_deconvolution;      /* 0/1 */
  /* TreePmSplitPolicy (from mesh spacing cbrt(hx*hy*hz)) */
  double asmth_cells;
  double rcut_cells;
  /* TreeGravityOptions */
  double opening_theta;
  int32_t opening_criterion;         /* chui_opening_criterion */
  int32_t multipole_order;           /* 0 monopole, 2 quadrupole */
  int32_t max_leaf_size;
  int32_t target_mapping;            /* chui_target_mapping (this library) */
  double softening_epsilon;          /* Plummer epsilon, comoving */
  double relative_force_tolerance;
  double relative_force_acceleration_floor;
  /* TreePmOptions */
  int32_t acceptance_policy;         /* chui_acceptance_policy */
  int32_t residual_block_size;       /* 16..256: visit positions per residual block, the span of
                                        group-walk target groups; the per-target walks run 64-wide */
  /* this library */
  int32_t precision;                 /* chui_precision */
  /* TreePmOptions::short_range_only (tree_pm_coupling.hpp:288-291): evaluate
   * only the short-range residual, with no PM solve, reuse or interpolation
   * (refresh_long_range must be 0 and the PM lanes NULL) */
  int32_t short_range_only;
  /* Identity of the long-range field (TreePmOptions::source_generation,
   * pm_field_version, force_epoch and allow_stale_long_range_field,
   * tree_pm_coupling.hpp:296-329). A refresh stores these tokens with the
   * field. A reuse (refresh_long_range = 0) is admitted only when the stored
   * field's pm_field_version, mesh and physical options match and, unless
   * allow_stale_long_range_field is set, its source_generation and force
   * epoch (sequence, scale factor) match too (tree_pm_coupling.cpp:2483-2506);
   * otherwise the solve fails with CHUI_ERR_RUNTIME. The caller owns the
   * tokens and must advance source_generation whenever positions, masses or
   * membership change. */
  uint64_t source_generation;
  uint64_t pm_field_version;
  uint64_t force_epoch_sequence;
  double force_epoch_scale_factor;
  int32_t allow_stale_long_range_field;  /* 0/1: core::TreePmPmReusePolicy::kStaleWithinCadence */
  int32_t reserved1;
} chui_treepm_options;

/* Per-solve counters and timings. Counter names follow TreePmTraversalCounters
 * (tree_pm_coupling.hpp:39-87) and the profile events (:505-524, pm_solver.hpp:256-302). */
typedef struct chui_treepm_diagnostics {
  uint32_t struct_size;
  uint32_t reserved0;
  uint64_t node_count;
  uint64_t leaf_count;
  uint64_t max_depth;
  uint64_t direct_pair_evaluations;
  uint64_t accepted_internal_multipoles;
  uint64_t accepted_leaves;
  uint64_t cutoff_pruned_nodes;
  uint64_t visited_nodes;
  double split_scale_comoving;
  double cutoff_radius_comoving;
  double mesh_spacing_comoving;
  double tree_root_half_size_comoving;
  double unwrap_anchor[3];
  /* device timings in milliseconds (HIP events on the library stream) */
  double ms_total;
  double ms_unwrap_keys_sort;
  double ms_tree_build;
  double ms_multipoles;
  double ms_pm_deposit;
  double ms_pm_fft_poisson;
  double ms_pm_interpolate;
  double ms_walk;
  /* wave-level work of the walk: node records processed per wave (union of
   * the 64 lanes' walks) and leaf source slots streamed per wave */
  uint64_t wave_node_visits;
  uint64_t wave_leaf_source_slots;
  /* mixed precision: times a wave folded its register pair sums (at most 64
   * sources since the previous fold) into the lanes' compensated totals */
  uint64_t wave_sum_flushes;
  /* FP64 per-lane walk (the walk of every FP64 solve with diagnostics): a
   * wrapping sum over its direct pairs inside r_cut of a 64-bit mix of
   * (target row, source row). Equal checksums mean, up to a 2^-64 chance, the
   * same pair set; a gradient check uses it to find finite-difference steps
   * that move a pair across r_cut. 0 for the other walks. */
  uint64_t pair_set_checksum;
} chui_treepm_diagnostics;

typedef struct chui_treepm_context chui_treepm_context;

/* ---- options ----------------------------------------------------------- */
chui_status chui_treepm_default_options(chui_treepm_options* options);
/* Applies one [numerics]/[cosmology] .param.txt key (config.cpp:1723-1766 names,
 * e.g. "treepm_asmth_cells", "treepm_tree_opening_theta", "box_size_mpc_comoving").
 * Lengths (box sizes, gravity_softening) read "<magnitude> [mpc|kpc]", a bare
 * magnitude in Mpc (config.cpp:300-362 with the default units.length_unit). */
chui_status chui_treepm_options_set(chui_treepm_options* options, const char* key, const char* value);
/* Same validation rules as the reference: residual_block_size in [16,256], leaf in
 * [1,1024] (config.cpp:3370-3372), eps/r_s <= 0.2 (:3669), r_cut < L_min/2 (:3682-3692). */
chui_status chui_treepm_options_validate(const chui_treepm_options* options, const int32_t mesh[3]);

/* ---- context ----------------------------------------------------------- */
/* mesh = {nx, ny, nz}; device = HIP ordinal. Each axis must be even with no prime
 * factor other than 2 and 3 (2^a 3^b, a >= 1); others return CHUI_ERR_UNSUPPORTED. */
chui_status chui_treepm_create(const int32_t mesh[3], int32_t device, chui_treepm_context** out);
chui_status chui_treepm_destroy(chui_treepm_context* ctx);
/* Stream the context enqueues on (hipStream_t; NULL = library-owned stream). */
chui_status chui_treepm_set_stream(chui_treepm_context* ctx, void* stream);
chui_status chui_treepm_synchronize(chui_treepm_context* ctx);
/* Launch shapes (no reference counterpart; the knobs of the ensemble's
 * autotuner): lines per block of the PM line passes (0 = the most the LDS
 * budget allows, else 1, 2, 4, 8 or 16, capped at that maximum) and waves per
 * block of the timed mixed walk (0 = 4, else 2, 4 or 8). Results do not
 * depend on them. */
chui_status chui_treepm_set_launch_shape(chui_treepm_context* ctx, int32_t fft_lines_per_block,
                                         int32_t walk_waves_per_block);

/* ---- capture-safe solves and HIP graphs ----------------------------------
 * No counterpart in the reference (its solves are driven from the host). For
 * callers that record force evaluations into a HIP graph: ensembles of boxes,
 * tight step loops. In capture-safe mode the first solve of a source count is
 * an ordinary one that also sizes the context for it and records the tree's
 * node capacity, depth and mass uniformity; every later solve of that source
 * count reads nothing back to the host and allocates nothing, so it can be
 * recorded. Capture-safe solves are strict solves (FP64 or mixed) without
 * diagnostics, counters, the nearest-pair lane or a history
 * (CHUI_ERR_UNSUPPORTED otherwise); their tree-only lanes equal the ordinary
 * solve's bit for bit. The FP64 force VJP follows the same mode. The library
 * detects a stream that records a graph by itself, so a caller may also record
 * the context's stream with its own capture (torch.cuda.graph); every
 * recorded solve and VJP must have been prepared by an eager one. What an
 * ordinary solve reports at once (non-finite inputs, an active index out of
 * range, a tree deeper than 21 levels), and sources that outgrow the prepared
 * node capacity, depth (prepared depth + 4) or mass uniformity
 * (CHUI_ERR_RUNTIME: solve once more outside a graph and record again), comes
 * back at the next synchronisation. */
typedef struct chui_treepm_graph chui_treepm_graph;
chui_status chui_treepm_set_capture_safe(chui_treepm_context* ctx, int32_t enable);
/* Reports the fault word of work replayed from a graph the caller recorded on
 * the context's stream (torch.cuda.graph, hipStreamBeginCapture): copies it
 * after the stream's work, synchronises and returns the failed solve's status.
 * Graphs launched through chui_treepm_graph_launch post it by themselves. */
chui_status chui_treepm_fault_check(chui_treepm_context* ctx);
/* Records everything issued on the context's stream (solves, kicks, drifts)
 * until graph_end into a graph. The context must be capture-safe, and a solve
 * of each source count to be recorded must have run before. */
chui_status chui_treepm_graph_begin(chui_treepm_context* ctx);
chui_status chui_treepm_graph_end(chui_treepm_context* ctx, chui_treepm_graph** out);
/* Replays the graph on the stream it was recorded on, after the member
 * contexts' earlier work; their later work waits for it. Faults of the
 * replayed solves come back at the members' next synchronisation. */
chui_status chui_treepm_graph_launch(chui_treepm_graph* graph);
chui_status chui_treepm_graph_destroy(chui_treepm_graph* graph);

/* ---- force evaluation ---------------------------------------------------- */
/* Host-pointer drop-in for solveActiveSetWithPmCadence. Positions/masses are
 * double[n]; active = uint32[n_active] source rows or NULL for all rows.
 * accel_* = double[n_active] total (PM + short-range residual) accelerations.
 * pm_accel_* (nullable) capture the PM-only part per active slot;
 * short_accel_* (nullable) the tree-only residual, indexed by source row
 * (double[n]): the active rows are written, the others left untouched, as the
 * accumulator view's source-indexed lanes (tree_pm_coupling.hpp:149-153,
 * tree_pm_coupling.cpp:1957-1966). refresh_long_range = 0 reuses the
 * previous PM mesh field (cadence) when its identity matches (see the identity
 * fields of chui_treepm_options), else returns CHUI_ERR_RUNTIME.
 * This call synchronises, so a walk that broke its traversal invariant is
 * reported here as CHUI_ERR_LOGIC. */
chui_status chui_treepm_solve(chui_treepm_context* ctx,
                              const double* pos_x, const double* pos_y, const double* pos_z,
                              const double* mass, uint64_t n,
                              const uint32_t* active, uint64_t n_active,
                              double* accel_x, double* accel_y, double* accel_z,
                              double* pm_accel_x, double* pm_accel_y, double* pm_accel_z,
                              double* short_accel_x, double* short_accel_y, double* short_accel_z,
                              const chui_treepm_options* options, int32_t refresh_long_range,
                              chui_treepm_diagnostics* diagnostics);

/* Device-pointer form: every array is device memory already resident in HBM.
 * Accelerations are double in FP64 precision and float in MIXED precision
 * (accel_* then point to float[n_active]). Work is enqueued on the context
 * stream; the call returns after enqueue unless diagnostics != NULL.
 * Every walk records whether it kept its traversal invariant (a wave whose
 * lanes all left the tree early); the flag is copied to pinned host memory at
 * the end of the solve and checked when the caller next synchronises through
 * the library (chui_treepm_synchronize, chui_treepm_profile_read) or starts
 * the next solve on this context, which then return CHUI_ERR_LOGIC. A solve
 * with diagnostics or counters checks it before it returns. */
chui_status chui_treepm_solve_device(chui_treepm_context* ctx,
                                     const double* pos_x, const double* pos_y, const double* pos_z,
                                     const double* mass, uint64_t n,
                                     const uint32_t* active, uint64_t n_active,
                                     void* accel_x, void* accel_y, void* accel_z,
                                     void* pm_accel_x, void* pm_accel_y, void* pm_accel_z,
                                     void* short_accel_x, void* short_accel_y, void* short_accel_z,
                                     const chui_treepm_options* options, int32_t refresh_long_range,
                                     chui_treepm_diagnostics* diagnostics);

/* ---- adaptive acceptance and the group walk ------------------------------- */
typedef enum chui_adaptive_estimator {
  CHUI_ESTIMATOR_RESIDUAL_DERIVATIVE = 0, /* kResidualDerivative (reference default) */
  CHUI_ESTIMATOR_RELATIVE_MONOPOLE = 1    /* kRelativeMonopole */
} chui_adaptive_estimator;

typedef enum chui_cutoff_node_policy {
  CHUI_CUTOFF_CONTAIN_NODE = 0, /* kContainNode (reference default) */
  CHUI_CUTOFF_COM_DISTANCE = 1  /* kComDistance */
} chui_cutoff_node_policy;

/* The TreePmOptions fields of the adaptive acceptance policy and the group walk
 * (tree_pm_coupling.hpp:248-276). chui_treepm_default_adaptive_options gives the
 * reference's defaults. The policy constants alpha (relative_force_tolerance),
 * the acceleration floor and theta stay in chui_treepm_options. */
typedef struct chui_treepm_adaptive_options {
  uint32_t struct_size;                     /* = sizeof(chui_treepm_adaptive_options) */
  int32_t error_estimator;                  /* chui_adaptive_estimator */
  double maximum_opening_angle;             /* adaptive_maximum_opening_angle (0.25) */
  int32_t cutoff_node_policy;               /* chui_cutoff_node_policy */
  int32_t gaussian_pair_lookup;             /* gaussian_pair_lookup_enabled: Hermite table for leaf pairs */
  int32_t radial_coefficients;              /* radial_coefficients_enabled: accepted-node derivatives */
  int32_t group_walk;                       /* group_walk_enabled */
  uint32_t group_target_size;               /* group_walk_target_group_size (0 = the active targets of one leaf) */
  uint32_t reserved0;
  double group_max_extent_cutoff_fraction;  /* group_walk_max_extent_cutoff_fraction (0.5) */
} chui_treepm_adaptive_options;

/* Group walk status of a solve (TreePmGroupWalkStatus, tree_pm_coupling.hpp:215-227). */
typedef enum chui_group_walk_status {
  CHUI_GROUP_WALK_NOT_REQUESTED = 0,
  CHUI_GROUP_WALK_ACTIVE = 1,
  CHUI_GROUP_WALK_NO_LOCAL_TREE = 2,
  CHUI_GROUP_WALK_INCOMPATIBLE_POLICY = 3,
  CHUI_GROUP_WALK_CUTOFF_OUTSIDE_TABLE = 6,
  CHUI_GROUP_WALK_LOOKUP_SCALING_UNREPRESENTABLE = 8
} chui_group_walk_status;

/* Traversal counters of an adaptive solve (the TreePmTraversalCounters fields of
 * tree_pm_coupling.hpp:40-90 that the group walk and the per-target walk share).
 * Per-target walk: visited / pruned / pairs / multipoles / leaves per target.
 * Group walk: visited nodes, pruned nodes, listed leaves and listed multipoles
 * per group; pairs and com-distance prunes per target. */
typedef struct chui_treepm_walk_counters {
  uint32_t struct_size;                 /* = sizeof(chui_treepm_walk_counters) */
  uint32_t group_walk_status;           /* chui_group_walk_status */
  uint64_t visited_nodes;
  uint64_t cutoff_pruned_nodes;
  uint64_t direct_pair_evaluations;
  uint64_t accepted_internal_multipoles;
  uint64_t accepted_leaves;
  uint64_t group_walk_groups;
  uint64_t group_walk_grouped_targets;
  uint64_t group_walk_target_group_size_max;
  uint64_t group_walk_fallback_targets;
  uint64_t group_walk_fallback_history;
  uint64_t group_walk_fallback_geometry;
  uint64_t group_walk_candidate_leaves;
  uint64_t group_walk_culled_leaves;
  uint64_t group_walk_candidate_particles;
  uint64_t group_walk_list_entries_max;
} chui_treepm_walk_counters;

chui_status chui_treepm_default_adaptive_options(chui_treepm_adaptive_options* adaptive);
/* Applies one [numerics] .param.txt key of these fields (config.cpp:1741-1754):
 * treepm_adaptive_maximum_opening_angle, treepm_adaptive_error_estimator
 * (residual_derivative | relative_monopole), treepm_cutoff_node_policy
 * (contain | com_distance), treepm_group_walk_enabled, treepm_group_walk_target_size,
 * treepm_gaussian_pair_lookup_enabled, treepm_radial_coefficients_enabled.
 * treepm_adaptive_acceptance_enabled selects the policy in chui_treepm_options_set. */
chui_status chui_treepm_adaptive_options_set(chui_treepm_adaptive_options* adaptive, const char* key,
                                             const char* value);

/* solveActiveSetWithPmCadence with options.acceptance_policy = CHUI_ACCEPT_ADAPTIVE
 * (tree_pm_coupling.cpp:4026-4546): the per-target walk with the adaptive
 * acceptance (tree_pm_acceptance.hpp:54-139) and, when adaptive->group_walk is
 * set and the policy allows it (tree_pm_coupling.cpp:3839-3869), the group walk
 * of tree_pm_group_walk.hpp:78-341 over groups of active targets inside each
 * residual block of options.residual_block_size visit positions; a group whose
 * targets lack a usable history or leave the bounded periodic metric falls back
 * to the per-target walk. previous_acceleration = double[n_active] magnitudes of
 * the previous total acceleration per active slot
 * (TreePmForceAccumulatorView::previous_acceleration_magnitude_code,
 * tree_pm_coupling.hpp:127-165), or NULL for none. short_nearest_pair_r2
 * (nullable, source-row indexed like the short-range lanes) receives the
 * squared separation of the nearest distinct direct-summed source inside
 * r_cut, +infinity when none (short_range_nearest_pair_r2_comoving,
 * tree_pm_coupling.hpp:154-159). Device arrays and element types as in
 * chui_treepm_solve_device. In mixed precision (FP32 lanes) the group walk
 * makes every decision in FP64 and evaluates the listed-source pairs in FP32;
 * the per-target adaptive walk runs in FP64 and its sums are rounded to the
 * FP32 lanes. With acceptance_policy = CHUI_ACCEPT_STRICT this is the strict
 * solve in either precision plus the nearest-pair lane (mixed: FP32, from the
 * counting walk).
 * counters may be NULL (no synchronisation then). */
chui_status chui_treepm_solve_adaptive_device(chui_treepm_context* ctx,
                                              const double* pos_x, const double* pos_y, const double* pos_z,
                                              const double* mass, uint64_t n,
                                              const uint32_t* active, uint64_t n_active,
                                              void* accel_x, void* accel_y, void* accel_z,
                                              void* pm_accel_x, void* pm_accel_y, void* pm_accel_z,
                                              void* short_accel_x, void* short_accel_y, void* short_accel_z,
                                              void* short_nearest_pair_r2,
                                              const double* previous_acceleration,
                                              const chui_treepm_options* options,
                                              const chui_treepm_adaptive_options* adaptive,
                                              int32_t refresh_long_range, chui_treepm_walk_counters* counters);

/* ---- ensembles of independent boxes ---------------------------------------
 * No counterpart in the reference (one box per run there). An ensemble holds
 * one context per box (own mesh, tree, workspaces and stream). */
typedef struct chui_treepm_ensemble chui_treepm_ensemble;

/* One box of a batched solve: device arrays, all n rows active, FP64 (double
 * accelerations) or mixed (float accelerations) by options->precision. */
typedef struct chui_treepm_box {
  const double* pos_x;
  const double* pos_y;
  const double* pos_z;
  const double* mass;
  uint64_t n;
  void* accel_x;
  void* accel_y;
  void* accel_z;
  const chui_treepm_options* options;
} chui_treepm_box;

/* meshes = int32[3 * box_count]; workers = host threads that issue the boxes'
 * solves (1..64). */
chui_status chui_treepm_ensemble_create(int32_t box_count, const int32_t* meshes, int32_t device, int32_t workers,
                                        chui_treepm_ensemble** out);
chui_status chui_treepm_ensemble_destroy(chui_treepm_ensemble* ensemble);
/* The context of one box (owned by the ensemble): kick, drift, adjoint and
 * parity entry points take it like any other context. */
chui_status chui_treepm_ensemble_box(chui_treepm_ensemble* ensemble, int32_t index, chui_treepm_context** box);
/* solveActiveSet for every box (all rows active), the boxes handed to the
 * workers largest cost (N log N + M log M) first; each box's solve runs on
 * its context's stream. Returns when every box is done. */
chui_status chui_treepm_ensemble_solve_device(chui_treepm_ensemble* ensemble, const chui_treepm_box* boxes,
                                              int32_t box_count, int32_t refresh_long_range);
/* Capture-safe mode for every box (see chui_treepm_set_capture_safe). */
chui_status chui_treepm_ensemble_set_capture_safe(chui_treepm_ensemble* ensemble, int32_t enable);
/* Records the work issued on every box's stream until graph_end (batched
 * solves, which are then issued from the calling thread, kicks, drifts) into
 * one graph in which each box is a branch: box 0's stream records, the others
 * join it. Every box must be capture-safe and prepared. */
chui_status chui_treepm_ensemble_graph_begin(chui_treepm_ensemble* ensemble);
chui_status chui_treepm_ensemble_graph_end(chui_treepm_ensemble* ensemble, chui_treepm_graph** out);

/* Per-stage device timing accumulated over solves with HIP events on the
 * context stream (no extra synchronisation inside a solve).
 * chui_treepm_profiling(ctx, 1) resets and enables, (ctx, 0) disables.
 * chui_treepm_profile_read synchronises and returns stage_ms[8] =
 * {total, tree build (unwrap, keys, sort, nodes, multipoles), source gather,
 *  PM (deposit + FFTs + Green + gradients), target set-up, walk (+ fused PM
 *  interpolation), reserved, reserved} summed over *solves solves. */
chui_status chui_treepm_profiling(chui_treepm_context* ctx, int32_t enable);
chui_status chui_treepm_profile_read(chui_treepm_context* ctx, double stage_ms[8], uint64_t* solves);

/* ---- KDK integrator pieces (FP64 on the device) -------------------------- */
typedef struct chui_kdk_factors {
  double scale_factor_begin, scale_factor_midpoint, scale_factor_end;
  double drift_factor;            /* int_{a0}^{a1} da / (a^2 H(a)), code units */
  double first_kick_factor;       /* int_{a0}^{am} da / (a^2 H(a)) / am */
  double second_kick_factor;      /* int_{am}^{a1} da / (a^2 H(a)) / a1 */
  double first_hubble_drag;       /* a0 / am */
  double second_hubble_drag;      /* am / a1 */
} chui_kdk_factors;

/* LambdaCDM factors for one global KDK step from a0 to a1. hubble0_code = H0 in
 * inverse code time, H(a) = H0 sqrt(omega_m a^-3 + omega_lambda) (flat,
 * cosmology.hpp:77-87). am is the scale factor at the cosmic-time midpoint; the
 * integrals use 64-sample midpoint quadrature (time_integration.cpp:242-266,3538-3556). */
chui_status chui_treepm_kdk_factors(double scale_factor_begin, double scale_factor_end, double hubble0_code,
                                    double omega_m, double omega_lambda, chui_kdk_factors* out);
/* v = drag * v + kick * a  (active rows or all rows when active == NULL). accel is
 * double (is_float_accel = 0) or float (= 1), indexed by active slot. */
chui_status chui_treepm_kick_device(chui_treepm_context* ctx, double* vel_x, double* vel_y, double* vel_z,
                                    const void* accel_x, const void* accel_y, const void* accel_z,
                                    int32_t is_float_accel, const uint32_t* active, uint64_t n_active,
                                    double kick_factor, double hubble_drag);
/* x += drift * v, then canonical wrap into [0, L) per axis. */
chui_status chui_treepm_drift_device(chui_treepm_context* ctx, double* pos_x, double* pos_y, double* pos_z,
                                     const double* vel_x, const double* vel_y, const double* vel_z,
                                     uint64_t n, double drift_factor, const double box_size[3]);
/* FP64 global reductions on the device: out[0] = sum m, out[1..3] = sum m v. */
chui_status chui_treepm_mass_momentum_device(chui_treepm_context* ctx, const double* mass,
                                             const double* vel_x, const double* vel_y, const double* vel_z,
                                             uint64_t n, double out_host[4]);
/* Allocation-free form of chui_treepm_mass_momentum_device. It has no
 * counterpart in the reference; it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops). Same sums, same
 * fixed grid and summation tree, so the result is bit-identical; it goes to
 * out_device[0..3] (device memory the caller owns) in the context's stream
 * order and the call returns without synchronising. No device memory is
 * allocated or freed: the partial sums live in a 32 KiB array the library
 * holds on each GPU (4 x 1024 FP64 partials, one per block of the fixed grid,
 * so every n fits it). Every context on one GPU shares that array: calls that
 * overlap on one GPU (several contexts, streams or host threads) race. */
chui_status chui_treepm_mass_momentum_noalloc_device(chui_treepm_context* ctx, const double* mass,
                                                     const double* vel_x, const double* vel_y,
                                                     const double* vel_z, uint64_t n, double* out_device);

/* ---- differentiation (no reference counterpart) --------------------------- */
/* The reference has no adjoint (no autograd or adjoint code in src/); these
 * entry points serve the differentiable forward of python/cosmosim/adjoint.py.
 *
 * Transpose of the Jacobian of the total acceleration a (PM + short-range
 * residual, as chui_treepm_solve_device computes it with every source active)
 * with respect to the source positions, applied to lam:
 *   out_k = sum_i (d a_i / d x_k)^T lam_i,  all arrays device FP64 in caller order.
 * The tree and both PM fields are rebuilt at (x, y, z). Requires FP64 precision
 * and the strict acceptance policy, with monopole (multipole_order 0: no node
 * is accepted, tree_pm_acceptance.hpp:62-66) or quadrupole moments
 * (multipole_order 2: the walk takes the forward's acceptance decisions and
 * differentiates each accepted node's expansion, tree_pm_coupling.cpp:743-827,
 * through the target position and the node's centre of mass and quadrupole);
 * otherwise CHUI_ERR_UNSUPPORTED. */
chui_status chui_treepm_force_vjp_device(chui_treepm_context* ctx, const double* pos_x, const double* pos_y,
                                         const double* pos_z, const double* mass, uint64_t n,
                                         const double* lam_x, const double* lam_y, const double* lam_z,
                                         double* out_x, double* out_y, double* out_z,
                                         const chui_treepm_options* options);
/* CIC density contrast of n unit-weight particles on a periodic mesh^3 grid of
 * side box (row-major, no padding): delta = count * mesh^3 / n - 1, with the
 * CIC stencil of pm_solver.cpp:925-936. */
chui_status chui_treepm_cic_density_device(chui_treepm_context* ctx, const double* pos_x, const double* pos_y,
                                          const double* pos_z, uint64_t n, double box, int32_t mesh,
                                          double* delta);
/* Its transpose: out_k = d/dx_k sum_c grad_delta(c) delta(c). */
chui_status chui_treepm_cic_density_vjp_device(chui_treepm_context* ctx, const double* pos_x,
                                              const double* pos_y, const double* pos_z, uint64_t n, double box,
                                              int32_t mesh, const double* grad_delta, double* out_x,
                                              double* out_y, double* out_z);

/* ---- parity / inspection entry points ----------------------------------- */
/* Periodic unwrap of one axis (device arrays): out = unwrapped coordinate, returns anchor. */
chui_status chui_treepm_unwrap_device(chui_treepm_context* ctx, const double* in, uint64_t n,
                                      double box_length, double* out, double* anchor_host);
/* Morton keys + stable order of (already unwrapped) device coordinates:
 * keys_sorted = uint64[n] keys in sorted order, order = uint32[n] source rows. */
chui_status chui_treepm_morton_device(chui_treepm_context* ctx, const double* x, const double* y,
                                      const double* z, uint64_t n, uint64_t* keys_sorted, uint32_t* order);
/* Size / copy-out of the tree built by the last solve (preorder node numbering). */
chui_status chui_treepm_tree_size(chui_treepm_context* ctx, uint64_t* node_count, uint64_t* particle_count);
/* Host arrays of node_count entries (child_index: 8*node_count; order: particle_count).
 * Any pointer may be NULL. geometry = double[node_count*4] (cx,cy,cz,half);
 * moments = double[node_count*11] (mass, comx, comy, comz, qxx, qxy, qxz, qyy, qyz, qzz, trace). */
chui_status chui_treepm_tree_copy(chui_treepm_context* ctx, double* geometry, double* moments,
                                  uint32_t* particle_begin, uint32_t* particle_count,
                                  uint32_t* child_index, uint32_t* order);

/* FP64 periodic Ewald accelerations at n_targets target rows (device arrays),
 * alpha in inverse length, image/mode limits inclusive. Used as the accuracy
 * oracle of the force error figures. */
chui_status chui_treepm_ewald_device(chui_treepm_context* ctx, const double* pos_x, const double* pos_y,
                                     const double* pos_z, const double* mass, uint64_t n,
                                     const uint32_t* targets, uint64_t n_targets, const double box_size[3],
                                     double gravitational_constant, double alpha, int32_t real_images,
                                     int32_t reciprocal_modes, double plummer_epsilon,
                                     double* accel_x, double* accel_y, double* accel_z);

/* Split-kernel diagnostics (S, L at r_s and r_cut; composition error). */
typedef struct chui_split_diagnostics {
  double split_scale_comoving, cutoff_radius_comoving;
  double short_range_factor_at_split, long_range_factor_at_split;
  double short_range_factor_at_cutoff, long_range_factor_at_cutoff;
  double composition_error_at_split, max_relative_composition_error;
} chui_split_diagnostics;
chui_status chui_treepm_split_diagnostics(double asmth_cells, double rcut_cells, double mesh_spacing,
                                          chui_split_diagnostics* out);

/* ---- testing --------------------------------------------------------------- */
/* No reference counterpart. Sets the context's walk traversal-fault word on
 * the device (enqueued on the context stream), as a walk whose wave lost its
 * traversal invariant would; the next synchronising call then reports
 * CHUI_ERR_LOGIC. Lets tests exercise that report without a faulty walk. */
chui_status chui_treepm_debug_inject_walk_fault(chui_treepm_context* ctx);

/* ---- misc ----------------------------------------------------------------- */
const char* chui_treepm_last_error(void);
int32_t chui_treepm_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif /* CHUI_TREEPM_H */
