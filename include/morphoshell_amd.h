/*
 * morphoshell_amd.h -- C-ABI boundary of the MI355X per-step force-and-dynamics
 * path of MorphoShell.
 *
 * Plain C types only: pointers, sizes, status codes.  Every entry point
 * returns an int status (MS_OK == 0); on failure ms_last_error() (or
 * ms_ctx_last_error() for a context) gives a message.  Host arrays passed in
 * are copied; device buffers are owned by the context.
 *
 * Layout conventions (those of the reference):
 *   - node positions/velocities/forces of one shell: "dofs" SoA
 *     [X_0..X_{N-1}, Y_0..Y_{N-1}, Z_0..Z_{N-1}]  (main.cpp:235-237,
 *     calc_deriv_mats.cpp:60);
 *   - per-triangle 2x2 symmetric tensors (a, b, abar, bbar): components
 *     (xx, xy, yy); as a T x 3 array they are SoA by component, like the
 *     reference's column-major Eigen::Matrix<double,Dynamic,3>
 *     (calc_a_comps_and_b_comps_and_normals.cpp:49-60);
 *   - patch ids: per triangle 6 node ids, slots 0-2 = the triangle's vertices
 *     in file order, slots 3-5 = non_vertex_patch_nodes_ids
 *     (set_up_patch_fitting.cpp:184-212, Tri_Class.hpp:105);
 *   - G: per triangle 2x3 row-major, G[k][v] = row k+1 of
 *     inv([1, x_v - c_x, y_v - c_y]) (calc_deriv_mats.cpp:70-91);
 *   - H: per triangle 6x3 row-major, H[n][k] = mat_for_patch_2nd_derivs
 *     (set_up_patch_fitting.cpp:232, Tri_Class.hpp:164).
 */
#ifndef MORPHOSHELL_AMD_H
#define MORPHOSHELL_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
enum {
    MS_OK = 0,
    MS_ERR_ARG = 1,          /* bad argument / size mismatch               */
    MS_ERR_HIP = 2,          /* HIP runtime error                          */
    MS_ERR_PATCH_SEARCH = 3, /* set_up_patch_fitting.cpp:289-295           */
    MS_ERR_MESH = 4,         /* calc_mesh_info.cpp:122-131 identities      */
    MS_ERR_HIGH_FORCE = 5,   /* advance_dynamics.cpp:58-73 ("implausibly_high_force") */
    MS_ERR_STATE = 6,        /* call not valid in the current state        */
    MS_ERR_NO_DEVICE = 7     /* no HIP device visible                      */
};

enum { MS_FP64 = 0, MS_MIXED = 1 };

/* dialling protocols of do_dialling.cpp: 1 = linear blend to abar_info /
 * bbar_info (the active one, do_dialling.cpp:83-84); 3 = LCE lambda dialling
 * from rows (director angle, final lambda, opto-thermal nu), bbar = 0
 * (do_dialling.cpp:104-124). */
enum { MS_DIAL_BLEND = 1, MS_DIAL_LCE = 3 };

/* Version string of the library. */
const char* ms_version(void);
/* Last error of a context-free call on this thread. */
const char* ms_last_error(void);

/* ============================================================ host set-up
 * These replace the reference's once-per-run precompute (host C++, OpenMP).
 */

typedef struct ms_mesh_info_out {
    double perimeter;        /* sample_char_length, calc_mesh_info.cpp:110-111 */
    double total_ref_area;   /* calc_mesh_info.cpp:135-143 */
    double approx_min_tri_size; /* smallest altitude, calc_mesh_info.cpp:149-167 */
    int32_t num_edges;
    int32_t num_boundary_edges;
    int32_t num_boundary_nodes;
    int32_t num_holes;
} ms_mesh_info_out;

/* calc_mesh_info (calc_mesh_info.cpp:60-184): reference areas (T), centroids
 * (3T, SoA cx.., cy.., cz..), boundary node flags (N, may be NULL) and mesh
 * scalars.  ref_pos is 3N dofs-SoA; tri_vertices 3T row-major. */
int ms_mesh_info(int32_t num_nodes, int32_t num_tris, const double* ref_pos,
                 const int32_t* tri_vertices, double* ref_areas, double* centroids,
                 uint8_t* boundary_nodes, ms_mesh_info_out* out);

/* set_up_patch_fitting + calc_deriv_mats (set_up_patch_fitting.cpp:57-297,
 * calc_deriv_mats.cpp:64-119): patch ids (6T), G (6T), H (18T) as described
 * above.  num_retried (may be NULL) receives the number of triangles whose
 * first patch choice was rejected.  Fails with MS_ERR_PATCH_SEARCH when a
 * triangle's search is exhausted. */
int ms_patch_fit(int32_t num_nodes, int32_t num_tris, const double* ref_pos,
                 const int32_t* tri_vertices, double conditioning_thresh,
                 int32_t* patch_ids, double* G, double* H, int32_t* num_retried);

/* calc_dof_masses (calc_dof_masses.cpp:45-58): node masses (N; the reference
 * stores the same value for the x, y and z dofs). */
int ms_dof_masses(int32_t num_nodes, int32_t num_tris, const int32_t* tri_vertices,
                  const double* ref_areas, const double* ref_thickness, double ref_density,
                  double* node_masses);

/* Inputs and outputs of the characteristic scales, long times, damping,
 * timestep, dialling and output cadence (calc_char_mechanics_scales.cpp:45-74,
 * calc_long_timescales.cpp:48-60, calc_damping_factor.cpp:50-57,
 * calc_timestep.cpp:52-71, calc_dialling_procedure_params.cpp:60-77,
 * calc_steps_between_outputs.cpp:48-69). */
typedef struct ms_scales {
    /* inputs */
    double poisson_ratio;
    double ref_density;
    double timestep_prefactor;
    double damping_prefactor;
    double dial_phase_time_prefactor;
    double time_between_equil_checks_prefactor;
    double outputs_per_char_long_time;
    int32_t gradient_descent;
    /* outputs */
    double min_ref_thickness;
    double min_ref_shear_modulus;
    double char_flexural_rigidity;
    double char_force_scale;
    double char_speed_scale;
    double char_stretch_energy_density_scale;
    double char_stretch_energy_scale;
    double stretch_long_time;
    double bend_long_time;
    double char_long_time;
    double damping_factor;
    double timestep;
    double dial_phase_time;
    int64_t timesteps_per_dial_phase;
    int64_t timesteps_between_equil_checks;
    int64_t steps_between_outputs;  /* -1 = no regular output */
} ms_scales;

/* Fill the output fields of *s from the mesh scalars and per-triangle
 * thickness/modulus/area arrays. */
int ms_calc_scales(int32_t num_tris, const double* ref_thickness, const double* ref_shear_modulus,
                   const double* ref_areas, double perimeter, double approx_min_tri_size,
                   ms_scales* s);

/* ========================================================= device context */

typedef struct ms_ctx ms_ctx;

/* Scalars the per-step kernels read, per shell. */
typedef struct ms_shell_params {
    double poisson_ratio;                 /* def_poisson_ratio */
    double bend_stiffening_scale_factor;  /* gamma */
    double ref_density;
    double min_ref_thickness;
    double timestep;
    double damping_factor;
    double char_force_scale;
    double char_speed_scale;
    int32_t gradient_descent;             /* using_gradient_descent_dynamics */
    int32_t dial_protocol;                /* MS_DIAL_BLEND or MS_DIAL_LCE */
    int32_t dial_from_ansatz;             /* dialling_from_ansatz_rather_than_ref_state */
    int32_t reserved;
} ms_shell_params;

/* One shell.  Host arrays are copied at ms_create. */
typedef struct ms_shell_desc {
    int32_t num_nodes;
    int32_t num_tris;
    const int32_t* patch_ids;      /* 6T */
    const double* G;               /* 6T */
    const double* H;               /* 18T */
    const double* ref_pos;         /* 3N dofs SoA: the reference positions G and H
                                      were fitted on (x and y are read) */
    const double* ref_areas;       /* T */
    const double* thickness;       /* T  def_thicknesses */
    const double* shear_modulus;   /* T  def_shear_moduli */
    const double* abar_info;       /* T x abar_info_cols, row-major */
    int32_t abar_info_cols;        /* 3 for MS_DIAL_BLEND, 3 for MS_DIAL_LCE */
    int32_t bbar_info_cols;        /* 3 (ignored for MS_DIAL_LCE) */
    const double* bbar_info;       /* T x bbar_info_cols, row-major, may be NULL for LCE */
    const double* a0;              /* 3T SoA a at zero dial factor, NULL = flat (1,0,1) */
    const double* b0;              /* 3T SoA b at zero dial factor, NULL = 0 */
    const double* node_masses;     /* N */
    const uint8_t* constrained;    /* N or NULL */
    ms_shell_params params;
} ms_shell_desc;

/* Create a context on the current HIP device for num_shells shells which
 * advance together (num_shells == 1 is the single-shell path).  precision is
 * MS_FP64 or MS_MIXED.  All node state starts at zero: call ms_set_state. */
int ms_create(const ms_shell_desc* shells, int32_t num_shells, int32_t precision, ms_ctx** out);
void ms_destroy(ms_ctx* ctx);
const char* ms_ctx_last_error(const ms_ctx* ctx);

int32_t ms_num_shells(const ms_ctx* ctx);

/* Node state of one shell (dofs SoA, 3N each); vel may be NULL (= zero on set). */
int ms_set_state(ms_ctx* ctx, int32_t shell, const double* dofs, const double* vel);
int ms_get_state(ms_ctx* ctx, int32_t shell, double* dofs, double* vel);

/* Dial clock of one shell.  The dial factor the force kernels use at a step is
 * min(1, last_eq + spacing * phase_steps / steps_per_phase), and phase_steps
 * advances by one per step while dialling != 0 (main.cpp:286-288, 478-488). */
int ms_set_dial_clock(ms_ctx* ctx, int32_t shell, double last_eq, double spacing,
                      int64_t steps_per_phase, int64_t phase_steps, int32_t dialling);
int ms_get_dial_factor(ms_ctx* ctx, int32_t shell, double* dial_factor, int64_t* phase_steps);
/* Per-shell step scalars changed at phase switches (main.cpp:293-304,501-507). */
int ms_set_shell_params(ms_ctx* ctx, int32_t shell, const ms_shell_params* p);
/* Replace a0/b0 (dial-from-ansatz: a, b of the state at step 0). */
int ms_set_zero_dial_geometry(ms_ctx* ctx, int32_t shell, const double* a0, const double* b0);
/* Switch the per-step force kernels between MS_FP64 and MS_MIXED (the FP64
 * per-triangle streams are always kept, so a mixed run can be finished in
 * FP64; energies and geometry are always evaluated in FP64). */
int ms_set_precision(ms_ctx* ctx, int32_t precision);
int ms_get_precision(const ms_ctx* ctx);

/* ---------------------------------------------------- the reference's calls
 * Each runs on all non-halted shells of the context, on device-resident
 * state, and mirrors one reference function.  The fused ms_step below is the
 * fast path; these are for the driver's event steps and for parity tests. */

/* calc_surface_derivs + calc_deformation_forces (calc_surface_derivs.cpp:27-34,
 * calc_deformation_forces.cpp:40-185): forces = -M^T dE/dq. */
int ms_calc_deformation_forces(ms_ctx* ctx);
/* calc_non_deformation_forces, damping part (calc_non_deformation_forces.cpp:44-65). */
int ms_calc_non_deformation_forces(ms_ctx* ctx);
/* impose_constraints (impose_constraints.cpp:56-68). */
int ms_impose_constraints(ms_ctx* ctx);

typedef struct ms_step_report {
    int32_t status;          /* MS_OK or MS_ERR_HIGH_FORCE */
    int32_t halted;          /* shell is halted (failed or frozen) */
    int64_t fail_step;       /* step index (ms_steps_taken clock) of the failure */
    int64_t bad_dof;         /* first offending dof (n + c*N) */
    double bad_force;
    double max_force_ratio;  /* is_at_equilibrium outputs of the last check */
    int32_t max_force_node;
    int32_t max_speed_node;
    double max_speed_ratio;
    int32_t at_equilibrium;
    int32_t checked;         /* an equilibrium check ran in this call */
} ms_step_report;

/* advance_dynamics (advance_dynamics.cpp:58-93): blow-up check of the current
 * forces, then the symplectic-Euler (or gradient-descent) update.  On a
 * blow-up nothing is updated and reports[s] names the first offending dof. */
int ms_advance_dynamics(ms_ctx* ctx, ms_step_report* reports);
/* is_at_equilibrium (is_at_equilibrium.cpp:58-98) on the current forces and
 * velocities: max |F_n|/F_char and |v_n|/c_char with the first argmax. */
int ms_is_at_equilibrium(ms_ctx* ctx, int32_t shell, double force_tol, double speed_tol,
                         ms_step_report* report);
/* calc_energy_densities + calc_energies (calc_energy_densities.cpp:46-82,
 * calc_energies.cpp:44-64): out[0] stretch, out[1] bend, out[2] kinetic,
 * compensated FP64 sums. */
int ms_calc_energies(ms_ctx* ctx, int32_t shell, double out[3]);
/* calc_a_comps_and_b_comps_and_normals (calc_a_comps_and_b_comps_and_normals.cpp:34-61):
 * a, b, n as 3T SoA each (any pointer may be NULL). */
int ms_calc_geometry(ms_ctx* ctx, int32_t shell, double* a, double* b, double* n);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): ms_calc_geometry of one shell, computed into a static
 * device scratch array of the library instead of the output buffer that
 * ms_calc_geometry allocates on its first call; this call allocates
 * nothing.  The array holds MS_NOALLOC_MAX_TRIS triangles (36 MiB on each
 * device the library runs on, held until the process ends); a larger shell
 * is refused with MS_ERR_ARG.  All calls on one device share the array, so
 * calls that overlap on one device (from several host threads) race. */
#define MS_NOALLOC_MAX_TRIS 524288
int ms_calc_geometry_noalloc(ms_ctx* ctx, int32_t shell, double* a, double* b, double* n);
/* Per-triangle energy densities (T x 2 SoA: stretch.., bend..), may be NULL. */
int ms_calc_energy_densities(ms_ctx* ctx, int32_t shell, double* densities);
/* Output-time diagnostics per triangle (T each, any pointer may be NULL):
 * Gauss and mean curvature of the shape operator (calc_curvatures.cpp:35-94)
 * and the strain measure sqrt(tr(E^T E)/2), E = (F^+)^T (a - abar) F^+ / 2
 * (calc_strains.cpp:35-78), both evaluated in their invariant forms
 * det(b)/det(a), tr(a^-1 b)/2 and sqrt(tr((a^-1 (a - abar))^2)/8). */
int ms_calc_diagnostics(ms_ctx* ctx, int32_t shell, double* gauss_curv, double* mean_curv, double* strain);
/* calc_stresses (calc_stresses.cpp:45-179) on the current state: per
 * triangle the tensor (A/J) sum_alpha dW/dR_alpha (x) R_alpha (the reference
 * area A times the membrane Cauchy stress; W the energy density per
 * reference area, J the area ratio), its symmetric part read from the lower
 * triangle, and of its eigenpairs the two the reference keeps: 8T SoA
 * (eigval_1.., eigvec_1 x.., y.., z.., eigval_2.., eigvec_2 x.., y.., z..),
 * eigenvectors with their largest-magnitude component positive.
 * write_output.cpp:136-156 divides the eigenvalues by the reference shear
 * modulus times thickness of each triangle. */
int ms_calc_stresses(ms_ctx* ctx, int32_t shell, double* stresses);
/* calc_angle_deficits (calc_angle_deficits.cpp:40-78) on the current state:
 * per node 2 pi (pi where boundary[n] != 0; N flags of the shell, e.g. from
 * ms_mesh_info) minus the corner angles of its triangles, added in ascending
 * triangle order as the reference's serial loop does.  deficits: N. */
int ms_calc_angle_deficits(ms_ctx* ctx, int32_t shell, const uint8_t* boundary, double* deficits);
/* Current force buffer of one shell (3N dofs-SoA). */
int ms_get_forces(ms_ctx* ctx, int32_t shell, double* forces);
int ms_set_forces(ms_ctx* ctx, int32_t shell, const double* forces);

/* ------------------------------------------------------------ fused path
 * nsteps steps of: elastic forces, damping, clamp, blow-up flag and update,
 * each step one triangle launch and one node launch, replayed from a HIP
 * graph.  If equil_every > 0 an equilibrium check (force_tol, speed_tol) runs
 * fused into every equil_every-th step of each shell's clock, and a shell
 * that passes is halted when halt_on_equilibrium != 0 (its state stays the
 * checked state).  reports (num_shells entries, may be NULL) receives per-shell
 * results.  Returns MS_ERR_HIGH_FORCE if any shell blew up (that shell is
 * halted at the failing step; the others ran all nsteps). */
int ms_step(ms_ctx* ctx, int64_t nsteps, int64_t equil_every, double force_tol, double speed_tol,
            int32_t halt_on_equilibrium, ms_step_report* reports);
/* Steps taken so far by the context clock. */
int64_t ms_steps_taken(const ms_ctx* ctx);

/* ------------------------------------------------- device-run relaxation
 * The reference's dynamics-loop state machine (main.cpp:269-507) run on the
 * device for each shell: dialling (damping eta1) for steps_per_dial_phase
 * steps, then waiting_for_equil (damping eta2) with an equilibrium check
 * every equil_interval steps (is_at_equilibrium.cpp:92-93: both ratios below
 * their thresholds), then the next dialling phase from the dial factor just
 * equilibrated at.  A shell halts at the first equilibrium with dial factor 1
 * (the reference's loop would re-dial forever, main.cpp:493), or at every
 * equilibrium if pause_on_eq is set, keeping the checked state. */
typedef struct ms_relax_params {
    double tol_f;                 /* force_to_char_force_ratio_equil_threshold */
    double tol_v;                 /* speed_to_char_speed_ratio_equil_threshold */
    int64_t equil_interval;       /* timesteps_between_equil_checks */
    int64_t steps_per_dial_phase; /* timesteps_per_dial_phase (> 0) */
    double dial_spacing;          /* dial_factors_to_hold_at_spacing */
    double dial_start;            /* dial factor of the first step */
    double eta1, eta2;            /* damping factors while dialling / waiting */
    int32_t pause_on_eq;
    int32_t reserved;
} ms_relax_params;

typedef struct ms_relax_report {
    int32_t status;      /* 0 dialling, 1 waiting, 2 final equilibrium */
    int32_t halted;
    int32_t failed;      /* blow-up (see fail_step) */
    int32_t num_eq;      /* equilibria reached */
    int64_t eq_step;     /* the shell's own step count (since ms_relax_setup) at its
                            last equilibrium check that passed; -1 if none */
    int64_t fail_step;
    double dial_factor;
    double max_force_ratio;
    double max_speed_ratio;
    int64_t steps;       /* the shell's own step count since ms_relax_setup (steps
                            applied to it; a halted shell's count stops) */
} ms_relax_report;

/* The protocol's log (main.cpp:303, 396-407; is_at_equilibrium.cpp:89-90):
 * with a check log enabled, each shell records every equilibrium check
 * (kind 0: step = the shell's own step count at the check, the dial factor
 * used, the force and speed ratios with their first-argmax nodes, the
 * verdict) and every end of a dialling phase (kind 1: the dial factor the
 * shell starts waiting at) into a device ring of `capacity` records;
 * ms_read_check_log returns the records since the previous read, oldest
 * first, and how many were overwritten before they were read. */
typedef struct ms_check_record {
    int64_t step;
    double dial_factor;
    double max_force_ratio;
    double max_speed_ratio;
    int32_t max_force_node;
    int32_t max_speed_node;
    int32_t at_equilibrium;
    int32_t kind;            /* 0 equilibrium check, 1 dialling phase ended */
} ms_check_record;
int ms_set_check_log(ms_ctx* ctx, int32_t capacity);  /* 0 disables */
int ms_read_check_log(ms_ctx* ctx, int32_t shell, ms_check_record* out, int32_t max_records, int32_t* count,
                      int64_t* dropped);
/* Arm the protocol for one shell (the current state is the start). */
int ms_relax_setup(ms_ctx* ctx, int32_t shell, const ms_relax_params* p);
/* Step all shells until every armed shell has halted or max_steps elapsed,
 * looking at the device every poll_every steps.  reports: num_shells entries
 * (may be NULL).  Returns MS_OK even if some shells blew up (see reports). */
int ms_relax_run(ms_ctx* ctx, int64_t max_steps, int64_t poll_every, ms_relax_report* reports);
/* Un-halt a shell (e.g. after the driver handled an equilibrium event). */
int ms_resume_shell(ms_ctx* ctx, int32_t shell);

/* ------------------------------------------------ device-path controls
 * The declarations from here to the warm start have no counterpart in the
 * reference: launch shapes, timing and measurement aids of the device path.
 *
 * Launch-shape knobs of the per-step kernels.  The autotuner in the Python
 * package searches these and caches the choice.  In ms_set_tuning a block
 * size or graph_steps of 0, or a grid_per_cu of -1, keeps the current value.
 * A fused shape whose tile does not fit the LDS is refused (MS_ERR_ARG); the
 * context then keeps the new block sizes in the two-launch form (fused 0). */
typedef struct ms_tuning {
    int32_t tri_block;        /* threads (= triangles) per block: 64, 128 or 256 */
    int32_t node_block;       /* threads (= nodes) per block: 64, 128, 256 or 512 */
    int32_t tri_grid_per_cu;  /* persistent grid: blocks per CU; 0 = one block per chunk */
    int32_t node_grid_per_cu;
    int32_t graph_steps;      /* steps captured per HIP graph */
    int32_t fused;            /* 1 = one owner-computes tile launch per step (a
                                 tile is tile_chunks node chunks; its workgroup recomputes
                                 the halo triangles and gathers from LDS);
                                 0 = triangle launch + node launch; -1 keeps.
                                 All give bit-identical states. */
    int32_t tile_chunks;      /* node chunks per tile of the fused step: 1 or 2
                                 (larger tiles recompute fewer halo triangles);
                                 0 keeps */
    int32_t tile_nodes;       /* nodes per tile of the fused step: each shell's
                                 nodes are cut into groups of this many (a tile),
                                 and each group into node_block chunks.
                                 > 0: that size, at most tile_chunks * node_block;
                                 -1: tile_chunks * node_block (the default);
                                 -2: balanced: when full tiles are fewer than
                                 the launch's workgroups, the smallest size (at
                                 least 64) whose tiles still fit one round of
                                 them, so that no workgroup idles; else full;
                                 0 keeps.  States are bit-identical for every
                                 choice.  ms_get_tuning reports the size in
                                 force. */
} ms_tuning;
int ms_set_tuning(ms_ctx* ctx, const ms_tuning* t);
int ms_get_tuning(const ms_ctx* ctx, ms_tuning* t);

/* Average device time (ms) of the last ms_step's kernels, measured with HIP
 * events on the context's stream: [0] triangle kernel, [1] node kernel,
 * [2] whole step.  Only filled when timing is enabled. */
int ms_enable_timing(ms_ctx* ctx, int32_t on);
int ms_last_step_timing(const ms_ctx* ctx, double out_ms[3]);

/* Bytes of device memory the context holds. */
int64_t ms_device_bytes(const ms_ctx* ctx);
/* Wait for all work of the context's stream (hipStreamSynchronize). */
int ms_synchronize(ms_ctx* ctx);

/* Measurement aid: reps launches of a plain device copy kernel of nbytes
 * (read nbytes, write nbytes), so that hardware traffic counters can be
 * scaled against a kernel of known traffic.  out_ms (may be NULL) receives
 * the average kernel time. */
int ms_calibration_copy(int64_t nbytes, int32_t reps, double* out_ms);

/* Measurement aid: reps launches of a read-only stream over nbytes (the
 * ceiling of a kernel that mostly reads).  out_ms as above. */
int ms_calibration_read(int64_t nbytes, int32_t reps, double* out_ms);

/* Measurement aid: the current device as the HIP runtime reports it: name
 * (its architecture name where the marketing name is empty; NUL-terminated,
 * at most name_len bytes), compute units, HBM bytes, and the memory clock
 * (kHz) and bus width (bits) of its properties.  Any output pointer may be
 * NULL. */
int ms_device_info(char* name, int32_t name_len, int32_t* num_cus, int64_t* hbm_bytes, int32_t* mem_clock_khz,
                   int32_t* mem_bus_bits);

/* ====================================================== learned warm start
 * A mesh message-passing network that predicts near-equilibrium node
 * positions from the flat reference mesh and the programmed metric.  It is
 * the source of start states that the reference takes from an ansatz file
 * (move_nodes_to_ansatz_positions.cpp:26-39, read_vtk_data.cpp:380-445); the
 * damped dynamics always finish the relaxation.  Forward and backward run on
 * the GPU; GEMMs use bfloat16 matrix-core operands with FP32 accumulation;
 * parameters, Adam moments and activations are FP32. */
typedef struct ms_nn ms_nn;

typedef struct ms_nn_arch {
    int32_t node_features;
    int32_t edge_features;
    int32_t width;       /* hidden width of every MLP */
    int32_t layers;      /* message-passing layers */
    int32_t out_dim;     /* 3: (u_x, u_y, z) per node */
    int32_t reserved;
    uint64_t seed;       /* parameter initialisation */
} ms_nn_arch;

/* A batch of meshes as one disjoint graph. Directed edges (both directions
 * of every mesh edge) sorted by destination; node ids are global. */
typedef struct ms_nn_graph_desc {
    int32_t num_graphs;
    int32_t num_nodes;
    int32_t num_edges;
    int32_t reserved;
    const int32_t* node_offsets;  /* num_graphs + 1 */
    const int32_t* edge_src;      /* num_edges */
    const int32_t* edge_dst;      /* num_edges, non-decreasing */
    const float* node_feat;       /* num_nodes x node_features */
    const float* edge_feat;       /* num_edges x edge_features */
} ms_nn_graph_desc;

int ms_nn_create(const ms_nn_arch* arch, ms_nn** out);
void ms_nn_destroy(ms_nn* nn);
const char* ms_nn_last_error(const ms_nn* nn);
int64_t ms_nn_num_params(const ms_nn* nn);
int ms_nn_get_params(ms_nn* nn, float* params);
/* Also resets the Adam moments. */
int ms_nn_set_params(ms_nn* nn, const float* params);
int ms_nn_set_graph(ms_nn* nn, const ms_nn_graph_desc* g);
/* Predictions for the current graph (num_nodes x 3); out may be NULL. */
int ms_nn_forward(ms_nn* nn, float* out);
/* Inference without the training state: the predictions (num_nodes x 3)
 * for graph g, equal bit for bit to ms_nn_set_graph + ms_nn_forward.  Each
 * half-layer is one fused launch (the concatenated rows gathered into the
 * first GEMM's A tile, the in-edge sums in CSR order, both weight matrices
 * in LDS, bias and SiLU in the epilogues); device memory is (2 E + N) x width
 * floats plus the graph, against some 36 edge-sized training buffers.  The
 * graph set by ms_nn_set_graph is kept.  Width 32 or 64.  Its output is the
 * start state that replaces the ansatz of move_nodes_to_ansatz_positions.cpp:26-39. */
int ms_nn_predict(ms_nn* nn, const ms_nn_graph_desc* g, float* out);
/* After ms_nn_forward: loss = sum over graphs of the smaller of the squared
 * errors against target and against target reflected in z, / (3 num_nodes),
 * and its gradient w.r.t. every parameter (grads may be NULL). */
int ms_nn_loss_backward(ms_nn* nn, const float* target, float* loss, float* grads);
int ms_nn_adam_step(ms_nn* nn, float lr, float beta1, float beta2, float eps, float weight_decay);
/* Measurement/test aid: C (M x N) = A @ B through the network's GEMM kernel
 * (A row-major M x K, or K x M read transposed; split-K over ksplit blocks). */
int ms_nn_gemm_test(int32_t M, int32_t N, int32_t K, int32_t transpose_a, int32_t ksplit, const float* A,
                    const float* B, float* C);

#ifdef __cplusplus
}
#endif

#endif /* MORPHOSHELL_AMD_H */
