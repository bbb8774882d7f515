/*
 * C-ABI of the MI355X SolverFEM substep library (libfem_hip.so).
 *
 * Plain pointers and sizes only.  "dev" pointers are HIP device pointers
 * (e.g. torch.Tensor.data_ptr() of cuda tensors), "host" pointers are CPU
 * memory.  Vectors of nodes are AoS float32 [P][3]; 3x3 blocks are row-major
 * float32 [9].  Streams are hipStream_t passed as void*.  Every function
 * returns FEM_OK (0) or an error code; fem_last_error() describes the last
 * failure of the calling thread.
 *
 * Entry points that run per substep (fem_restore_fixed, fem_assemble,
 * fem_step, fem_eval_forces) allocate nothing, never synchronize and are
 * safe to record into a HIP graph.
 *
 * Reference: nvidia-beer/newton-solver-fem, src/newton_solver_fem/_src/
 * solvers/fem/ (paths below are relative to that directory).
 */
#ifndef FEM_HIP_H
#define FEM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fem_system fem_system;

enum {
  FEM_OK = 0,
  FEM_ERR_ARG = 1,         /* invalid argument (shape, range, pointer)        */
  FEM_ERR_HIP = 2,         /* a HIP runtime call failed                       */
  FEM_ERR_UNSUPPORTED = 3, /* mesh outside the library's limits               */
  FEM_ERR_STATE = 4        /* call order misuse (e.g. buffers not bound)      */
};

/* Force terms for fem_eval_forces (solver.py:1357-1556 evaluators). */
enum {
  FEM_TERM_TET = 1,
  FEM_TERM_HEX = 2,
  FEM_TERM_GROUND = 4,
  FEM_TERM_GRAVITY = 8,
  FEM_TERM_DAMPING = 16,
  FEM_TERM_SPRING = 32,
  FEM_TERM_TRI = 64,
  FEM_TERM_PARTICLE = 128, /* hash-grid particle-particle contact (fem_set_particle_contact) */
  FEM_TERM_SOFT = 256      /* particle-shape soft contacts of a contact list (fem_set_soft_contacts) */
};

/* Device-side solve status bits (fem_solve_stats.error_flags). */
enum { FEM_STATUS_BARRIER_TIMEOUT = 1 };

/* Model arrays, all HOST pointers, copied at creation.  Mirrors the Model /
 * LinearFEMData fields SolverFEM.__init__ consumes (solver.py:212-341,
 * builder.py:12-20): particle q/mass/inv_mass/radius/flags, tet
 * indices/poses/materials, hex indices/inv_J0/det_J0_w/materials, the
 * sorted unique fixed_particle_indices (solver.py:244-249), linear springs
 * (Model.spring_indices/rest_length/stiffness/damping, eval_springs
 * kernels.py:176-219) and triangle membranes (Model.tri_indices/poses/
 * materials, eval_triangles kernels.py:876-970).  Independent meshes may be
 * packed back to back: instance_offsets[k]..[k+1] are the particles of
 * instance k (instance_count 0 or 1 means one instance). */
typedef struct fem_model_desc {
  int32_t particle_count;
  const float* particle_q;        /* [P][3] construction positions (anchors) */
  const float* particle_mass;     /* [P] */
  const float* particle_inv_mass; /* [P] */
  const float* particle_radius;   /* [P] */
  const int32_t* particle_flags;  /* [P], bit 0 = ACTIVE */
  int32_t tet_count;
  const int32_t* tet_indices;     /* [T][4] */
  const float* tet_poses;         /* [T][9] inverse rest shape matrix */
  const float* tet_materials;     /* [T][3] mu, lambda, k_damp */
  int32_t hex_count;
  const int32_t* hex_indices;     /* [H][8] VTK order */
  const float* hex_inv_J0;        /* [H][8][9] */
  const float* hex_det_J0_w;      /* [H][8] */
  const float* hex_materials;     /* [H][3] */
  int32_t fixed_count;
  const int32_t* fixed_indices;   /* [fixed_count], sorted, unique */
  int32_t instance_count;
  const int32_t* instance_offsets; /* [instance_count+1] or NULL */
  int32_t spring_count;
  const int32_t* spring_indices;  /* [S][2]; springs with a -1 endpoint are ignored */
  const float* spring_rest_length; /* [S] */
  const float* spring_stiffness;  /* [S] */
  const float* spring_damping;    /* [S] */
  int32_t tri_count;
  const int32_t* tri_indices;     /* [R][3] */
  const float* tri_poses;         /* [R][2][2] inverse rest edge matrix, row-major */
  const float* tri_materials;     /* [R][5] k_mu, k_lambda, k_damp, k_drag, k_lift */
} fem_model_desc;

/* Physical and solver parameters (SolverFEM.__init__ arguments,
 * solver.py:212-240; model.gravity). */
typedef struct fem_params {
  float min_stretch, max_stretch;
  int32_t has_ground;
  float ground[4];
  float ground_ke, ground_kd, ground_kf, ground_mu;
  float linear_damping;
  float gravity[3];
  int32_t enable_tet_activations;
  float tolerance;          /* linear_solver_tolerance (relative) */
  int32_t maxiter;          /* linear_solver_maxiter */
  int32_t early_exit;       /* 1: stop at the tolerance; 0: full budget, residual-masked updates */
} fem_params;

/* Replaces SolverFEM.__init__'s pattern construction (solver.py:312-341:
 * triplet layout + bsr_set_from_triplets) and sparse_update.map_triplet_blocks
 * (sparse_update.py:9-30): builds the BSR pattern of A = M + h^2 K and the
 * per-node element incidence/slot lists on the host, uploads model data. */
int fem_create(const fem_model_desc* desc, fem_system** out);
int fem_destroy(fem_system* sys);
const char* fem_last_error(void);

/* Number of stored 3x3 blocks (A_bsr.nnz) and the largest row length. */
int fem_pattern_size(const fem_system* sys, int64_t* nnz, int32_t* max_row_blocks);

/* Device allocations the library holds for this system and their bytes (no
 * reference counterpart: lets a caller check that refreshes allocate
 * nothing). */
int fem_memory_stats(const fem_system* sys, int64_t* allocations, int64_t* bytes);

/* Binds caller-owned device buffers that stay valid for the system's life:
 * offsets [P+1] int32, columns [capacity] int32 (offsets and the first nnz
 * columns filled here, synchronously), values [capacity][9], inv_diag [P][9]
 * (block-Jacobi inverse, SolverFEM _pc_inv_diag) and increment [P][3]
 * (SolverFEM.dv warm start).  block_capacity (0: nnz) is the number of blocks
 * the columns/values storage holds; a topology refresh without a reserved
 * pool re-sorts the pattern within it, as A_bsr's nnz_capacity of
 * P + 2S + 3R + 16T + 64H blocks does in the reference (solver.py:318-325). */
int fem_bind(fem_system* sys, int32_t* offsets_dev, int32_t* columns_dev, float* values_dev, float* inv_diag_dev,
             float* increment_dev, int64_t block_capacity);

int fem_set_params(fem_system* sys, const fem_params* params);

/* Particle-particle contact, the "particle" term of _accumulate_forces
 * (solver.py:570-589): SolverFEM.eval_particle_particle_forces
 * (solver.py:1477-1510) builds a hash grid of cell width 2 max_radius +
 * cohesion at the step's positions and eval_particle_forces
 * (kernels.py:1069-1114, particle_force :1043-1066) sums penalty (ke),
 * damping (kd) and Coulomb friction (kf, mu) over every active pair with
 * distance - r_i - r_j <= cohesion.  Replaces the Model.particle_ke/kd/kf/mu/
 * cohesion fields and Model.particle_grid: after this call every step and
 * fem_eval_forces with FEM_TERM_PARTICLE evaluate the term, with the grid's
 * storage allocated here (once).  ke <= 0, fewer than two particles or no
 * positive radius switch the term off, as the reference skips it.  Radii
 * and flags are the ones given to fem_create. */
int fem_set_particle_contact(fem_system* sys, float ke, float kd, float kf, float mu, float cohesion);

/* Soft contacts, the "soft_contact" term of _accumulate_forces
 * (solver.py:570-589): SolverFEM.eval_soft_contact_forces (solver.py:1540-1566)
 * launches eval_soft_contacts (kernels.py:1199-1252) over the collision
 * pipeline's contacts.soft_contact_* arrays with Model.soft_contact_ke/kd/kf/mu
 * and the particle radii.  All arrays are DEVICE pointers that stay bound
 * until the next call: count (int[1], read on the device at every
 * evaluation, so a recorded graph follows it), particle (int[max_contacts];
 * negative entries are skipped), body_pos, body_vel and normal
 * (float[max_contacts][3], normal pointing from the body to the particle).
 * Every later step and fem_eval_forces with FEM_TERM_SOFT add the term;
 * scratch for the largest max_contacts seen is allocated here.  A NULL count
 * or max_contacts <= 0 switches the term off.  Each particle sums its
 * contacts in contact order (the reference adds them with atomics). */
int fem_set_soft_contacts(fem_system* sys, const int* count, const int* particle, const float* body_pos,
                          const float* body_vel, const float* normal, int max_contacts, float ke, float kd, float kf,
                          float mu);

/* Spring bend torque, SolverFEM(torque_stiffness, torque_damping,
 * spring_rest_direction) (solver.py:227-229, :390-401; the torque part of
 * eval_springs_linear_and_torque kernels.py:222-292): rest_direction is a
 * HOST float[spring_count][3] copied here (springs whose direction has
 * length < 0.5 get no torque); stiffness <= 0 (or a NULL direction array)
 * switches the torque off.  The spring term of every later force evaluation
 * adds the torque force; it has no tangent block, as in the reference. */
int fem_set_spring_torque(fem_system* sys, const float* rest_direction, float stiffness, float damping);

/* Linear solver of the step, SolverFEM(solver_type=..., preconditioner_type=
 * ...) (solver.py:212-218, _make_linear_solver :1219-1252, _build_
 * preconditioner :1142-1173).  method: 0 cg (the persistent PCG, default),
 * 1 bicgstab, 2 cr, 3 gmres (restarted after `restart` iterations, 1..64;
 * warp.optim.linear's default is 31); preconditioner: 0 block_diag (default),
 * 1 diag, 2 diag_abs (inverse of the diagonal entries of A's diagonal blocks,
 * or of their magnitudes; 1 where zero or absent), 3 id (none).  The
 * methods besides cg run a fixed sequence of launches per iteration with
 * residual-masked updates (no host synchronisation, graph-capturable).
 * Batched instances, reserved pools, component scaling and
 * fem_step_solve_static need cg with block_diag.  Allocates the work
 * storage; call between steps. */
int fem_set_solver(fem_system* sys, int32_t method, int32_t preconditioner, int32_t restart);

/* Gravity read from caller-owned device memory (float[3]) by every later
 * kernel instead of fem_params.gravity, so that a recorded graph follows
 * changes made between replays: the reference's Model.gravity is a device
 * array read by eval_gravity_from_array and the ground friction cap
 * (kernels.py:1132-1195, :1273-1286).  NULL returns to fem_params.gravity. */
int fem_bind_gravity(fem_system* sys, const float* gravity_dev);

/* Launch shapes (0 keeps the current value): lanes per block row of the
 * force/RHS kernel and the maximum rows per workgroup of the assembly (powers
 * of two in [4, 64]; consecutive rows are packed up to that count while their
 * element incidences fit one pass of the workgroup), and the persistent CG
 * workgroup count (0 = automatic). */
int fem_set_launch(fem_system* sys, int32_t elem_lanes, int32_t assemble_rows, int32_t cg_workgroups);
/* Current launch shapes.  cg_workgroups counts the CG chunks (one workgroup
 * solves a chunk); the solve is one persistent launch of
 * cg_launch_workgroups workgroups: one per chunk when they all fit on the
 * device at once, else as many as fit, taking chunks of whole instances from
 * a device queue as their instances converge.  cg_variant is
 * FEM_CG_STREAMED or FEM_CG_RESIDENT. */
int fem_get_launch(const fem_system* sys, int32_t* elem_lanes, int32_t* assemble_rows, int32_t* cg_workgroups,
                   int32_t* cg_rows_per_thread, int32_t* cg_launch_workgroups, int32_t* cg_variant);

/* Persistent CG kernel variant (the same iteration as warp.optim.linear.cg's
 * block-Jacobi PCG, linear.py cg, as called by SolverFEM._solve_linear_system
 * solver.py:1206-1286; the choice changes only the fp32 summation order):
 * FEM_CG_STREAMED reads the matrix from HBM in every iteration;
 * FEM_CG_RESIDENT keeps each workgroup's blocks in registers and LDS for the
 * whole solve (one workgroup per CU; FEM_ERR_UNSUPPORTED when a chunk does
 * not fit); FEM_CG_AUTO (default) picks resident whenever it fits. */
enum { FEM_CG_AUTO = 0, FEM_CG_STREAMED = 1, FEM_CG_RESIDENT = 2 };
int fem_set_cg_variant(fem_system* sys, int32_t variant);

/* FixedParticles.restore / restore_positions (fixed_particles.py:11-21,
 * 152-168): q[i] = anchor[i] (and qd[i] = 0 when qd_dev is not NULL). */
int fem_restore_fixed(fem_system* sys, float* q_dev, float* qd_dev, void* stream);

/* refresh_tangent's numeric assembly (solver.py:769-857 ->
 * build_system_matrix_tet_kernel kernels.py:635-850,
 * build_system_matrix_hex_kernel :1609-1747, mass diagonal :141-153,
 * accumulate_triplet_blocks sparse_update.py:33-59,
 * invert_block_diagonal_kernel kernels.py:1836-1849): writes A's values at
 * positions q and the block-Jacobi inverse.  q is read only; callers restore
 * fixed positions first (fem_restore_fixed). */
int fem_assemble(fem_system* sys, const float* q_dev, float dt, void* stream);

/* SolverFEM.step fast path (solver.py:942-1002, implicit_integration
 * :1254-1351): restores fixed q/qd in the input state, writes the projected
 * right-hand side h*sum(f) + h*ext - (A-M)v into f_out (state_in.particle_f),
 * solves (PAP+I-P) dv = P b with block-Jacobi PCG (relative tolerance,
 * atol 0, warm start from the bound increment) and integrates
 * v' = v + dv, q' = q + h v' into q_out/qd_out.  ext_dev and act_dev may be
 * NULL. */
int fem_step(fem_system* sys, float* q_in_dev, float* qd_in_dev, float* f_out_dev, const float* ext_dev,
             const float* act_dev, float* q_out_dev, float* qd_out_dev, float dt, void* stream);

/* The two halves of fem_step, for callers that overlap or time them
 * separately.  fem_step_rhs: fixed restore of the input state, forces and
 * predictor -> f_out (solver.py:987-999, 1268-1299).  fem_step_solve: RHS
 * normalization, PCG and integration (solver.py:1301-1351); f_in is the
 * f_out of the preceding fem_step_rhs. */
int fem_step_rhs(fem_system* sys, float* q_in_dev, float* qd_in_dev, float* f_out_dev, const float* ext_dev,
                 const float* act_dev, float dt, void* stream);
int fem_step_solve(fem_system* sys, const float* q_in_dev, const float* qd_in_dev, const float* f_in_dev,
                   float* q_out_dev, float* qd_out_dev, float dt, void* stream);

/* SolverFEM.implicit_integration's right-hand side (solver.py:1254-1307):
 * f_dev holds the caller's h-scaled force sum (state_in.particle_f, forces
 * and external loads already multiplied by dt); it is overwritten, in place,
 * with that sum minus the elastic predictor (subtract_elastic_predictor
 * implicit_predictor.py:9-23), plus the spring-damping restore (:26-41),
 * fixed rows zeroed.  fem_step_solve(f_in = f_dev) then finishes the
 * integration.  The input state is not touched. */
int fem_integrate_rhs(fem_system* sys, const float* qd_in_dev, float* f_dev, float dt, void* stream);

/* fem_step_solve without library-allocated scratch, for callers that cannot
 * allocate device memory (code under HIP graph capture that must not touch
 * the allocator, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  The persistent CG's partial sums,
 * publication flags and exchange records live in a static device array of
 * 32 MiB per GPU (held until the process ends) instead of the system's own
 * buffers: 72 bytes per row plus 132 bytes per CG chunk (and the chunk
 * queue word), so up to about 460,000 rows; a larger system returns
 * FEM_ERR_UNSUPPORTED before any
 * launch.  The call allocates and frees nothing; the flags it needs zeroed
 * are cleared by a kernel on the stream.  Every call on a GPU shares that
 * GPU's array: calls that overlap on one GPU (other streams, other systems)
 * race and must be serialized by the caller.  Results equal fem_step_solve's
 * bit for bit. */
int fem_step_solve_static(fem_system* sys, const float* q_in_dev, const float* qd_in_dev, const float* f_in_dev,
                          float* q_out_dev, float* qd_out_dev, float dt, void* stream);

/* Sum of the selected unscaled force terms (FEM_TERM_* mask) at (q, qd):
 * eval_tetrahedra + eval_tet_activations (kernels.py:484-632),
 * eval_hexahedra (:1481-1601), eval_particle_ground_contacts (:1132-1195),
 * eval_gravity_from_array (:1273-1286), eval_linear_damping_kernel (:1289-1307). */
int fem_eval_forces(fem_system* sys, int32_t terms, const float* q_dev, const float* qd_dev, const float* act_dev,
                    float* out_dev, void* stream);

/* Last solve's per-instance statistics (synchronizes the stream): CG
 * iterations performed, final squared residual, squared absolute tolerance,
 * and FEM_STATUS_* error flags.  Arrays have instance_count entries.  A solve
 * whose persistent launch was not fully co-resident (another kernel or
 * process holding CUs) ends its bounded waits with
 * FEM_STATUS_BARRIER_TIMEOUT and writes NaN into the increment and the
 * output state (or the solution of fem_solve) instead of a partial solve, so
 * an unsynchronized or recorded step cannot return a plausible wrong state;
 * the warm start stays NaN until it is cleared. */
int fem_solve_stats(fem_system* sys, void* stream, int32_t* iterations, float* residual_sq, float* atol_sq,
                    int32_t* error_flags);

/* ---- Prepared affine mechanics (PreparedFEMStep, prepared_step.py:64-302) --
 * The prepared right-hand side without extra load is fem_step_rhs on a copy of
 * the input state (solver.prepare_step, solver.py:483-500). */

/* apply_operator (prepared_step.py:233-241): y = A x, or (PAP + I - P) x with
 * fixed particles.  x and y must differ. */
int fem_apply_operator(fem_system* sys, const float* x_dev, float* y_dev, void* stream);

/* apply_preconditioner (:243-251): y = M^-1 x with the block-Jacobi inverse. */
int fem_apply_preconditioner(fem_system* sys, const float* x_dev, float* y_dev, void* stream);

/* Bound of every wait of the persistent CG, in polls of the awaited flag (a
 * device-scope load and a short sleep each; 0 restores the default 2^22,
 * which bounds a wait to seconds).
 * No counterpart in the reference, whose CG needs no cross-workgroup waits;
 * a small bound forces the barrier-timeout path of a multi-workgroup solve
 * from the host (failure-handling tests). */
int fem_set_barrier_spin_limit(fem_system* sys, int64_t spins);

/* The preconditioner operator SolverFEM.M_bsr (solver.py:1175-1201) applied as
 * LinearOperator.matvec(x, y, z, alpha, beta): z = alpha M^-1 x + beta y with
 * the block-Jacobi inverse (block_jacobi_mv_kernel kernels.py:1852-1868).  y
 * may be NULL when beta is 0; z may alias x or y. */
int fem_block_jacobi_mv(fem_system* sys, const float* x_dev, const float* y_dev, float* z_dev, float alpha,
                        float beta, void* stream);

/* diagonal_blocks (_diagonal :46-61): out[i] = A_ii (zero when the row has no
 * diagonal block), the identity on fixed rows; [P][9]. */
int fem_diagonal_blocks(fem_system* sys, float* out_dev, void* stream);

/* rhs(external_forces) (:197-231): out = base + h ext with fixed rows zero
 * (ext may be NULL; out may alias base but not ext). */
int fem_add_load(fem_system* sys, const float* base_dev, const float* ext_dev, float dt, float* out_dev,
                 void* stream);

/* solve (:268-289): the step's normalised block-Jacobi PCG (fem_set_params
 * tolerance, budget and exit rule) on b with the warm start and result in
 * dv_dev (a caller-owned increment, not the bound one); no state update. */
int fem_solve_increment(fem_system* sys, const float* b_dev, float* dv_dev, void* stream);

/* reconstruct (:291-302): v' = qd_in + dv, q' = q_in + h v' (update_state
 * kernels.py:115-128), fixed particles restored; outputs must not alias inputs. */
int fem_reconstruct(fem_system* sys, const float* dv_dev, const float* q_in_dev, const float* qd_in_dev,
                    float* q_out_dev, float* qd_out_dev, float dt, void* stream);

/* ---- Differentiable substep (differentiation.py:149-394) -----------------
 * The matrix of a recorded step is the current assembly (a fixed input, as in
 * the reference's MatrixSnapshot).  The forward right-hand side of a
 * differentiable step is fem_step_rhs's (no fixed particles in this mode).
 *
 * MatrixSnapshot.solve (:208-232): x = A^-1 b by the block-Jacobi PCG from a
 * zero start (maximum_rhs_component / normalize_linear_rhs /
 * restore_linear_increment), relative tolerance `tolerance`, atol 0, device
 * residual exit; then the FP64 residual ||A x - b|| / ||b|| and its ratio to
 * the budget 2e-5 ||b|| + 1e-10 update running maxima for `kind`
 * (0 = forward, 1 = adjoint; residual_norms / record_residual :118-146).
 * A is symmetric, so the adjoint solve (bsr_transposed) uses the same call.
 * With fixed particles the solve (and its residual) uses PAP + I - P, as the
 * fast path does; the reference rejects fixed particles when differentiable,
 * so that case is an extension (b must be zero on fixed rows). */
int fem_solve(fem_system* sys, const float* b_dev, float* x_dev, float tolerance, int32_t maxiter, int32_t kind,
              void* stream);

/* Reverse products of a differentiable step for the adjoint z = A^-1 dv_bar
 * (what the reference's tape computes through sum_forces / form_rhs and the
 * force kernels, :62-111, with clamp_deformation_stretch_grad kernels.py:447-481):
 *   q_bar  = h (d terms / d q)^T z                       (written, P x 3)
 *   qd_bar = h (d terms / d qd)^T z - A z + rowsum(A)^T z  (written, P x 3)
 *   act_bar[e] = h (d t_tet / d act_e)^T z               (written if not NULL)
 *   mat_bar[e] = d/d(mu, lambda, k_damp)_e of  h z^T terms  - z^T (dA/d(mu, lambda))(dv + qd),
 *     with A assembled at q_asm (written if not NULL; needs dv and q_asm).
 * The reference has no material gradients (tet_materials must not require
 * gradients, differentiation.py:244-262); mat_bar extends it.  Springs
 * (force and spring-damping predictor restore) and triangle membranes are
 * differentiated; models with triangle drag or lift return
 * FEM_ERR_UNSUPPORTED. */
int fem_vjp(fem_system* sys, const float* q_dev, const float* qd_dev, const float* act_dev, const float* z_dev,
            const float* dv_dev, const float* q_asm_dev, float dt, float* q_bar_dev, float* qd_bar_dev,
            float* act_bar_dev, float* mat_bar_dev, void* stream);

/* Replace the per-tet (mu, lambda, k_damp) copied at creation with the T x 3
 * device array mat_dev (a material-optimisation step); the next fem_assemble
 * uses them.  Tets must keep their zero / non-zero material status (the
 * pattern excluded zero-material tets).  No reference counterpart: there the
 * materials are fixed model inputs. */
int fem_set_tet_materials(fem_system* sys, const float* mat_dev, void* stream);

/* check() (:269-279): synchronizes and copies the running maxima
 * {forward ratio, adjoint ratio, forward violation, adjoint violation};
 * reset != 0 clears them. */
int fem_diff_stats(fem_system* sys, void* stream, double* out4, int32_t reset);

/* ---- Topology updates (solver.py:695-720, :870-935; streaming.py) ---------
 *
 * fem_prepare_topology replaces prepare_tetrahedral_updates
 * (solver.py:695-720, StreamingTetAssembly.__init__ streaming.py:487-597): the
 * P particles are P / original_node_count copies of the original nodes (pool
 * node i stands for original node i % original_node_count); the pattern
 * becomes the reserved one -- row i holds i and every copy of every original
 * neighbour of i % original_node_count in the current tets -- with room for the
 * incidences of any T tets over those rows.  Tets only, no fixed particles,
 * before component scaling.  The pattern changes, so fem_pattern_size /
 * fem_bind must be called again, then fem_assemble at the current positions.
 * Captured launches made after that stay valid across fem_update_topology.
 * Errors: FEM_ERR_ARG for a pool that is not whole copies, tet nodes outside
 * the pool or a tet with a repeated original node; FEM_ERR_STATE when
 * already prepared or component scaling is bound. */
int fem_prepare_topology(fem_system* sys, int32_t original_node_count);

/* Replaces refresh_topology (solver.py:870-935): publishes edited tets
 * (T x 4 indices, T x 9 rest inverses, T x 3 materials; zero materials for
 * unused slots), particle masses / inverse masses / flags (P each), all
 * device arrays of the model's fixed capacities, and assembles A and M^-1 at
 * q.  With a reserved pool (StreamingTetAssembly.update streaming.py:624-716)
 * the edit is validated first and rejected without changing anything
 * (FEM_ERR_ARG with the reference's message: an element source outside
 * [-1, T), a node outside the pool, an edge outside the reserved adjacency, a
 * repeated original node); element_sources (T int32, NULL: identity) names
 * each slot's previous element or -1.  stats (8 int32, may be NULL) receives
 * the reference's update counters {dirty, removed, reused, active, touched
 * rows, topology-inserted, topology-removed, full numeric refresh}.  The pool's
 * pattern, matrix storage and launch shapes do not change (fem_set_launch and
 * fem_set_cg_variant return FEM_ERR_STATE once a pool is prepared); the CG runs
 * over the live slots -- diagonals and slots with element contributions --
 * when the resident kernel holds the reserved rows.  Without a pool the
 * exact pattern of the edited model is built again (within the bound block
 * capacity; offsets and columns are rewritten and fem_pattern_size reports the
 * new nnz), as the reference re-sorts into the same A_bsr storage
 * (solver.py:920-935).  When the bound capacity exceeds the initial pattern
 * (fem_bind), the pattern arrays, the resident CG plan's storage and the
 * assembly launch are sized for it, and an edit that fits them is written in
 * place: the same device arrays and launches, so launches captured before the
 * edit replay the edited topology, and nothing is allocated.  An edit that
 * does not fit (the streamed CG plan, more assembly groups or CG columns than
 * reserved) gets new arrays; the old ones stay allocated until fem_destroy
 * (launches captured before still name them and step the old topology, so
 * they must be captured again) and stats[0] is set to 1 (else 0; the other
 * counters are 0 without a pool).  An unchanged topology allocates nothing.
 * Synchronizes the stream.  Not for batched systems. */
int fem_update_topology(fem_system* sys, const int32_t* tet_indices_dev, const float* tet_poses_dev,
                        const float* tet_materials_dev, const float* particle_mass_dev,
                        const float* particle_inv_mass_dev, const int32_t* particle_flags_dev, const float* q_dev,
                        const int32_t* element_sources_dev, float dt, int32_t* stats, void* stream);

/* ---- Component scaling (component_scaling.py; solver.py:722-767) ---------
 *
 * fem_bind_component_scaling replaces prepare_component_scaling: caller-owned
 * device storage for the per-component fp64 total masses [capacity], the
 * per-row factors [P], and the scaled copies of the matrix values (the bound
 * block capacity x 9, sharing the matrix's offsets and columns), of the
 * inverse diagonal blocks [P][9] and of the right-hand side [P][3].  From then
 * on every fem_step_solve solves (S A) dv = S b with preconditioner M^-1 S^-1,
 * S = diag(factor), whose exact solution is the physical dv
 * (ComponentScaling.prepare, solver.py:1300-1304).  No fixed particles.
 * fem_refresh_component_scaling replaces refresh_component_scaling: factors
 * from the P int32 labels (-1: unused row; a label outside [0, capacity)
 * keeps factor 1) and the P particle masses: 1 / the label's total mass, 1
 * where that total is not positive. */
int fem_bind_component_scaling(fem_system* sys, int32_t capacity, double* totals_dev, float* scale_dev,
                               float* values_dev, float* inv_diag_dev, float* rhs_dev);
int fem_refresh_component_scaling(fem_system* sys, const int32_t* labels_dev, const float* particle_mass_dev,
                                  void* stream);

#ifdef __cplusplus
}
#endif

#endif /* FEM_HIP_H */
