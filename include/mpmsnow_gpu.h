/* mpmsnow_gpu.h -- C-ABI of the MI355X (gfx950) implicit MPM snow step.
 *
 * This is the drop-in boundary under the MpmSim library interfaces of
 * davidsminor/mpmsnow (the headers in include/MpmSim of the reference): every
 * entry point below carries out one member function of those classes on the
 * GPU, and is cited to the reference declaration it replaces. Plain C types
 * only: float/int arrays in the reference's host layouts (Vector3f = 3
 * floats, Matrix3f = 9 floats column-major, dense grid vectors in node order
 * i + n0 (j + n1 k) with xyz interleaved), copied in and out (no aliasing).
 *
 * Every function returns MPM_OK (0) or an error code; mpm_last_error()
 * returns the message, which for the reference's own failures is the
 * reference's exact std::runtime_error text (e.g. "grid is too big",
 * "nan in explicit momenta!"). One context = one HIP stream = one caller
 * thread; calls return when their results are complete.
 */
#ifndef MPMSNOW_GPU_H
#define MPMSNOW_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  MPM_OK = 0,
  MPM_ERR_INVALID = 1,   /* bad argument / unknown name / wrong state */
  MPM_ERR_RUNTIME = 2,   /* a reference runtime_error (message is the reference's text) */
  MPM_ERR_HIP = 3,       /* HIP runtime failure: "hip: <name>" */
  MPM_ERR_NO_DEVICE = 4
};

typedef struct mpm_context mpm_context;
typedef struct mpm_grid mpm_grid;

/* Collision object (reference include/MpmSim/CollisionObject.h:10-45):
 * type 0 = CollisionPlane (phi = c0 x + c1 y + c2 z + c3, CollisionPlane.h:14-40),
 * type 1 = analytic sphere (phi = |x - center| - radius),
 * type 2 = level-set volume, the reference's VDBCollisionObject
 *   (houdiniPlugin/VDBCollisionObject.cpp:11-37, which samples OpenVDB grids; that object uses
 *   friction 0.5 and is not sticky): phi at the points origin + (i, j, k) voxel, i < dims[0],
 *   j < dims[1], k < dims[2], point (i, j, k) at phi[i + dims[0] (j + dims[1] k)], sampled
 *   trilinearly (clamped to the volume); grad = central difference of the sampled phi over one
 *   voxel per axis; velocity = phi_velocity sampled the same way (3 floats per point), or the
 *   uniform `velocity` when phi_velocity is null. The arrays are copied by mpm_set_colliders. */
typedef struct {
  int type;
  float coeffs[4];
  float center[3];
  float radius;
  float velocity[3];
  float friction; /* coulombFriction() */
  int sticky;     /* sticky() */
  int dims[3];              /* type 2 */
  float origin[3];          /* type 2 */
  float voxel;              /* type 2 */
  const float* phi;         /* type 2: dims[0] dims[1] dims[2] floats */
  const float* phi_velocity; /* type 2, optional: 3 dims[0] dims[1] dims[2] floats */
} mpm_collider;

/* Solver settings: SquareMagnitudeTermination(maxIters, tolError)
 * (reference include/MpmSim/SquareMagnitudeTermination.h:10-26) plus optional
 * host hooks standing in for a custom TerminationCriterion / LinearSolver::Debug
 * (reference TerminationCriterion.h:10-16, LinearSolver.h:14-21). When a hook
 * is set the solve synchronises every iteration and passes dense vectors. */
typedef struct {
  int max_iters;
  float tolerance;
  /* termination(r, n, i): nonzero stops the solve; replaces the built-in test */
  int (*termination)(void* user, const float* r, int64_t n, int iteration);
  /* debug(x, n): called after every x update with the velocities relative to the colliders */
  void (*debug)(void* user, const float* x, int64_t n);
  /* cancelled(): polled after each body's solve (TerminationCriterion::cancelled) */
  int (*cancelled)(void* user);
  void* user;
} mpm_solver_params;

typedef struct {
  int iterations;          /* x updates (ConjugateResiduals.cpp:84-128) */
  int early_exit;          /* stopped on a zero denominator */
  double rhs_norm2;        /* |b|^2 */
  double residual_norm2;   /* last |r|^2 (r = M z) */
  int history_length;      /* entries written to history */
  int history_capacity;    /* caller-provided capacity of history */
  double* history;         /* caller-provided: |r|^2 per iteration, may be NULL */
  int particles;           /* particles of the body (grid API: of the grid) */
  int active_nodes;        /* grid nodes with m > 0 */
  int fallbacks;           /* learned preconditioner: 1 if the body switched to the diagonal one */
  int restarts;            /* CR restarts (p = z) where the expanded denominator lost its significance */
} mpm_solve_stats;

typedef struct {
  int num_bodies;          /* bodies for the next step (after the re-partition) */
  int num_ballistic;
  int total_iterations;
  int max_iterations;
  int cancelled;
  int num_solved;          /* bodies solved in this step (valid indices of mpm_sim_body_solve_stats) */
} mpm_step_stats;

/* ---- context ------------------------------------------------------------ */
int mpm_create(mpm_context** ctx, int device);
int mpm_destroy(mpm_context* ctx);
const char* mpm_last_error(const mpm_context* ctx);
int mpm_version(void);

/* ---- particles: MaterialPointData (reference MaterialPointData.h:14-44) ---
 * mpm_particles_create gives the default variables p, v (0), F (identity),
 * m (0), volume (0) (MaterialPointData.cpp:8-16, Sim.cpp:34-38).
 * Names: p v F m volume, and after mpm_model_snow: Fp FinvTrans R Ginv J mu
 * lambda, plus the per-particle material mu0 lambda0 xi. count = number of
 * floats (3n, 9n or n). */
int mpm_particles_create(mpm_context* ctx, int64_t n);
int64_t mpm_particles_count(const mpm_context* ctx);
int mpm_particles_num_variables(const mpm_context* ctx);
int mpm_particles_set(mpm_context* ctx, const char* name, const float* data, int64_t count);
int mpm_particles_get(mpm_context* ctx, const char* name, float* data, int64_t count);

/* ---- constitutive model: SnowConstitutiveModel (SnowConstitutiveModel.h:13-32) */
/* constructor + setParticles (SnowConstitutiveModel.cpp:9-25, 123-169) */
int mpm_model_snow(mpm_context* ctx, float youngs_modulus, float poisson_ratio, float hardening,
                   float compressive_strength, float tensile_strength);
/* updateParticleData over all particles (SnowConstitutiveModel.cpp:27-121) */
int mpm_model_update_particle_data(mpm_context* ctx);
/* energyDensity / dEnergyDensitydF / dEdFDifferential for a list of particles
 * (SnowConstitutiveModel.cpp:171-256); out: n, 9n, 9n floats */
int mpm_model_energy_density(mpm_context* ctx, const int32_t* particles, int64_t n, float* out);
int mpm_model_stress(mpm_context* ctx, const int32_t* particles, int64_t n, float* out);
int mpm_model_stress_differential(mpm_context* ctx, const int32_t* particles, const float* dF, int64_t n,
                                  float* out);
/* No counterpart in the reference. For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops): mpm_model_stress_differential with particles (n
 * int32), dF (9n floats) and out (9n floats) in device memory of the context's device, calling no
 * hipMalloc/hipFree. Its scratch (the slot list) is a fixed per-GPU device array, so n is at most
 * 1048576 (larger n: MPM_ERR_INVALID, nothing computed) and calls that overlap on one GPU race.
 * An index outside [0, particles) fails the call (MPM_ERR_INVALID, its rows of out are NaN). */
int mpm_model_stress_differential_noalloc(mpm_context* ctx, const int32_t* particles, const float* dF, int64_t n,
                                          float* out);

/* ---- colliders and force fields (CollisionObjectSet, ForceFieldSet) -------- */
int mpm_set_colliders(mpm_context* ctx, const mpm_collider* colliders, int n);
/* GravityField list: n vectors of 3 floats (GravityField.cpp:9-17) */
int mpm_set_gravity(mpm_context* ctx, const float* gravity, int n);

/* ---- grid of one body: Grid (reference include/MpmSim/Grid.h:23-217) ----- */
/* Grid::Grid (Grid.cpp:156-263): bbox, binning, mass and velocity splat */
int mpm_grid_create(mpm_context* ctx, const int32_t* particles, int64_t count, float grid_size,
                    const float frame_velocity[3], int dimension, mpm_grid** out);
int mpm_grid_destroy(mpm_grid* grid);
/* m_n, m_min */
int mpm_grid_dims(const mpm_grid* grid, int n[3], float min[3]);
/* names: "masses" (n floats), "velocities", "prev_velocities" (3n floats),
 * "node_collided" (n int8: -1 none, -2 several, >= 0 object);
 * get only: "base_cells" (3 int32 per particle, in the order of the particle list given to
 * mpm_grid_create): the stencil base node floor((x - m_min) / h) - 1 of each particle as the
 * binning kernel computed it (ShapeFunctionIterator::initialize, Grid.cpp:1105-1136) */
int mpm_grid_get(mpm_grid* grid, const char* name, void* out);
int mpm_grid_set(mpm_grid* grid, const char* name, const void* in);
/* computeParticleVolumes (Grid.cpp:265-292) */
int mpm_grid_compute_particle_volumes(mpm_grid* grid);
/* calculateForces (Grid.cpp:438-463) */
int mpm_grid_calculate_forces(mpm_grid* grid, float* forces);
/* calculateForceDifferentials (Grid.cpp:620-632) */
int mpm_grid_force_differentials(mpm_grid* grid, const float* dx, float* df);
/* dForceidXi (Grid.cpp:539-546) */
int mpm_grid_dforce_dxi(mpm_grid* grid, float* out);
/* calculateExplicitMomenta (Grid.cpp:634-690); collided: n int8 */
int mpm_grid_explicit_momenta(mpm_grid* grid, float dt, float* momenta, int8_t* collided);
/* collisionVelocities (Grid.cpp:692-724) */
int mpm_grid_collision_velocities(mpm_grid* grid, const int8_t* collided, float* vc);
/* updateGridVelocities (Grid.cpp:889-1016): implicit solve by diagonal-preconditioned CR */
int mpm_grid_update_grid_velocities(mpm_grid* grid, float dt, const mpm_solver_params* params,
                                    mpm_solve_stats* stats);
/* updateParticleVelocities (Grid.cpp:1049-1081) */
int mpm_grid_update_particle_velocities(mpm_grid* grid);
/* updateDeformationGradients (Grid.cpp:1018-1047) */
int mpm_grid_update_deformation_gradients(mpm_grid* grid, float dt);

/* ---- ConjugateResiduals over a caller's matrix --------------------------------
 * ConjugateResiduals(termination, preconditioner, log)(mat, rhs, x, debug)
 * (ConjugateResiduals.h:20-60, ConjugateResiduals.cpp:19-160) for a ProceduralMatrix the caller
 * implements (ProceduralMatrix.h:8-19): the recurrences run on the GPU over n-vectors (FP32, FP64
 * inner products in a fixed order); the matrix, the preconditioner and the projection are the
 * caller's callbacks, called with the library's work complete and expected complete on return. */
typedef struct {
  /* result = M x (multVector) */
  void (*mult)(void* user, const float* x, float* result, int64_t n);
  /* result = M^-1 x (multInverseVector); required of a preconditioner, unused for the matrix */
  void (*mult_inverse)(void* user, const float* x, float* result, int64_t n);
  /* x = P x in place (subspaceProject); NULL: the identity */
  void (*project)(void* user, float* x, int64_t n);
  void* user;
  int device_pointers; /* 1: the callbacks get device pointers on the context's device; 0: host */
} mpm_procedural_matrix;
/* b (n floats) is the right hand side; x (n floats) the initial guess, overwritten by the result.
 * params: SquareMagnitudeTermination(max_iters, tolerance) or the termination hook, and the debug
 * hook (params->cancelled is not used). precond may be NULL (no preconditioner). With
 * log_capacity > 0 the residual r and search direction p of the setup and of every iteration that
 * continues are written to residuals and directions (log_capacity x n floats each; the
 * reference's m_residuals / m_searchDirections), their count to *logged (may be NULL). */
int mpm_conjugate_residuals(mpm_context* ctx, const mpm_procedural_matrix* A, const mpm_procedural_matrix* precond,
                            const float* b, float* x, int64_t n, const mpm_solver_params* params,
                            mpm_solve_stats* stats, float* residuals, float* directions, int log_capacity,
                            int* logged);

/* ---- simulation driver: Sim (reference include/MpmSim/Sim.h:32-55) -------- */
/* Sim::Sim after the particles and model are set (Sim.cpp:17-50): bodies + volumes */
int mpm_sim_init(mpm_context* ctx, float grid_size, int dimension);
/* Sim::advance (Sim.cpp:57-134) */
int mpm_sim_advance(mpm_context* ctx, float dt, const mpm_solver_params* params, mpm_step_stats* stats);
int mpm_sim_num_bodies(const mpm_context* ctx);
/* layout of the last step's batched grids (no reference counterpart; measurement): out[7] =
 * bodies, tiles, particle chunks, node slots, particles in bodies, tiles and chunks of the
 * bodies the batched CR loop iterates (laid out first) */
int mpm_sim_grid_layout(const mpm_context* ctx, int64_t* out);
/* particle indices of body b (ascending); returns the count (out may be NULL) */
int64_t mpm_sim_body(const mpm_context* ctx, int body, int32_t* out);
int64_t mpm_sim_ballistic(const mpm_context* ctx, int32_t* out);
/* solve statistics of body b in the last advance */
int mpm_sim_body_solve_stats(const mpm_context* ctx, int body, mpm_solve_stats* stats);

/* ---- learned preconditioner (no counterpart in the reference) ----------------
 * A BF16 3-D convolutional network (MFMA, FP32 accumulation) maps a body's grid residual to
 * a velocity correction inside a flexible conjugate-residual solve; it attaches where
 * ConjugateResiduals takes a preconditioner (ConjugateResiduals.h:20, ProceduralMatrix.h:16).
 * An iteration that used the network and did not lower |r|^2 below `ratio` times its previous
 * value (mpm_lpc_set_fallback; 0.95 by default), or by less than the body's last iteration on
 * the diagonal preconditioner did, is a fallback: the body takes its next iteration with the
 * diagonal preconditioner, and after mpm_lpc_set_strikes fallbacks (1 by default) keeps the
 * diagonal for the rest of its solve; the termination test is unchanged. */
int64_t mpm_lpc_num_params(void);
/* random (He) initialisation; the output layer starts at zero */
int mpm_lpc_init(mpm_context* ctx, uint64_t seed);
int mpm_lpc_set_params(mpm_context* ctx, const float* params, int64_t n);
int mpm_lpc_get_params(mpm_context* ctx, float* params, int64_t n);
/* 0 = diagonal-preconditioned CR (the reference's solver), 1 = learned flexible CR */
int mpm_lpc_set_mode(mpm_context* ctx, int mode);
/* fallback ratio in (0, 1]: 1 = a fallback only when |r|^2 rises */
int mpm_lpc_set_fallback(mpm_context* ctx, double ratio);
/* fallbacks after which a body keeps the diagonal preconditioner for the rest of its solve, 1-16
 * (default 1): each earlier fallback costs the body one diagonal iteration before the network
 * is tried again */
int mpm_lpc_set_strikes(mpm_context* ctx, int strikes);
/* directions the flexible CR orthogonalises against, 1-8 (default 2; with the diagonal
 * preconditioner every m >= 1 gives the same Krylov iteration in exact arithmetic) */
int mpm_lpc_set_directions(mpm_context* ctx, int m);
/* 1 (default): once no unfinished body of the solve uses the network (each has taken its third
 * fallback or none had one), the rest of the solve runs the preconditioned CR recurrences of
 * ConjugateResiduals.cpp:84-151 from the flexible CR's iterate, with r = b - A x recomputed; the
 * iteration count, residual history and iteration cap continue. 0: the flexible CR runs to the
 * end on the diagonal. */
int mpm_lpc_set_handoff(mpm_context* ctx, int on);
/* CR iterations at which training solves keep their residuals: 1-16 increasing indices
 * (default 0 1 2 3 4 6 8 11 15 20 26 33 41 50 60 72); a solve keeps the first `samples` of them */
int mpm_lpc_set_capture_iterations(mpm_context* ctx, const int32_t* iters, int n);
/* training: the next mpm_sim_advance (diagonal CR) keeps up to `samples` residuals
 * r_k = b - A x_k of its solve (at chosen iterations k) and, before the particle update, takes
 * `epochs` Adam passes over them with the library's forward and backward kernels, minimising
 * per body the residual left by the optimal step along the correction,
 * 1 - <r, A z>_D^2 / (|r|_D^2 |A z|_D^2) (D = diagonal, z = r / diag + scale * network(r)):
 * the exact correction A^-1 r gives 0. */
int mpm_lpc_train_next(mpm_context* ctx, int samples, int epochs, float learning_rate);
/* training loss of mpm_lpc_train_next: 0 (default) the optimal-step residual loss above; 1 the
 * exact-correction loss: the solve also keeps its iterates x_k, and with its final iterate x*
 * each sample's exact correction is e = x* - x_k (A e = r_k up to the final residual); per body
 * the network minimises the fraction of the error energy e^T A e that the best step along z
 * leaves, 1 - (z^T r)^2 / ((z^T A z)(e^T r)) (samples at least 4 iterations before the solve's
 * end). Solve the training steps to a tighter tolerance than the production ones so that x* is
 * close to the exact solution. */
int mpm_lpc_set_training_loss(mpm_context* ctx, int mode);
/* mean per-body loss of that training's first and last epoch, and of the diagonal step alone */
int mpm_lpc_last_loss(mpm_context* ctx, double* first, double* last, double* diagonal);
/* (body, sample) training pairs that training used (bodies of 256+ particles still iterating) */
int mpm_lpc_last_pairs(const mpm_context* ctx);
/* grid API (one body; dense arrays in node order): network output for residual r, diagonal
 * diag and input scale, and the loss w |out - e / scale|^2 with its parameter gradient */
int mpm_grid_lpc_apply(mpm_grid* grid, const float* r, const float* diag, float scale, float* out);
int mpm_grid_lpc_gradients(mpm_grid* grid, const float* r, const float* diag, const float* e, float scale,
                           float weight, float* grad, double* loss);

/* ---- full-FP64 step (precision studies) ------------------------------------------
 * Sim (reference include/MpmSim/Sim.h:32-55; Sim.cpp:17-50 and 57-134) with every scalar in
 * double on the GPU: particle state, grid quantities, shape weights, stresses, the
 * conjugate-residual solve and its reductions. It has no counterpart in the reference, whose
 * types are single precision; it is the full-FP64 run that the FP32-bulk step above is measured
 * against, on the same GPU at full size. Semantics are those of mpm_sim_*: bodies by the flood
 * fill |x_i - x_j|^2 < h^2 (Sim.cpp:136-203), one dense grid per body, the diagonal
 * preconditioned CR of ConjugateResiduals.cpp:19-160 with SquareMagnitudeTermination.
 * Sums are FP64 atomic adds: results agree run to run to the rounding of double, not bitwise.
 * Arrays are in the reference's host layouts, in double: Vector3 = 3 doubles, Matrix3 = 9
 * doubles column-major. */
typedef struct mpm_sim64 mpm_sim64;
/* collision object as mpm_collider, in double */
typedef struct {
  int type;
  double coeffs[4];
  double center[3];
  double radius;
  double velocity[3];
  double friction;
  int sticky;
  int dims[3];
  double origin[3];
  double voxel;
  const double* phi;
  const double* phi_velocity;
} mpm_collider64;
/* n particles: v = 0, F = I, volume 0 (MaterialPointData, MaterialPointData.h:14-44) */
int mpm_sim64_create(mpm_sim64** sim, int device, int64_t n);
int mpm_sim64_destroy(mpm_sim64* sim);
const char* mpm_sim64_last_error(const mpm_sim64* sim);
/* the variables of mpm_particles_set / mpm_particles_get, in double */
int mpm_sim64_set(mpm_sim64* sim, const char* name, const double* data, int64_t count);
int mpm_sim64_get(mpm_sim64* sim, const char* name, double* data, int64_t count);
/* SnowConstitutiveModel(E, nu, xi, theta_c, theta_s) + setParticles (SnowConstitutiveModel.cpp:9-25,
 * 123-169); per-particle mu0, lambda0 and xi may then be set by name */
int mpm_sim64_model_snow(mpm_sim64* sim, double youngs_modulus, double poisson_ratio, double hardening,
                         double theta_c, double theta_s);
int mpm_sim64_set_colliders(mpm_sim64* sim, const mpm_collider64* colliders, int n);
int mpm_sim64_set_gravity(mpm_sim64* sim, const double* gravity, int n);
/* Sim::Sim after the particles and model are set (Sim.cpp:17-50): bodies + volumes */
int mpm_sim64_init(mpm_sim64* sim, double grid_size);
/* Sim::advance (Sim.cpp:57-134) with SquareMagnitudeTermination(max_iters, tolerance) */
int mpm_sim64_advance(mpm_sim64* sim, double dt, int max_iters, double tolerance, mpm_step_stats* stats);
int mpm_sim64_num_bodies(const mpm_sim64* sim);
int mpm_sim64_body_solve_stats(const mpm_sim64* sim, int body, mpm_solve_stats* stats);

/* ---- small bodies (no counterpart in the reference) -----------------------------
 * Bodies of at most 8 node blocks (2048 grid nodes) and max_particles particles
 * run their whole conjugate residual solve in one workgroup each, on a second
 * stream, concurrently with the batched iteration of the larger bodies (default
 * 2048; 0 = every body iterates in the batched loop). The iterates are bitwise
 * the same either way; host termination/debug hooks and training captures
 * always use the batched loop. */
int mpm_set_small_bodies(mpm_context* ctx, int max_particles);

/* ---- measurement (no counterpart in the reference) ----------------------------
 * Kernel timing with HIP events on the context's streams. Kernel classes, in bit
 * order: p2g_mass p2g_vel forces hvp diag g2p cr_update_x cr_ar cr_update_px lpc_forward
 * cr_small (cr_update_x: a solve's first x update; cr_update_px: each later iteration's
 * p and x updates in one pass; cr_small: the small-body solver, one launch per solve).
 * mask selects the classes to time (bit k = class k; -1 = all, 0 = off); each
 * timed launch costs two event records, so time only what is needed.
 * algorithmic_bytes sums, over the timed launches, bytes/particle * particles +
 * bytes/node * nodes with m > 0 of that kernel's minimal HBM traffic. */
int mpm_profile_enable(mpm_context* ctx, int mask); /* also resets the counters */
/* Time one launch in `period` of each enabled class (1: every launch; the default): the
 * event records around a timed launch idle the stream for a few microseconds, so sampling
 * keeps the measurement from slowing what it measures. Resets the sampling counters. */
int mpm_profile_period(mpm_context* ctx, int period);
int mpm_profile_query(mpm_context* ctx, const char* kernel, double* total_ms, int64_t* launches,
                      double* algorithmic_bytes);

#ifdef __cplusplus
}
#endif

#endif /* MPMSNOW_GPU_H */
