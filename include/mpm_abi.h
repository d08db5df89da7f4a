/*
 * C ABI of the MI355X MLS-MPM substep engine (libmpm_hip.so).
 *
 * This is the drop-in boundary for taichi_elements' MPMSolver
 * (engine/mpm_solver.py).  The reference has no FFI of its own: its host
 * Python calls @ti.kernel methods that Taichi JIT-compiles.  Each entry point
 * below replaces the reference function cited beside it; a Python host
 * (taichi_elements_amd/mpm_solver.py, via ctypes) keeps the reference's
 * method signatures on top of it.
 *
 * Conventions
 *  - Plain C types only; host pointers are borrowed for the duration of the
 *    call and copied.  Particle arrays at this boundary are row-major AoS in
 *    insertion order: x/v (n, dim) float32, C/F (n, dim, dim) float32,
 *    Jp (n,) float32, material/color (n,) int32 -- the layout
 *    MPMSolver.particle_info() returns (mpm_solver.py:1172-1191).
 *  - Every function returns 0 on success or a negative MPM_E* code; the
 *    message of the last failure on the calling thread is mpm_last_error().
 *  - A handle owns its device buffers and one HIP stream; it is not
 *    thread-safe, but any single thread may drive it (the reference is driven
 *    from a worker thread by its Blender add-on, blender/operators.py:403).
 *  - mpm_substeps() only enqueues work; device faults and particles that
 *    leave the simulated window are reported by the next call that
 *    synchronises (mpm_read_max_velocity, mpm_particle_info, mpm_get_state,
 *    mpm_sync).
 */
#ifndef MPM_ABI_H
#define MPM_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPM_OK 0
#define MPM_EINVAL (-1)   /* bad argument (reference: assert / ValueError) */
#define MPM_EHIP (-2)     /* HIP runtime error or device fault */
#define MPM_EDOMAIN (-3)  /* particles left the simulated window */
#define MPM_ENOMEM (-4)

/* Materials (mpm_solver.py:16-20) and surfaces (mpm_solver.py:32-36). */
#define MPM_WATER 0
#define MPM_ELASTIC 1
#define MPM_SNOW 2
#define MPM_SAND 3
#define MPM_STATIONARY 4
#define MPM_STICKY 0
#define MPM_SLIP 1
#define MPM_SEPARATE 2

/* Fields for mpm_get_state / mpm_set_state. */
#define MPM_FIELD_X 0
#define MPM_FIELD_V 1
#define MPM_FIELD_C 2
#define MPM_FIELD_F 3
#define MPM_FIELD_JP 4
#define MPM_FIELD_MATERIAL 5
#define MPM_FIELD_COLOR 6

typedef struct mpm_solver mpm_solver;

/* Solver constants.  The host computes them exactly as
 * MPMSolver.__init__ does with Python floats (mpm_solver.py:82-87,
 * 200-210) and the library rounds them to float32 for the kernels, as
 * Taichi does with kernel literals. */
typedef struct {
  int32_t dim;              /* 2 or 3 = len(res) (mpm_solver.py:67)      */
  int32_t res[3];           /* cells per axis; res[2] ignored in 2D      */
  int32_t padding;          /* wall thickness in cells (mpm_solver.py:52) */
  int32_t device;           /* HIP device ordinal                         */
  int32_t reorder_interval; /* substeps between physical particle
                               reorders (0: default, <0: never)          */
  int32_t chunk;            /* particles per transfer work item (0: default) */
  double dx, inv_dx, p_vol, p_mass;
  double mu_0, lambda_0, alpha; /* Lame parameters, sand friction coefficient */
  double water_density;     /* mpm_solver.py:61, 571-573                  */
} mpm_config;

/* MPMSolver.__init__ (mpm_solver.py:44-310) minus the Python-side state:
 * allocates the solver on cfg->device with gravity (0,-9.8[,0]) and an
 * empty grid post-process list (the Python host adds the bounding box
 * first, as mpm_solver.py:300-302 does). */
int mpm_create(const mpm_config* cfg, mpm_solver** out);
/* Batched scenes (no reference counterpart: the reference simulates one
 * scene per MPMSolver).  nscenes independent scenes of the same dim, each
 * with its own constants (res, dx, dt, Lame parameters...), are advanced
 * together: every stage of a substep is one launch over all of them.
 * Configuration calls (gravity, post-processes, mpm_add_particles) apply to
 * the scene chosen with mpm_set_active_scene (0 after creation). */
int mpm_create_batch(const mpm_config* cfgs, int32_t nscenes, mpm_solver** out);
int mpm_set_active_scene(mpm_solver* s, int32_t scene);
int mpm_destroy(mpm_solver* s);
/* Message of the last failing call on this thread ("" if none). */
const char* mpm_last_error(void);

/* MPMSolver.set_gravity (mpm_solver.py:316-319); g has dim entries. */
int mpm_set_gravity(mpm_solver* s, const double* g);

/* grid_postprocess list (mpm_solver.py:300, 642, 687-692), applied in the
 * order added after gravity.  The bounding box's walls sit `padding` cells
 * inside [0, res), or, with unbounded = 1, inside [-grid_size/2,
 * grid_size/2) (mpm_solver.py:604-616; grid_size as set by
 * mpm_set_unbounded, else 4096). */
int mpm_add_bounding_box(mpm_solver* s, int32_t unbounded);
/* add_surface_collider (mpm_solver.py:647-687): normal is normalised here
 * as the reference does on the host; sticky with friction != 0 is
 * MPM_EINVAL (the reference raises ValueError). */
int mpm_add_surface_collider(mpm_solver* s, const double* point, const double* normal,
                             int32_t surface, double friction);
/* add_sphere_collider (mpm_solver.py:618-642). */
int mpm_add_sphere_collider(mpm_solver* s, const double* center, double radius,
                            int32_t surface);
/* clear_grid_postprocess (mpm_solver.py:644-645): removes every entry,
 * the bounding box included. */
int mpm_clear_grid_postprocess(mpm_solver* s);

/* Appends n particles: seed_particle (mpm_solver.py:823-838) as called by
 * seed_from_external_array / recover_from_external_array
 * (mpm_solver.py:1081-1126): F = I, C = 0, Jp = 0 for sand else 1.
 * v may be NULL (zero velocity); color may be NULL (0xFFFFFF). */
int mpm_add_particles(mpm_solver* s, int64_t n, const float* x, const float* v,
                      const int32_t* material, const int32_t* color);
/* n_particles[None] (mpm_solver.py:81). */
int mpm_num_particles(mpm_solver* s, int64_t* n);

/* n iterations of the substep body of step() (mpm_solver.py:788-799):
 * build_pid, p2g, grid_normalization_and_gravity, grid_postprocess, g2p,
 * compute_max_velocity -- all with the same dt; t0 is the time handed to
 * the post-process of the first substep.  Asynchronous. */
int mpm_substeps(mpm_solver* s, int32_t n, float dt, double t0);
/* Batched form of mpm_substeps: one substep of every scene per iteration,
 * scene k with time step dts[k].  dts[k] == 0 means scene k sits these
 * substeps out: its particles and grid are left untouched, which lets scenes
 * with different substep counts per frame share one batch.  The maximum
 * velocity of mpm_read_max_velocity is then taken over the scenes that
 * moved. */
int mpm_substeps_batch(mpm_solver* s, int32_t n, const float* dts);
/* Per-substep max over particles of max_i |v_i| (compute_max_velocity,
 * mpm_solver.py:726-735) for the substeps of the last mpm_substeps call;
 * out has room for that call's n entries.  Synchronises. */
int mpm_read_max_velocity(mpm_solver* s, int32_t n, float* out);
/* compute_max_grid_velocity (mpm_solver.py:737-746), which step() calls
 * before every substep when use_adaptive_dt (mpm_solver.py:762-770): max
 * over grid nodes of max_i |v_i| on the grid the last substep left (the
 * g2p2g input grid of the next substep with use_g2p2g); 0 before the first
 * substep and without particles.  One-scene solvers.  Synchronises. */
int mpm_max_grid_velocity(mpm_solver* s, float* out);

/* One substep in two halves, for hosts that run their own grid
 * post-processes: the callables of grid_postprocess (mpm_solver.py:300,
 * 794-795; the reference calls each as p(t, dt, grid_v) between
 * grid_normalization_and_gravity and g2p).  One-scene solvers.
 *
 * mpm_substep_begin runs build_pid, p2g, grid_normalization_and_gravity and
 * the built-in post-processes (bounding box, colliders) of one substep with
 * time step dt (use_g2p2g: the whole fused substep, whose output grid is
 * the next substep's input).  Until mpm_substep_end, the nodes of the
 * active grid blocks (4^dim nodes each, as the reference's leaf blocks) can
 * be read and written; every other call on the handle fails with
 * MPM_EINVAL.  mpm_substep_end runs g2p and compute_max_velocity (read with
 * mpm_read_max_velocity(s, 1, ...)).  Asynchronous. */
int mpm_substep_begin(mpm_solver* s, float dt);
/* Number of grid nodes of the active blocks of the substep in flight.
 * Synchronises. */
int mpm_grid_node_count(mpm_solver* s, int64_t* count);
/* Copies the count nodes (count from mpm_grid_node_count) into DEVICE
 * arrays of the handle's device: coords (count, dim) int32 global node
 * indices I (node position I * dx, the reference's grid_v index), v
 * (count, dim) float32 velocities, m (count,) float32 masses.  Waits for
 * the copy before returning. */
int mpm_grid_gather(mpm_solver* s, int32_t* coords, float* v, float* m, int64_t count);
/* Writes the count velocities v (count, dim) float32 (a DEVICE array, same
 * node order as mpm_grid_gather) back into the grid; masses are unchanged.
 * The caller's own work on v must be complete.  Waits for the copy before
 * returning. */
int mpm_grid_scatter(mpm_solver* s, const float* v, int64_t count);
/* Second half of the substep begun by mpm_substep_begin. */
int mpm_substep_end(mpm_solver* s);

/* particle_info (mpm_solver.py:1172-1191): position (n,dim), velocity
 * (n,dim), material (n,), color (n,) in insertion order.  Any pointer may
 * be NULL to skip that field.  Synchronises. */
int mpm_particle_info(mpm_solver* s, float* x, float* v, int32_t* material, int32_t* color);

/* Full particle state access in insertion order (no reference counterpart:
 * the reference exposes its Taichi fields x, v, C, F, Jp, material directly,
 * mpm_solver.py:101-136).  Used by parity tests, checkpoints and restarts. */
int mpm_get_state(mpm_solver* s, int32_t field, void* host);
int mpm_set_state(mpm_solver* s, int32_t field, const void* host);

/* Grid after the last grid operator (no reference counterpart: the
 * reference's grid_v / grid_m fields, mpm_solver.py:162-186).  Copies the
 * node window [lo, lo+n) (n[a] nodes along axis a, x fastest) into
 * m (prod n) and v (prod n * dim); nodes outside the active blocks read 0. */
int mpm_read_grid(mpm_solver* s, const int32_t* lo, const int32_t* n, float* m, float* v);

/* Binning of the last substep (build_pid, mpm_solver.py:344-361): perm is the
 * order of the physical slots sorted by block key -- a stable sort of the
 * previous substep's order (of the slots themselves on the first substep and
 * after positions are set), keys[k] the block key of slot perm[k], and
 * slot_to_insertion[i] the insertion index of the particle in slot i.
 * Arrays hold mpm_num_particles entries. */
int mpm_read_binning(mpm_solver* s, int32_t* keys, int32_t* perm, int32_t* slot_to_insertion);

/* Debug switches: bit 0 = grid operator writes raw p2g output (mass and
 * momentum, no normalisation / gravity / post-process). */
int mpm_set_debug(mpm_solver* s, int32_t flags);

/* Waits for the solver's stream; reports device faults and domain escapes. */
int mpm_sync(mpm_solver* s);

/* Kernel timing with HIP events on the solver's stream.  While enabled each
 * kernel launch of mpm_substeps is bracketed by events; mpm_kernel_times
 * returns, for every kernel name (newline-separated into names), the number
 * of launches and the summed milliseconds since the last reset. */
int mpm_set_timing(mpm_solver* s, int32_t enabled);
int mpm_kernel_times(mpm_solver* s, char* names, int32_t names_cap, int32_t* counts,
                     double* ms, int32_t cap, int32_t* n_kernels);

/* The engine's 2x2 / 3x3 SVD (the replacement of ti.svd, mpm_solver.py:525)
 * evaluated on the host CPU with the same code the kernels inline, for tests:
 * A = U diag(sig) V^T, row-major dim x dim. */
int mpm_host_svd(int32_t dim, const float* A, float* U, float* sig, float* V);
/* The same SVD as the kernels run it on the GPU of `device` (where the
 * Jacobi sweeps use the hardware reciprocal square root), for n matrices:
 * A, U, V (n, dim, dim) and sig (n, dim), host arrays.  Synchronous. */
int mpm_device_svd(int32_t device, int32_t dim, int64_t n, const float* A, float* U, float* sig,
                   float* V);

/* use_g2p2g mode (MPMSolver(use_g2p2g=True, v_clamp_g2p2g, g2p2g_allowed_cfl),
 * mpm_solver.py:363-485, 586-598, 773-787): every substep is one fused
 * g2p2g transfer -- G2P from the previous grid, advection, p2g at the new
 * positions into the other grid -- followed by the grid operator with each
 * velocity component clamped to dx * allowed_cfl / dt when clamp != 0 and
 * allowed_cfl > 0.  Particles added since the previous substep skip the G2P
 * once (they keep their velocity).  Must be chosen before the first substep.
 * A particle that moves more than one cell in a substep (only possible
 * without the clamp) is reported as MPM_EDOMAIN. */
int mpm_set_g2p2g(mpm_solver* s, int32_t enabled, double allowed_cfl, int32_t clamp);

/* HIP graphs (on by default): inside mpm_substeps / mpm_substeps_batch,
 * every pair of steady-state substeps (neighbour-sort binning, no physical
 * reorder due) is captured once as a graph and replayed, so a substep costs
 * one graph launch instead of six kernel launches.  Results are bitwise the
 * same as with enabled = 0 (kernel by kernel). */
int mpm_set_graphs(mpm_solver* s, int32_t enabled);

/* Launch tuning (no reference counterpart; the reference leaves launch
 * shapes to Taichi).  chunk: particles per transfer work item (16-1024);
 * groups: workgroups of the persistent transfer kernels (64-16384); 0 keeps
 * a value.  The next substep re-bins from scratch; results change only by
 * float32 summation order.  MPMSolver.autotune() times candidates on the
 * live scene and keeps the fastest. */
int mpm_set_tuning(mpm_solver* s, int32_t chunk, int32_t groups);
int mpm_get_tuning(mpm_solver* s, int32_t* chunk, int32_t* groups);
/* g2p's occupancy (no reference counterpart): waves = 0 lets the compiler
 * allocate registers freely (5 waves per SIMD), 6 holds g2p to six waves
 * per SIMD (more grid-tile gathers in flight, which pays in sparse scenes
 * whose blocks hold few particles).  Results are bitwise the same either
 * way; MPMSolver.autotune() times both at the chosen launch shape. */
int mpm_set_g2p_waves(mpm_solver* s, int32_t waves);
int mpm_get_g2p_waves(mpm_solver* s, int32_t* waves);

/* Device-side snapshot of the simulated state (no reference counterpart;
 * MPMSolver.autotune uses it to put the scene back after timing launch
 * shapes).  op = MPM_SNAPSHOT_SAVE copies, on the device, every particle
 * slot and, in use_g2p2g mode, the grid the next substep's G2P reads and
 * which particles are new; MPM_SNAPSHOT_RESTORE puts them back (the next
 * substep re-bins from scratch), and fails with MPM_EINVAL when particles
 * were added or the storage mode changed since the save; an unbounded
 * scene's block window is moved back first.  MPM_SNAPSHOT_RELEASE frees the
 * copy.  Colours, gravity, post-processes and tuning are not part of it. */
#define MPM_SNAPSHOT_SAVE 0
#define MPM_SNAPSHOT_RESTORE 1
#define MPM_SNAPSHOT_RELEASE 2
int mpm_snapshot(mpm_solver* s, int32_t op);

/* Bookkeeping of the last substep: number of valid particles, transfer
 * work items (chunks), active grid blocks, and particle blocks. */
int mpm_last_counts(mpm_solver* s, int64_t* out4);

/* Differentiable mode (no reference counterpart: the reference's p2g is not
 * differentiable -- it mutates F and Jp in place -- and it has no per-particle
 * stiffness).  One substep is a function of a caller-owned device state:
 * state arrays are float32 SoA of NF = 2d + 2d^2 + 1 dense rows of n entries
 * (x[d], v[d], C[d*d] row-major, F[d*d], Jp; row r starts at element r n),
 * in particle order; youngs is
 * the per-particle Young's modulus E_p (device, n floats) and nu the global
 * Poisson ratio: mu_0 = E_p / (2 (1 + nu)), lambda_0 = E_p nu / ((1 + nu)(1 -
 * 2 nu)), scaled by h = 0.3 for elastic particles as the reference does.
 * Materials are water, elastic, snow (clamp and hardening) and sand
 * (Drucker-Prager return map, whose coefficient (d lambda_0 + 2 mu_0) /
 * (2 mu_0) follows nu); stationary particles are refused.  Every
 * post-process (bounding box, surface and sphere colliders) has its adjoint;
 * the derivative is taken on the branch the forward took (the clamp, the
 * cone and the colliders' contact tests are kinks).  stream: the
 * hipStream_t to enqueue on (NULL: the null stream,
 * e.g. torch's default stream).  The handle must hold no particles of its
 * own.
 * mpm_diff_setup: n particles with these host materials.
 * mpm_diff_binning_bytes: size of the binning record of one substep.
 * mpm_diff_substep: state_out = substep(state_in) (may alias); binning
 *   (device, may be NULL) receives the substep's binning record.
 * mpm_diff_substep_backward: given grad_out = dL/d state_out, writes
 *   grad_in = dL/d state_in, grad_youngs = dL/dE (may be NULL) and
 *   grad_nu[0] = dL/dnu (device).  The grid is recomputed from state_in, on
 *   the forward's binning record when given (else the binning is redone).
 * mpm_diff_sync: waits for stream and reports domain / CFL errors. */
int mpm_diff_setup(mpm_solver* s, int64_t n, const int32_t* material);
int mpm_diff_binning_bytes(mpm_solver* s, int64_t* bytes);
int mpm_diff_substep(mpm_solver* s, float dt, const float* state_in, float* state_out,
                     const float* youngs, double nu, void* binning, void* stream);
int mpm_diff_substep_backward(mpm_solver* s, float dt, const float* state_in,
                              const float* youngs, double nu, const void* binning,
                              const float* grad_out, float* grad_in, float* grad_youngs,
                              float* grad_nu, void* stream);
int mpm_diff_sync(mpm_solver* s, void* stream);
/* Allocation-free form of mpm_diff_substep (no counterpart in the reference:
 * it is for callers that cannot allocate device memory -- code under HIP
 * graph capture, tight serving loops).  Same state_out as mpm_diff_substep
 * on the same inputs (bitwise: the same kernels in the same order), without
 * the binning record.  Its workspace (binning, chunk list, scan and sort
 * state, node tiles, grid) lives in a static 64 MiB device array that each
 * GPU the library runs on holds until the process ends; the call performs
 * no hipMalloc / hipFree.  A scene whose workspace exceeds 64 MiB is
 * MPM_EINVAL (its size is in the message).  Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU (other streams, other
 * threads, other handles) race.  The scene table is uploaded (and the stream
 * synchronised) only when dt, youngs or nu change from the handle's last
 * call. */
int mpm_diff_substep_noalloc(mpm_solver* s, float dt, const float* state_in, float* state_out,
                             const float* youngs, double nu, void* stream);

/* The quantised particle file of ParticleIO.write_particles
 * (engine/particle_io.py:12-76), computed on the GPU: ranges (2, dim, 2)
 * float32 = [x | v][axis][min, max] with too-narrow ranges widened to
 * min + 1e-5; x_and_v (n, dim) uint32 = (x quantised to 24 bits << 8) +
 * (v quantised to 8 bits); color (n, 3) uint8 = R, G, B of the 0xRRGGBB
 * colour.  Insertion order, every particle of the handle (all scenes of a
 * batch).  No particles is MPM_EINVAL (the reference's np.min raises). */
int mpm_quantize_particles(mpm_solver* s, float* ranges, uint32_t* x_and_v, uint8_t* color);

/* MPMSolver(quant=..., support_plasticity=...) (mpm_solver.py:47, 62,
 * 106-136, 216-262).  quant = 1 stores each particle packed: x as three
 * signed 21-bit fixed-point values (step 2^-19, range [-2, 2)), v as a
 * 7-bit shared exponent and three 19-bit sign-magnitude mantissas, F as
 * signed 16-bit fixed-point values (step 4.1 / 32768); every store rounds to
 * nearest and saturates, and the kernels compute with what a load returns
 * (x advances by the stored v; in use_g2p2g mode the new F is clamped to
 * +-4 and rounded before the SVD, mpm_solver.py:413-416).  C (absent in
 * use_g2p2g mode) and Jp stay float32.  support_plasticity = 0 drops Jp: no
 * hardening, no Jp updates, water restarts each substep from F = I, sand
 * has no stress.  Only before particles are added; not for the
 * differentiable mode.  mpm_get_state / mpm_particle_info return the stored
 * (rounded) values. */
int mpm_set_storage(mpm_solver* s, int32_t quant, int32_t support_plasticity);
/* MPMSolver(unbounded=True) (mpm_solver.py:74, 143-148, 604-611) for a
 * one-scene solver: grid_size = 4096 doubled until it exceeds twice the
 * largest resolution, and the dense block window follows the particles --
 * every 8 substeps the particles' bounding box (one small reduction and a
 * host synchronisation) is checked against the window, which is moved and
 * resized when a particle could otherwise leave it within the next 8
 * substeps (the grid a use_g2p2g substep reads is carried over).  The
 * window holds at most 2^24 blocks; a scene spread wider raises
 * MPM_EDOMAIN.  Call before mpm_add_bounding_box. */
int mpm_set_unbounded(mpm_solver* s, int32_t enabled);
/* Bytes one particle occupies (device state plus its host-side colour) in
 * the current storage mode (the reference's figure is 40 B for quant with
 * use_g2p2g and no plasticity, README.md:60). */
int mpm_particle_bytes(mpm_solver* s, int32_t* bytes);

/* Voxelizer(res, dx, super_sample, padding=padding).voxelize(triangles)
 * (engine/voxelizer.py:18-38, 111-126), run on the GPU of `device`.
 * voxel_res (3) receives the voxel grid: every axis of res rounded up to a
 * power of two, times super_sample (voxel size dx / super_sample).  With
 * voxels non-NULL the (ntri, 9) float64 triangles (a, b, c) are voxelized
 * and voxels receives the (voxel_res[0], voxel_res[1], voxel_res[2]) int32
 * counts, C order: +1 per triangle facing up (normal z > 0) and -1 per
 * triangle facing down above the voxel; positive inside a closed,
 * outward-oriented mesh.  Heights are capped at voxel_res[2]. */
int mpm_voxelize(int32_t device, const int32_t* res, double dx, int32_t super_sample,
                 int32_t padding, int64_t ntri, const double* triangles, int32_t* voxel_res,
                 int32_t* voxels);
/* MPMSolver.add_mesh (mpm_solver.py:1049-1079) with seed_from_voxels
 * (mpm_solver.py:1017-1047): voxelizes the triangles on the active scene's
 * grid (res, dx, padding of its configuration) and seeds, in every voxel
 * with a positive count, particle l = 0..sample_density when
 * r + l < sample_density / super_sample^3, at ((r0, r1, r2) + voxel index) *
 * (dx / super_sample) + translation, with velocity (3, NULL: zero), the
 * material and the color.  r are uniform float32 in [0, 1) hashed from
 * (seed, voxel, l, component) in place of ti.random; particles are appended
 * in voxel order (C order of the voxel grid, then l).  *added (may be NULL)
 * receives the number seeded.  3D only; sample_density 0..4096. */
int mpm_add_mesh(mpm_solver* s, int64_t ntri, const double* triangles, int32_t material,
                 int32_t color, int32_t sample_density, const float* velocity,
                 const float* translation, int32_t super_sample, uint32_t seed, int64_t* added);

#ifdef __cplusplus
}
#endif
#endif
