/*
 * CPU restatement of the taichi_elements MLS-MPM substep -- TEST INFRASTRUCTURE.
 *
 * This library is the oracle that the HIP product path is checked against.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker (or as the timed CPU baseline); the
 * product path (taichi_elements_amd) never links or calls it.
 *
 * It restates, as plain C99 compiled once for float and once for double,
 * the reference kernels of engine/mpm_solver.py (cited per function in
 * mpm_oracle_impl.inc):
 *   constants                      mpm_solver.py:82-87, 200-210
 *   build_pid (block keys)         mpm_solver.py:344-361
 *   p2g + constitutive models      mpm_solver.py:487-584
 *   sand_projection                mpm_solver.py:321-342
 *   ti.svd (third-party, Taichi)   call sites mpm_solver.py:525
 *   grid_normalization_and_gravity mpm_solver.py:586-598
 *   grid_bounding_box              mpm_solver.py:600-616
 *   sphere / surface colliders     mpm_solver.py:618-687
 *   g2p                            mpm_solver.py:694-724
 *   compute_max_velocity           mpm_solver.py:726-735
 *
 * Pinned by tests/test_oracle_pins.py (closed forms, invariants, LAPACK),
 * tests/test_oracle_g2p2g.py, tests/test_oracle_quant.py and
 * tests/test_golden.py: whole substeps against the golden fixtures of an
 * independent float64 brute-force restatement (tests/golden/make_golden.py).
 *
 * The grid is a dense node window [grid_lo, grid_lo + grid_n) per axis,
 * x fastest.  Particle arrays are AoS: x[p*dim + a], C[p*dim*dim + a*dim + b].
 */
#ifndef MPM_ORACLE_H
#define MPM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_MAX_POST 16

enum { ORACLE_POST_BBOX = 0, ORACLE_POST_PLANE = 1, ORACLE_POST_SPHERE = 2 };
enum { ORACLE_WATER = 0, ORACLE_ELASTIC = 1, ORACLE_SNOW = 2, ORACLE_SAND = 3,
       ORACLE_STATIONARY = 4 };

typedef struct {
  int32_t kind;     /* ORACLE_POST_* */
  int32_t surface;  /* 0 sticky, 1 slip, 2 separate (mpm_solver.py:32-36) */
  double point[3];  /* plane point or sphere centre */
  double normal[3]; /* plane normal, already normalised on the host (mpm:654) */
  double radius;
  double friction;
} oracle_post_op;

typedef struct {
  int32_t dim;
  int32_t res[3];
  int32_t padding;
  int32_t n_post;
  double dx, inv_dx, p_vol, p_mass;
  double mu_0, lambda_0, alpha, water_density;
  double gravity[3];
  int32_t grid_lo[3];
  int32_t grid_n[3];
  oracle_post_op post[ORACLE_MAX_POST];
  /* Optional per-particle Young's modulus (NULL: mu_0 / lambda_0 above):
   * mu_0 = E_p / (2 (1 + nu)), lambda_0 = E_p nu / ((1 + nu)(1 - 2 nu)),
   * the constants of mpm_solver.py:201-206 with E made per particle (the
   * engine's differentiable mode; the reference has one global E). */
  const double* youngs;
  double nu;
  /* g2p2g mode's grid velocity clamp (mpm:589, 596-598): each component
   * limited to dx * v_clamp_cfl / dt; 0 = no clamp (the p2g/g2p path). */
  double v_clamp_cfl;
  /* MPMSolver(quant=True) (mpm_solver.py:106-114, 216-262): x, v and F are
   * stored quantised -- every store is replaced by the value a load of the
   * packed field returns (oracle_quant_*); in g2p2g the new F is clamped to
   * +-F_bound first (mpm:413-416). */
  int32_t quant;
  /* MPMSolver(support_plasticity=False) (mpm_solver.py:62, 135-136, 419-421,
   * 436-437, 456, 510-518, 533-534, 541-542, 553). */
  int32_t no_plasticity;
  /* MPMSolver(unbounded=True) (mpm_solver.py:143-148, 604-611): walls at
   * -grid_size/2 + padding and grid_size/2 - padding instead of padding and
   * res - padding; this holds grid_size/2 (0: bounded). */
  int32_t unbounded_half;
} oracle_cfg;

/* The packed particle formats, in place: the value each element reads back
 * after a store (round to nearest even, saturating).  x: signed 21-bit
 * fixed point, step 2^-19.  F: signed 16-bit fixed point, step 4.1/32768.
 * v (n vectors of dim): one 7-bit exponent shared by the vector's
 * components and 18-bit magnitudes (+ sign). */
void oracle_quant_x(int64_t n, float* x);
void oracle_quant_f(int64_t n, float* f);
void oracle_quant_v(int dim, int64_t n, float* v);

/* ---- float instantiation ---- */
void oracle_svd_f32(int dim, const float* A, float* U, float* sig, float* V);
void oracle_sand_projection_f32(int dim, float* sig, float* Jp, double mu_0,
                                double lambda_0, double alpha);
/* Per-particle constitutive update of p2g (mpm:507-568): updates F and Jp in
 * place and writes the Kirchhoff stress (before the -dt*p_vol*4*inv_dx^2
 * scale) into stress (dim*dim). */
void oracle_constitutive_f32(const oracle_cfg* cfg, int material, float dt,
                             float* F, const float* C, float* Jp,
                             float* stress);
/* The same update inside g2p2g (mpm:411-473). */
void oracle_constitutive_g2p2g_f32(const oracle_cfg* cfg, int material, float dt,
                                    float* F, const float* C, float* Jp,
                                    float* stress);
void oracle_p2g_f32(const oracle_cfg* cfg, int64_t n, const float* x,
                    const float* v, const float* C, float* F, float* Jp,
                    const int32_t* material, float dt, float* grid_m,
                    float* grid_v, int nthreads);
void oracle_grid_op_f32(const oracle_cfg* cfg, float dt, double t,
                        float* grid_m, float* grid_v, int nthreads);
void oracle_g2p_f32(const oracle_cfg* cfg, int64_t n, float* x, float* v,
                    float* C, const int32_t* material, float dt,
                    const float* grid_v, float* max_v, int nthreads);
void oracle_substep_f32(const oracle_cfg* cfg, int64_t n, float* x, float* v,
                        float* C, float* F, float* Jp,
                        const int32_t* material, float dt, double t,
                        float* grid_m, float* grid_v, float* max_v,
                        int nthreads);
/* g2p2g (mpm:363-485): G2P from grid_v_in, advection, p2g into the zeroed
 * output grid; particles with index >= n_final skip the G2P. */
void oracle_g2p2g_f32(const oracle_cfg* cfg, int64_t n, float* x, float* v, float* F,
                      float* Jp, const int32_t* material, float dt, int64_t n_final,
                      const float* grid_v_in, float* grid_m_out, float* grid_v_out);

/* ---- double instantiation ---- */
void oracle_svd_f64(int dim, const double* A, double* U, double* sig,
                    double* V);
void oracle_sand_projection_f64(int dim, double* sig, double* Jp, double mu_0,
                                double lambda_0, double alpha);
void oracle_constitutive_f64(const oracle_cfg* cfg, int material, double dt,
                             double* F, const double* C, double* Jp,
                             double* stress);
/* The same update inside g2p2g (mpm:411-473). */
void oracle_constitutive_g2p2g_f64(const oracle_cfg* cfg, int material, double dt,
                                    double* F, const double* C, double* Jp,
                                    double* stress);
void oracle_p2g_f64(const oracle_cfg* cfg, int64_t n, const double* x,
                    const double* v, const double* C, double* F, double* Jp,
                    const int32_t* material, double dt, double* grid_m,
                    double* grid_v, int nthreads);
void oracle_grid_op_f64(const oracle_cfg* cfg, double dt, double t,
                        double* grid_m, double* grid_v, int nthreads);
void oracle_g2p_f64(const oracle_cfg* cfg, int64_t n, double* x, double* v,
                    double* C, const int32_t* material, double dt,
                    const double* grid_v, double* max_v, int nthreads);
void oracle_substep_f64(const oracle_cfg* cfg, int64_t n, double* x,
                        double* v, double* C, double* F, double* Jp,
                        const int32_t* material, double dt, double t,
                        double* grid_m, double* grid_v, double* max_v,
                        int nthreads);
void oracle_g2p2g_f64(const oracle_cfg* cfg, int64_t n, double* x, double* v, double* F,
                      double* Jp, const int32_t* material, double dt, int64_t n_final,
                      const double* grid_v_in, double* grid_m_out, double* grid_v_out);

/* ---- binning (integer work; f32 positions as in the reference) ----
 * Block key of a particle: base = floor(x*inv_dx - 0.5) computed in f32
 * without contraction (mpm:357), block = floor(base / block_size), key =
 * linear index of the block inside the block window [blo, blo + bn) with x
 * fastest; a particle whose block falls outside the window gets key
 * bn[0]*bn[1]*bn[2] (sorted last). */
void oracle_block_keys(int dim, int64_t n, const float* x, float inv_dx,
                       int block_size, const int32_t* blo, const int32_t* bn,
                       int32_t* keys);
/* Stable sort of particle indices 0..n-1 by key (counting sort). */
void oracle_sort_by_key(int64_t n, const int32_t* keys, int32_t n_keys,
                        int32_t* perm);

#ifdef __cplusplus
}
#endif
#endif
