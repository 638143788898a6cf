/* ORACLE — test infrastructure, not product code.
 *
 * A plain C (+OpenMP) restatement of OpenCAL-OMP's per-step hot path for the
 * three models on the accelerated path (SciddicaT, Life, SCIARA-fv2).  It keeps
 * the reference loop structure: one sweep per elementary process, then the
 * next->current copy of every double-layer substate (full grid, or active
 * cells after a row-major rebuild of the active set), then steering and its
 * copy.  Only tests/, the smoke check and the bench CPU-baseline leg may load
 * it; the product library never links or calls it.
 *
 * Each function names the reference lines it restates.  How each is pinned
 * (reference build in oracle/_ref, shipped result files, closed forms) is
 * listed in the test files that check it.
 */
#ifndef OPENCAL_ORACLE_H
#define OPENCAL_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* SciddicaT, full grid (CAL_NO_OPT) or active cells (CAL_OPT_ACTIVE_CELLS).
 * z, h: initial altitude and debris thickness (row-major, rows*cols);
 * the init step z -= h where h > 0 is applied inside, as the reference init.
 * active_counts: optional, A.size_current after each step (active mode). */
int orc_sciddicat_run(int rows, int cols, const double* z, const double* h,
                      double p_r, double p_eps, int steps, int active,
                      double* out_z, double* out_h, int* active_counts);

/* The unsafe SciddicaT of cal_sciddicaT-unsafe-glut on the serial engine
 * (active cells, next-layer read-modify-writes to the neighbours).  order 0:
 * the serial visit order; 1: the device's colour phases.  z, h before the
 * init (which lowers z where h > 0); out_active: the final list length. */
int orc_sciddicat_unsafe_run(int rows, int cols, const double* z0, const double* h0, double p_r, double p_eps,
                             int steps, int order, double* out_h, int* out_active);
/* The same; track 0: cal_sciddicaT-unsafe (ACTIVE_CELLS not defined, the
 * list stays empty and every sweep covers the grid). */
int orc_sciddicat_unsafe_run_track(int rows, int cols, const double* z0, const double* h0, double p_r,
                                   double p_eps, int steps, int order, int track, double* out_h, int* out_active);

/* Conway Life, Moore, toroidal, uint8 0/1 cells. */
int orc_life_run(int rows, int cols, const unsigned char* alive0, int steps,
                 unsigned char* out);

/* SCIARA-fv2 (same argument meaning as the reference harness). math_mode:
 * 0 = host libm (bit-identical to the reference build), 1 = the correctly
 * rounded elementary functions shared with the device kernels (crmath.h). */
int orc_sciara_run(int rows, int cols, const double* z0, double cell_size,
                   const double* params, int n_vents, const int* vent_rc,
                   int n_rates, const double* rates, unsigned emission_time,
                   int steps, int use_active, int math_mode,
                   double* out_z, double* out_h, double* out_T, double* out_Msl,
                   int* out_steps, int* out_active);

/* The correctly rounded functions of crmath.h elementwise, their binary128
 * values rounded to double, and the host libm calls of the reference (fn: 0
 * pow(10, x), 1 atan, 2 cos, 3 pow(x, 3.0), 4 pow(x, 1.0/3.0)); -1 for an
 * unknown fn.  orc_cr_check counts mismatches over n seeded points of
 * [lo, hi] (out: vs binary128, vs libm, max ulp vs libm, max ulp vs
 * binary128, libm vs binary128). */
int orc_cr_eval(int fn, const double* x, int n, double* y);
int orc_quad_eval(int fn, const double* x, int n, double* y);
int orc_libm_eval(int fn, const double* x, int n, double* y);
int orc_cr_check(int fn, double lo, double hi, int log_scale, long n, unsigned long long seed, double* out);

/* One SciddicaT full-grid step h -> h_next (flows with the reference's
 * minimisation + width update), used by the adjoint oracle. */
void orc_sciddicat_step(int rows, int cols, const double* z, const double* h,
                        double p_r, double p_eps, double* h_next);

/* Reverse-mode adjoint of `steps` full-grid SciddicaT steps from h0 (after
 * init) with the elimination sets frozen at their forward values.
 * lam_T: dL/dh_T.  Outputs dL/dh0, dL/dz (all cells), dL/dr, dL/deps. */
int orc_sciddicat_adjoint(int rows, int cols, const double* z, const double* h0,
                          double p_r, double p_eps, int steps, const double* lam_T,
                          double* lam_h0, double* lam_z, double* d_r, double* d_eps);

int orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
