/*
 * lcb_oracle.h - CPU restatement of the reference lcbinint binary-lens hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the parity tests,
 * the golden-vector generator and bench.py's cpu_baseline leg compare the HIP
 * product against.  Nothing in the product (lcbinint_amd/) links, loads or calls
 * it.  Every function cites the reference file:line whose algorithm it restates.
 *
 * Pinning (see tests/test_oracle.py):
 *   - point-source layer: bitwise against the compiled reference
 *     (oracle/_ref) on randomized inputs, and against the literal golden values
 *     of tests/unit/test_core.cpp:86,94 and tests/regression/test_vbm_consistency.py:8-13;
 *     the q->0 single-lens closed form A=(u^2+2)/(u sqrt(u^2+4)).
 *   - root solver: residual |p(z)| on the polynomial cases of
 *     tests/regression/test_solver_vbm_consistency.py:4-33 and bitwise vs oracle/_ref.
 *   - limb-darkening table: closed form I(r^2) on the table nodes.
 *   - hexadecapole: bitwise vs oracle/_ref (through the dispatcher) and the
 *     rho->0 limit A_hex -> A_ps.
 */
#ifndef LCB_ORACLE_H
#define LCB_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Skowron & Gould general solver, restated (SkowronGould.cpp:79-168).
 * coeffs: interleaved (re, im), constant first, degree+1 entries.
 * roots: interleaved, `degree` entries; used as starting points if use_start. */
void lcbo_cmplx_roots_gen(const double* coeffs, int degree, int polish, int use_start,
                          double* roots);

/* binary_mag0 (point_source_magnifier.cpp:567 -> 652), VBM convention. */
double lcbo_binary_mag0(double s, double q, double x, double y, int* image_count);

/* binary_mag0_with_derivatives (point_source_magnifier.cpp:583-614). */
double lcbo_binary_mag0_with_derivatives(double s, double q, double x, double y,
                                         int* image_count, double* indicator);

/* binary_image_candidates (point_source_magnifier.cpp:739-778): rows of
 * (x, y, jacobian, residual, physical) sorted by residual; returns count. */
int lcbo_binary_image_candidates(double s, double q, double x, double y, double* out);

/* Point-source batch: n solves of binary_mag0(s[i], q[i], (x[i], y[i])). */
void lcbo_binary_mag0_many(const double* s, const double* q, const double* x,
                           const double* y, long n, double* out, int* image_count);

/* Static VBM trajectory (lcbinint_pybind.cpp:1143-1148) + point-source light
 * curve with effective q = 1/q (lcbinint_pybind.cpp:1125-1139). */
void lcbo_light_curve_ps(const double* times, long n, double t0, double tE, double u0,
                         double alpha, double s, double q, double* out);

/* Limb-darkening node table (finite_source_magnifier.cpp:3979-3999) and lookup
 * (finite_source_magnifier.cpp:4001-4012). table must hold 20001 doubles. */
void lcbo_ld_table(double c, double d, double* table);
double lcbo_ld_lookup(const double* table, double normalized_radius2);

/* Hexadecapole (finite_source_magnifier.cpp:792-847) with known A0; returns the
 * magnification, relative error through rel_err. */
double lcbo_hexadecapole(double s, double q, double x, double y, double rho, double c,
                         double d, double a0, double* rel_err);

/* Critical curve at phase (cos_phi, sin_phi): the quartic of
 * critical_curve_polynomial_coefficients (finite_source_magnifier.cpp:528-545)
 * solved as PolynomialRootSolver::solve (polynomial_roots.cpp:35-73), then
 * mapped to the source plane by the complex lens equation binary_lens_equation
 * (point_source_magnifier.cpp:780-792, as caustic_points_at_phase :547-567 does)
 * into caustic[8], and by the real map map_binary_lens_real
 * (finite_source_magnifier.cpp:278-291, as critical_sources_at_phase
 * :1622-1641 does) into critical[8]; interleaved (x, y). */
void lcbo_critical_curve(double s, double q, double cos_phi, double sin_phi, double* caustic,
                         double* critical);

/* Number of floating-point operations counted by the restatement since the last
 * reset (+,-,*,/,sqrt = 1 each; compares 0).  Used to state algorithmic work. */
double lcbo_flop_count(void);
void lcbo_flop_reset(void);

#ifdef __cplusplus
}
#endif

#endif
