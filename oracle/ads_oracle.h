/*
 * ads_oracle.h -- CPU restatement of the IGA-ADS heat-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product library (ads_mpi_amd/csrc) never links it.
 *
 * Every function restates one routine of macwozni/ADS_MPI (Fortran) and cites
 * the file:line it follows.  Arrays use the reference's layouts:
 *   - knot vector U(0:n+p+1)                        -> U[k]
 *   - basis table NN(0:d, 0:p, q, r)                -> NN[k + (d+1)*(a + (p+1)*(iq + q*ir))]
 *   - quadrature points X(q, r)                     -> X[iq + q*ir]
 *   - coefficients / RHS, x fastest                 -> c[x + (nx+1)*(y + (ny+1)*z)]
 *   - band storage (LAPACK dgbtrf layout, ldab=2kl+ku+1): A(i,j) -> ab[(kl+ku+i-j) + ldab*j]
 *
 * Parity status of each function is recorded in the header of ads_oracle.c.
 */
#ifndef ADS_ORACLE_H
#define ADS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- knots, quadrature, basis ---------------------------------------- */
void   orc_fill_open_knot(int n, int p, double* U);
int    orc_count_spans(int n, int p, const double* U);
int    orc_prepare_knot(int n, int p, double* U, int* nelem);
int    orc_gauss_rule(int q, double* X, double* W);
void   orc_ders_basis_funs(int i, double uu, int p, int d, const double* U, double* ders);
int    orc_find_span(int n, int p, double uu, const double* U);
int    orc_basis_data(int p, int m, const double* U, int d, int q, int r,
                      int* O, double* J, double* W, double* X, double* NN);
double orc_eval_spline(int d,
                       const double* Ux, int px, int nx,
                       const double* Uy, int py, int ny,
                       const double* Uz, int pz, int nz,
                       const double* coeffs, double x, double y, double z);

/* ---- 1-D matrices and banded LU --------------------------------------- */
int    orc_mkbbt_band(const double* U, int p, int n, int nelem, const double mix[4],
                      double* ab, int ldab);
int    orc_dgbtrf(int n, int kl, int ku, double* ab, int ldab, int* ipiv);
void   orc_dgbtrs(char trans, int n, int kl, int ku, int nrhs, const double* ab, int ldab,
                  const int* ipiv, double* b, int ldb);

/* ---- reorders (reorderRHS.F90) ---------------------------------------- */
void   orc_reorder_for_y(const int s[3], const double* F, double* F2);
void   orc_reorder_for_z(const int s[3], const double* F, double* F2);
void   orc_reorder_for_x(const int s[3], const double* F, double* F2);

/* ---- the heat step ---------------------------------------------------- */
enum { ORC_F0_CONST = 0, ORC_F0_SINPROD = 1, ORC_F0_BUMP3D = 2 };
enum { ORC_MODE_FAITHFUL = 0, ORC_MODE_SUMFACT = 1 };

typedef struct orc_setup orc_setup;

orc_setup* orc_setup_create(const int n[3], const int p[3]);
void       orc_setup_destroy(orc_setup* s);
int        orc_setup_dofs(const orc_setup* s, int dim);
int        orc_setup_nelem(const orc_setup* s, int dim);

/* F = Form3DRHS after FormUn (heat integrand, t > 0): kappa per element (ex
 * fastest) or NULL for kappa = 1.  elem_begin/elem_end restrict the element
 * range along ez (for bounded CPU-baseline samples); pass 0, -1 for all. */
int orc_form_rhs_heat(const orc_setup* s, int mode, const double* c, const double* kappa,
                      double dt, double* F, double* l2norm, int ez_begin, int ez_end);
/* F = Form3DRHS at t = 0 (projection of f0, RHS_eq.F90:82-86). */
int orc_form_rhs_project(const orc_setup* s, int mode, int f0_kind, double* F, double* l2norm);
/* the three 1-D solves with reorders (ADS.F90:514-779); F in/out x fastest.
 * line_limit > 0 solves only the first line_limit lines of each direction
 * (bounded timing samples); 0 solves all. */
int orc_solve_kron(const orc_setup* s, int faithful, double* F, int line_limit);
/* the three solves with A = mix1 M + mix2 K + mix3 B + mix4 B^T per direction */
int orc_solve_kron_mix(const orc_setup* s, const double mix[4], double* F);
/* MultiStep with the heat integrand (mix column-major 4 x 3) */
int orc_multistep(const orc_setup* s, const double mix[12], const double* c_in, const double* kappa, double dt,
                  double* c_out, double* l2norm);
/* transposed solve M^T x = F (adjoint of orc_solve_kron) */
int orc_solve_kron_T(const orc_setup* s, double* F);
/* one explicit heat step c_out = M^-1 (M - dt K(kappa)) c_in */
int orc_heat_step(const orc_setup* s, int mode, const double* c_in, const double* kappa,
                  double dt, double* c_out, double* l2norm);
/* t = 0 projection c = M^-1 F(f0) */
int orc_project(const orc_setup* s, int mode, int f0_kind, double* c_out, double* l2norm);
/* adjoint of one heat step: given c_n and lam_next = dL/dc_{n+1}, returns
 * lam_out = dL/dc_n and accumulates dkappa[e] += dL/dkappa_e. */
int orc_heat_step_adjoint(const orc_setup* s, const double* c_n, const double* kappa, double dt,
                          const double* lam_next, double* lam_out, double* dkappa);

/* ---- the oil problem (problems/oil) ------------------------------------
 * Integrand (problems/oil/RHS_eq.F90:69-91), t > 0:
 *   ret = J W (v u + Dt (-kq(X) exp(mi u) grad u . grad v + v (pump(X) - max(0, drain(X) u))))
 *   l2norm += J W v u^2,  drained += J W v Dt max(0, drain(X) u)
 * t = 0: ret = J W v u0(X) with u0 = initial_state (input_data.F90:317-329).
 * drained is the step's share of the reference's running total (RHS_eq.F90:86;
 * module variable, never reset, printed by ComputeResults).
 * Curves, pumps and drains are inputs (the reference draws the curves from
 * an unseeded random_number and reads pumps/drains from the command line). */
typedef struct {
    const double* cx;  /* curve points, cN * cL each (input_data.F90:8-9) */
    const double* cy;
    const double* cz;
    int cN, cL;
    const double* pumps;  /* x, y, z per pump */
    int npumps;
    const double* drains; /* x, y, z per drain */
    int ndrains;
    double mi, kq_min, kq_max, radius, pump_strength, drain_strength;
} orc_oil;

/* math.F90:44-51, 74-88 */
double orc_oil_lerp(double t, double x, double y);
double orc_oil_falloff(double r, double R, double t);
/* squared distance of p from segment [a, b] (input_data.F90:171-205) */
double orc_oil_dist_from_segment(const double p[3], const double a[3], const double b[3]);
/* min squared distance from the curves' segments, at most 1e3 (:212-238) */
double orc_oil_dist_from_curves(const orc_oil* o, const double p[3]);
double orc_oil_kq(const orc_oil* o, double x, double y, double z);          /* :290-303 */
double orc_oil_pumping(const orc_oil* o, double x, double y, double z);     /* :244-262 */
double orc_oil_draining(const orc_oil* o, double u, double x, double y, double z); /* :269-287 */
double orc_oil_initial_state(const orc_oil* o, double x, double y, double z);      /* :317-329 */
/* Kqvals(kx, ky, kz, ex, ey, ez) at the Gauss points (CacheKqValues,
 * input_data.F90:333-380), kx fastest */
int orc_oil_kq_cache(const orc_setup* s, const orc_oil* o, double* kq);
/* the oil right-hand side (Form3DRHS with the oil integrand; tzero selects
 * the t = 0 branch), faithful loops */
int orc_oil_form_rhs(const orc_setup* s, const orc_oil* o, const double* kq, const double* c, double dt, int tzero,
                     double* F, double* l2norm, double* drained);
/* one oil step with the faithful loops (Step with problems/oil RHS at t > 0);
 * drained (may be NULL) receives this step's sum J W v Dt vdrain */
int orc_oil_step(const orc_setup* s, const orc_oil* o, const double* kq, const double* c_in, double dt,
                 double* c_out, double* l2norm, double* drained);
/* t = 0: c = M^-1 F(initial_state) */
int orc_oil_project(const orc_setup* s, const orc_oil* o, double* c_out, double* l2norm);
/* curves by the reference's random walk (InitInputData, input_data.F90:139-165)
 * with uniforms from SplitMix64(seed) in the same call order */
void orc_oil_curves(uint64_t seed, int cN, int cL, double* cx, double* cy, double* cz);

/* ---- synthetic inputs ------------------------------------------------- */
uint64_t orc_splitmix64(uint64_t x);
void     orc_uniform(uint64_t seed, int64_t n, double lo, double hi, double* out);
void     orc_lognormal(uint64_t seed, int64_t n, double mu, double sigma, double* out);

#ifdef __cplusplus
}
#endif
#endif
