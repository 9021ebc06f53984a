/*
 * ads_mi355x.h -- C ABI of the MI355X-native IGA-ADS heat-step hot path.
 *
 * Plain extern "C": pointers and sizes only.  Every function returns 0 on
 * success or one of the ADS_ERR_* codes; ads_last_error() gives the text.
 * Nothing here calls exit()/stop: the reference's fatal stops become codes.
 *
 * Layouts follow the reference (macwozni/ADS_MPI):
 *   - coefficients and right-hand sides: doubles, x fastest,
 *       c[x + (nx+1)*(y + (ny+1)*z)]          (projection_engine.F90:230)
 *   - per-element fields (kappa, dkappa): ex fastest, E = nex*ney*nez
 *   - basis tables as BasisData returns them   (basis.F90:74-112)
 *
 * "_d" entry points take device pointers and run on the context's stream
 * (ads_set_stream); they do not synchronise.  Host-pointer entry points copy
 * and synchronise.  Calls on one context are not reentrant.
 *
 * Device buffers must be ordinary (coarse-grained) device memory: hipMalloc,
 * or a torch tensor on the GPU.  The assembly adds into its output with
 * floating-point global atomics (built with -munsafe-fp-atomics), which are
 * not guaranteed on fine-grained memory (hipHostMalloc, coherent
 * hipMallocManaged, fine-grained hipExtMallocWithFlags): results there may
 * be wrong without an error.
 */
#ifndef ADS_MI355X_H
#define ADS_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ADS_OK 0
#define ADS_ERR_INVALID 1     /* bad argument or size                           */
#define ADS_ERR_HIP 2         /* HIP runtime failure (message has the HIP error) */
#define ADS_ERR_NOMEM 3       /* device or host allocation failed               */
#define ADS_ERR_SINGULAR 4    /* zero pivot in a 1-D factorisation (ADS.F90:357-361) */
#define ADS_ERR_UNSUPPORTED 5 /* order/form outside the compiled kernels        */
#define ADS_ERR_NUMERIC 6     /* NaN/Inf in a checked scalar                    */

/* initial states for ads_project (integrand of the t = 0 step) */
#define ADS_F0_CONST 0   /* f = 1: problems/L2Projection/RHS_eq.F90:66-70          */
#define ADS_F0_SINPROD 1 /* f = sin(pi x) sin(pi y) sin(pi z)                      */
#define ADS_F0_BUMP3D 2  /* f = 2 bump3d(0.05, 0.4, x, y, z): problems/heat/input_data.F90:68 */

/* precision forms */
#define ADS_PREC_FP64 0    /* FP64 assembly and sweeps                            */
#define ADS_PREC_FP32_IR 1 /* FP32 assembly and sweeps + FP64 residual refinement */
#define ADS_PREC_FP32_IR_KRON 2 /* FP32 sweeps on FP64 residuals from a zero start */
#define ADS_PREC_FP64_KRON 3    /* FP64 sweeps on the fused FP64 Kronecker right-hand side */
#define ADS_PREC_FP32_IR_SEG 4  /* FP32 segmented solves on FP64 residuals from a zero start */

/* kernels for ads_time_kernel */
#define ADS_KERNEL_HEAT_FP64 0      /* heat RHS assembly, FP64 */
#define ADS_KERNEL_KRON_RESIDUAL 1  /* fused FP64 Kronecker residual, first pass (x = 0, FP32 out, l2) */
#define ADS_KERNEL_OIL 2            /* oil right-hand side (after ads_oil_setup) */
#define ADS_KERNEL_SOLVE_FP64 3     /* the FP64 mass solve in x, y, z (partitioned when the lines fit) */
#define ADS_KERNEL_SOLVE_FP64_LU 4  /* the same by the FP64 LU sweeps */
#define ADS_KERNEL_SOLVE_SEG_FP32 5 /* the segmented FP32 solve of ADS_PREC_FP32_IR_SEG */

/* right-hand-side forms (ads_form_rhs_sub, ads_sub_step) */
#define ADS_FORM_HEAT 0         /* problems/heat/RHS_eq.F90:69-81 (t > 0), diffusivity of ads_set_kappa */
#define ADS_FORM_PROJECT 1      /* J W v f0(X) with f0 of f0_kind (the t = 0 branch, RHS_eq.F90:82-86) */
#define ADS_FORM_OIL 2          /* problems/oil/RHS_eq.F90:76-91 at t > 0 (after ads_oil_setup) */
#define ADS_FORM_OIL_PROJECT 3  /* the oil problem's t = 0 branch, J W v initial_state(X) */

/* 1-D mass solves (ads_set_solver) */
#define ADS_SOLVER_AUTO 0        /* partitioned where every direction's lines fit it, else LU sweeps */
#define ADS_SOLVER_LU 1          /* banded LU with partial pivoting (dgbtrf/dgbtrs form), line-sequential */
#define ADS_SOLVER_PARTITIONED 2 /* no-interchange LU of M, lines cut into partitions solved in parallel */

typedef struct ads_ctx ads_ctx;

/* ---- set-up and teardown ------------------------------------------------ */

/* ADSS::initialize(n, p, ads, ads_data, mierr) -- src/ADS.F90:11-107.
 * n = number of basis functions minus one per direction (the heat CLI SIZE),
 * p = B-spline degree per direction (1..5 each; directions may differ, as
 * in the reference's p(3): the kernels run at the largest degree on tables
 * widened with exact zeros).
 * Builds the knots (PrepareKnot, knot_vector.F90:43-61), the basis tables
 * (BasisData, basis.F90:74-112), the 1-D mass and stiffness matrices
 * (MKBBT, projection_engine.F90:63-132) and their FP64 banded LU
 * (dgbtrf form: replaces the MUMPS analysis+factorisation in
 * SolveOneDirection, ADS.F90:294-361), and allocates device state. */
int ads_init(const int n[3], const int p[3], ads_ctx** out);

/* ADSS::Cleanup -- src/ADS.F90:802-863. */
int ads_destroy(ads_ctx* ctx);

/* Text of the last error on this context (or of the last failed ads_init
 * when ctx is NULL). */
const char* ads_last_error(const ads_ctx* ctx);

/* Use this HIP stream (hipStream_t passed as void*) for all device work;
 * NULL restores the context's own stream. */
int ads_set_stream(ads_ctx* ctx, void* hip_stream);
int ads_synchronize(ads_ctx* ctx);

/* ---- host mirrors of ADS_setup (src/Setup.F90:29-112) -------------------- */

/* dofs[d] = n(d)+1, nelem[d] = ads%nelem(d) */
int ads_get_sizes(const ads_ctx* ctx, int dofs[3], int nelem[3], int p[3]);
/* ads%Ux/Uy/Uz: n+p+2 doubles */
int ads_get_knots(const ads_ctx* ctx, int dim, double* U);
/* ads%Ox (0-based first DOF per element), ads%Jx, ads%Wx, ads%Xx(q, nelem),
 * ads%NNx(0:1, 0:p, q, nelem); any pointer may be NULL. */
int ads_get_basis(const ads_ctx* ctx, int dim, int* O, double* J, double* W, double* X, double* NN);
/* 1-D matrix A = mix1*M + mix2*K + mix3*B + mix4*BT as a dense row-major
 * (n+1)x(n+1) array: ComputeMatrix/MKBBT (projection_engine.F90:531-544). */
int ads_get_matrix(const ads_ctx* ctx, int dim, const double mix[4], double* dense);

/* ---- state held by the context (ads_data%F / R(:,2,2,2)) ----------------- */

int ads_set_coeffs(ads_ctx* ctx, const double* host);
int ads_get_coeffs(ads_ctx* ctx, double* host);
/* device pointer of the context's coefficient vector; the context alternates
 * two buffers, so the pointer is valid until the next ads_step/ads_project */
int ads_coeffs_device(ads_ctx* ctx, double** dptr);
/* per-element diffusivity (E doubles, ex fastest); NULL restores kappa = 1 */
int ads_set_kappa(ads_ctx* ctx, const double* host);

/* ---- the step ------------------------------------------------------------ */

/* ADSS::Step at t = 0 with the problem's t == 0 integrand
 * (problems/heat/RHS_eq.F90:82-86): c = M^-1 F(f0).  l2norm (may be NULL)
 * receives sum J*W*f0^2 = ||f0||^2 by quadrature (RHS_eq.F90:85). */
int ads_project(ads_ctx* ctx, int f0_kind, double* l2norm);

/* ADSS::Step(iter, RHS_fun, ...) at t > 0 (src/ADS.F90:437-461) with the heat
 * integrand (problems/heat/RHS_eq.F90:69-81, f = 0, kq = kappa_e):
 *   F_a = sum_e sum_q J W (v_a u - dt kappa_e grad u . grad v_a)
 *   c  <- (Mz (x) My (x) Mx)^-1 F      (Sub_Step, ADS.F90:471-794)
 * prec = ADS_PREC_FP64: FP64 throughout (refine ignored).
 * prec = ADS_PREC_FP32_IR: FP32 assembly and an FP32 solve, then refine =
 *   1..4 corrections, each an FP64 residual and an FP32 solve (the FP32
 *   solves are the segmented ones of ADS_PREC_FP32_IR_SEG).
 * prec = ADS_PREC_FP32_IR_KRON: refine = 1..4 FP32 sweeps in all, on FP64
 *   residuals from a zero start (the first residual is the right-hand side).
 * prec = ADS_PREC_FP32_IR_SEG: as ADS_PREC_FP32_IR_KRON, with each FP32 solve
 *   cut into segments of at most 32 entries per line, each solved by the LU
 *   recurrences started w entries outside it from zero state (the
 *   recurrences forget their state geometrically; w, at most 48, covers a
 *   decay to 1e-8), so a line's segments run in parallel.  Its factors are
 *   an FP64 LU of M without interchanges (M is symmetric positive definite),
 *   rounded to FP32.  With a per-element diffusivity, b = A(kappa) c comes
 *   from the FP64 element assembly and the passes solve M x = b.
 * prec = ADS_PREC_FP64_KRON: FP64 throughout; for a constant diffusivity the
 *   right-hand side is the Kronecker operator (Mz(x)My(x)Mx - dt K) c applied
 *   with the 1-D matrices (equal to the Gauss-point assembly, whose sum over
 *   elements and points factorises per direction), else the element assembly.
 * The FP64 residual is the Kronecker operator A c - M x for a constant
 * diffusivity, and b - M x with b = A(kappa) c from the FP64 element
 * assembly for a per-element one.  l2norm (may be NULL: no host sync)
 * receives sum J W u_n^2 of the state before the step (RHS_eq.F90:81),
 * always computed in FP64. */
int ads_step(ads_ctx* ctx, double dt, int prec, int refine, double* l2norm);

/* ADSS::MultiStep(iter, mix, RHS_fun, ...) (src/ADS.F90:391-429) with the
 * heat integrand, on the context's state: three substeps s = 1..3.  Substep
 * s assembles (Form3DRHS, projection_engine.F90:154-284)
 *   F_a = sum_e sum_q J W (v_a u_n - dt kappa_e grad u_{s-1} . grad v_a)
 * with the point values of the step's initial state (Un of FormUn(1)) and
 * the gradients of the latest coefficients (dUn of FormUn(s)), then solves
 * with A_s (x) A_s (x) A_s, A_s = mix(1,s) M + mix(2,s) K + mix(3,s) B +
 * mix(4,s) B^T in every direction (ComputeMatrix -> MKBBT,
 * projection_engine.F90:531-544, 63-132; B(i,j) = sum N_i' N_j J W).  mix is
 * column-major 4 x 3, Fortran's mix(4,3).  The factors of each distinct mix
 * column are built once and kept.  l2norm (may be NULL) as ads_step. */
int ads_multistep(ads_ctx* ctx, const double mix[12], double dt, double* l2norm);

/* FormUn(subun, ads, ads_data) -- src/projection_engine.F90:306-441.  The
 * point values and gradients are formed inside the assembly kernels, so this
 * fixes only which coefficients they come from: subun = 1 keeps a device copy
 * of the current coefficients as U_n, the point values (the reference's Un)
 * of every later heat right-hand side of ads_form_rhs_sub / ads_sub_step;
 * subun = 2, 3 (Un13, Un23) keep U_n.  The gradients (dUn) always come from
 * the current coefficients.  Other subun: ADS_ERR_INVALID (the reference
 * prints "wrong substep", :336-337). */
int ads_form_un(ads_ctx* ctx, int subun);

/* Form3DRHS(ads, ads_data, direction, substep, RHS_fun, l2norm) --
 * src/projection_engine.F90:154-284, with RHS_fun a device form (ADS_FORM_*):
 * F_host (x fastest, the reference's ads_data%F) receives the right-hand side
 * of the context's coefficients without advancing them, l2norm (may be NULL)
 * its l2 sum (sum J W v u^2, with u from U_n; J W v f0^2 for the projection
 * forms).  ADS_FORM_HEAT: sum J W (v u - dt kappa grad u . grad v), u of U_n
 * (of the current coefficients when ads_form_un(1) was not called or the state
 * has not changed since), grad u of the current coefficients.  ADS_FORM_OIL
 * needs U_n to be the current coefficients (ADS_ERR_UNSUPPORTED otherwise).
 * direction and substep (1..3) are RHS_fun's arguments; the device forms do
 * not depend on them. */
int ads_form_rhs_sub(ads_ctx* ctx, int form, int f0_kind, int direction, int substep, double dt, double* F_host,
                     double* l2norm);

/* Sub_Step(ads, iter, mix, direction, substep, RHS_fun, ads_data, l2norm,
 * mierr) -- src/ADS.F90:471-794: the right-hand side of ads_form_rhs_sub,
 * then the solves with A = mix(1) M + mix(2) K + mix(3) B + mix(4) B^T in x,
 * y and z (ComputeMatrix + SolveOneDirection per direction; each distinct mix
 * is factored once and kept, mix = (1, 0, 0, 0) is the mass solve), the result
 * the new coefficients (F -> R(:,2,2,2), :777-779).  l2norm (may be NULL) as
 * ads_form_rhs_sub. */
int ads_sub_step(ads_ctx* ctx, int form, int f0_kind, const double mix[4], int direction, int substep, double dt,
                 double* l2norm);

/* The drained sum of the last oil right-hand side (ads_oil_step, or
 * ADS_FORM_OIL in ads_sub_step / ads_form_rhs_sub); synchronises. */
int ads_last_drained(ads_ctx* ctx, double* drained);

/* The algorithm of the FP64 mass solves (ADS_SOLVER_*; no reference
 * counterpart: the reference's MUMPS solves one column at a time).  AUTO
 * (the default) takes the partitioned solve when every direction's lines
 * have at most 512 entries (16 partitions of 24, or of 32 for lines longer
 * than 384) and an x-y plane stays below 4 GiB; PARTITIONED is refused
 * beyond that. */
int ads_set_solver(ads_ctx* ctx, int solver);

/* Form3DRHS after FormUn (projection_engine.F90:154-284) on the context's
 * coefficients and diffusivity, heat integrand at t > 0, without advancing:
 * F (host, x fastest) = A(kappa) c; l2norm (may be NULL) = sum J W u^2. */
int ads_form_rhs(ads_ctx* ctx, double dt, double* F_host, double* l2norm);

/* l2norm of the last ads_step/ads_project (synchronises) */
int ads_last_l2norm(ads_ctx* ctx, double* l2norm);

/* ---- device-buffer forms (caller-owned buffers, context stream) ---------- */

/* Form3DRHS after FormUn (projection_engine.F90:154-284, :306-441), heat
 * integrand: F = A(kappa) c with A = M - dt K(kappa).  kappa_d may be NULL
 * (kappa = 1).  l2_d (may be NULL) receives one double. */
int ads_form_rhs_d(ads_ctx* ctx, const double* c_d, const double* kappa_d, double dt,
                   double* F_d, double* l2_d);
/* The three 1-D solves of Sub_Step (ADS.F90:514-779; SolveOneDirection
 * job 3 per column, :363-368): F <- M^-1 F in place.  trans = 1 solves with
 * M^T (dgbtrs 'T' semantics) for the adjoint. */
int ads_solve_d(ads_ctx* ctx, double* F_d, int trans);
/* One heat step c_out = M^-1 A(kappa) c_in (c_in may equal c_out). */
int ads_step_d(ads_ctx* ctx, const double* c_in_d, double* c_out_d, const double* kappa_d,
               double dt, int prec, int refine, double* l2_d);
/* Adjoint of ads_step_d (no reference counterpart; the reference has no
 * gradients).  Given c_n and lam_next = dL/dc_{n+1}:
 *   mu = M^-T lam_next,  lam_out = A(kappa) mu = dL/dc_n,
 *   dkappa[e] += -dt sum_{q in e} J W grad c_n . grad mu   (if dkappa_d != NULL).
 * prec = ADS_PREC_FP64 or ADS_PREC_FP64_KRON: all in FP64 (refine ignored).  ADS_PREC_FP32_IR or
 * ADS_PREC_FP32_IR_KRON: mu by FP32 transposed sweeps with refine-1 FP64
 * mass-residual corrections (refine = 1..4); ADS_PREC_FP32_IR_SEG: the same
 * with segmented solves (M is symmetric); then the assembly applied to
 * mu and the diffusivity gradient in FP32 arithmetic (dkappa accumulated in
 * FP64).  lam_out may equal lam_next, not c_n. */
int ads_step_adjoint_d(ads_ctx* ctx, const double* c_n_d, const double* lam_next_d,
                       const double* kappa_d, double dt, int prec, int refine, double* lam_out_d,
                       double* dkappa_d);

/* ---- the oil problem (problems/oil) ---------------------------------------
 * Integrand (problems/oil/RHS_eq.F90:76-91) at t > 0:
 *   J W (v u + Dt (-kq(X) exp(mi u) grad u . grad v + v (pump(X) - max(0, drain(X) u))))
 * with kq(X) = lerp(falloff(0, 0.06, dist(X)), Kqmin, Kqmax), dist the
 * distance from the curves, pump/drain sums of strength * falloff(0, radius,
 * |source - X|) (input_data.F90:171-303); at t = 0 the projection of
 * initial_state (:317-329).  The reference draws the curves from an unseeded
 * random_number (InitInputData, :139-165) and reads pumps and drains from the
 * command line (InitPumps, :99-136); here all of them are inputs. */

/* InitInputData's curves, PrecomputeKq (input_data.F90:384-395) and the
 * problem constants.  cx, cy, cz: cN curves of cL points (point j of curve i
 * at [i*cL + j]); pumps, drains: x, y, z triples; params = {mi, Kqmin, Kqmax,
 * radius, pumping_strength, draining_strength} (input_data.F90:6-13).  Caches
 * kq at every Gauss point on the device.  Host arrays may be freed on return. */
int ads_oil_setup(ads_ctx* ctx, const double* cx, const double* cy, const double* cz, int cN, int cL,
                  const double* pumps, int npumps, const double* drains, int ndrains,
                  const double params[6]);
/* Step(0, ...) of problems/oil/main.F90:66-77: c = M^-1 F(initial_state). */
int ads_oil_project(ads_ctx* ctx, double* l2norm);
/* Step(iter > 0, ...) with the oil integrand: c <- M^-1 F(c).  l2norm and
 * drained (either may be NULL: no host sync) receive sum J W v u^2 and this
 * step's sum J W v Dt vdrain (the reference's module variable `drained`
 * accumulates these over all steps, RHS_eq.F90:86). */
int ads_oil_step(ads_ctx* ctx, double dt, double* l2norm, double* drained);
/* the cached kq values (PrecomputeKq's Kqvals, problems/oil/input_data.F90:384-395)
 * in Kqvals(kx, ky, kz, ex, ey, ez) order (kx fastest): E (p+1)^3 doubles.
 * The device copy is point-major (point index times E plus element), so this
 * call copies it to the host and permutes it there (a host buffer of the same
 * size; ADS_ERR_NOMEM when that allocation fails). */
int ads_oil_get_kq(ads_ctx* ctx, double* kq_host);
/* Host only: the factor tables of the segmented form for one direction of
 * SIZE n, degree p (no reference counterpart): tab = (n+2) x (2p+1) FP32 rows
 * (n + 1 basis functions), row j = L(j, j-1..j-p), 1/U(j, j), U(j, j+1..j+p)
 * of the mass matrix's LU factors (zero outside the matrix), then the middle
 * row, which rows [*lo, *hi) match to 1e-7 (L U = M, no interchanges);
 * *w = the warm-up distance. */
int ads_host_seg_tables(int n, int p, int* w, int* lo, int* hi, float* tab);

/* Host only: the tables of the partitioned FP64 solve for one direction of
 * SIZE n, degree p (no reference counterpart): the line's n + 1 entries cut
 * into *K partitions of *S (the last may be shorter; S = 24, or 32 for lines
 * of 385 to 512 entries; longer lines: ADS_ERR_UNSUPPORTED); rows j = 0 .. K*S + p
 * (zero beyond n): lrow[j*p + t] = L(j, j-1-t), urow[j*(p+1)] = 1/U(j, j) and
 * urow[j*(p+1) + t] = U(j, j+t) of the no-interchange LU of the mass matrix,
 * gsp[j*p + i] / hsp[j*p + i] the forward / backward recurrence's response at
 * j to a unit state i at the edge of j's partition.  Arrays may be NULL (to
 * learn K and S first). */
int ads_host_part_tables(int n, int p, int* K, int* S, double* lrow, double* urow, double* gsp, double* hsp);

/* curves by InitInputData's random walk with uniforms from SplitMix64(seed)
 * (the reference's own stream is unseeded): cN*cL points each of cx, cy, cz */
int ads_oil_curves(uint64_t seed, int cN, int cL, double* cx, double* cy, double* cz);

/* ---- output ---------------------------------------------------------------- */

/* The spline with the context's coefficients on a tensor grid, for
 * PrintSolution (src/ADS.F90:955-989): SaveSplinePlot (src/plot.F90:97-145)
 * samples x_i = lerp(i/(res_x-1), lo_x, hi_x) (math.F90:44-51), likewise y
 * and z, and evaluates EvalSpline with d = 0 (src/basis.F90:323-369) at each
 * point.  vals (host, res_x*res_y*res_z doubles) is x fastest, the order
 * VtkOutput writes (src/vtk.F90:55-60).  res = 2..65536 per direction. */
int ads_eval_grid(ads_ctx* ctx, const int res[3], const double lo[3], const double hi[3], double* vals);

/* No counterpart in the reference: the output grid of ads_eval_grid for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same points and
 * values as ads_eval_grid, written to the caller's device buffer vals_d
 * (res_x*res_y*res_z doubles, x fastest) on the context stream, without
 * synchronising.  It calls no hipMalloc/hipFree: its tables go to a static
 * device array of 540,672 bytes that each GPU holds until the process ends,
 * so res is limited to 2..4096 per direction (ADS_ERR_INVALID beyond) and
 * the degree to 4.  All calls on one GPU share that array: calls that
 * overlap on one GPU (other streams, other threads, other contexts) race. */
int ads_eval_grid_noalloc_d(ads_ctx* ctx, const int res[3], const double lo[3], const double hi[3], double* vals_d);

/* ---- 1-D matrices (ComputeMatrix / SolveOneDirection) -------------------- */

typedef struct ads_matrix ads_matrix;

/* ComputeMatrix(U, p, n, nelem, mix, sprsmtrx) -- src/projection_engine.F90:
 * 531-544 -> MKBBT (:63-132): the (n+1) x (n+1) matrix mix(1) M + mix(2) K +
 * mix(3) B + mix(4) B^T of the degree-p B-splines on the knot vector
 * U(0:n+p+1) (non-decreasing; nelem = its non-empty spans in [U(p), U(n+1)]),
 * entries by Gauss quadrature as MKBBT sums them, and its banded LU with
 * partial pivoting (dgbtrf form: the MUMPS analysis and factorisation of
 * SolveOneDirection, src/ADS.F90:306-361, done once here).  Host work; p =
 * 1..9.  ADS_ERR_SINGULAR for a zero pivot (the reference stops there). */
int ads_matrix_create(const double* U, int n, int p, int nelem, const double mix[4], ads_matrix** out);
int ads_matrix_destroy(ads_matrix* m);
/* Text of the matrix's last error (of the last failed ads_matrix_create when m is NULL). */
const char* ads_matrix_last_error(const ads_matrix* m);
/* the matrix as a dense row-major (n+1) x (n+1) array */
int ads_matrix_dense(const ads_matrix* m, double* dense);
/* SolveOneDirection(RHS, eqnum, n, p, sprsmtrx) -- src/ADS.F90:294-379:
 * rhs (host, column-major (n+1) x eqnum, each column one right-hand side) is
 * overwritten by the solutions, computed on the GPU (one lane per column, the
 * LU substitution sweeps; p = 1..9, as ads_matrix_create).  Copies in and out
 * and synchronises. */
int ads_matrix_solve(ads_matrix* m, double* rhs, int eqnum);

/* ---- host-only helpers (no GPU needed) ----------------------------------- */

/* The tables ads_init builds for one direction, without a device: knots U
 * (n+p+2), O (nelem), J (nelem), W (p+1), X ((p+1)*nelem), NN
 * (2*(p+1)*(p+1)*nelem) in the layouts of ads_get_knots/ads_get_basis.
 * nelem = n - p + 1.  Any pointer may be NULL. */
int ads_host_basis(int n, int p, double* U, int* O, double* J, double* W, double* X, double* NN);

/* The per-direction tables ads_eval_grid uses, without a device: for the res
 * points u_i = lerp(i/(res-1), lo, hi), first[i] = FindSpan - p
 * (basis.F90:226-259) and vals[i*(p+1)+a] = the p+1 B-splines at u_i. */
int ads_host_grid_tables(int n, int p, int res, double lo, double hi, int* first, double* vals);

/* ---- synthetic inputs (host) --------------------------------------------- */

/* SplitMix64 counter generator: u_i = (splitmix64(seed*0xD1B54A32D192ED03 + i) >> 11) * 2^-53,
 * out[i] = lo + (hi - lo) u_i. */
int ads_fill_uniform(uint64_t seed, int64_t n, double lo, double hi, double* out);
/* lognormal(mu, sigma) by Box-Muller on the same generator (key offset
 * 0x8CB92BA72F3D8DD7, u1 = u_{2i}, u2 = u_{2i+1}). */
int ads_fill_lognormal(uint64_t seed, int64_t n, double mu, double sigma, double* out);

/* ---- measurement ------------------------------------------------------------ */

/* No reference counterpart (measurement support).  Times `launches`
 * back-to-back launches of one kernel of the step (ADS_KERNEL_*) on the
 * context's current coefficients with HIP events recorded on the context
 * stream around the launches only (no fills, reductions or copies inside the
 * events); results go to internal scratch.  ms_per_launch = elapsed / launches. */
int ads_time_kernel(ads_ctx* ctx, int which, double dt, int launches, double* ms_per_launch);

#ifdef __cplusplus
}
#endif
#endif /* ADS_MI355X_H */
