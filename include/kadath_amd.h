/*
 * kadath_amd.h -- C-ABI of the MI355X Newton-Raphson path for Kadath/FUKA.
 *
 * Plain pointers and sizes only.  Every buffer argument is DEVICE memory unless
 * stated.  Every function returns an int status: 0 = ok, < 0 = error
 * (KAD_ERR_*), and for the factorisation i > 0 = LAPACK-style info (exactly
 * zero pivot U(i-1,i-1)).
 *
 * The reference has no C ABI: its Newton step is the C++ member
 *   bool System_of_eqs::do_newton(double precision, double& error, SOLVER)
 *   (reference include/system_of_eqs.hpp:871, src_par/do_newton.cpp:65-430)
 * swapped at link time (build_release/CMakeLists.txt:24-31).  The functions below
 * are the thin layer under that member, one per reference subsystem that moves to
 * the GPU; include/kadath_amd.hpp restates the member itself (a C++
 * System_of_eqs / Space_spheric / Scalar / Vector over this layer).
 */
#ifndef KADATH_AMD_H
#define KADATH_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KAD_OK 0
#define KAD_ERR_ARG (-1)
#define KAD_ERR_HIP (-2)
#define KAD_ERR_ALLOC (-3)
#define KAD_ERR_NOCONV (-4)
#define KAD_ERR_SINGULAR (-5)

/* Factor precisions. */
#define KAD_PREC_F64 0  /* fp64 storage, fp64 MFMA                                  */
#define KAD_PREC_F32 1  /* fp32 storage, fp32 MFMA                                  */
#define KAD_PREC_BF16 2 /* fp32 storage, bf16 MFMA operands, fp32 accumulation      */

typedef struct kad_ctx kad_ctx;

/* ---- context ------------------------------------------------------------ */
/* One context per GPU: owns a HIP stream and all scratch buffers.
 * Replaces the process-global state of the reference (FFTW plan maps
 * src/Coef/coef_1d.cpp:34-44, transform scratch include/memory.hpp:55-69). */
int kad_ctx_create(kad_ctx** ctx, int device);
/* Systems built on a context (kad_xcts_create, kad_system_build) must be destroyed
 * before it. */
int kad_ctx_destroy(kad_ctx* ctx);
void* kad_ctx_stream(kad_ctx* ctx);
int kad_sync(kad_ctx* ctx);

/* Live timing of one kernel class with HIP events on the stream of each launch: kind
 * 0 = the MFMA GEMM launches (work in executed flops: tile pairs skipped as structural
 * zeros are not counted; the device counts them, no host synchronisation), kind 1 =
 * the fused 3-D spectral passes of the residual / Jacobian programs (work in
 * algorithmic HBM bytes).  on = 1 starts a new measurement window (at most 2^18
 * launches; read returns KAD_ERR_ALLOC when a window held more); read returns total
 * work, kernel milliseconds and launches. */
int kad_kernel_timing(kad_ctx* ctx, int kind, int on);
int kad_kernel_timing_read(kad_ctx* ctx, int kind, double* work, double* ms, int* launches);

/* ---- dense linear algebra (replaces pdgesv_, src_par/do_newton.cpp:317-320,
 *      and Matrice::set_lu/solve, src/Matrice/matrice.cpp:192-250) ---------- */

/* C -= A*B (op 0) or C = A*B (op 1); column-major; prec selects the MFMA type. */
int kad_gemm(kad_ctx* ctx, int prec, int op, int m, int n, int k, const void* A, int64_t lda, const void* B,
             int64_t ldb, void* C, int64_t ldc);

/* In-place LU with partial pivoting of the n x n column-major A (fp64 for
 * KAD_PREC_F64, fp32 otherwise).  ipiv: n ints, 0-based interchanges (LAPACK
 * dgetrf semantics shifted by one); perm (may be NULL): n ints, composed row
 * order used by kad_lu_solve.  nb: outer panel width (0 = default: KAD_LU_NB, else
 * 256, 512 for BF16).  Returns info (> 0: first exact zero pivot, 1-based), or
 * KAD_ERR_HIP when a wait of the persistent panel reached its spin limit (the
 * factors are then not valid). */
int kad_lu_factor(kad_ctx* ctx, int prec, int n, void* A, int64_t lda, int* ipiv, int* perm, int nb);

/* nbat LUs of order n (a multiple of 256, nbat <= 16) factored together: matrix z
 * at A + z strideA (same lda), interchanges at ipiv + z n (0-based, local to the
 * matrix), composed row orders at perm + z n (may be NULL); info: HOST, nbat words
 * (first exact zero pivot, 1-based, or 0).  Every launch covers all matrices, so the
 * panel chain is one matrix's: the diagonal domain blocks of the domain-decomposed
 * solve below, which replaces the single pdgesv_ factorisation
 * (src_par/do_newton.cpp:317-320) by independent per-domain ones.  The matrices are
 * taken as dense; each pivot search is LAPACK's within its matrix. */
int kad_lu_factor_batched(kad_ctx* ctx, int prec, int n, int nbat, void* A, int64_t lda, int64_t strideA, int* ipiv,
                          int* perm, int* info);

/* As kad_lu_factor_batched without row interchanges (static pivoting; no reference
 * counterpart, ScaLAPACK pivots every column): the caller has already put the rows of
 * every matrix in the order of an earlier factorisation's pivots, so row j is the
 * pivot of column j and no pivot search is made -- each 32-column panel is one launch
 * with no exchange between workgroups.  One matrix may have any order; a batch,
 * orders that are multiples of 256.  ipiv receives the identity; info (HOST, nbat)
 * the first exactly-zero pivot; growth (HOST, nbat) the largest |multiplier| of each
 * matrix (partial pivoting keeps it at 1; a large one means the order no longer suits
 * the matrix and the factors should not be used). */
int kad_lu_factor_batched_static(kad_ctx* ctx, int prec, int n, int nbat, void* A, int64_t lda, int64_t strideA,
                                 int* ipiv, int* info, double* growth);

/* B_z <- A_z^{-1} B_z (nrhs columns each, B_z at B + z strideB) with the FP64
 * factors of kad_lu_factor_batched (n a multiple of 32). */
int kad_lu_solve_batched(kad_ctx* ctx, int n, int nbat, const double* LU, int64_t lda, int64_t strideA, const int* perm,
                         double* B, int64_t ldb, int64_t strideB, int nrhs);

/* Domain-decomposed solve of J x = b, replacing the single pdgesv_ of
 * src_par/do_newton.cpp:317-320 for a system whose unknowns and rows split into
 * blocks (its domains) coupled only through a few rows (the interface matching rows,
 * take_into_account, src/System_of_eqs/solver.cpp:181,203).  Block b holds the J rows
 * rows[o_b .. o_b + blk_size[b]) and the unknowns cols[o_b ..) (o_b = sum of the
 * earlier sizes; HOST arrays, every row and column of J once, each block square).
 * kad_dd_factor factors the diagonal blocks together (kad_lu_factor_batched), then
 * the Woodbury coupling: D^{-1} of the coupling rows' unit vectors and the
 * capacitance matrix I + C D^{-1} P (C: the coupling rows outside their own block),
 * with its own LU.  Returns 0, < 0 an error, > 0 a singular block or capacitance
 * matrix.  kad_dd_solve runs flexible GMRES on the fp64 J (device, ld >= n),
 * preconditioned by that solve, to ||b - J x||_inf <= tol ||b||_inf (KAD_ERR_NOCONV
 * otherwise; iters as kad_ir_solve).  kad_dd_apply applies the preconditioner alone
 * (v: rows in, unknowns out).  kad_dd_info: sizes [padded block order, coupling rows,
 * the most per block, blocks], times (s) of the last factorisation [block gather,
 * block LU (the scan of J for coupling rows runs beside it on a second stream),
 * D^{-1} P, capacitance product, capacitance LU]. */
typedef struct kad_dd_s kad_dd;
int kad_dd_create(kad_ctx* ctx, int n, int nblk, const int* blk_size, const int* rows, const int* cols, kad_dd** out);
int kad_dd_factor(kad_dd* dd, const double* J, int64_t ldj);
/* kad_dd_factor with static pivoting (kad_lu_factor_batched_static): the rows of every
 * block and of the capacitance matrix are taken in the pivot orders of the last
 * kad_dd_factor, for a J close to the one it factored (the next Newton step's).
 * Returns 0 (factored), < 0 an error, or > 0 when the orders do not apply: 1 no
 * earlier kad_dd_factor, 2 a zero pivot or a multiplier above max_growth (>= 1) in a
 * block, 3 other coupling rows, 4 the same in the capacitance matrix; the factors are
 * then not usable and kad_dd_factor should follow. */
int kad_dd_factor_static(kad_dd* dd, const double* J, int64_t ldj, double max_growth);
/* kad_dd_factor (partial pivoting in every block) that keeps the capacitance matrix's
 * pivot order: I + C D^{-1} P does not depend on the blocks' pivot orders, so when the
 * coupling rows are those of the last factorisation its rows are taken in that order
 * and factored without interchanges (kad_lu_factor_batched_static), kept while every
 * multiplier stays within max_growth (>= 1); otherwise it is factored with partial
 * pivoting as in kad_dd_factor.  Returns as kad_dd_factor; kad_dd_pivoting then gives
 * static_orders 0 and the kept order's largest |multiplier| (0 when it was refactored).
 * The Newton solve of do_newton.cpp:256-330 factors every step; this is the same
 * solve with less work on the coupling. */
int kad_dd_factor_keep(kad_dd* dd, const double* J, int64_t ldj, double max_growth);
int kad_dd_solve(kad_dd* dd, const double* J, int64_t ldj, const double* b, double* x, double tol, int max_outer,
                 int* iters, double* resid);
int kad_dd_apply(kad_dd* dd, double* v);
int kad_dd_info(const kad_dd* dd, int* sizes, double* times);
/* whether the last factorisation took the kept pivot orders (kad_dd_factor_static)
 * and its largest |multiplier| (0 after kad_dd_factor; the capacitance matrix's after
 * kad_dd_factor_keep kept its order); HOST out, either may be NULL */
int kad_dd_pivoting(const kad_dd* dd, int* static_orders, double* growth);
int kad_dd_destroy(kad_dd* dd);

/* Solve op(A) x = b in place with the factors (b fp64).  trans = 1: A^T x = b
 * (adjoint solve; the reference's Matrice::solve uses trans 'T' on row-major
 * storage, matrice.cpp:221-250). */
int kad_lu_solve(kad_ctx* ctx, int prec, int n, const void* LU, int64_t lda, const int* perm, double* b, int trans);

/* As kad_lu_solve, for n <= 8192, allocating nothing: no reference counterpart (the
 * reference solves inside ScaLAPACK / LAPACK); for callers that cannot allocate device
 * memory (code captured into a HIP graph, serving loops, harnesses that forbid
 * hipMalloc).  Its tables and work vectors live in the library's static per-GPU
 * scratch (34 MB, held until the process ends, its address taken at
 * kad_ctx_create); n > 8192 returns KAD_ERR_ARG before any device work.  Calls that
 * overlap on one GPU race on that scratch.  Work is queued on the context's stream. */
int kad_lu_solve_noalloc(kad_ctx* ctx, int prec, int n, const void* LU, int64_t lda, const int* perm, double* b,
                         int trans);

/* y = b - op(A) x (b != NULL) or y = op(A) x, fp64, column-major A. */
int kad_gemv64(kad_ctx* ctx, int n, const double* A, int64_t lda, const double* x, const double* b, double* y,
               int trans);

/* A <- diag(R) A diag(C) in place; R, C (device, n) are powers of two chosen
 * as LAPACK dgeequb does (row max, then column max of the row-scaled matrix),
 * so the scaling is exact in fp64. */
int kad_equilibrate(kad_ctx* ctx, int n, double* A, int64_t lda, double* R, double* C);
/* B = (float)(diag(R) A diag(C)) with R, C computed as by kad_equilibrate; A unchanged. */
int kad_cast_equilibrated(kad_ctx* ctx, int n, const double* A, int64_t lda, float* B, int64_t ldb, double* R,
                          double* C);
/* v <- s .* v (device) */
int kad_vec_scale(kad_ctx* ctx, int n, const double* s, double* v);

/* B = (float)A, n x n column-major. */
int kad_cast_f64_to_f32(kad_ctx* ctx, int n, const double* A, int64_t lda, float* B, int64_t ldb);

/* Mixed-precision solve: fp64 residuals with A64, corrections from the
 * low-precision factors (prec_lo, LUlo, perm).  mode 0 = classical IR,
 * mode 1 = GMRES-IR (flexible GMRES right-preconditioned by the factors, so the
 * attainable accuracy is that of the fp64 products).  Stops at
 * ||b - op(A) x||_inf <= tol ||b||_inf, or when two outer steps in a row gain less
 * than a factor 2 (a step that increases the residual is undone).  Products with A
 * skip its exactly-zero 256 x 1024 tiles, found by one scan per call.
 * iters[0] = outer steps, iters[1] = inner GMRES steps.  Returns KAD_ERR_NOCONV if
 * tol is not reached (x holds the best iterate; x = 0 if no step helped). */
int kad_ir_solve(kad_ctx* ctx, int n, const double* A64, int64_t lda, int prec_lo, const void* LUlo, int64_t ldlu,
                 const int* perm, const double* b, double* x, double tol, int max_outer, int mode, int trans,
                 int* iters, double* resid);

/* GMRES-IR Krylov dimension per restart of the context (default 100, at most 4096;
 * the basis takes (2 m + 1) n doubles). */
int kad_ir_set_restart(kad_ctx* ctx, int m);
/* Residual history of the last refinement (HOST out): kind 0 = inner GMRES
 * residual 2-norms, kind 1 = ||b||_inf then ||b - op(A) x||_inf after each outer
 * step.  Returns the entry count. */
int kad_ir_history(kad_ctx* ctx, int kind, double* out, int cap);

/* ---- spectral transforms (replace Base_spectral::coef / coef_i,
 *      src/Coef/coef.cpp:96-104, src/Coef/coef_i.cpp:96-104, and the 1-D
 *      FFTW-backed dispatch coef_1d / coef_i_1d, src/Coef/coef_1d.cpp:607-636,
 *      src/Coef/coef_i_1d.cpp:527-556) ----------------------------------- */

/* Batched 3-D transform of nbatch domain arrays of one spherical domain.
 * conf: nbatch x (nr*nt*np) doubles, r fastest; coef: nbatch x (nr*nt*(np+2)).
 * base_r: HOST, nt*(np+2) basis ids of the r lines, index j + nt*k;
 * base_t: HOST, np+2 basis ids of the theta lines (per phi coefficient k);
 * base_p: basis of the phi lines (COSSIN = 4).  Ids as include/base_spectral.hpp. */
int kad_coef_3d(kad_ctx* ctx, int nr, int nt, int np, const int* base_r, const int* base_t, int base_p, int nbatch,
                const double* conf, double* coef);
int kad_coef_i_3d(kad_ctx* ctx, int nr, int nt, int np, const int* base_r, const int* base_t, int base_p,
                  int nbatch, const double* coef, double* conf);

/* Batched 1-D transforms of nlines contiguous lines (one basis for all lines):
 * conf lines of nconf values <-> coef lines of nconf (+2 for COSSIN) values. */
int kad_coef_1d(kad_ctx* ctx, int base, int nconf, int64_t nlines, const double* conf, double* coef);
int kad_coef_i_1d(kad_ctx* ctx, int base, int nconf, int64_t nlines, const double* coef, double* conf);

/* Host-side tables (no device work): analysis (kind 0) / synthesis (kind 1)
 * matrix of one 1-D basis, row-major; returns the element count or < 0. */
int kad_host_transform_matrix(int kind, int base, int nconf, double* out, int cap);
/* Matrix of a coefficient-space 1-D operator (op: 0 der_1d, 1 mult_sin_1d,
 * 2 mult_cos_1d, 3 div_sin_1d, 4 mult_x_1d, 5 mult_xm1_1d, 6 div_xm1_1d;
 * src/Ope_1d/<op>_1d.cpp); *out_base receives the output basis. */
int kad_host_op1d_matrix(int op, int base, int n, double* out, int cap, int* out_base);

/* ---- the Newton system (replaces System_of_eqs::do_newton and what it calls:
 *      sec_member / check_equations src/System_of_eqs/solver.cpp:24-85,
 *      do_col_J solver.cpp:120-214, xx_to_vars system_of_eqs.cpp:597-612) -- */
typedef struct kad_system kad_system;

/* XCTS black-hole system (codes/Par_version/Kerr/kerr.cpp:69-99 in the Cartesian
 * basis) on Space_spheric(CHEB_TYPE, 0, (nr,nt,np), bounds[0..nbounds)): solved
 * from shell dmin to the compactified domain dmax = nbounds.  n0: lapse on the
 * excision sphere r = a; ome: horizon rotation.  order 0: the reference's unknown
 * and row order (system_of_eqs.cpp:597-612, kerr.cpp:84-99); order 1: domain-major
 * (unknowns by domain then field, rows by the equation's domain with matching
 * d|d+1 rows in d+1), which makes J block-tridiagonal by domain -- same system,
 * rows and columns permuted.  batch: Jacobian columns per tangent launch chain. */
int kad_xcts_create(kad_ctx* ctx, int nr, int nt, int np, const double* bounds, int nbounds, int dmin, int dmax,
                    double n0, double a, double ome, int order, int batch, kad_system** out);
int kad_xcts_destroy(kad_system* s);
/* stats (may be NULL): value instructions, tangent instructions, value slots, tangent slots */
int kad_xcts_sizes(kad_system* s, int* n_unknowns, int* n_rows, int* stats);
/* 4 ints per (field, domain): field (0 P, 1 N, 2..4 bet^x,y,z), domain, first unknown, count */
int kad_xcts_layout(kad_system* s, int* out, int cap);
/* 5 ints per (equation, component) in row order: kind (0 inside, 1 matching
 * d|d+1, 2 outer bc, 3 inner bc), domain, component, first row, row count
 * (the per-equation row blocks of sec_member, solver.cpp:49-85) */
int kad_xcts_row_layout(kad_system* s, int* out, int cap);
/* fields := flat guess + affecte_tau(X0); X0 HOST (n_unknowns) or NULL */
int kad_xcts_set_guess(kad_system* s, const double* X0);
/* F(u) (device, n_rows): System_of_eqs::sec_member */
int kad_xcts_residual(kad_system* s, double* F);
/* J (device, column-major, ld >= n_rows): every do_col_J column */
int kad_xcts_jacobian(kad_system* s, double* J, int64_t ldj);
/* fields -= affecte_tau(X), X device */
int kad_xcts_update(kad_system* s, const double* X);
/* coefficients of field f in domain d, HOST out (nr*nt*(np+2)) */
int kad_xcts_get_coef(kad_system* s, int f, int d, double* out);
/* overwrite them, HOST in (Scalar::set_domain of the reference's field API) */
int kad_xcts_set_coef(kad_system* s, int f, int d, const double* in);
/* ADM mass, Komar mass and ADM angular momentum about z of the current fields,
 * surface integrals at infinity (codes/Par_version/Kerr/reader.cpp:114-125 with
 * Domain_compact::integ, domain_compact.cpp:89-121), HOST out[3] */
int kad_xcts_observables(kad_system* s, double* out);
/* set a parameter: which 0 = Ome, the horizon rotation of the inner shift condition */
int kad_xcts_set_param(kad_system* s, int which, double value);
/* adjoint sensitivities at a solution: out[q] = dQ_q/dOme for Q = (M_ADM, M_Komar,
 * J_ADM), from J^T lambda_q = dQ_q/du (transposed solve with the LU factors of J,
 * refilled at the current fields) and dQ_q/dOme = -lambda_q . dF/dOme.  HOST out[3];
 * dQdu (HOST 3 x n, may be NULL) receives dQ/du. */
int kad_xcts_sensitivity(kad_system* s, double* out, double* dQdu);
/* One Newton iteration.  solver: 0 fp64 LU, 1 fp32 LU + fp64 IR, 2 fp32 LU +
 * GMRES-IR, 3 bf16-update LU + GMRES-IR, 4 Gauss-Newton (J^T J X = J^T F in fp64,
 * the reference's SOLVER GAUSS_NEWTON, do_newton.cpp:264-311; it holds J^T and J^T J
 * beside J, 3 n^2 doubles), 5 fp64 LU factors of an earlier J as the preconditioner of
 * fp64 flexible GMRES on the current J, refactored (and solved directly) when the
 * previous refinement took more than the reuse limit (kad_system_set_option 1, default
 * 30) inner steps or did not converge; it holds two n^2 fp64 buffers.  6: the
 * domain-decomposed fp64 factors (kad_dd_factor over the blocks of
 * kad_system_dd_blocks: the domains, each block factored with partial pivoting inside
 * it, the blocks together; once a step of the solve has done that, later steps factor
 * with its pivot orders, kad_dd_factor_static, kad_system_set_option 4) precondition
 * fp64 flexible GMRES on the same J; 7: the same,
 * the factors of an earlier J kept as for 5.  6 and 7 need square domain blocks (no
 * integral equations or double unknowns); a singular domain block or capacitance
 * matrix, or a refinement that does not converge, falls back to the fp64 LU of the
 * same J (reported by kad_xcts_last_solve iters[2]).  Returns 1 if ||F||_inf < precision at
 * entry (no step), 0 after a step, KAD_ERR_SINGULAR on an exactly singular J.
 * times[5]: residual, Jacobian fill, factor + solve (incl. refinement), update,
 * factorisation alone (seconds). */
int kad_xcts_do_newton(kad_system* s, double precision, double* error, int solver, double* times);
/* seeded perturbation amp 2^-(i+j+k) U(-1,1) of every unknown (splitmix64, drawn in
 * the reference unknown order whatever the system's order), HOST out */
int kad_xcts_perturbation(kad_system* s, uint64_t seed, double amp, double* out);
/* Statistics of the last do_newton step's solve.  iters[4] (HOST): outer refinement
 * steps; inner GMRES steps; 1 if the low-precision refinement did not reach a
 * relative residual of max(1e-6, its target) and the step was solved with an fp64
 * factorisation of the same J instead (0 otherwise, and for solver 0); 1 if the step
 * factored J (0 when it reused earlier low-precision factors, kad_system_set_option).
 * resid: the refinement's final relative residual ||F - J X||_inf / ||F||_inf. */
int kad_xcts_last_solve(kad_system* s, int* iters, double* resid);
/* host-only tracing of the same system (no device): info[6] = n_unknowns, n_rows,
 * value instr, tangent instr, value slots, tangent programs; returns #layout entries */
int kad_xcts_host_info(int nr, int nt, int np, const double* bounds, int nbounds, int dmin, int dmax, double n0,
                       double a, double ome, int order, int* info, int* layout, int cap);
/* kad_system_dd_blocks of the same system, traced on the host only (no device) */
int kad_xcts_host_dd_blocks(int nr, int nt, int np, const double* bounds, int nbounds, int dmin, int dmax, int order,
                            int* sizes, int cap, int* rows, int* cols);

/* As kad_ir_solve when LUlo holds the factors of diag(Rs) A diag(Cs) (see
 * kad_cast_equilibrated): the preconditioner is v -> Cs .* LU^{-1} (Rs .* v);
 * residuals and the stopping test use the unscaled A. */
int kad_ir_solve_scaled(kad_ctx* ctx, int n, const double* A64, int64_t lda, int prec_lo, const void* LUlo,
                        int64_t ldlu, const int* perm, const double* Rs, const double* Cs, const double* b, double* x,
                        double tol, int max_outer, int mode, int trans, int* iters, double* resid);

/* ---- systems given as equation strings (the reference's System_of_eqs front end:
 *      add_var / add_cst / add_def / add_eq_*, src/System_of_eqs/system_of_eqs.cpp,
 *      add_eq.cpp, the parser give_ope.cpp:559-1135; Space_spheric::add_eq /
 *      add_inner_bc / add_outer_bc, src/Space/Spheric/spheric_add_eq.cpp:26-52) ---
 *
 * A declaration (kad_sysdef, host only) collects unknowns, constants, definitions and
 * equations; kad_system_build traces it into device programs.  Expressions follow the
 * reference's grammar: split at the last top-level '+', else '-', else a leading '-',
 * else the last '*', else the last '/'; then Lap / lap / dn / dr / multr / divr /
 * multsint / divsint (x), D_i x / D^i x / partial_i x (flat metric, Cartesian basis),
 * x^n (integer n > 0), numbers, and names with indices (A^ij, bet_i).  A repeated
 * index is summed.  Unknowns and constant fields are scalars (valence 0) or Cartesian
 * vectors (valence 1).  Operators without a device form (sqrt, exp, integ, ...) are
 * rejected with KAD_ERR_ARG and a message on stderr.  bound: 0 INNER_BC, 1 OUTER_BC. */
typedef struct kad_sysdef kad_sysdef;

/* Space_spheric(CHEB_TYPE, center 0, (nr, nt, np), bounds[0..nbounds)) with the
 * System_of_eqs(space, dmin, dmax) constructor's domain range (dmax = nbounds, the
 * compactified domain). */
int kad_sysdef_create(int nr, int nt, int np, const double* bounds, int nbounds, int dmin, int dmax,
                      kad_sysdef** out);
int kad_sysdef_destroy(kad_sysdef* def);
int kad_sysdef_add_var(kad_sysdef* def, const char* name, int valence);   /* add_var */
int kad_sysdef_add_cst(kad_sysdef* def, const char* name, double value);  /* add_cst(name, double) */
/* add_var(name, double&) (system_of_eqs.cpp:382-400): a global unknown; double unknowns
 * are unknowns 0 .. ndouble-1, before every field, as in the reference; returns its
 * index.  It is used in equations like a constant (its Jacobian column is its dF/dp). */
int kad_sysdef_add_var_double(kad_sysdef* def, const char* name, double value);
/* Integral equation in domain dom at boundary bound, e.g. "integ(intS) - chi * CM * CM = 0":
 * integ(x) is Domain::integ of x there (domain_shell.cpp:862-897); its terms combine
 * linearly with numbers (Eq_int, eq_int.cpp:38-51; Space_adapted_bh::add_eq_int_bh,
 * adapted_bh_add_eq.cpp:88-119).  One row, before every other equation's rows. */
int kad_sysdef_add_eq_int(kad_sysdef* def, int dom, int bound, const char* eq);
/* Global quantity: an integral expression (as in kad_sysdef_add_eq_int, without '=')
 * evaluated in domain dom at boundary bound, e.g. M_ADM = "-integ(dn(P)) / 2. / PI" at
 * the outer boundary (codes/Par_version/Kerr/reader.cpp:114-125); returns its index. */
int kad_sysdef_add_global(kad_sysdef* def, const char* name, int dom, int bound, const char* expr);
/* a double constant that carries a tangent: a parameter p of dF/dp and of the adjoint
 * sensitivities; returns its parameter index (>= 0; parameters and double unknowns are
 * numbered together in declaration order) */
int kad_sysdef_add_param(kad_sysdef* def, const char* name, double value);
/* add_cst(name, Tensor): conf values, component-major, then domains dmin..dmax, each
 * nr*nt*np values in the reference's Array order (r fastest); count = the total */
int kad_sysdef_add_cst_field(kad_sysdef* def, const char* name, int valence, const double* conf, int64_t count);
/* Metric_flat::set_system(syst, name): name^ij = name_ij = delta in the Cartesian basis */
int kad_sysdef_set_flat_metric(kad_sysdef* def, const char* name);
int kad_sysdef_add_def(kad_sysdef* def, const char* def_text);                       /* add_def */
int kad_sysdef_add_eq_inside(kad_sysdef* def, int dom, const char* eq, int order);  /* add_eq_inside / add_eq_order */
int kad_sysdef_add_eq_bc(kad_sysdef* def, int dom, int bound, const char* eq);      /* add_eq_bc */
int kad_sysdef_add_eq_matching(kad_sysdef* def, int dom, int bound, const char* expr); /* add_eq_matching */
int kad_sysdef_space_add_eq(kad_sysdef* def, const char* eq, const char* m1, const char* m2);
int kad_sysdef_space_add_bc(kad_sysdef* def, int bound, const char* eq);  /* add_inner_bc / add_outer_bc */
/* Domain d is a shell of an adapted space whose boundary spheres are round (the
 * homothetic shells of Space_adapted_bh, space_adapted_bh.cpp:59-75): its surface
 * integrals use Domain_shell_inner_adapted::integ's configuration-space weight
 * (domain_shell_inner_adapted.cpp:1134-1208) instead of Domain_shell::integ's
 * coefficient-space mult_r (domain_shell.cpp:862-897).  Call before building. */
int kad_sysdef_set_adapted(kad_sysdef* def, int d);
/* collocation r, theta, phi of domain d's conf grid (HOST out, nr*nt*np each; r = +inf
 * at infinity); any pointer may be NULL */
int kad_sysdef_coloc(kad_sysdef* def, int d, double* r, double* theta, double* phi);
/* Host-only trace of the declaration (no device work), for checks without a GPU:
 * info[6] = n_unknowns, n_rows, value instructions, tangent instructions, value slots,
 * tangent programs; row_layout (may be NULL): 5 ints per (equation, component) as
 * kad_xcts_row_layout.  Returns the entry count.  The declaration can then only be
 * destroyed. */
int kad_sysdef_host_trace(kad_sysdef* def, int order, int* info, int* row_layout, int cap);
/* Trace and upload (order 0: reference unknown / row order, 1: domain-major; batch:
 * Jacobian columns per tangent launch chain).  Consumes def (freed, also on error). */
int kad_system_build(kad_ctx* ctx, kad_sysdef* def, int order, int batch, kad_system** out);

/* ---- entry points of every system (XCTS or built from strings) ---------- */
int kad_system_destroy(kad_system* s);
int kad_system_sizes(kad_system* s, int* n_unknowns, int* n_rows, int* nfields);
int kad_system_layout(kad_system* s, int* out, int cap);      /* as kad_xcts_layout */
int kad_system_row_layout(kad_system* s, int* out, int cap);  /* as kad_xcts_row_layout */
/* field component f (declaration order; vectors x, y, z) of domain d from / to its
 * values at the collocation points (HOST, nr*nt*np) */
int kad_system_set_field(kad_system* s, int f, int d, const double* conf);
int kad_system_get_field(kad_system* s, int f, int d, double* conf);
int kad_system_get_coef(kad_system* s, int f, int d, double* out);
int kad_system_set_coef(kad_system* s, int f, int d, const double* in);
/* fields += affecte_tau(X0), X0 HOST (n_unknowns) */
int kad_system_add_unknowns(kad_system* s, const double* X0);
int kad_system_update(kad_system* s, const double* X);  /* fields -= affecte_tau(X), X device */
int kad_system_perturbation(kad_system* s, uint64_t seed, double amp, double* out);
int kad_system_set_param(kad_system* s, int which, double value);
/* Global quantities at the current fields (HOST out, one per kad_sysdef_add_global). */
int kad_system_globals(kad_system* s, double* out);
/* Adjoint sensitivities dQ_q/dp of the global quantities with respect to parameter
 * `which` at the current (converged) fields: J^T lambda_q = dQ_q/du with the LU factors
 * of J refilled there, dQ_q/dp = dQ_q/dp|_u - lambda_q . dF/dp.  HOST out (one per
 * quantity); dQdu (HOST, nq x n_unknowns, may be NULL) receives dQ/du. */
int kad_system_sensitivity(kad_system* s, int which, double* out, double* dQdu);
/* number of double unknowns; the current value of double unknown k (HOST out) */
int kad_system_ndouble(kad_system* s);
int kad_system_get_double(kad_system* s, int k, double* value);
/* Options of the iterative solvers (1-3, 5-7) of do_newton: which 0 = forcing factor
 * c >= 0, the refinement's target becomes max(1e-13, c min(1, ||F||_inf)) relative
 * (0, the default: 1e-13, i.e. its attainable accuracy); which 1 = reuse the
 * low-precision factors of an earlier step while the previous refinement took at most
 * `value` inner steps (0, the default: factor every step).  Reused factors that stop
 * converging are replaced by a factorisation of the current J within the step.
 * which 2 = forget the factorisations kept for reuse (solvers 5 and 7 too): the next
 * step factors its own J.  which 3 = streams of the Jacobian fill (1-4, default 3): the
 * tangent programs are dealt to that many streams, each with its own tangent arena
 * (nphys x batch x slot doubles more per extra stream); J is the same bit for bit.
 * which 4 = static pivoting in solvers 6-7 (1, the default; 0: partial pivoting at
 * every factorisation): after a factorisation with partial pivoting, later steps of
 * the solve factor their J with its pivot orders (kad_dd_factor_static), falling
 * back to partial pivoting when a multiplier exceeds 1e8 or the refinement needs more
 * than 6 inner steps; after a step whose multipliers exceed 1e3 the next step factors
 * with partial pivoting again (the orders drift with the Jacobian); those
 * factorisations keep the capacitance matrix's order (kad_dd_factor_keep) while its
 * multipliers stay within 1e8, and factor it with partial pivoting too when the
 * refinement then needs more than 6 inner steps.  which 5 = an
 * element-wise instruction whose operand is made by the spectral pass just before it,
 * and used nowhere else, is applied in that pass's store phase (1, the default; 0:
 * launched on its own); J is the same bit for bit. */
int kad_system_set_option(kad_system* s, int which, double value);
/* as kad_xcts_do_newton / kad_xcts_last_solve */
int kad_system_do_newton(kad_system* s, double precision, double* error, int solver, double* times);
/* Newton iteration with backtracking line search:
 * System_of_eqs::do_newton_with_linesearch(precision, error, ntrymax, stepmax)
 * (src_par/do_newton.cpp:454-517; reference defaults ntrymax 10, stepmax 1.0).  The
 * step p solves J p = F with `solver` (as kad_system_do_newton) and is scaled by
 * stepmax / max(p) when its signed maximum exceeds stepmax; u = u_old - lambda p with
 * lambda from the quadratic, then cubic, model until sum F^2 <= f_old + 1e-4 lambda
 * slope (slope = -2 f_old), sum F^2 <= 5e-13, lambda <= 1e-10, or ntrymax tries.
 * Returns 1 if ||F||_inf < precision at entry (no step), 0 after a step,
 * KAD_ERR_NOCONV where the reference aborts with "roundoff error in
 * do_newton_with_linesearch".  lambda, ntry (HOST, may be NULL): the lambda applied,
 * the residual evaluations of the search. */
int kad_system_do_newton_linesearch(kad_system* s, double precision, double* error, int ntrymax, double stepmax,
                                    int solver, double* lambda, int* ntry);
int kad_system_last_solve(kad_system* s, int* iters, double* resid);
/* The domain blocks of solvers 6-7 (no reference counterpart: the reference factors J
 * whole): per domain, its unknowns and the rows of the equations living there, an
 * interface's value-matching rows in the inner domain and its derivative-matching
 * rows in the outer one.  sizes (HOST, cap), rows and cols (HOST, n each, may be
 * NULL) as kad_dd_create takes them.  Returns the number of blocks, or KAD_ERR_ARG
 * when they are not square. */
int kad_system_dd_blocks(kad_system* s, int* sizes, int cap, int* rows, int* cols);
/* kad_dd_info of the system's last domain-decomposed factorisation */
int kad_system_dd_info(kad_system* s, int* sizes, double* times);
/* kad_dd_pivoting of the system's last domain-decomposed factorisation */
int kad_system_dd_pivoting(kad_system* s, int* static_orders, double* growth);

/* F(u) (device, n_rows): System_of_eqs::sec_member (solver.cpp:49-85) */
int kad_residual(kad_system* s, double* F);
/* per (equation, component) max |F| (HOST out, row-layout order):
 * System_of_eqs::check_equations (solver.cpp:24-47); returns the entry count */
int kad_check_equations(kad_system* s, double* errors, int cap);
/* Columns [col_begin, col_begin + col_count) of J into J_tile (device, column-major,
 * ld >= n_rows), dtype 0 fp64 / 1 fp32: the do_col_J loop of do_newton.cpp:132-174
 * handing a column range to the GPU (System_of_eqs::do_col_J, solver.cpp:120-214). */
int kad_jvp_fill(kad_system* s, int col_begin, int col_count, void* J_tile, int64_t ld, int dtype);
/* out = J^T w (device; w n_rows, out n_unknowns), matrix-free: the tangent programs with
 * their row exports transposed; J is never formed (no reference counterpart: the
 * reference forms J^T only inside Matrice::solve with 'T', matrice.cpp:221-250). */
int kad_vjp(kad_system* s, const double* w, double* out);
/* dF/dp of parameter `which` at the current fields (device out, n_rows) */
int kad_param_tangent(kad_system* s, int which, double* out);

#ifdef __cplusplus
}
#endif
#endif
