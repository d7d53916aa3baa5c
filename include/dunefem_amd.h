/*
 * dunefem_amd.h -- C ABI of the MI355X-native matrix-free Galerkin apply and
 * Krylov solve library.
 *
 * Every entry point returns an int status: 0 on success, < 0 on error (the
 * message is available from fem_last_error()).  Nothing throws across this
 * boundary.  Non-convergence of a solver is not an error: it is reported as
 * a negative iteration count, as the reference does
 * (dune/fem/solver/linear/cg.hh:116, dune/fem/solver/linear/gmres.hh:298).
 *
 * Handles are opaque and owned by the caller (create/destroy).  All device
 * work of one context is ordered on that context's HIP stream; calls that
 * return scalars to the host synchronise that stream.  A context is used by
 * one host thread at a time.
 */
#ifndef DUNEFEM_AMD_H
#define DUNEFEM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------- */
#define FEM_OK 0
#define FEM_ERR_ARG (-1)      /* invalid argument / shape mismatch           */
#define FEM_ERR_HIP (-2)      /* HIP runtime error                           */
#define FEM_ERR_STATE (-3)    /* wrong call order (e.g. solve before bind)   */
#define FEM_ERR_NOMEM (-4)    /* allocation failed                           */
#define FEM_ERR_UNSUPPORTED (-5)

/* ---- precisions of DoF vectors and operator arithmetic -------------------- */
#define FEM_F64 0            /* IEEE double                                  */
#define FEM_F32 1            /* IEEE single                                  */
#define FEM_BF16MFMA 2       /* operator only: FP32 data, BF16 MFMA contractions, FP32 accumulation */

typedef struct fem_ctx fem_ctx;
typedef struct fem_vec fem_vec;
typedef struct fem_op fem_op;
typedef struct fem_precond fem_precond;
typedef struct fem_solver fem_solver;

const char *fem_last_error(void);
int fem_version(void);

/* ---- context ---------------------------------------------------------------
 * One device + stream (replaces the MPIManager thread pool as the place where
 * parallel work runs: dune/fem/misc/mpimanager.hh:155-198). */
int fem_ctx_create(int device, fem_ctx **out);
/* Drops the caller's reference: every handle made on the context (vectors,
 * operators, preconditioners, solvers, matrices, networks, batches) holds
 * its own, and the stream is released with the last of them. */
int fem_ctx_destroy(fem_ctx *ctx);
int fem_ctx_synchronize(fem_ctx *ctx);
/* elapsed milliseconds between two event marks recorded on the context stream */
int fem_ctx_mark(fem_ctx *ctx, int which /* 0 start, 1 stop */);
int fem_ctx_elapsed_ms(fem_ctx *ctx, float *ms);

/* ---- DoF vectors -----------------------------------------------------------
 * Device-resident contiguous DoF array; replaces the DoF storage of
 * AdaptiveDiscreteFunction / MutableBlockVector<DynamicArray<double>>
 * (dune/fem/function/adaptivefunction/adaptivefunction.hh:45-61,
 *  dune/fem/function/blockvectors/defaultblockvectors.hh:444-505). */
int fem_vec_create(fem_ctx *ctx, int64_t n, int precision, fem_vec **out);
/* A view of n DoFs the caller already holds in device memory (`device_ptr`
 * from hipMalloc, 256-byte aligned), as the AdaptiveDiscreteFunction
 * constructor over existing DoF storage
 * (dune/fem/function/adaptivefunction/adaptivefunction.hh:81-87): no copy;
 * fem_vec_destroy releases the handle, not the memory, which must outlive it. */
int fem_vec_wrap(fem_ctx *ctx, int64_t n, int precision, void *device_ptr, fem_vec **out);
int fem_vec_destroy(fem_vec *v);
int64_t fem_vec_size(const fem_vec *v);
int fem_vec_precision(const fem_vec *v);
void *fem_vec_device_ptr(fem_vec *v);
/* host <-> device copies with conversion from/to double */
int fem_vec_upload(fem_vec *v, const double *host);
int fem_vec_download(const fem_vec *v, double *host);
int fem_vec_upload_f32(fem_vec *v, const float *host);
/* DoF vectors in the reference's binary stream format:
 * DiscreteFunctionDefault::write / read (function/common/discretefunction_inline.hh:213-305)
 * through a BinaryFileOutStream / StandardOutStream (io/streams/standardstreams.hh:60-170):
 * little-endian unsigned version id (written: DUNE_VERSION_ID(2,12,0) =
 * 2<<24 | 12<<16; files older than 0.9.1 are refused), int space id (checked
 * for files from 1.5.0 on), the name as unsigned length + bytes, int32 DoF
 * count, then the DoFs as doubles in DoF order.  The space ids are the
 * DFSpaceIdentifier values (space/common/discretefunctionspace.hh:95-108).
 * Legendre DG DoFs are in the reference's order (element-major, the element
 * index x fastest); the Q2 DoFs are in this library's node-lattice order, a
 * permutation of the reference's IndexSetDofMapper order.
 * fem_dofs_encode / fem_dofs_decode act on host arrays (no device work);
 * *len receives the encoded size (also when cap is too small: FEM_ERR_ARG). */
#define FEM_SPACE_LAGRANGE 6
#define FEM_SPACE_LEGENDRE_DG 8
int fem_dofs_encode(const double *dofs, int64_t n, const char *name, int space_id, unsigned char *buf, int64_t cap,
                    int64_t *len);
int fem_dofs_decode(const unsigned char *buf, int64_t len, int space_id, double *dofs, int64_t n, char *name,
                    int64_t name_cap);
int fem_vec_write(const fem_vec *v, const char *path, const char *name, int space_id);
int fem_vec_read(fem_vec *v, const char *path, int space_id, char *name, int64_t name_cap);
/* DiscreteFunctionInterface BLAS-1 (dune/fem/function/common/discretefunction.hh:731-835,
 * defaultblockvectors.hh:110-135): clear, assign, axpy, scaling, scalarProductDofs,
 * normSquaredDofs (scalarproducts.hh:101-127).  Mixed precisions convert. */
int fem_vec_fill(fem_vec *v, double value);               /* clear() = fill 0 */
int fem_vec_copy(fem_vec *dst, const fem_vec *src);       /* assign()         */
int fem_vec_axpy(fem_vec *y, double a, const fem_vec *x); /* y += a x         */
int fem_vec_scal(fem_vec *x, double a);                   /* x *= a           */
int fem_vec_dot(const fem_vec *x, const fem_vec *y, double *out);
int fem_vec_norm2sq(const fem_vec *x, double *out);
/* fem_vec_dot with nothing allocated: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The reduction
 * scratch is a fixed 4,352-byte device array per GPU, independent of the
 * vector length (every length is accepted), held until the process ends.
 * All calls on one GPU share it: calls that overlap on one GPU race. */
int fem_vec_dot_noalloc(const fem_vec *x, const fem_vec *y, double *out);

/* ---- operators ---------------------------------------------------------------
 * Matrix-free Galerkin operator on a structured Cartesian mesh of the unit
 * square/cube.  Replaces GalerkinOperator<Integrands,DF,DF>::operator()(u,w)
 * (dune/fem/schemes/galerkin.hh:1430-1433,1460-1496; the element/intersection
 * loop :811-917) for the integrands of the SIPG/upwind form
 * (dune/fempy/test/parallelTest.py:113-129, pydemo/advectiondiffusion.py:30-56)
 * and of the continuous Lagrange form with strong Dirichlet constraints
 * (DirichletWrapperOperator, dune/fem/schemes/dirichletwrapper.hh:101-105). */
#define FEM_OP_DG_SIPG_3D 1   /* discontinuous Legendre Q_p on hexes, weak Dirichlet  */
#define FEM_OP_CG_Q2_2D 2     /* continuous Lagrange Q2 on quads, strong u = 0        */

#define FEM_KAPPA_CONST 0     /* constant diffusion coefficient               */
#define FEM_KAPPA_QP 1        /* pointwise: volume + face quadrature values   */
#define FEM_KAPPA_CELL 2      /* one value per element                        */

typedef struct {
    int kind;            /* FEM_OP_*                                          */
    int n[3];            /* cells per direction (n[2] ignored in 2D)          */
    int order;           /* polynomial order p (CG: must be 2)                */
    double beta;         /* SIPG penalty factor, e.g. 10 p^2                  */
    double b[3];         /* constant advection field                          */
    double c;            /* reaction coefficient (u v term)                   */
    int kappa_mode;      /* FEM_KAPPA_*                                       */
    double kappa_const;
    int precision;       /* FEM_F64, FEM_F32 or FEM_BF16MFMA                   */
} fem_op_desc;

int fem_op_create(fem_ctx *ctx, const fem_op_desc *desc, fem_op **out);
int fem_op_destroy(fem_op *op);
int64_t fem_op_ndofs(const fem_op *op);
/* coefficient data (host arrays, copied).  Layouts:
 *  volume qp  [E][m^d], qp index dim 0 fastest
 *  face qp    per normal direction d: faces on the element grid extended by
 *             one in d (x fastest), [faces_d][m^(d-1)], tangential dims in
 *             increasing order, lower fastest
 *  cell       [E] */
int fem_op_set_kappa_qp(fem_op *op, const double *kvol, const double *kface_x,
                        const double *kface_y, const double *kface_z);
int fem_op_set_kappa_cell(fem_op *op, const double *kcell);
/* Operator::operator()(u, w): w = A u, w overwritten
 * (dune/fem/operator/common/operator.hh:55; w.clear() at galerkin.hh:1463).
 * u and w must have the operator's precision class (F64 for FEM_F64, F32 otherwise). */
int fem_op_apply(fem_op *op, const fem_vec *u, fem_vec *w);
/* Affine part: load vector F_i = int f phi_i from f at the volume quadrature
 * points [E][m^d] (host array); Dirichlet entries are zero.  The reference's
 * GalerkinOperator applies R(u) = A u - F (galerkin.hh:333-360, source term). */
int fem_op_assemble_rhs(fem_op *op, const double *fq, fem_vec *F);
/* operator diagonal, matrix-free (takes the place of assembling the Jacobian
 * and SparseRowMatrix::extractDiagonal, dune/fem/operator/matrix/spmatrix.hh:846,
 * used by FemPreconditioning, dune/fem/solver/fempreconditioning.hh:85-107) */
int fem_op_diagonal(fem_op *op, fem_vec *d);
/* Boundary data u = g (the reference's boundary integrands with the exact
 * data, pydemo/advectiondiffusion.py:43-49 and dune/fempy/test/parallelTest.py:127-129
 * for DG; DirichletConstraints with a boundary function,
 * dune/fem/schemes/dirichletconstraints.hh:237-262, for CG).  Host arrays, copied.
 *  DG (weak): g at the face quadrature points in the face layout of
 *     fem_op_set_kappa_qp; only boundary faces are read.  fem_op_assemble_rhs
 *     then adds the g-dependent Nitsche and inflow terms
 *     int (kappa beta/h + max(-b.n, 0)) g v - kappa g dv/dn.
 *  CG Q2 (strong): nodal values [N] in g_x (g_y, g_z unused); only the
 *     Dirichlet nodes are read.  fem_op_assemble_rhs then puts F_D = g_D, so
 *     that R(u) = A u - F has the Dirichlet rows u_D - g_D of the reference's
 *     DirichletWrapperOperator (dirichletwrapper.hh:101-105). */
int fem_op_set_boundary_data(fem_op *op, const double *g_x, const double *g_y, const double *g_z);
/* CG Q2: x_D = g_D + add_D (FemScheme::_solve's setConstraints and
 * addConstraints, dune/fem/schemes/femscheme.hh:247-253; g = 0 when no data
 * was set, add may be NULL); DG: no effect (weak boundary conditions). */
int fem_op_constrain(fem_op *op, fem_vec *x, const fem_vec *add);
/* DG operators: the element (diagonal) blocks of the linear operator, the
 * local matrices that DifferentiableGalerkinOperator::jacobian assembles on
 * the diagonal (dune/fem/schemes/galerkin.hh:1548-1551, 1195-1314) and that
 * block-Jacobi preconditioning inverts.  `blocks` is an FP64 vector of
 * E * nb * nb entries: block of element e at e * nb^2, row-major, rows and
 * columns in the local DoF order.  Every coefficient mode and advection are
 * supported (computed by probing the apply). */
int fem_op_block_diagonal(fem_op *op, fem_vec *blocks);

/* ---- pointwise and nonlinear reaction (DG and Q2 Lagrange operators) -------------
 * Integrands whose source term is pointwise at the volume quadrature points,
 * as the reference's GalerkinOperator evaluates every integrand
 * (dune/fem/schemes/galerkin.hh:333-360, interior term of the integrands,
 * dune/fem/schemes/integrands.hh:412-664): the volume term of the operator
 * becomes int (c + cvol(x) + coef u^pw) u v, so that -div(kappa grad u) +
 * coef u^(pw+1) = f is the nonlinear problem of a Newton solve.  DG operators
 * (any kappa mode) and the 2D Q2 Lagrange operator (any kappa;
 * its Dirichlet rows keep S[u]_D = u_D, the Jacobian's rows are the identity,
 * as DirichletWrapperOperator, dune/fem/schemes/dirichletwrapper.hh:101-112),
 * FP64 or FP32 (FEM_ERR_UNSUPPORTED otherwise).
 *
 * fem_op_set_reaction_qp: cvol at the volume quadrature points [E][m^3]
 *   (DG) or [E][9] (Q2: qx + 3 qy), x fastest, host array, copied; NULL
 *   clears.
 * fem_op_set_reaction_law: coef u^pw, pw in 1..4 (coef = 0 clears).  With a
 *   law set the operator is nonlinear (the reference's Operator::nonlinear(),
 *   dune/fem/operator/common/operator.hh:64): fem_op_apply evaluates S[u],
 *   while the diagonal, the element blocks, the CSR assembly and binding to
 *   a linear solver are refused (FEM_ERR_STATE): they belong to the Jacobian.
 * fem_op_nonlinear: 1 if a law is set, else 0 (negative: error).
 * fem_op_eval_qp: u at the volume quadrature points [E][m^3] (DG) or
 *   [E][9] (Q2), x fastest, vectors of the operator's precision.
 * fem_op_jacobian_create: a linear operator of the same description and
 *   coefficient data (device copies), the Jacobian operator object
 *   (NewtonInverseOperator's jOp, newtoninverseoperator.hh:704).
 * fem_op_jacobian: DifferentiableOperator::jacobian(u, jOp)
 *   (dune/fem/operator/common/differentiableoperator.hh:54): jop = S'(u),
 *   the linear operator with cvol_J = cvol + (pw + 1) coef u^pw; a linear
 *   operator's Jacobian is the operator itself.  Matrix-free: jop is applied
 *   by the same kernel, and fem_op_assemble_csr(jop) gives the assembled
 *   matrix when one is needed. */
int fem_op_set_reaction_qp(fem_op *op, const double *cvol);
int fem_op_set_reaction_law(fem_op *op, double coef, int pw);

/* A compiled pointwise reaction r(u, x): the volume term int r(u, x) v, with
 * r any expression of u and the coordinates (the counterpart of the
 * reference's generated integrands, python/dune/models/integrands/ufl.py:
 * 197-348 and dune/fem/schemes/integrands.hh:412-664, for the source part of
 * the interior integrand).  The Python module dunefem_amd.forms writes r and
 * dr/du (symbolic derivative) as HIP code and compiles it into a shared
 * library exporting
 *     int fem_reaction_eval(const fem_qp_geometry *g, const double *uq,
 *                           double *r, double *dr, void *stream);
 * which writes r(u(x_q), x_q) and dr/du at every volume quadrature point q
 * (either output may be NULL; launched on `stream`; 0 on success).
 * fem_op_set_reaction_plugin loads it (NULL unloads): the operator is then
 * nonlinear, S[u] = A u + int r(u, x) v (plus any law above), and
 * fem_op_jacobian adds dr/du(u) to the Jacobian's cvol.  FP64 DG operators
 * and FP64 Q2 Lagrange operators (2D: dim
 * = 2, n[2] = 1, h[2] = 0, m^2 points per element, z = 0). */
typedef struct {
    int n[3];          /* cells per direction                               */
    int m;             /* points per direction (p + 1; 3 for Q2)            */
    double h[3];       /* cell size per direction                           */
    double gx[8];      /* Gauss points on [0, 1]                            */
    int64_t nq;        /* E m^dim quadrature points, [E][m^dim], x fastest  */
    int dim;           /* 3 (DG hexahedra) or 2 (Q2 quadrilaterals)         */
} fem_qp_geometry;
typedef int (*fem_reaction_eval_fn)(const fem_qp_geometry *g, const double *uq, double *r, double *dr,
                                    void *stream);
int fem_op_set_reaction_plugin(fem_op *op, const char *path);
int fem_op_nonlinear(const fem_op *op);
int fem_op_eval_qp(fem_op *op, const fem_vec *u, fem_vec *uq);
int fem_op_jacobian_create(const fem_op *op, fem_op **jop);
int fem_op_jacobian(const fem_op *op, const fem_vec *u, fem_op *jop);

/* ---- assembled Jacobian -------------------------------------------------------
 * DifferentiableGalerkinOperator::jacobian / assemble into a SparseRowMatrix
 * (dune/fem/schemes/galerkin.hh:1195-1314, 1548-1643;
 * dune/fem/operator/matrix/spmatrix.hh:39-860), on the device in CSR form:
 * int64 row offsets, int32 column indices increasing within a row, FP64
 * values.  Rows of a DG element hold the blocks of the element and its face
 * neighbours in element order; Q2 rows the nodes of the elements around the
 * node, Dirichlet rows unit rows.  Assembled from the operator's own apply
 * (the operators are linear), so A x equals fem_op_apply to rounding. */
typedef struct fem_csr fem_csr;
int fem_op_assemble_csr(fem_op *op, fem_csr **out);
int fem_csr_destroy(fem_csr *A);
int64_t fem_csr_rows(const fem_csr *A);
int64_t fem_csr_nnz(const fem_csr *A);
/* y = A x (SparseRowMatrix::apply, spmatrix.hh:148-178); FP64 or FP32 vectors */
int fem_csr_apply(const fem_csr *A, const fem_vec *x, fem_vec *y);
/* d = diag(A) (SparseRowMatrix::extractDiagonal, spmatrix.hh:846), FP64 */
int fem_csr_diagonal(const fem_csr *A, fem_vec *d);
/* host copies: rowptr [rows + 1], col [nnz], val [nnz] (any may be NULL) */
int fem_csr_download(const fem_csr *A, int64_t *rowptr, int *col, double *val);

/* ---- preconditioners -------------------------------------------------------- */
#define FEM_PRE_NONE 0
#define FEM_PRE_JACOBI 1        /* d^-1 with |d| < 16 eps -> 1 (fempreconditioning.hh:99-106) */
#define FEM_PRE_BLOCK_JACOBI 2  /* DG element blocks A_KK^-1                        */
#define FEM_PRE_LEARNED 3       /* CNN coarse correction + block Jacobi (FGMRES only) */
#define FEM_PRE_MULTIGRID 4     /* DG, b = 0: V-cycle over DG p-levels, the conforming trilinear space and vertex h-levels */

int fem_precond_create(fem_op *op, int kind, fem_precond **out);
int fem_precond_destroy(fem_precond *p);
/* y = M^-1 x */
int fem_precond_apply(fem_precond *p, const fem_vec *x, fem_vec *y);

/* ---- inverse operators -------------------------------------------------------
 * Replaces InverseOperatorInterface / KrylovInverseOperator
 * (dune/fem/solver/inverseoperatorinterface.hh:33-239,
 *  dune/fem/solver/krylovinverseoperators.hh:46-295) with the CG and GMRES
 * recurrences of dune/fem/solver/linear/{cg,gmres}.hh and a flexible GMRES. */
#define FEM_SOLVER_CG 0
#define FEM_SOLVER_GMRES 1
#define FEM_SOLVER_FGMRES 2
#define FEM_SOLVER_BICGSTAB 3   /* LinearSolver::bicgstab (dune/fem/solver/linear/bicgstab.hh:56-222) */

#define FEM_ERRMEASURE_ABSOLUTE 0
#define FEM_ERRMEASURE_RELATIVE 1
#define FEM_ERRMEASURE_REDUCTION 2

typedef struct {
    int method;          /* fem.solver.method                                 */
    double tolerance;    /* fem.solver.tolerance (default 1e-8)               */
    int errormeasure;    /* fem.solver.errormeasure (default absolute)        */
    int maxiterations;   /* fem.solver.maxiterations                          */
    int restart;         /* fem.solver.gmres.restart (default 20)             */
    int verbose;         /* fem.solver.verbose                                */
    /* mixed precision: 0 = plain (solve in the operator's precision);
     * 1 = FP64 iterative refinement around an inner solve in `inner_op`'s
     * precision; the true residual, inner products of the outer loop and the
     * update are FP64 (errormeasure residualreduction: relative to the
     * first FP64 residual) */
    int refinement;
    double inner_tolerance;   /* relative reduction of each inner solve        */
    int inner_maxiterations;
    /* CG stopping rule: 0 = the reference rule <r, M^-1 r> <= eps^2 ref
     * (dune/fem/solver/linear/cg.hh:62-70); 1 = on the unpreconditioned
     * residual ||r||_2^2 <= eps^2 ref, and at convergence on the true
     * residual b - A x (CG restarts from x, at most twice, while it is above
     * the tolerance); GMRES always stops on ||r||_2 */
    int residual_check;
    /* fem.solver.preconditioning.method (dune/fem/solver/parameter.hh:206-217):
     * the preconditioner the solver builds for itself at bind when no
     * preconditioner object is passed (and for the inner operator of a
     * refinement solve when no inner one is passed): FEM_PRE_NONE,
     * FEM_PRE_JACOBI, FEM_PRE_BLOCK_JACOBI or FEM_PRE_MULTIGRID */
    int preconditioning;
    /* fem.solver.preconditioning.relaxation (parameter.hh:219-222, default
     * 1.1) and .iterations (:224-233, default 1); as in the reference's
     * FemPreconditioning<jacobi> (fempreconditioning.hh:178-180) the Jacobi
     * preconditioner applies d^-1 once and does not use them */
    double relaxation;
    int precond_iterations;
    /* CG: capture one iteration (operator, preconditioner, updates) in a HIP
     * graph and replay it; for preconditioners of many small launches (the
     * multigrid V-cycle).  No reference counterpart; default 0 */
    int graph;
} fem_solver_params;

typedef struct {
    int converged;
    int iterations;          /* signed: < 0 not converged (reference semantics) */
    int outer_iterations;    /* refinement steps (0 for a plain solve)         */
    double true_rel_residual; /* ||b - A x||_2 / ||b||_2 in FP64                */
    double time_s;
    int fallback;            /* 1: the learned preconditioner stalled and the
                                solve continued with block Jacobi            */
} fem_solver_info;

void fem_solver_default_params(fem_solver_params *p);
/* key/value setter mirroring the fem.solver.* keys of SolverParameter
 * (dune/fem/solver/parameter.hh:71-250).  Accepted (prefix "fem.solver.",
 * optionally followed by "linear.", "newton.linear." or "nonlinear.linear.",
 * or one of those prefixes alone, as scheme parameters are written):
 *   method                 cg | gmres | fgmres | bicgstab   (parameter.hh:170-195)
 *   tolerance, absolutetol, reductiontol   (:118-146; the last two deprecated)
 *   errormeasure           absolute | relative | residualreduction (:108-116)
 *   maxiterations, maxlineariterations (deprecated)           (:148-162)
 *   gmres.restart                                            (:197-201)
 *   verbose                true | false                      (:90-97)
 *   preconditioning.method none | jacobi | blockjacobi | multigrid (:206-217;
 *                          the reference's sor, ssor, gauss-seidel, ilu, amg-*,
 *                          ildl, oas and icc need an assembled matrix and are
 *                          rejected with FEM_ERR_UNSUPPORTED)
 *   preconditioning.relaxation, preconditioning.iterations,
 *   preconditioning.iteration (deprecated)                   (:219-233)
 *   preconditioning.level  0 only                            (:235-238)
 *   threading              true | false (accepted; one stream per context)
 *   knolltrick             false only                        (:245-248)
 *   residualcheck          true | false (this library: CG stops on ||r||, and
 *                          restarts from x while ||b - A x|| is above the tolerance)
 * Newton keys (nonlinear.*, newton.* without linear.) are not linear-solver
 * keys: FEM_ERR_ARG. */
int fem_solver_params_set(fem_solver_params *p, const char *key, const char *value);
int fem_solver_create(fem_ctx *ctx, const fem_solver_params *params, fem_solver **out);
int fem_solver_destroy(fem_solver *s);
/* bind(op[, precond]) / unbind() (inverseoperatorinterface.hh:127-151).
 * For refinement, inner_op is the low-precision operator (may be NULL).
 * Without refinement an FP64 operator also takes a preconditioner built on
 * the FP32 (or BF16-MFMA) operator of the same problem: each application
 * rounds its argument to FP32 and widens the result, so the Krylov vectors,
 * dots and the stopping test stay FP64 while the V-cycle runs in FP32.
 * Without a preconditioner argument the solver builds its own from
 * fem.solver.preconditioning.method (krylovinverseoperators.hh:86-97);
 * binding the same operator again, after its coefficients changed,
 * recomputes that one in place (same result as a new one, no allocation
 * for block Jacobi and multigrid), as each Newton step does with S'(w). */
int fem_solver_bind(fem_solver *s, fem_op *op, fem_precond *pre, fem_op *inner_op,
                    fem_precond *inner_pre);
int fem_solver_unbind(fem_solver *s);
/* operator()(b, x): x is the initial guess on entry, the solution on exit */
int fem_solver_solve(fem_solver *s, const fem_vec *b, fem_vec *x, fem_solver_info *info);

/* ---- Newton ---------------------------------------------------------------------
 * Replaces NewtonInverseOperator (dune/fem/solver/newtoninverseoperator.hh:
 * 423-802) for the operators of this library: solve S[w] = u.  Each step
 * binds S'(w) (fem_op_jacobian, matrix-free) to an owned linear solver with
 * the preconditioner its fem.solver.preconditioning.method builds (rebuilt
 * per step), or the one given at bind, solves S'(w) dw = S[w] - u from
 * dw = 0 and sets w -= dw (:688-802).  A linear S takes one step and counts
 * no nonlinear iteration (:786-790).  Stopping and failure follow
 * NewtonInverseOperator::failed (:568-584) with the NewtonFailure codes
 * (:389-400).  With a linear solver in FP64 refinement (`refinement` = 1;
 * FP64 DG operators) each Newton system is
 * solved by FP32 inner solves on an FP32 twin of S'(w), the residual and the
 * update staying FP64; a preconditioner given at bind must then be FP32 and
 * serves the inner solves. */
#define FEM_NEWTON_SUCCESS 0
#define FEM_NEWTON_INVALID_RESIDUAL 1
#define FEM_NEWTON_ITERATIONS_EXCEEDED 2
#define FEM_NEWTON_LINEAR_ITERATIONS_EXCEEDED 3
#define FEM_NEWTON_LINESEARCH_FAILED 4
#define FEM_NEWTON_TOO_MANY_ITERATIONS 5
#define FEM_NEWTON_TOO_MANY_LINEAR_ITERATIONS 6
#define FEM_NEWTON_LINEAR_SOLVER_FAILED 7

#define FEM_LINESEARCH_NONE 0
#define FEM_LINESEARCH_SIMPLE 1
#define FEM_FORCING_NONE 0
#define FEM_FORCING_EISENSTATWALKER 1

typedef struct fem_newton fem_newton;

/* NewtonParameter (newtoninverseoperator.hh:101-380), keys under
 * fem.solver.nonlinear. */
typedef struct {
    double tolerance;              /* nonlinear.tolerance, default 1e-6 (:203-207)      */
    int maxiterations;             /* nonlinear.maxiterations, default INT_MAX (:231-235) */
    int maxlineariterations;       /* total over all steps; < 0: the linear solver's
                                      maxiterations (:246-250)                          */
    int linesearch;                /* nonlinear.linesearch none | simple (:273-288)       */
    int maxlinesearchiterations;   /* default INT_MAX (:259-264)                         */
    int forcing;                   /* nonlinear.forcing none | eisenstatwalker (:294-315);
                                      eisenstatwalker needs the linear errormeasure
                                      residualreduction (:173-181)                      */
    double eisenstatwalker_etamax; /* default 0.99 (:326-331)                            */
    double eisenstatwalker_gamma;  /* default 0.1 (:343-348)                             */
    int simplified;                /* nonlinear.simplified: the Jacobian of the first
                                      step only (:359-362)                              */
    int verbose;                   /* nonlinear.verbose                                  */
} fem_newton_params;

typedef struct {
    int converged;
    int failure;             /* FEM_NEWTON_*                                          */
    int iterations;          /* nonlinear steps (0 for a linear operator)            */
    int linear_iterations;   /* total; negative when a linear solve failed           */
    double residual;         /* |S[w] - u|_2 when last evaluated (the initial one for
                                a linear operator)                                   */
    double timing[3];        /* seconds: total, Jacobian (and bind), solve (:798-801) */
} fem_newton_info;

void fem_newton_default_params(fem_newton_params *p);
/* keys (prefix "fem.solver." optional, then "nonlinear." or, deprecated in
 * the reference and accepted here, "newton."): tolerance, maxiterations,
 * maxlineariterations, linesearch (or lineSearch), maxlinesearchiterations,
 * forcing (or linear.tolerance.strategy), eisenstatwalker.etamax,
 * eisenstatwalker.gamma, simplified, verbose.  Linear-solver keys
 * (nonlinear.linear.*) go to fem_solver_params_set: FEM_ERR_ARG here. */
int fem_newton_params_set(fem_newton_params *p, const char *key, const char *value);
int fem_newton_create(fem_ctx *ctx, const fem_newton_params *params, const fem_solver_params *linear,
                      fem_newton **out);
int fem_newton_destroy(fem_newton *n);
/* bind(op[, precond]) / unbind() (newtoninverseoperator.hh:540-551) */
int fem_newton_bind(fem_newton *n, fem_op *op, fem_precond *pre);
int fem_newton_unbind(fem_newton *n);
/* operator()(u, w): u the right-hand side (NULL: zero), w the initial guess
 * on entry, the solution on exit */
int fem_newton_solve(fem_newton *n, const fem_vec *u, fem_vec *w, fem_newton_info *info);

/* ---- learned coarse-space preconditioner ----------------------------------------
 * Two-level preconditioner for SIPG DG operators with an element-mean coarse
 * space (A_c = P^T A P, a 7-point stencil).  The coarse solve A_c c = R v is
 * started from the output of a 4-layer 3D CNN (3x3x3 dilated convolutions,
 * 32 channels, no biases, ReLU; BF16 weights and activations, FP32
 * accumulation, BF16 MFMA for the hidden layers) and finished by a few
 * Jacobi-preconditioned CG steps on the device; then z1 = alpha P c with the
 * energy-optimal alpha and z = z1 + BJ(v - A z1) with the exact block-Jacobi
 * smoother.  Nonlinear in v: FGMRES only.  FGMRES falls back to block Jacobi
 * when a restart cycle stalls.  No reference counterpart. */
typedef struct fem_cnn fem_cnn;

int fem_cnn_create(fem_ctx *ctx, int nx, int ny, int nz, const int *dilations /* 4 */, fem_cnn **out);
int fem_cnn_destroy(fem_cnn *c);
int64_t fem_cnn_nparams(const fem_cnn *c);
/* weights: layer blocks [27][cin][cout], tap t = (dx+1) + 3(dy+1) + 9(dz+1) */
int fem_cnn_set_weights(fem_cnn *c, const float *w);
int fem_cnn_get_weights(const fem_cnn *c, float *w);
/* host arrays: x [V][2] (voxel x fastest), y [V] */
int fem_cnn_forward(fem_cnn *c, const float *x, float *y);
/* MSE loss mean((y - t)^2) and its gradient w.r.t. the weights (host arrays) */
int fem_cnn_loss_grad(fem_cnn *c, const float *x, const float *t, double *loss, float *grad);
/* per-element kappa and the SIPG penalty factor: sets the second input
 * channel (1 / kappa) and the Galerkin coarse operator A_c = P^T A P of
 * element means (7-point stencil: only the penalty terms act on piecewise
 * constants) */
int fem_cnn_set_coefficient(fem_cnn *c, const double *kcell, double beta);
/* y = A_c x on element means (host arrays [V]) */
int fem_cnn_coarse_apply(fem_cnn *c, const float *x, float *y);
/* `steps` Jacobi-preconditioned CG steps on A_c x = b from x = 0, the coarse
 * solve of the learned preconditioner (host arrays [V]); rz (may be null):
 * rz[0] = final (r, D^-1 r), rz[1] = its start value */
int fem_cnn_coarse_solve(fem_cnn *c, const float *b, int steps, float *x, double *rz);
/* plain FP64 CG solve of op64 x = b from x = 0 with `pre64` (FP64 block
 * Jacobi or multigrid); stores coarse (residual, error) pairs of every
 * `every`-th iterate (at most max_pairs).  With the multigrid the stored
 * residual is r - A MG(r), what one V-cycle leaves (the input of the
 * multigrid form of the learned preconditioner). */
int fem_cnn_collect_pairs(fem_cnn *c, fem_op *op64, fem_precond *pre64, const fem_vec *b, int max_pairs, int every,
                          int *npairs);
/* Adam on the energy-norm distance between the coarse correction and the
 * solver's error, ||e_k - P c||_A^2 (normalised per pair; 1 for c = 0).
 * val_rel_energy: mean over the validation pairs of the relative coarse
 * energy error (c - c*)^T A_c (c - c*) / (c*^T A_c c*), c* = A_c^-1 R r_k
 * (-1 without validation pairs). */
int fem_cnn_train(fem_cnn *c, int epochs, int batch, float lr, float val_frac, unsigned seed, double *train_loss,
                  double *val_rel_energy);

#define FEM_LEARNED_OFF 0          /* block Jacobi only (fallback) */
#define FEM_LEARNED_NET 1          /* coarse CG started from the network output */
#define FEM_LEARNED_COARSE_ONLY 2  /* coarse CG started from zero: the method without the network */
/* coarse_steps: Jacobi-CG steps on A_c per application (0..1000) */
int fem_precond_create_learned(fem_op *op, fem_cnn *cnn, int coarse_steps, fem_precond **out);
/* the same with a choice of smoother: FEM_PRE_BLOCK_JACOBI (the form above:
 * coarse correction, then block Jacobi on what remains) or FEM_PRE_MULTIGRID
 * (a multigrid V-cycle, the learned coarse correction of the residual it
 * leaves, then one damped block-Jacobi sweep of the fine level;
 * FEM_LEARNED_OFF is the V-cycle and the sweep).  Train such a network on
 * pairs collected with a multigrid preconditioner (fem_cnn_collect_pairs). */
int fem_precond_create_learned_ex(fem_op *op, fem_cnn *cnn, int coarse_steps, int smoother, fem_precond **out);
int fem_precond_set_learned_mode(fem_precond *p, int mode);

/* ---- batched solves -----------------------------------------------------------
 * Many independent continuous-Q2 problems -div(kappa grad u) + c u = f on
 * [0,1]^2 (strong u = 0), of different mesh sizes, solved together: every
 * launch covers all unconverged instances.  No reference counterpart (the
 * reference solves one problem per FemScheme::solve,
 * dune/fem/schemes/femscheme.hh:195-253); each instance follows the CG
 * recurrence of dune/fem/solver/linear/cg.hh with Jacobi preconditioning in
 * FP32, inside an FP64 iterative refinement loop on the true residual.
 * DoFs of all instances are concatenated in instance order; coefficient and
 * forcing arrays are concatenated [sum_i E_i][9]. */
typedef struct fem_batch fem_batch;

typedef struct {
    int count;
    int converged;                 /* instances with FP64 ||b-Ax|| <= tol ||b||  */
    int outer_iterations;          /* refinement steps                           */
    int max_inner_iterations;      /* sum over steps of the slowest instance     */
    int total_inner_iterations;
    double max_true_rel_residual;
    double time_s;
    int tile_x, tile_y;            /* tile shape in use (autotuner)              */
} fem_batch_info;

int fem_batch_create(fem_ctx *ctx, int count, const int *nx, const int *ny, double c, fem_batch **out);
int fem_batch_destroy(fem_batch *b);
int64_t fem_batch_ndofs(const fem_batch *b);
/* count + 1 DoF offsets */
int fem_batch_dof_offsets(const fem_batch *b, int64_t *off);
int fem_batch_set_kappa_qp(fem_batch *b, const double *kvol);
/* w = A u for all instances (FP64 or FP32 vectors of fem_batch_ndofs) */
int fem_batch_apply(fem_batch *b, const fem_vec *u, fem_vec *w);
int fem_batch_assemble_rhs(fem_batch *b, const double *fq, fem_vec *F);
/* time every tile shape on the batch and keep the fastest; returns its shape */
int fem_batch_autotune(fem_batch *b, int *tile_x, int *tile_y, float *ms_per_shape /* 5 entries or NULL */);
/* x: initial guess in, solution out (FP64).  use_graph: replay one captured
 * inner step as a HIP graph */
int fem_batch_solve(fem_batch *b, const fem_vec *rhs, fem_vec *x, double tol, double inner_tol, int inner_maxit,
                    int max_outer, int use_graph, fem_batch_info *info);
/* per-instance results of the last solve: inner iterations and FP64 relative residual */
int fem_batch_instance_results(const fem_batch *b, int *iterations, double *rel_residual);

/* ---- quadrature (shared with hosts that generate coefficient data) ---------- */
/* m-point Gauss-Legendre rule on [0,1], ascending points */
int fem_gauss_points(int m, double *x, double *w);

#ifdef __cplusplus
}
#endif
#endif
