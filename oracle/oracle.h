/*
 * oracle.h -- CPU restatement of dune-fem's matrix-free Galerkin apply and
 * Krylov solvers.  TEST INFRASTRUCTURE ONLY: the product library never links,
 * loads or calls anything declared here.  Tests, the smoke check and the CPU
 * baseline leg of the benchmark use it as the checker.
 *
 * The oracle keeps the reference algorithm's structure on purpose:
 *   - dense per-element basis tables (nq x nb values and gradients), as
 *     DefaultBasisFunctionSet::evaluateAll/jacobianAll/axpy
 *     (dune/fem/space/basisfunctionset/default.hh:276-487);
 *   - element loop, skeleton integral once per neighbour pair with
 *     index(inside) < index(outside), boundary integral on domain faces
 *     (dune/fem/schemes/galerkin.hh:811-917, :879);
 *   - quadrature orders 2p (volume) and 2p+1 (faces)
 *     (dune/fem/schemes/galerkin.hh:131-132) on tensor Gauss rules
 *     (dune/fem/quadrature/femquadratures_inline.hh:59-93);
 *   - CG and restarted GMRES recurrences of dune/fem/solver/linear/{cg,gmres}.hh.
 *
 * Parity pins (what fixes each function without running the reference) are
 * listed next to each declaration; the tests under tests/ implement them.
 */
#ifndef DUNEFEM_ORACLE_H
#define DUNEFEM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- 1D quadrature -------------------------------------------------------
 * m-point Gauss-Legendre rule on [0,1] (points ascending), computed by Newton
 * iteration on P_m in long double.  Follows GaussPts
 * (dune/fem/quadrature/gausspoints_implementation.hh:12-160) in meaning.
 * Pinned by: weight sum 1, exactness for monomials up to degree 2m-1, and the
 * quad_test known answers (dune/fem/quadrature/test/quad_test.cc:246-278). */
int ora_gauss(int m, double *x, double *w);
/* smallest m with 2m-1 >= order (CubeQuadratureBase, femquadratures_inline.hh:59-67) */
int ora_gauss_m_for_order(int order);

/* ---- 1D shape functions --------------------------------------------------
 * Orthonormal shifted Legendre polynomial sqrt(2n+1) P_n(2x-1) on [0,1]
 * (LegendrePolynomials, dune/fem/space/shapefunctionset/legendrepolynomials.hh:24-60).
 * Pinned by orthonormality and the closed-form coefficient formula. */
double ora_legendre(int n, double x);
double ora_legendre_d(int n, double x);
/* Quadratic Lagrange basis on nodes {0, 1/2, 1}
 * (LagrangeShapeFunctionSet, dune/fem/space/lagrange/shapefunctionset.hh:78-330).
 * Pinned by the nodal property and partition of unity. */
double ora_lagrange2(int i, double x);
double ora_lagrange2_d(int i, double x);

/* ---- SIPG DG operator on a structured hexahedral mesh of [0,1]^3 ----------
 * Form (the UFL SIPG + upwind form of dune/fempy/test/parallelTest.py:113-129
 * and pydemo/advectiondiffusion.py:30-56, with weak u = 0 on the boundary):
 *   sum_K int_K (kappa grad u - b u) . grad v + c u v
 * + sum_F int_F kp beta/h_F [u][v] - {kappa grad u}.n [v] - [u]{kappa grad v}.n
 *              + (max(b.n,0) u+ - max(-b.n,0) u-) [v]
 * + sum_dF int_F kappa beta/h_F u v - kappa grad u.n v - kappa u grad v.n
 *              + max(b.n,0) u v
 * h_F = avg|K|/|F| (= h along the face normal on the uniform mesh).
 * kp = kappa at the face point (pointwise kappa) or avg(kappa) (cellwise kappa).
 *
 * Layouts: element e = ix + nx*(iy + ny*iz); DoF = e*nb + local, local index
 * (a*(p+1)+b)*(p+1)+c with a the x-degree (dim 0 outermost,
 * dune/fem/space/shapefunctionset/legendre.hh:178-193).
 * Volume qp q = qx + m*(qy + m*qz) (dim 0 fastest).
 * Faces normal to d are numbered on the element grid extended by one layer
 * in d (x fastest); face qp = t1 + m*t2 over the tangential dims, lower dim
 * first. */
enum { ORA_KAPPA_CONST = 0, ORA_KAPPA_QP = 1, ORA_KAPPA_CELL = 2 };

typedef struct {
    int n[3];
    int p;
    double beta;          /* penalty factor, e.g. 10 p^2 */
    double b[3];          /* constant advection field */
    double c;             /* reaction coefficient */
    int kappa_mode;
    double kappa_const;
    const double *kvol;   /* ORA_KAPPA_QP: [E][m^3] */
    const double *kface[3]; /* ORA_KAPPA_QP: [faces_d][m^2] */
    const double *kcell;  /* ORA_KAPPA_CELL: [E] */
    /* pointwise reaction coefficient at the volume quadrature points
     * [E][m^3] (x fastest), added to c; NULL for none.  With c(x) =
     * c3 u_lin(x)^2 it gives the residual of the nonlinear term c3 u^3
     * (u_lin = u) and, with 3 c3 u_lin^2, its Jacobian at u_lin */
    const double *cvol;
} ora_dg_t;

int64_t ora_dg_ndofs(const ora_dg_t *op);
/* w = A u (w overwritten).  nthreads <= 1: serial. */
void ora_dg_apply(const ora_dg_t *op, const double *u, double *w, int nthreads);
/* the same w = A u evaluated by sum factorisation (1D contractions, faces
 * owner-computes); the labelled sum-factorised CPU baseline */
void ora_dg_apply_sf(const ora_dg_t *op, const double *u, double *w, int nthreads);
/* F_i = int f phi_i with f given at the volume quadrature points [E][m^3] */
void ora_dg_rhs(const ora_dg_t *op, const double *fq, double *F);
/* F += the boundary-data part of the load vector (Nitsche and inflow terms
 * for u = g), g at the boundary face quadrature points [faces_d][m^2] */
void ora_dg_rhs_bnd(const ora_dg_t *op, const double *const gface[3], double *F);
/* element blocks A_KK, [E][nb][nb] row-major (row = test function) */
void ora_dg_block_diag(const ora_dg_t *op, double *blocks);

/* ---- continuous Q2 Lagrange operator on a structured quad mesh of [0,1]^2 --
 * Form: int kappa grad u . grad v + c u v, strong u = 0 on the boundary
 * (DirichletWrapperOperator: w_D = u_D, dune/fem/schemes/dirichletwrapper.hh:101-105,
 * dune/fem/schemes/dirichletconstraints.hh:132-160).
 * Nodes on the (2nx+1) x (2ny+1) lattice, node = j*(2nx+1) + i, i along x.
 * Element local node (i,j) i fastest.  Volume qp as for DG (3x3 Gauss). */
typedef struct {
    int n[2];
    double c;
    int kappa_mode;       /* ORA_KAPPA_CONST or ORA_KAPPA_QP */
    double kappa_const;
    const double *kvol;   /* [E][9] */
    const double *cvol;   /* pointwise reaction at the volume qps [E][9], added to c, or NULL */
} ora_cg_t;

int64_t ora_cg_ndofs(const ora_cg_t *op);
/* u at the volume quadrature points: out[e*9 + qx + 3*qy] */
void ora_cg_eval_qp(const ora_cg_t *op, const double *u, double *out);
void ora_cg_apply(const ora_cg_t *op, const double *u, double *w);
/* load vector with f at volume qps; Dirichlet entries set to 0 */
void ora_cg_rhs(const ora_cg_t *op, const double *fq, double *F);
/* operator diagonal (Dirichlet entries 1) */
void ora_cg_diag(const ora_cg_t *op, double *d);

/* ---- Krylov solvers --------------------------------------------------------
 * Operator and preconditioner are callbacks y = op(x). */
typedef void (*ora_apply_fn)(void *ctx, const double *x, double *y);

/* LinearSolver::cg (dune/fem/solver/linear/cg.hh:19-117), including its
 * stopping rule on <r, M^-1 r> and the signed iteration count.
 * criteria: 0 absolute, 1 relative, 2 residual reduction. */
int ora_cg(int64_t n, ora_apply_fn op, void *opctx, ora_apply_fn pre, void *prectx,
           double *x, const double *b, double eps, int maxit, int criteria);
/* LinearSolver::gmres (dune/fem/solver/linear/gmres.hh:117-299): right
 * preconditioned, restarted, classical Gram-Schmidt, Givens on the host. */
/* BiCGStab (bicgstab.hh:56-222), right preconditioned; signed iterations */
int ora_bicgstab(int64_t n, ora_apply_fn op, void *opctx, ora_apply_fn pre, void *prectx,
                 double *x, const double *b, double eps, int maxit, int criteria);
int ora_gmres(int64_t n, ora_apply_fn op, void *opctx, ora_apply_fn pre, void *prectx,
              double *x, const double *b, int restart, double eps, int maxit, int criteria);
/* Flexible GMRES(m) (Saad 1993): stores z_j = M_j^-1 v_j; the same
 * restart/Givens/stopping structure as the GMRES above. */
int ora_fgmres(int64_t n, ora_apply_fn op, void *opctx, ora_apply_fn pre, void *prectx,
               double *x, const double *b, int restart, double eps, int maxit, int criteria);

/* adapters so callers can pass the operators above as callbacks */
void ora_dg_apply_cb(void *op, const double *x, double *y);
void ora_cg_apply_cb(void *op, const double *x, double *y);
/* y = d * x elementwise; ctx = struct { int64_t n; const double *d; } */
typedef struct { int64_t n; const double *d; } ora_diag_t;
void ora_diag_cb(void *ctx, const double *x, double *y);
/* block-diagonal preconditioner, ctx holds inverted blocks [E][nb][nb] */
typedef struct { int64_t nblocks; int nb; const double *inv; } ora_block_t;
void ora_block_cb(void *ctx, const double *x, double *y);

#ifdef __cplusplus
}
#endif
#endif
