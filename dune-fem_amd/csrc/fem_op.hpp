// Operator and preconditioner objects behind the C ABI.
#pragma once

#include "fem_cg.hpp"
#include "fem_dg.hpp"

namespace femamd {

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    void alloc(size_t b)
    {
        free();
        if (b) {
            hipError_t e = hipMalloc(&p, b);
            if (e != hipSuccess) throw Error(FEM_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
        }
        bytes = b;
    }
    // allocate unless it already holds exactly b bytes (contents kept then)
    void ensure(size_t b)
    {
        if (b != bytes || (b && !p)) alloc(b);
    }
    void free()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    ~DevBuf() { free(); }
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    template <typename T> T *as() const { return (T *)p; }
};

struct Operator {
    Context *ctx = nullptr;
    fem_op_desc desc{};
    int k = 0;         // points/modes per direction (DG: p+1; CG: 3)
    int64_t E = 0, N = 0;
    int vec_prec = FEM_F64;  // precision of the vectors it acts on
    // DG
    DGArgs<double> a64{};
    DGArgs<float> a32{};
    DevBuf tab64, tab32;
    DevBuf kvol64, kface64[3], kcell64, kvol32, kface32[3], kcell32;
    // CG
    CGInst inst{};
    DevBuf inst_dev, cgtab;
    int tiles = 0;

    void init(Context *c, const fem_op_desc &d);
    void set_kappa_qp(const double *kvol, const double *kx, const double *ky, const double *kz);
    void set_kappa_cell(const double *kc);
    // boundary data u = g: DG (weak, Nitsche and inflow terms): g at the face
    // quadrature points of each normal direction [faces_d][m^2]; CG Q2
    // (strong): nodal values [N], only the Dirichlet nodes are read
    void set_boundary_data(const double *g0, const double *g1, const double *g2);
    // CG Q2: x_D = g_D (+ add_D) (the reference's setConstraints and
    // addConstraints); DG: nothing
    void constrain(Vec &x, const Vec *add = nullptr) const;
    DevBuf gbnd64[3];   // boundary data as set (DG: per direction; CG: gbnd64[0])
    bool has_bnd = false;
    // DG reaction beyond the constant c: a pointwise coefficient at the
    // volume quadrature points (cvol, FP64 and FP32 copies) and the
    // nonlinear law nl_coef u^nl_pow (the volume term int (c + cvol +
    // nl_coef u^nl_pow) u v); constant or cellwise kappa only
    DevBuf cvol64, cvol32;
    double nl_coef = 0;
    int nl_pow = 0;
    // a compiled pointwise reaction r(u, x) (shared library, owned handle)
    // and its scratch: u and r at the volume quadrature points
    void *plugin_dl = nullptr;
    fem_reaction_eval_fn plugin = nullptr;
    mutable DevBuf pq, ps;
    bool nonlinear() const { return nl_pow > 0 || plugin != nullptr; }
    // volume quadrature points: E (p+1)^3 (DG), E 9 (Q2 in 2D)
    int64_t nqp() const { return is_dg() ? E * k * k * k : E * 9; }
    bool has_reaction() const { return cvol64.p != nullptr || nonlinear(); }
    void set_reaction_plugin(const char *path);
    fem_qp_geometry qp_geometry() const;
    Operator() = default;
    Operator(const Operator &) = delete;
    Operator &operator=(const Operator &) = delete;
    ~Operator();
    // cvol from host values [E][k^3] (null clears)
    void set_reaction_qp(const double *cvol);
    // the law nl_coef u^pw, pw in 1..4 (coef = 0 clears)
    void set_reaction_law(double coef, int pw);
    // u (operator precision) at the volume quadrature points, into `out`
    // of the same precision [E][k^3]
    void eval_qp(const Vec &u, Vec &out) const;
    // the Jacobian at u into jop (an operator made by make_jacobian from
    // this one, in its precision or FP32): cvol_J = cvol + (pw + 1) nl_coef
    // u^pw (+ dr/du of a compiled reaction), no law
    void jacobian(const Vec &u, Operator &jop) const;
    // jop: a linear operator of the same description and coefficient data
    // (kappa, boundary data, cvol), device copies; precision >= 0 makes it
    // in that precision (FP32 for an FP64 operator: the inner operator of a
    // refinement solve)
    void make_jacobian(Operator &jop, int precision = -1) const;
    // y = A x on raw device pointers of the operator's vector precision;
    // `dot` (DG only, may be null): per-workgroup FP64 partials of <x, y>,
    // dot_blocks() entries
    void apply_raw(const void *x, void *y, const int *skip, double *dot = nullptr) const;
    int64_t dot_blocks() const;
    void apply(const Vec &x, Vec &y, const int *skip = nullptr) const;
    void rhs(const double *fq_host, Vec &F) const;
    void diagonal(Vec &d) const;  // FP64 or FP32 output
    // DG: the element blocks A_ee, [E][nb][nb] row-major in the local DoF
    // order, FP64 (by coloured probing with the apply kernel)
    void block_diagonal(Vec &blocks) const;
    bool is_dg() const { return desc.kind == FEM_OP_DG_SIPG_3D; }
};

struct Learned;
struct Multigrid;

struct Precond {
    Operator *op = nullptr;
    int kind = FEM_PRE_NONE;
    int prec = FEM_F64;
    DevBuf dinv;       // Jacobi
    DevBuf Q, lam;     // block Jacobi (tensor-product eigen-decomposition)
    Learned *learned = nullptr;   // FEM_PRE_LEARNED (owned)
    Multigrid *mg = nullptr;      // FEM_PRE_MULTIGRID (owned)
    // an FP32 preconditioner applied to the FP64 vectors of an FP64 Krylov
    // solve (not owned): x is rounded to FP32, M32 applied, the result
    // widened back; cx, cy hold the FP32 vectors
    const Precond *lowp = nullptr;
    DevBuf cx, cy;
    bool bound = false;   // block Jacobi: a pointwise reaction by its largest value per element
    fem_op_desc desc{};   // the operator's, at init
    // bound (block Jacobi): a pointwise reaction enters the element blocks by
    // its largest value, not its mean (the multigrid smoothers)
    void init(Operator *o, int kind, bool bound = false);
    // recompute the data built from the operator's coefficients in place
    // (block Jacobi, Jacobi, multigrid), for new coefficient values on the
    // same operator
    void update();
    // whether update() serves operator o as preconditioner kind k: the same
    // operator object with the same description and structure
    bool updatable_for(const Operator *o, int k) const;
    void init_widening(Operator *o64, const Precond *p32);
    Learned *learned_part() const { return lowp ? lowp->learned : learned; }
    // `dot` (block Jacobi only, may be null): per-workgroup FP64 partials of
    // <x, y> and then of <x, x>, op->dot_blocks() entries each
    void apply_raw(const void *x, void *y, const int *skip, double *dot = nullptr) const;
    bool pointwise() const { return kind == FEM_PRE_JACOBI; }
    bool fused_dot() const { return kind == FEM_PRE_BLOCK_JACOBI; }
    ~Precond();
};

}  // namespace femamd

struct fem_op { femamd::Operator o; femamd::Context *held = nullptr; };
struct fem_precond { femamd::Precond p; femamd::Context *held = nullptr; };
