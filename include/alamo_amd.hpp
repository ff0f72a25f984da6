// C++ façade over the C ABI of libalamo_amd.so, mirroring the reference's class and
// method names so that code written against
//   Operator::Elastic<SYM>             (src/Operator/Elastic.H:21-190)
//   Operator::Operator<Grid::Node>     (src/Operator/Operator.H:27-125)
//   Solver::Nonlocal::Linear           (src/Solver/Nonlocal/Linear.H:18-389)
//   BC::Operator::Elastic::Constant    (src/BC/Operator/Elastic/Constant.H:31-649)
// keeps its call sequence (construct, SetBC, SetModel, solver.Define, solver.solve, Clear).
// Fields are device pointers to the single-box nodal layout described in alamo_amd.h.
// Where the reference calls Util::Abort, this header throws alamo::Abort.
#ifndef ALAMO_AMD_HPP
#define ALAMO_AMD_HPP

#include <array>
#include <stdexcept>
#include <string>

#include "alamo_amd.h"

namespace alamo {
struct Abort : std::runtime_error {
    int code;
    Abort(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
inline void check(int rc) {
    if (rc != ALAMO_OK) throw Abort(rc, alamo_last_error());
}
}  // namespace alamo

namespace BC {
namespace Operator {
namespace Elastic {
// BC::Operator::Elastic::Constant: 26 regions x 3 components of (type, value)
class Constant {
public:
    enum Type { Displacement = ALAMO_BC_DISPLACEMENT, Traction = ALAMO_BC_TRACTION, Periodic = ALAMO_BC_PERIODIC,
                Neumann = ALAMO_BC_NEUMANN, None = ALAMO_BC_NONE };
    enum Face { XLO, YLO, ZLO, XHI, YHI, ZHI, YLO_ZLO, YLO_ZHI, YHI_ZLO, YHI_ZHI, ZLO_XLO, ZLO_XHI, ZHI_XLO,
                ZHI_XHI, XLO_YLO, XLO_YHI, XHI_YLO, XHI_YHI, XLO_YLO_ZLO, XLO_YLO_ZHI, XLO_YHI_ZLO, XLO_YHI_ZHI,
                XHI_YLO_ZLO, XHI_YLO_ZHI, XHI_YHI_ZLO, XHI_YHI_ZHI, INT };
    enum Direction { X = 0, Y = 1, Z = 2 };
    Constant() {
        for (int f = 0; f < 26; f++)
            for (int d = 0; d < 3; d++) { type[f][d] = Displacement; val[f][d] = 0.0; }
    }
    void Set(Face face, Direction dir, Type t, double value) {
        type[face][dir] = t;
        val[face][dir] = value;
    }
    int type[26][3];
    double val[26][3];
};
}  // namespace Elastic
}  // namespace Operator
}  // namespace BC

namespace Operator {

enum SymKind { Isotropic = ALAMO_SYM_ISOTROPIC, MajorMinor = ALAMO_SYM_MAJORMINOR, Diagonal = ALAMO_SYM_DIAGONAL };

// Operator::Elastic<SYM> over one nodal box [0, n_cell] with cell size dx.
template <int SYM>
class Elastic {
public:
    Elastic(const std::array<int, 3>& n_cell, const std::array<double, 3>& dx, int max_coarsening_level = -1) {
        alamo::check(alamo_mg_desc_default(&desc_));
        for (int d = 0; d < 3; d++) { desc_.n_cell[d] = n_cell[d]; desc_.dx[d] = dx[d]; }
        desc_.sym = SYM;
        desc_.max_coarsening_level = max_coarsening_level;
    }
    ~Elastic() { if (h_) alamo_mg_destroy(h_); }
    Elastic(const Elastic&) = delete;
    Elastic& operator=(const Elastic&) = delete;

    // Elastic::SetBC (Elastic.H:61-65); must precede SetModel, which builds the hierarchy
    void SetBC(const BC::Operator::Elastic::Constant* bc) {
        for (int f = 0; f < 26; f++)
            for (int d = 0; d < 3; d++) { desc_.bc_type[f][d] = bc->type[f][d]; desc_.bc_val[f][d] = bc->val[f][d]; }
    }
    void SetOmega(double omega) { desc_.omega = omega; UpdateOpts(); }    // Operator.H:57
    void SetPrecision(int p) { desc_.precision = p; }
    // Elastic::SetModel(amrlev, MultiTab) (Elastic.cpp:87-116); C: device, AoS Matrix4 records
    void SetModel(int amrlev, const double* C, int layout = ALAMO_LAYOUT_AOS) {
        ensure();
        alamo::check(alamo_mg_set_model(h_, amrlev, C, layout));
    }
    // Elastic::SetPsi (Elastic.H:55-57): psi a device cell field, n_cell product values
    void SetPsi(int amrlev, const double* psi, double psi_small = 1e-8) {
        ensure();
        alamo::check(alamo_mg_set_psi(h_, amrlev, psi, psi_small));
    }
    void prepareForSolve() { ensure(); alamo::check(alamo_mg_prepare(h_)); }
    // Operator<Grid::Node>::Apply -> Elastic::Fapply (Operator.H:55, Elastic.cpp:139-292)
    void Apply(int amrlev, int mglev, double* out, const double* in) const {
        alamo::check(alamo_mg_fapply(h_, amrlev, mglev, out, in, nullptr));
    }
    void Fsmooth(int amrlev, int mglev, double* x, const double* b) const {
        alamo::check(alamo_mg_fsmooth(h_, amrlev, mglev, x, b, nullptr));
    }
    void Diagonal(int amrlev, int mglev, double* diag) const {
        alamo::check(alamo_mg_diagonal(h_, amrlev, mglev, diag, nullptr));
    }
    // Elastic::Strain / Stress / Energy (Elastic.H:71-89, Elastic.cpp:430-603); device fields
    void Strain(int amrlev, double* eps, const double* u, bool voigt = false) const {
        alamo::check(alamo_mg_strain(h_, amrlev, u, eps, voigt ? 1 : 0, nullptr));
    }
    void Stress(int amrlev, double* sigma, const double* u, bool voigt = false) const {
        alamo::check(alamo_mg_stress(h_, amrlev, u, sigma, voigt ? 1 : 0, nullptr));
    }
    void Energy(int amrlev, double* energy, const double* u) const {
        alamo::check(alamo_mg_energy(h_, amrlev, u, energy, nullptr));
    }
    // Elastic::Error0x (Elastic.cpp:382-404): R0x = x - A (x / D)
    void Error0x(int amrlev, int mglev, double* R0x, const double* x) const {
        alamo::check(alamo_mg_error0x(h_, amrlev, mglev, R0x, x, nullptr));
    }
    // Elastic::SetUniform (Elastic.H:105): true drops grad(C) . grad(u) from the interior rows
    void SetUniform(bool uniform) {
        uniform_ = uniform;
        if (h_) alamo::check(alamo_mg_set_uniform(h_, uniform ? 1 : 0));
    }
    // stored flags the reference keeps without reading them here (Elastic.H:48,104,106-107;
    // Operator.H:59): homogeneity, testing, the coarse-AMR coefficient overwrite (one AMR level
    // here), DDW normalisation
    void SetHomogeneous(bool a) { homogeneous_ = a; }
    void SetTesting(bool a) { testing_ = a; }
    void SetAverageDownCoeffs(bool a) { average_down_coeffs_ = a; }
    void SetNormalizeDDW(bool a) { normalize_ddw_ = a; }
    // Operator<Grid::Node>::isSingular / isBottomSingular / getNGrow / getNComp (Operator.H:96-105)
    bool isSingular(int /*amrlev*/) const { return false; }
    bool isBottomSingular() const { return false; }
    int getNGrow(int = 0, int = 0) const { return 2; }
    int getNComp() const { return 3; }
    int NMGLevels() const { int n = 0; alamo::check(alamo_mg_num_levels(h_, &n)); return n; }
    // push changed solver options of desc() to a built handle (before the build they are read at creation)
    void UpdateOpts() {
        if (h_) alamo::check(alamo_mg_update_opts(h_, &desc_));
    }
    alamo_mg* handle() { ensure(); return h_; }
    alamo_mg_desc& desc() { return desc_; }

private:
    void ensure() {
        if (h_) return;
        alamo::check(alamo_mg_create(&desc_, &h_));
        if (uniform_) alamo::check(alamo_mg_set_uniform(h_, 1));
    }
    alamo_mg_desc desc_{};
    alamo_mg* h_ = nullptr;
    bool uniform_ = false, homogeneous_ = false, testing_ = false, average_down_coeffs_ = false,
         normalize_ddw_ = false;
};

}  // namespace Operator

namespace Solver {
namespace Nonlocal {
// Solver::Nonlocal::Linear (Linear.H:18-389) over Operator::Elastic<SYM>
template <int SYM>
class Linear {
public:
    Linear() = default;
    explicit Linear(::Operator::Elastic<SYM>& op) { Define(op); }
    void Define(::Operator::Elastic<SYM>& op) {
        if (linop_) throw alamo::Abort(ALAMO_ERR_STATE, "Solver cannot be re-defined");
        linop_ = &op;
    }
    void Clear() {
        if (!linop_) throw alamo::Abort(ALAMO_ERR_STATE, "Solver cannot be cleared if not defined");
        linop_ = nullptr;
    }
    // Linear::setMaxIter / ... (Linear.H:178-184): take effect on the operator's built handle too
    void setMaxIter(int n) { opt().max_iter = n; push(); }
    void setBottomMaxIter(int n) { opt().bottom_max_iter = n; push(); }
    void setFixedIter(int n) { opt().fixed_iter = n; push(); }
    void setVerbose(int v) { opt().verbose = v; push(); }
    void setPreSmooth(int n) { opt().pre_smooth = n; push(); }
    void setPostSmooth(int n) { opt().post_smooth = n; push(); }
    void setBottomTolerance(double t) { opt().bottom_tol_rel = t; push(); }
    // amrex::MLMG::setMaxFmgIter: F-cycles for the first n iterations (the reference sets 1e8)
    void setMaxFmgIter(int n) { opt().max_fmg_iter = n; push(); }
    // solver.bottom_solver (Linear.H:280-282): "bicgstab", "smoother" or "dense" ("cg" is not built)
    void setBottomSolver(const std::string& name) {
        int b = name == "bicgstab" ? ALAMO_BOTTOM_BICGSTAB
                : name == "smoother" ? ALAMO_BOTTOM_SMOOTHER
                : name == "dense"    ? ALAMO_BOTTOM_DENSE
                                     : -1;
        if (b < 0) throw alamo::Abort(ALAMO_ERR_ARG, "unknown bottom_solver " + name);
        opt().bottom_solver = b;
        push();
    }
    // smoother of the cycles: "jacobi" (the reference's damped point Jacobi) or "l1" (l1-Jacobi;
    // no reference counterpart)
    void setSmoother(const std::string& name) {
        int v = name == "jacobi" ? ALAMO_SMOOTHER_JACOBI : name == "l1" ? ALAMO_SMOOTHER_L1 : -1;
        if (v < 0) throw alamo::Abort(ALAMO_ERR_ARG, "unknown smoother " + name);
        opt().smoother = v;
        push();
    }
    // operators of the coarse levels: "restrict" (the reference's re-discretisation) or "galerkin"
    // (R A P; no reference counterpart)
    void setCoarseOp(const std::string& name) {
        int v = name == "restrict" ? ALAMO_COARSE_RESTRICT : name == "galerkin" ? ALAMO_COARSE_GALERKIN : -1;
        if (v < 0) throw alamo::Abort(ALAMO_ERR_ARG, "unknown coarse_op " + name);
        opt().coarse_op = v;
        push();
    }
    // Linear::apply (Linear.H:169-174): rhs = A sol on the finest level
    void apply(double* rhs, const double* sol) { defined().Apply(0, 0, rhs, sol); }
    // Linear::solveaffine (Linear.H:53-130) subtracts A 0 under inhomogeneous boundaries first;
    // Elastic stores its homogeneity flag without reading it (Elastic.H:48,172), so A 0 = 0 and
    // this is solve
    double solveaffine(double* sol, const double* rhs, double tol_rel, double tol_abs) {
        return solve(sol, rhs, tol_rel, tol_abs);
    }
    // Linear::solve (Linear.H:131-147): sol (initial guess in, solution out), rhs; returns |r|_inf
    double solve(double* sol, const double* rhs, double tol_rel, double tol_abs) {
        if (!linop_) throw alamo::Abort(ALAMO_ERR_STATE, "Solver not defined");
        alamo_solve_stats st{};
        alamo::check(alamo_linear_solve(linop_->handle(), sol, rhs, tol_rel, tol_abs, &st));
        last_ = st;
        return st.final_res_inf;
    }
    const alamo_solve_stats& stats() const { return last_; }

protected:
    ::Operator::Elastic<SYM>& defined() {
        if (!linop_) throw alamo::Abort(ALAMO_ERR_STATE, "Solver not defined");
        return *linop_;
    }
    alamo_mg_desc& opt() { return defined().desc(); }
    void push() { linop_->UpdateOpts(); }
    ::Operator::Elastic<SYM>* linop_ = nullptr;
    alamo_solve_stats last_{};
};

// Solver::Nonlocal::Newton<T> (Newton.H:14-654) for an affine model: stiffness C (the model's
// DDW, device, any SetModel layout) and eigenstrain F0 (9 AoS doubles per node, device, or null).
template <int SYM>
class Newton : public Linear<SYM> {
public:
    Newton() = default;
    explicit Newton(::Operator::Elastic<SYM>& op) : Linear<SYM>(op) {}
    void setNRIters(int n) { nriters_ = n; }           // newton.nriters (Newton.H:646)
    void setNRTolerance(double t) { nrtolerance_ = t; } // newton.nrtolerance (Newton.H:649)
    // Newton::setPsi (Newton.H:46-50)
    void setPsi(const double* psi, double psi_small = 1e-8) { this->defined().SetPsi(0, psi, psi_small); }
    // Newton::solve (Newton.H:279-358): u updated in place; returns max |du| on an early stop, else 0
    double solve(double* u, const double* b, const double* C, const double* F0, double tol_rel, double tol_abs,
                 int layout = ALAMO_LAYOUT_AOS) {
        this->defined().SetModel(0, C, layout);
        alamo_solve_stats st{};
        double rel = 0.0;
        alamo::check(alamo_newton_solve(this->defined().handle(), u, b, F0, nriters_, nrtolerance_, tol_rel, tol_abs,
                                        &st, &rel));
        this->last_ = st;
        return rel;
    }
    // Newton::compResidual (Newton.H:445-500): res = b - div DW(grad u), boundary rows b - BC(...)
    void compResidual(double* res, const double* u, const double* b, const double* C, const double* F0,
                      int layout = ALAMO_LAYOUT_AOS) {
        this->defined().SetModel(0, C, layout);
        alamo::check(alamo_newton_rhs(this->defined().handle(), u, b, F0, res));
    }
    // Newton::compLinearSolverResidual (Newton.H:502-526): res = b - A u
    void compLinearSolverResidual(double* res, const double* u, const double* b) {
        alamo::check(alamo_mg_residual(this->defined().handle(), 0, 0, res, u, b, nullptr));
    }
    // Newton::W / DW (Newton.H:528-611) of the operator's model
    void W(double* w, const double* u, const double* F0) {
        alamo::check(alamo_newton_w(this->defined().handle(), 0, u, F0, w, nullptr));
    }
    void DW(double* dw, const double* u, const double* F0) {
        alamo::check(alamo_newton_dw(this->defined().handle(), 0, u, F0, dw, nullptr));
    }

private:
    int nriters_ = 1;
    double nrtolerance_ = 0.0;
};

// Outer flexible GMRES around the operator's cycle (no reference counterpart; same Define /
// solve / Clear sequence as Linear, so a caller swaps one class name).  With Learned the
// operator needs an attached network (alamo::Net::Attach).
template <int SYM>
class Krylov {
public:
    enum Precond { None = ALAMO_PRECOND_NONE, Cycle = ALAMO_PRECOND_MG, Learned = ALAMO_PRECOND_LEARNED };
    Krylov() { alamo::check(alamo_krylov_opts_default(&opts_)); }
    explicit Krylov(::Operator::Elastic<SYM>& op) : Krylov() { Define(op); }
    void Define(::Operator::Elastic<SYM>& op) {
        if (linop_) throw alamo::Abort(ALAMO_ERR_STATE, "Solver cannot be re-defined");
        linop_ = &op;
    }
    void Clear() {
        if (!linop_) throw alamo::Abort(ALAMO_ERR_STATE, "Solver cannot be cleared if not defined");
        linop_ = nullptr;
    }
    void setRestart(int m) { opts_.restart = m; }
    void setMaxIter(int n) { opts_.max_iter = n; }
    void setPrecond(Precond p) { opts_.precond = p; }
    void setStagnationRatio(double s) { opts_.stagnation_ratio = s; }
    double solve(double* sol, const double* rhs, double tol_rel, double tol_abs) {
        if (!linop_) throw alamo::Abort(ALAMO_ERR_STATE, "Solver not defined");
        alamo_solve_stats st{};
        alamo::check(alamo_krylov_solve(linop_->handle(), sol, rhs, tol_rel, tol_abs, &opts_, &st));
        last_ = st;
        return st.final_res_inf;
    }
    const alamo_solve_stats& stats() const { return last_; }
    alamo_krylov_opts& opts() { return opts_; }

private:
    ::Operator::Elastic<SYM>* linop_ = nullptr;
    alamo_krylov_opts opts_{};
    alamo_solve_stats last_{};
};
}  // namespace Nonlocal
}  // namespace Solver

namespace alamo {
// The learned correction network (no reference counterpart): create, train on an operator's
// own problem, save / load, attach to an operator for Krylov<SYM>::Learned.
class Net {
public:
    explicit Net(uint64_t seed = 0) { check(alamo_net_create(&n_, seed)); }
    ~Net() { if (n_) alamo_net_destroy(n_); }
    Net(const Net&) = delete;
    Net& operator=(const Net&) = delete;
    template <int SYM>
    alamo_train_stats Train(::Operator::Elastic<SYM>& op, const alamo_train_opts& o) {
        alamo_train_stats st{};
        check(alamo_net_train(n_, op.handle(), &o, &st));
        return st;
    }
    template <int SYM>
    void Attach(::Operator::Elastic<SYM>& op) { check(alamo_mg_set_net(op.handle(), n_)); }
    void Save(const std::string& path) { check(alamo_net_save(n_, path.c_str())); }
    void Load(const std::string& path) { check(alamo_net_load(n_, path.c_str())); }
    int64_t NumParams() const { int64_t p = 0; check(alamo_net_num_params(n_, &p)); return p; }
    alamo_net* handle() { return n_; }

private:
    alamo_net* n_ = nullptr;
};
}  // namespace alamo

#endif
