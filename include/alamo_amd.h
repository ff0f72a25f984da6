/* C ABI of the MI355X nodal elastic multigrid library (libalamo_amd.so).
 *
 * This is the drop-in boundary beneath Integrator::Mechanics: each entry point
 * replaces one member of the reference's operator/solver plugin surface, cited
 * as file:line into the alamo source tree.  Plain pointers and sizes only.
 *
 * Conventions
 *   - A handle owns one multigrid hierarchy of a single-box, non-periodic nodal
 *     domain (amrlev 0 only).  mglev 0 is the finest level.
 *   - Node (i,j,k) of a level with n[d] cells sits at flat index
 *     i + (n0+1)*(j + (n1+1)*k); vector fields are 3 SoA planes of that size
 *     (component slowest, the amrex::Array4 order).
 *   - Pointers passed to entry points without a `_host` suffix are device
 *     pointers (HBM); `stream` is a hipStream_t (NULL = default stream).
 *   - Every function returns 0 on success and a nonzero ALAMO_ERR_* code on
 *     failure; alamo_last_error() gives the message.  No exceptions cross the ABI.
 *   - One handle per host thread; calls are stream-ordered.
 */
#ifndef ALAMO_AMD_H
#define ALAMO_AMD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Set::Sym subset (src/Set/Base.H) */
enum { ALAMO_SYM_ISOTROPIC = 0, ALAMO_SYM_MAJORMINOR = 1, ALAMO_SYM_DIAGONAL = 2 };
/* BC::Operator::Elastic::Elastic::Type, same numbering (src/BC/Operator/Elastic/Elastic.H:26) */
enum { ALAMO_BC_DISPLACEMENT = 0, ALAMO_BC_TRACTION = 1, ALAMO_BC_PERIODIC = 2, ALAMO_BC_NEUMANN = 3,
       ALAMO_BC_NONE = 4 };
/* stiffness layouts accepted by alamo_mg_set_model */
enum { ALAMO_LAYOUT_AOS = 0, /* BaseFab<Set::Matrix4<3,SYM>>: ncoef doubles per node */
       ALAMO_LAYOUT_SOA = 1  /* ncoef planes of (n0+1)(n1+1)(n2+1) doubles */ };
enum { ALAMO_PREC_FP64 = 0, ALAMO_PREC_MIXED = 1 /* FP32 cycles inside FP64 iterative refinement */ };
enum { ALAMO_CYCLE_V = 0, ALAMO_CYCLE_F = 1 };
/* solver.bottom_solver (Solver/Nonlocal/Linear.H:241,258,280-282) */
enum { ALAMO_BOTTOM_BICGSTAB = 0, /* Jacobi-preconditioned BiCGStab on the coarsest level (the default) */
       ALAMO_BOTTOM_SMOOTHER = 1, /* 8 Fsmooth calls from zero (amrex::MLMG's final_smooth default) */
       ALAMO_BOTTOM_DENSE = 2     /* direct: the coarsest operator assembled and inverted at prepare
                                     (at most ALAMO_DENSE_BOTTOM_MAX unknowns), one product per cycle; a
                                     singular coarsest level (2 cells with traction-free sides) is dropped
                                     while the next one fits (alamo_mg_num_levels reports the result) */ };
#define ALAMO_DENSE_BOTTOM_MAX 1200
#define ALAMO_NOALLOC_MAX_DOUBLES 8388608 /* 64 MiB: the device scratch of alamo_mg_fsmooth_noalloc */
enum { ALAMO_SMOOTHER_JACOBI = 0, ALAMO_SMOOTHER_L1 = 1 };
enum { ALAMO_COARSE_RESTRICT = 0, ALAMO_COARSE_GALERKIN = 1 };
enum { ALAMO_OK = 0, ALAMO_ERR_ARG = 1, ALAMO_ERR_HIP = 2, ALAMO_ERR_STATE = 3, ALAMO_ERR_NOCONV = 4,
       ALAMO_ERR_NAN = 5 };

typedef struct alamo_mg alamo_mg;
typedef struct alamo_net alamo_net;
typedef struct alamo_amr alamo_amr;

/* Geometry, boundary table and solver options of one hierarchy.  Field names
 * follow the reference's input keys (solver.* of Solver/Nonlocal/Linear.H:296-382,
 * bc.constant.* of BC/Operator/Elastic/Constant.H:439-648). */
typedef struct {
    int n_cell[3];             /* cells per direction at mglev 0 (amr.n_cell) */
    double dx[3];              /* cell size (Geometry::CellSize) */
    int sym;                   /* ALAMO_SYM_* */
    int bc_type[26][3];        /* per Face (Elastic.H:35-43) and component, ALAMO_BC_* */
    double bc_val[26][3];      /* prescribed values, written into b by alamo_bc_init */
    int max_coarsening_level;  /* -1: coarsen while every direction keeps >= 2 cells */
    int precision;             /* ALAMO_PREC_* */
    int cycle;                 /* ALAMO_CYCLE_*; the reference runs F-cycles (Linear.H:261) */
    int pre_smooth, post_smooth; /* Fsmooth calls per level (solver.pre_smooth/post_smooth) */
    int sweeps_per_smooth;     /* Jacobi sweeps per Fsmooth call (reference: 2, Operator.cpp:107) */
    double omega;              /* damping (solver.omega; Operator.H:123 default 2/3) */
    int max_iter;              /* solver.max_iter */
    int bottom_max_iter;       /* solver.bottom_max_iter */
    double bottom_tol_rel;     /* solver.bottom_tol_rel */
    int verbose;
    int smoother;              /* ALAMO_SMOOTHER_*: the sweeps of Fsmooth and of the cycles.  JACOBI:
                                  the reference's damped point Jacobi (Operator.cpp:109-146).  L1:
                                  x += (b - A x) / d with d_i = sign(a_ii) sum_j |a_ij| (l1-Jacobi, no
                                  damping), a smoother that stays stable on rows whose off-diagonal
                                  sum exceeds the diagonal, as the strong-form rows next to a sharp
                                  stiffness jump do; no reference counterpart, not with psi. */
    int coarse_op;             /* ALAMO_COARSE_*: the operators of the levels >= 1 of the cycles.
                                  RESTRICT: re-discretised from the restricted coefficients (the
                                  reference, Elastic.cpp:746-836).  GALERKIN: A_l = R A_{l-1} P with
                                  the cycle's own restriction and prolongation, a 27-point stencil of
                                  3 x 3 blocks per coarse node built at prepare; robust where the
                                  restricted coefficients misrepresent sharp stiffness jumps.  Needs
                                  bottom_solver dense or smoother; not with psi.  No reference
                                  counterpart.  Level 0 keeps the reference operator either way. */
    int fixed_iter;            /* solver.fixed_iter (Linear.H:181,267): > 0 runs exactly that many
                                  multigrid iterations and returns without a convergence failure
                                  (amrex::MLMG::setFixedIter); 0 off */
    int bottom_solver;         /* ALAMO_BOTTOM_*; the reference's "cg" choice is not built (the
                                  operator is not symmetric) */
    int max_fmg_iter;          /* solver.max_fmg_iter (Linear.H:261,264): with cycle F, the first
                                  max_fmg_iter iterations are F-cycles and the rest V-cycles
                                  (amrex::MLMG::setMaxFmgIter); the reference's 1e8 is the default */
    int sweeps_finest;         /* > 0: Jacobi sweeps per Fsmooth call of the cycles on level 0 (the
                                  other levels keep sweeps_per_smooth); 0: sweeps_per_smooth
                                  everywhere.  No reference counterpart (Fsmooth runs the same 2
                                  sweeps on every level, Operator.cpp:107); alamo_mg_fsmooth keeps
                                  sweeps_per_smooth. */
} alamo_mg_desc;

typedef struct {
    int iterations;            /* multigrid (or refinement) iterations performed */
    int converged;             /* 1 if |b - A x|_inf <= target */
    double initial_res_inf, final_res_inf, final_res_l2, rhs_inf, target;
    double seconds;            /* device-synchronised wall time of the solve */
    int fallback_events;       /* learned solve: plain V-cycle fallbacks taken */
} alamo_solve_stats;

const char* alamo_last_error(void);
int alamo_version(void);

/* Fill `d` with the reference defaults (Dirichlet-0 on all 26 regions, F-cycle,
 * 2 pre / 2 post smooths of 2 sweeps, omega 2/3, max_iter 200, bottom 200 / 1e-4). */
int alamo_mg_desc_default(alamo_mg_desc* d);

/* Operator::Elastic<SYM>::Elastic(geom, grids, dmap, info) + define()
 * (src/Operator/Elastic.cpp:9-55) together with Solver::Nonlocal::Linear::Define
 * (src/Solver/Nonlocal/Linear.H:36-43): allocates the hierarchy on the current device. */
int alamo_mg_create(const alamo_mg_desc* d, alamo_mg** h);
/* ~Elastic / Linear::Clear (Linear.H:44-51) */
int alamo_mg_destroy(alamo_mg* h);
int alamo_mg_set_stream(alamo_mg* h, void* stream);
int alamo_mg_num_levels(alamo_mg* h, int* nlev);
int alamo_mg_level_ncell(alamo_mg* h, int mglev, int* ncell3);
/* Node layout of level mglev: node (i,j,k) of component c is element i + pitch[0] j + pitch[1] k
 * + pitch[2] c; pitch[2] elements per component (the size of one component's array).  Levels are
 * compact: pitch = (N0, N0 N1, N0 N1 N2) with N = n_cell + 1. */
int alamo_mg_level_layout(alamo_mg* h, int mglev, int64_t* pitch3);

/* Elastic::SetModel(amrlev, MultiTab) (Elastic.cpp:87-116): copy the per-node stiffness of
 * mglev 0.  `C` is a device pointer (alamo_mg_set_model_host: host pointer). */
int alamo_mg_set_model(alamo_mg* h, int amrlev, const double* C, int layout);
int alamo_mg_set_model_host(alamo_mg* h, int amrlev, const double* C, int layout);
/* Polycrystal fast path: per-node grain id (uint16) and a grain table of ngrains
 * 21-slot MajorMinor records (host pointers).  Same operator as SetModel with
 * C(node) = table[gid[node]]; the finest level then reads 2 bytes per node of stiffness. */
int alamo_mg_set_model_grains_host(alamo_mg* h, const uint16_t* gid, const double* table, int ngrains);

/* Operator<Grid::Node>::prepareForSolve (Operator.cpp:253-260): coefficient restriction
 * to every mglev (Elastic::averageDownCoeffs, Elastic.cpp:605-626, 746-836) and the
 * diagonal (Operator::Diagonal(true)). */
/* Post-processing on the finest level (device fields, work queued on `stream`):
 * Elastic::Strain (src/Operator/Elastic.cpp:430-484): eps = (grad u + grad u^T)/2;
 * Elastic::Stress (:487-550): sigma = (C : grad u) psi_avg;
 * Elastic::Energy (:552-603): energy += (C : grad u) : eps / 2 (accumulates, as the reference).
 * GetStencil gradients (one-sided on the domain's faces).  Tensors: 9 planes row-major, or with
 * voigt = 1 the 6 planes 00, 11, 22, 12, 20, 01. */
int alamo_mg_strain(alamo_mg* h, int amrlev, const double* u, double* eps, int voigt, void* stream);
int alamo_mg_stress(alamo_mg* h, int amrlev, const double* u, double* sigma, int voigt, void* stream);
int alamo_mg_energy(alamo_mg* h, int amrlev, const double* u, double* energy, void* stream);
/* Solver::Nonlocal::Newton::DW / W (src/Solver/Nonlocal/Newton.H:528-611) for the affine model of
 * the handle's stiffness and eigenstrain F0 (Model/Solid/Affine/Isotropic.H:38-45, kinematic variable
 * grad u, GetStencil gradients): dw = C : (grad u - F0) as 9 row-major planes; w = (grad u - F0) : dw / 2
 * per node (written).  F0: 9 AoS doubles per node, or NULL for F0 = 0.  Device pointers. */
int alamo_newton_dw(alamo_mg* h, int amrlev, const double* u, const double* F0, double* dw, void* stream);
int alamo_newton_w(alamo_mg* h, int amrlev, const double* u, const double* F0, double* w, void* stream);
/* Elastic::SetUniform (src/Operator/Elastic.H:105, used at Elastic.cpp:245): nonzero drops the
 * grad(C) . grad(u) term of the interior rows on every level (the reference's default and the
 * integrators' setting is 0). */
int alamo_mg_set_uniform(alamo_mg* h, int uniform);

/* Integrator::Base::Mechanics::Integrate, trac_hi[0] (src/Integrator/Base/Mechanics.H:391-448, the
 * thermo output of tension tests): the x-hi face integral of sigma e_x, each face cell taking the mean
 * of its 4 nodes' sigma = C : (grad u - F0) times dy dz.  u, F0 (9 AoS per node, or NULL) device;
 * trac3 host; synchronous on the handle's stream. */
int alamo_mg_trac_xhi(alamo_mg* h, const double* u, const double* F0, double* trac3);

/* Integrator::Mechanics::UpdateModel blending (src/Integrator/Mechanics.H:242-254): per node
 * out = sum_n eta_n table_n over nmodels models of nrec values (a stiffness record: 2, 21 or 9;
 * an affine model's F0: 9).  eta: nmodels device planes of nnodes values; table: host,
 * nmodels x nrec; out: device, ALAMO_LAYOUT_AOS (nrec per node: SetModel's and
 * alamo_newton_rhs's F0 records) or ALAMO_LAYOUT_SOA (nrec planes). */
int alamo_blend_models(int64_t nnodes, int nmodels, int nrec, const double* eta, const double* table, double* out,
                       int layout, void* stream);

/* Linear::setMaxIter / setBottomMaxIter / setFixedIter / setVerbose / setPreSmooth /
 * setPostSmooth and Operator::SetOmega (Linear.H:178-184, Operator.H:57) on a built handle: copies
 * cycle, pre_smooth, post_smooth, sweeps_per_smooth, omega, max_iter, bottom_max_iter,
 * bottom_tol_rel, fixed_iter, verbose, bottom_solver, max_fmg_iter, sweeps_finest and smoother
 * from d; the fields that shape the hierarchy (n_cell, dx, sym, bc, precision,
 * max_coarsening_level) must be unchanged. */
int alamo_mg_update_opts(alamo_mg* h, const alamo_mg_desc* d);

/* Elastic::SetPsi(amrlev, psi, psi_small) (src/Operator/Elastic.H:55-57, Elastic.cpp:118-137): the
 * cell field psi (n_cell[0] n_cell[1] n_cell[2] doubles, i fastest).  From then on every row carries the psi terms of Elastic.cpp:191-195,243-260 and the
 * diagonal its psi_avg (:340-363); prepare restricts psi to the coarse levels (:839-907) and
 * alamo_newton_rhs uses the psi variant of Newton.H:234-266.  psi_small in [0, 1) (the
 * reference's default is 1e-8).  Device pointer; _host copies from host memory. */
int alamo_mg_set_psi(alamo_mg* h, int amrlev, const double* psi, double psi_small);
int alamo_mg_set_psi_host(alamo_mg* h, int amrlev, const double* psi, double psi_small);
int alamo_mg_prepare(alamo_mg* h);

/* Copy the stiffness of level mglev out (SoA planes, device pointer). */
int alamo_mg_get_coeff(alamo_mg* h, int mglev, double* C_soa);

/* Elastic::Fapply (Elastic.cpp:139-292): out = A in on level mglev. */
int alamo_mg_fapply(alamo_mg* h, int amrlev, int mglev, double* out, const double* in, void* stream);
/* Elastic::Diagonal (Elastic.cpp:296-379). */
int alamo_mg_diagonal(alamo_mg* h, int amrlev, int mglev, double* diag, void* stream);
/* Operator<Grid::Node>::normalize (Operator.cpp:156-165): x /= D on level mglev (the diagonal
 * computed by prepareForSolve; error if it has not been computed). */
int alamo_mg_normalize(alamo_mg* h, int amrlev, int mglev, double* x, void* stream);
/* Operator<Grid::Node>::Fsmooth (Operator.cpp:88-154): sweeps_per_smooth damped-Jacobi sweeps. */
int alamo_mg_fsmooth(alamo_mg* h, int amrlev, int mglev, double* x, const double* b, void* stream);
/* alamo_mg_fsmooth without any allocation: no reference counterpart.  For callers that cannot
 * allocate device memory (code under HIP graph capture, a serving loop, an external multigrid
 * driver calling the smoother once per level and cycle): the sweeps ping-pong through a static
 * device scratch array of ALAMO_NOALLOC_MAX_DOUBLES doubles (64 MiB) that every GPU the library
 * runs on holds until the process ends, instead of a buffer allocated per call.  Same kernels and
 * results as alamo_mg_fsmooth.  The handle must be prepared (alamo_mg_prepare) beforehand, since
 * preparing allocates; levels with 3 x nodes > ALAMO_NOALLOC_MAX_DOUBLES are rejected
 * (ALAMO_ERR_ARG).  All calls on one GPU share its scratch: calls that overlap on one GPU (other
 * streams, other threads) race. */
int alamo_mg_fsmooth_noalloc(alamo_mg* h, int amrlev, int mglev, double* x, const double* b, void* stream);
/* Elastic::Error0x (src/Operator/Elastic.cpp:382-404): R0x = x - A (x / D) on level mglev. */
int alamo_mg_error0x(alamo_mg* h, int amrlev, int mglev, double* R0x, const double* x, void* stream);
/* solutionResidual / correctionResidual (Operator.cpp:635-657): r = b - A x. */
int alamo_mg_residual(alamo_mg* h, int amrlev, int mglev, double* r, const double* x, const double* b,
                      void* stream);
/* Operator<Grid::Node>::restriction (Operator.cpp:262-363): crse (level cmglev) = R fine. */
int alamo_mg_restrict(alamo_mg* h, int amrlev, int cmglev, double* crse, const double* fine, void* stream);
/* Operator<Grid::Node>::interpolation (Operator.cpp:365-434): fine (level fmglev) += P crse. */
int alamo_mg_interpolate(alamo_mg* h, int amrlev, int fmglev, double* fine, const double* crse, void* stream);
/* max-norm and L2 norm of a level vector (MLMG's ResNormInf; 3 components). Synchronises. */
int alamo_mg_norms(alamo_mg* h, int mglev, const double* x, double* norm_inf, double* norm_l2, void* stream);
/* the two halves of alamo_mg_norms, under the names a caller binding MLMG's norm calls uses
 * (MultiFab::norm0 / norm2 over the 3 components of a level vector). */
int alamo_mg_norm_inf(alamo_mg* h, int mglev, const double* x, double* norm, void* stream);
int alamo_mg_norm_l2(alamo_mg* h, int mglev, const double* x, double* norm, void* stream);

/* BC::Operator::Elastic::Constant::Init (BC/Operator/Elastic/Constant.H:197-286):
 * write the prescribed boundary values into b (device pointer, level 0). */
int alamo_bc_init(alamo_mg* h, double* b, int homogeneous);
/* Newton<T>::prepareForSolve residual form (Solver/Nonlocal/Newton.H:144-231) for a
 * linear/affine model DW = C:(grad u - F0): rhs = b - div(DW) inside, b - BC(u, grad u, DW)
 * on the boundary.  F0 may be NULL (9 doubles per node, row-major, device pointer). */
int alamo_newton_rhs(alamo_mg* h, const double* u, const double* b, const double* F0, double* rhs);
/* Solver::Nonlocal::Newton<T>::solve (src/Solver/Nonlocal/Newton.H:279-358) for the affine model of
 * the handle's stiffness (set before) and eigenstrain F0 (9 AoS per node, or NULL): nriters times
 * rhs = alamo_newton_rhs(u, b, F0), du = alamo_linear_solve(du, rhs) (du zeroed once, so each later
 * solve starts from the previous du), u += du, stopping early when max |du| < nrtolerance.
 * *relnorm (may be NULL) receives that max |du| on an early stop, else 0, as the reference
 * returns.  stats (may be NULL): the last linear solve.  u, b device; scratch is allocated on the
 * handle's stream and freed before return. */
int alamo_newton_solve(alamo_mg* h, double* u, const double* b, const double* F0, int nriters, double nrtolerance,
                       double tol_rel, double tol_abs, alamo_solve_stats* stats, double* relnorm);

/* Solver::Nonlocal::Linear::solve (Linear.H:131-147) -> amrex::MLMG::solve: sol is the
 * initial guess on entry and the solution on exit (device pointers, level 0).  Stops when
 * |rhs - A sol|_inf <= max(tol_abs, tol_rel * max(|rhs|_inf, |r0|_inf)).  Returns
 * ALAMO_ERR_NOCONV if max_iter is reached (stats still filled). */
int alamo_linear_solve(alamo_mg* h, double* sol, const double* rhs, double tol_rel, double tol_abs,
                       alamo_solve_stats* stats);

/* Outer Krylov solve (no reference counterpart; the reference stops at amrex::MLMG).
 * Flexible GMRES(restart) in FP64, right-preconditioned by one multigrid cycle of the
 * handle's kind (ALAMO_PRECOND_MG: cycle, smoothing and precision of the descriptor, so a
 * mixed-precision handle preconditions with FP32 cycles) or by nothing.  The stopping test
 * is the reference's: |rhs - A sol|_inf of the FP64 true residual <= max(tol_abs,
 * tol_rel * max(|rhs|_inf, |r0|_inf)).  stats->iterations counts preconditioner applications. */
enum { ALAMO_PRECOND_NONE = 0, ALAMO_PRECOND_MG = 1, ALAMO_PRECOND_LEARNED = 2 };
typedef struct {
    int restart;               /* Krylov vectors per restart cycle (1..31) */
    int precond;               /* ALAMO_PRECOND_* */
    int max_iter;              /* total preconditioner applications */
    double stagnation_ratio;   /* learned preconditioner: fall back to the cycle when a restart
                                  reduces |r|_2 by less than this factor */
    int verbose;
} alamo_krylov_opts;
int alamo_krylov_opts_default(alamo_krylov_opts* o);
int alamo_krylov_solve(alamo_mg* h, double* sol, const double* rhs, double tol_rel, double tol_abs,
                       const alamo_krylov_opts* opts, alamo_solve_stats* stats);
/* The learned solve: alamo_krylov_solve with the learned preconditioner (the network attached by
 * alamo_mg_set_net), opts may be NULL (the defaults); opts->precond is ignored. */
int alamo_learned_solve(alamo_mg* h, double* sol, const double* rhs, double tol_rel, double tol_abs,
                        const alamo_krylov_opts* opts, alamo_solve_stats* stats);

/* Two-level AMR helpers of Operator<Grid::Node> (src/Operator/Operator.cpp).  The coarse level
 * is one box covering the non-periodic domain of ncell_crse[3] cells (nodes (n+1)^3, i fastest).
 * A fine box is an inclusive cell box in fine index space (refinement ratio 2; lo even, hi odd,
 * inside [0, 2 n - 1]); fine fields live on its nodal box grown by one node, [lo - 1, hi + 2]^3,
 * as 3 SoA planes (the valid nodes plus the one ghost layer the reflux stencil reads).
 * Device pointers; work is queued on `stream`. */
/* m_nd_fine_mask of the coarse level (read at Operator.cpp:236,553; built by
 * amrex::MLNodeLinOp::buildMasks): per coarse node 0 crse_node, 1 crse_fine_node, 2 fine_node;
 * fine_lo/fine_hi are host arrays of nbox (<= 64) boxes x 3 ints. */
int alamo_amr_fine_mask(const int* ncell_crse, int nbox, const int* fine_lo, const int* fine_hi, int* mask,
                        void* stream);
/* Operator<Grid::Node>::reflux (Operator.cpp:526-633) for one fine box: the coarse residual on the
 * box's crse_fine / fine nodes becomes the restricted fine residual. */
int alamo_amr_reflux(const int* ncell_crse, const int* fine_lo, const int* fine_hi, const int* mask, double* crse_res,
                     const double* fine_res, void* stream);
/* Operator<Grid::Node>::reflux under its operator name: alamo_amr_reflux for the coarse level of
 * the handle (ncell_crse = the handle's n_cell at mglev 0). */
int alamo_mg_reflux(alamo_mg* h, const int* fine_lo, const int* fine_hi, const int* mask, double* crse_res,
                    const double* fine_res, void* stream);
/* Operator<Grid::Node>::averageDownSolutionRHS (Operator.cpp:435-447): nodal injection of the fine
 * box's solution onto the coarse nodes it covers. */
int alamo_amr_average_down(const int* ncell_crse, const int* fine_lo, const int* fine_hi, double* crse_sol,
                           const double* fine_sol, void* stream);
/* Operator<Grid::Node>::fixUpResidualMask (Operator.cpp:230-251): resmsk = 1 on crse_fine nodes. */
int alamo_amr_fixup_residual_mask(const int* ncell_crse, const int* mask, int* resmsk, void* stream);

/* Learned correction network (no reference counterpart).  A 3-level convolutional network
 * (16/32/64 channels, 3x3x3 kernels, BF16 weights and activations, FP32 accumulation on the
 * matrix cores) that maps the residual left by one multigrid cycle, together with log(lambda),
 * log(mu), to that cycle's remaining error.  Used by alamo_krylov_solve with
 * ALAMO_PRECOND_LEARNED: z = cycle(r); z += net(r - A z); the plain cycle takes over for good
 * when a restart reduces |r|_2 by less than opts.stagnation_ratio.  Isotropic operators only;
 * level-0 grids need (n+1-1) divisible by 4. */
typedef struct {
    int npairs;                /* (residual, error) pairs generated with the handle's operator */
    int steps;                 /* Adam steps (one pair per step) */
    double lr;                 /* Adam learning rate */
    uint64_t seed;             /* training-set seed (stored in checkpoints) */
    int solve_max_iter;        /* FGMRES iterations allowed for each pair's error */
    double solve_tol;          /* relative tolerance of those solves (ALAMO_TRAIN_LOSS_ERROR) */
    int verbose;
    int loss;                  /* ALAMO_TRAIN_LOSS_* (0, the error loss, when zero-initialised) */
} alamo_train_opts;
/* ALAMO_TRAIN_LOSS_ERROR: mean |y - e / rms(e)|^2 against e = A^-1 q, solved per pair by FGMRES.
 * ALAMO_TRAIN_LOSS_RESIDUAL: 1 - <q, A y>^2 / (|q|^2 |A y|^2), the squared residual ratio
 * |q - alpha A y| / |q| the stage reaches with its residual-minimising alpha; needs no solves,
 * back-propagated through A^T (alamo_mg_apply_transpose's row table; no psi terms). */
enum { ALAMO_TRAIN_LOSS_ERROR = 0, ALAMO_TRAIN_LOSS_RESIDUAL = 1 };
typedef struct {
    int pairs, pairs_converged, steps;
    double first_loss, last_loss, seconds_data, seconds_train;
} alamo_train_stats;
int alamo_net_create(alamo_net** net, uint64_t seed);
int alamo_net_destroy(alamo_net* net);
int alamo_net_num_params(alamo_net* net, int64_t* nparam);
int alamo_net_get_params(alamo_net* net, float* host);
int alamo_net_set_params(alamo_net* net, const float* host);
int alamo_net_get_grads(alamo_net* net, float* host);
int alamo_net_zero_grads(alamo_net* net);
/* x, y, dy, dx: host [V][16] float arrays on a grid of nnodes[3] nodes (inputs rounded to bf16) */
int alamo_net_forward_host(alamo_net* net, const int* nnodes, const float* x, float* y);
int alamo_net_backward_host(alamo_net* net, const int* nnodes, const float* x, const float* dy, float* dx);
/* one layer's weight gradient alone: dW[27][cout][cin] += sum_v dout[v][co] in[v + o][ci] (zero
 * outside the grid); in [V][cin], dout [V][cout] bf16 bit patterns, host arrays */
int alamo_net_wgrad_host(alamo_net* net, const int* nnodes, int cin, int cout, const uint16_t* in,
                         const uint16_t* dout, float* dW);
int alamo_net_adam_step(alamo_net* net, double lr);
int alamo_net_train(alamo_net* net, alamo_mg* h, const alamo_train_opts* opts, alamo_train_stats* stats);
int alamo_net_save(alamo_net* net, const char* path);
int alamo_net_load(alamo_net* net, const char* path);
int alamo_mg_set_net(alamo_mg* h, alamo_net* net);
/* z = P(r): one application of the preconditioner kind (ALAMO_PRECOND_MG: one multigrid cycle
 * from zero, the operator's MLMG settings; ALAMO_PRECOND_LEARNED: that cycle plus the network
 * stage), device pointers, synchronous */
int alamo_mg_precond(alamo_mg* h, int kind, const double* r, double* z);
/* out = A^T r for the level-0 operator (no reference counterpart: the reference applies A only).
 * Builds an explicit table of A's rows (243 doubles per node, kept with the handle; no psi terms)
 * from the current coefficients and boundary table on every call, then gathers.  Device vectors
 * of 3 planes, r and out distinct; synchronous.  Used by ALAMO_TRAIN_LOSS_RESIDUAL. */
int alamo_mg_apply_transpose(alamo_mg* h, const double* r, double* out);
/* mean forward-kernel time of each of the 6 convolution layers at h's level-0 size, HIP events on
 * h's stream, reps launches each: ms[6] (for throughput and roofline reports) */
int alamo_net_layer_times(alamo_net* net, alamo_mg* h, int reps, double* ms);
/* mean time of each layer's backward-data convolution and weight gradient (the two-pass reduction
 * included, the bias gradient not) at h's level-0 size, HIP events, reps launches each: ms[6] each */
int alamo_net_layer_times_bwd(alamo_net* net, alamo_mg* h, int reps, double* ms_data, double* ms_wgrad);
/* one 16x16x16 bf16 matrix-core product of row-major host matrices (lane-layout check) */
int alamo_net_mfma_probe(const uint16_t* A, const uint16_t* B, float* D);
/* one 16x16x32 bf16 product (gfx950): A 16 x 32, B 32 x 16 row-major host matrices, D 16 x 16 */
int alamo_net_mfma_probe32(const uint16_t* A, const uint16_t* B, float* D);

/* Attainable-bandwidth calibration for roofline reports (no counterpart in the reference): o[v n + q]
 * = a[v n + q] a[3 n + q] + a[4 n + q], v = 0..2, for q < n -- the isotropic FP64 Fapply's five
 * reads and three writes per node as a plain stream.  a: 5 n doubles, o: 3 n doubles (device). */
int alamo_stream_probe(const double* a, double* o, int64_t n, void* stream);

/* ---------------------------------------------------------------------------------------------
 * Composite multi-level (AMR) solve: Solver::Nonlocal::Linear::solve over several AMR levels
 * (Linear.H:131-168 with amrex::MLMG's composite cycle) on an explicit mesh
 * (Integrator.cpp:121-139: explicitmesh.lo<l> / hi<l>).  Level l has n_cell = base n_cell 2^l and
 * dx = base dx / 2^l; nbox[l - 1] boxes for each level l >= 1, inclusive cell boxes in that level's
 * own index space, concatenated in box_lo / box_hi (3 ints each), covering whole coarse cells (lo
 * even, hi odd, at most 64 per level), one coarse cell inside the domain and properly nested (each
 * box, coarsened and grown by one cell, inside the union of the next coarser level's boxes).
 * Every level is a full-domain field at its resolution (3 SoA planes of
 * (n_cell + 1)^3 nodes); a level's values outside its region follow the coarser level.  The
 * composite equations and the V-cycle are described in the library source (capi.hip).  Options
 * (cycle, smoothing, smoother, tolerances, max_iter) come from `base`.
 * --------------------------------------------------------------------------------------------- */
int alamo_amr_create(const alamo_mg_desc* base, int nlevels, const int* nbox, const int* box_lo, const int* box_hi,
                     alamo_amr** amr);
int alamo_amr_destroy(alamo_amr* amr);
/* the operator handle of AMR level amrlev (Elastic::SetModel(amrlev, ...), alamo_bc_init,
 * alamo_newton_rhs, Fapply, post-processing all work on it) */
int alamo_amr_level(alamo_amr* amr, int amrlev, alamo_mg** h);
/* device int arrays of the level's nodes: region (2 inside the level's region, 1 on its boundary,
 * 0 outside; level 0: all 2) and coverage by the next level (the nodal fine mask: 2 fine, 1
 * coarse-fine, 0 coarse; the finest level: all 0); either may be null */
int alamo_amr_masks(alamo_amr* amr, int amrlev, int* region, int* fine_mask, void* stream);
/* sol[l] (initial guess in, solution out) and rhs[l]: device fields of every level.  Stops when the
 * composite residual's max norm is <= max(tol_abs, tol_rel max(|b|_inf, |r0|_inf)); covered coarse
 * nodes receive the fine solution (averageDownSolutionRHS, Operator.cpp:435-447). */
int alamo_amr_solve(alamo_amr* amr, double* const* sol, const double* const* rhs, double tol_rel, double tol_abs,
                    alamo_solve_stats* stats);

/* AMReX plotfile of nodal fields: Integrator::WritePlotFile's "<prefix>node" output
 * (src/Integrator/Integrator.cpp:808-1054, through amrex::WriteMultiLevelPlotfile), in the
 * HyperCLaw-V1.1 layout: <dir>/Header, <dir>/Level_<l>/Cell_H and Cell_D_00000 (one I/O rank), and
 * the reference's <dir>/Checkpoint list of the level box arrays.  Level l has n_cell0 2^l cells
 * per direction on [prob_lo, prob_hi]; nbox[l] inclusive cell boxes of level l in its own index
 * space, concatenated in box_lo / box_hi (3 ints each; nbox[0] = 0: the whole domain), written as
 * nodal boxes.  fields[l]: device array of ncomp SoA planes of the level's (n + 1)^3 nodes (the
 * full-domain levels of alamo_amr_level); names[ncomp] the component names; level_steps[l] the
 * step numbers recorded per level.  Creates `dir` and its Level_<l> subdirectories. */
int alamo_write_plotfile(const char* dir, int nlevels, const int* n_cell0, const double* prob_lo,
                         const double* prob_hi, const int* nbox, const int* box_lo, const int* box_hi, int ncomp,
                         const char* const* names, const double* const* fields, double time, const int* level_steps);
/* the same with host arrays fields[l] */
int alamo_write_plotfile_host(const char* dir, int nlevels, const int* n_cell0, const double* prob_lo,
                              const double* prob_hi, const int* nbox, const int* box_lo, const int* box_hi, int ncomp,
                              const char* const* names, const double* const* fields, double time,
                              const int* level_steps);

#ifdef __cplusplus
}
#endif
#endif
