// Drives the library through the C++ façade exactly as Integrator::Base::Mechanics drives the
// reference (Base/Mechanics.H:193-208): build the operator, SetBC, SetModel, Define, solve, Clear.
// Writes the solution (3 SoA planes, float64) to argv[1] and prints iterations and |r|inf; then
// solves again from zero with the Krylov class (FGMRES around the same cycle) into argv[2], and
// creates a network to check the Net class.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

#include "alamo_amd.hpp"

int main(int argc, char** argv) {
    if (argc < 3) { fprintf(stderr, "usage: facade_solve linear.bin krylov.bin\n"); return 2; }
    const int n = 16;
    const int N = (n + 1) * (n + 1) * (n + 1);
    BC::Operator::Elastic::Constant bc;  // Dirichlet 0 everywhere (Constant.H:36-45)
    bc.Set(BC::Operator::Elastic::Constant::XHI, BC::Operator::Elastic::Constant::X,
           BC::Operator::Elastic::Constant::Displacement, 0.01);
    Operator::Elastic<Operator::Isotropic> op({n, n, n}, {1.0 / n, 1.0 / n, 1.0 / n});
    op.SetBC(&bc);
    std::vector<double> C(2 * N), b(3 * N, 0.0), x(3 * N, 0.0);
    for (int m = 0; m < N; m++) {
        const int i = m % (n + 1);
        C[2 * m] = 1.0 + 0.5 * i / n;  // lambda
        C[2 * m + 1] = 1.0;            // mu
    }
    double *dC, *db, *dx;
    if (hipMalloc(&dC, C.size() * 8) || hipMalloc(&db, b.size() * 8) || hipMalloc(&dx, x.size() * 8)) return 3;
    hipMemcpy(dC, C.data(), C.size() * 8, hipMemcpyHostToDevice);
    hipMemcpy(dx, x.data(), x.size() * 8, hipMemcpyHostToDevice);
    try {
        op.SetModel(0, dC);
        alamo::check(alamo_bc_init(op.handle(), db, 0));  // BC::Operator::Elastic::Constant::Init
        Solver::Nonlocal::Linear<Operator::Isotropic> solver;
        solver.Define(op);
        // options set after the operator is built take effect (Linear::setFixedIter, setMaxIter)
        solver.setFixedIter(2);
        solver.solve(dx, db, 1e-12, 0.0);
        printf("fixed iterations %d\n", solver.stats().iterations);
        solver.setFixedIter(0);
        solver.setMaxIter(300);
        hipMemset(dx, 0, x.size() * 8);
        double r = solver.solve(dx, db, 1e-12, 0.0);
        printf("iterations %d res_inf %.3e converged %d\n", solver.stats().iterations, r, solver.stats().converged);
        solver.Clear();
        hipMemcpy(x.data(), dx, x.size() * 8, hipMemcpyDeviceToHost);
        FILE* f = fopen(argv[1], "wb");
        fwrite(x.data(), 8, x.size(), f);
        fclose(f);

        hipMemset(dx, 0, x.size() * 8);
        Solver::Nonlocal::Krylov<Operator::Isotropic> krylov(op);
        krylov.setRestart(16);
        krylov.setPrecond(Solver::Nonlocal::Krylov<Operator::Isotropic>::Cycle);
        r = krylov.solve(dx, db, 1e-12, 0.0);
        printf("krylov iterations %d res_inf %.3e converged %d\n", krylov.stats().iterations, r,
               krylov.stats().converged);
        krylov.Clear();
        // Newton from u = 0 with F0 = 0: its right-hand side is b, so it repeats the linear solve
        double* du;
        if (hipMalloc(&du, x.size() * 8)) return 3;
        hipMemset(du, 0, x.size() * 8);
        Solver::Nonlocal::Newton<Operator::Isotropic> newton(op);
        newton.setMaxIter(300);
        newton.solve(du, db, dC, nullptr, 1e-12, 0.0);
        std::vector<double> xn(x.size());
        hipMemcpy(xn.data(), du, x.size() * 8, hipMemcpyDeviceToHost);
        double dmax = 0.0;
        for (size_t q = 0; q < x.size(); q++) dmax = std::max(dmax, std::fabs(xn[q] - x[q]));
        printf("newton iterations %d max diff to linear %.3e\n", newton.stats().iterations, dmax);
        newton.Clear();
        hipFree(du);
        printf("ngrow %d ncomp %d singular %d\n", op.getNGrow(), op.getNComp(), (int)op.isSingular(0));
        alamo::Net net(7);
        printf("net params %lld\n", (long long)net.NumParams());
    } catch (const alamo::Abort& e) {
        fprintf(stderr, "abort: %s\n", e.what());
        return 1;
    }
    hipMemcpy(x.data(), dx, x.size() * 8, hipMemcpyDeviceToHost);
    FILE* f = fopen(argv[2], "wb");
    fwrite(x.data(), 8, x.size(), f);
    fclose(f);
    return 0;
}
