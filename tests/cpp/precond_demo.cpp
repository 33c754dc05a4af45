// Newton-GMRES solves of F(x) = Phi_T(x) - x with and without the diffusion
// preconditioner (include/stratiflow/DiffusionPreconditioner.h), on the
// held-out problems of learned_demo.cpp (64x1x64, T = 100 steps of 0.02).
// FP32 time stepping, FP64 Krylov algebra.
//   argv: ntest K tol Re delta kmax advect [newton]
// Prints, one per line:
//   test <problem> <plain matvecs> <plain s> <precond matvecs> <precond s> <plain res> <precond res>
//        <|dx_precond - dx_plain| / |dx_plain|>
//   total <plain matvecs> <precond matvecs> <plain converged> <precond converged>
// With newton > 0, full Newton iterations instead (at most `newton` steps, GMRES
// tolerance tol per step) until |F(x)| <= 1e-8 |F(x0)|:
//   newton <problem> <plain steps> <plain matvecs> <plain s> <plain |F|/|F0|>
//          <precond steps> <precond matvecs> <precond s> <precond |F|/|F0|>
//   newton_total <plain s> <precond s> <plain converged> <precond converged>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "stratiflow/DiffusionPreconditioner.h"
#include "stratiflow/NewtonKrylov.h"

class FixedNewton : public NewtonKrylov<StateVector> {
public:
    stratifloat dt = stratifloat(0.02);
    int nsteps = 100;

protected:
    StateVector EvalFunction(const StateVector& at) override {
        StateVector s = at;
        std::vector<StateVector> traj;
        s.FixedEvolve(dt, nsteps + 1, traj);
        StateVector r = traj[nsteps];
        r -= at;
        return r;
    }
};

static double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static StateVector problem(unsigned seed) {
    SeedRandom(seed);
    StateVector x;
    x.ExciteLowWavenumbers(stratifloat(1e-3 * (1.0 + 0.5 * (seed % 3))));
    return x;
}

int main(int argc, char** argv) {
    setvbuf(stdout, nullptr, _IOLBF, 0);
    const int ntest = argc > 1 ? std::atoi(argv[1]) : 4;
    const int K = argc > 2 ? std::atoi(argv[2]) : 61;
    const double tol = argc > 3 ? std::atof(argv[3]) : 2e-2;
    gridParams = {64, 1, 64, Dimensionality::TwoDimensional};
    flowParams.L1 = stratifloat(2 * 3.14159265358979 / 0.44);
    flowParams.L2 = 1;
    flowParams.L3 = 10;
    flowParams.Re = argc > 4 ? std::atof(argv[4]) : 100;
    flowParams.Ri = stratifloat(0.1);
    flowParams.Pr = 1;
    const double delta = argc > 5 ? std::atof(argv[5]) : 1e-3;
    const double kmax = argc > 6 ? std::atof(argv[6]) : 0.0;
    const bool advect = argc > 7 && std::atoi(argv[7]) != 0;
    StateVector::ResetForParams();
    StateVector::solver().SetBackground(InitialU);

    typedef WithDiffusionPreconditioner<FixedNewton> Pre;
    Pre nk(DiffusionPreconditioner(0.02 * 100, delta, kmax, advect));
    nk.verbose = false;
    nk.SetKrylovDimension(K);
    nk.gmresTolerance = tol;
    nk.maxNewtonSteps = 1;

    const int newton = argc > 8 ? std::atoi(argv[8]) : 0;
    if (newton > 0) {
        double sp = 0, sl = 0;
        int okp = 0, okl = 0;
        for (int i = 0; i < ntest; ++i) {
            const StateVector x0 = problem(1000 + i);
            double r0 = 0;
            {
                StateVector y = x0;
                FixedNewton probe;
                probe.verbose = false;
                probe.SetKrylovDimension(2);
                probe.maxNewtonSteps = 0;
                probe.Run(y);  // residual evaluation only
                r0 = probe.residualHistory.front();
            }
            nk.targetResidual = 1e-8 * r0;
            nk.maxNewtonSteps = newton;
            double res[2], sec[2];
            int steps[2];
            long mv[2];
            for (int pass = 0; pass < 2; ++pass) {
                StateVector x = x0;
                nk.enabled = pass == 1;
                const long m0 = nk.matVecs;
                const double t = now();
                nk.Run(x);
                sec[pass] = now() - t;
                mv[pass] = nk.matVecs - m0;
                steps[pass] = int(nk.residualHistory.size()) - 1;
                res[pass] = nk.residualHistory.back() / r0;
            }
            printf("newton %d %d %ld %.4f %.3e %d %ld %.4f %.3e\n", i, steps[0], mv[0], sec[0], res[0], steps[1], mv[1],
                   sec[1], res[1]);
            sp += sec[0];
            sl += sec[1];
            okp += res[0] <= 1e-8;
            okl += res[1] <= 1e-8;
        }
        printf("newton_total %.4f %.4f %d %d\n", sp, sl, okp, okl);
        return 0;
    }

    long tp_mv = 0, tl_mv = 0;
    int cp = 0, cl = 0;
    for (int i = 0; i < ntest; ++i) {
        const StateVector x0 = problem(1000 + i);
        StateVector xp = x0, xl = x0;
        nk.enabled = false;
        long m0 = nk.matVecs;
        double t = now();
        nk.Run(xp);
        const double tp = now() - t;
        const long mp = nk.matVecs - m0;
        const double rp = nk.lastGmresResidual;
        nk.enabled = true;
        m0 = nk.matVecs;
        t = now();
        nk.Run(xl);
        const double tl = now() - t;
        const long ml = nk.matVecs - m0;
        const double rl = nk.lastGmresResidual;
        StateVector dp = xp - x0, dl = xl - xp;
        printf("test %d %ld %.4f %ld %.4f %.3e %.3e %.3e\n", i, mp, tp, ml, tl, rp, rl,
               sfkrylov::norm(dl) / sfkrylov::norm(dp));
        tp_mv += mp;
        tl_mv += ml;
        cp += rp < tol;
        cl += rl < tol;
    }
    printf("total %ld %ld %d %d\n", tp_mv, tl_mv, cp, cl);
    return 0;
}
