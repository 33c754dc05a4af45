// Trains the learned preconditioner (include/stratiflow/LearnedPreconditioner.h)
// on the operator applications of plain GMRES solves, then solves held-out
// problems with and without it (argv[7] > 0: dilated layers 1, 2, 4, 8;
// argv[8]: rounds of retraining on the directions of preconditioned solves;
// argv[9] > 0: the network corrects the advected diffusion preconditioner with
// that delta instead of the identity; argv[10] = n > 0: training solves and
// tests are full Newton runs of up to n steps, the tests to |F| <= 1e-8 |F(x0)|,
// printed as  newton <problem> <plain steps> <plain matvecs> <plain s> <plain |F|/|F0|>
// <learned steps> <learned matvecs> <learned s> <learned |F|/|F0|>).
// FP32 time stepping (the facade's default
// precision), FP64 Krylov algebra.  Prints, one per line:
//   train <problem> <gmres iterations> <seconds>
//   loss <first epoch> <last epoch> <samples> <train seconds>
//   test <problem> <plain matvecs> <plain s> <learned matvecs> <learned s> <plain res> <learned res>
//        <rel diff of the Newton updates> <fallbacks so far>
//   hist <problem> <GMRES residual history of the plain solve>
//   learned <problem> <Krylov dimension> <residual> of the preconditioned attempt
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "stratiflow/LearnedPreconditioner.h"
#include "stratiflow/NewtonKrylov.h"

// F(x) = Phi(x) - x, Phi = nsteps fixed-dt steps (FixedEvolve start-up)
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

typedef WithLearnedPreconditioner<FixedNewton> Learned;

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
    setvbuf(stdout, nullptr, _IOLBF, 0);  // one line at a time, also into files
    const int ntrain = argc > 1 ? std::atoi(argv[1]) : 6;
    const int ntest = argc > 2 ? std::atoi(argv[2]) : 3;
    const int K = argc > 3 ? std::atoi(argv[3]) : 61;
    const double tol = argc > 4 ? std::atof(argv[4]) : 1e-3;
    const int epochs = argc > 5 ? std::atoi(argv[5]) : 150;
    gridParams = {64, 1, 64, Dimensionality::TwoDimensional};
    flowParams.L1 = stratifloat(2 * 3.14159265358979 / 0.44);
    flowParams.L2 = 1;
    flowParams.L3 = 10;
    flowParams.Re = argc > 6 ? std::atof(argv[6]) : 100;
    flowParams.Ri = stratifloat(0.1);
    flowParams.Pr = 1;
    StateVector::ResetForParams();
    StateVector::solver().SetBackground(InitialU);

    LearnedPreconditioner pre(64, 4, 1);
    if (argc > 7 && std::atoi(argv[7]) > 0) pre.SetDilations({1, 2, 4, 8});
    const double bdelta = argc > 9 ? std::atof(argv[9]) : 0.0;
    const DiffusionPreconditioner base(0.02 * 100, bdelta > 0 ? bdelta : 1.0, 1e9, true);
    if (bdelta > 0) pre.SetBase(&base);
    Learned nk(pre);
    nk.verbose = false;
    nk.SetKrylovDimension(K);
    nk.gmresTolerance = tol;
    nk.maxNewtonSteps = 1;
    const int newton = argc > 10 ? std::atoi(argv[10]) : 0;
    if (newton > 0) {
        nk.maxNewtonSteps = newton;
        nk.targetResidual = 0;
        nk.fallback = false;
    }

    // harvest (A x, x) from plain solves
    nk.harvest = true;
    nk.enabled = false;
    for (int i = 0; i < ntrain; ++i) {
        StateVector x = problem(100 + i);
        const double t0 = now();
        nk.Run(x);
        printf("train %d %d %.4f\n", i, nk.lastKrylovDimension, now() - t0);
    }
    nk.harvest = false;
    double l0 = 0, l1 = 0;
    const double t0 = now();
    pre.Train(epochs, 64, 1e-3, &l0, &l1);
    printf("loss %.6e %.6e %d %.3f\n", l0, l1, pre.Samples(), now() - t0);
    // further rounds: harvest the directions of preconditioned solves of the
    // training problems (the inputs the network meets in use), then retrain
    const int rounds = argc > 8 ? std::atoi(argv[8]) : 0;
    for (int r = 0; r < rounds; ++r) {
        nk.harvest = true;
        nk.enabled = true;
        for (int i = 0; i < ntrain; ++i) {
            StateVector x = problem(100 + i);
            nk.Run(x);
        }
        nk.harvest = false;
        const double t1 = now();
        pre.Train(epochs, 64, 1e-3, &l0, &l1);
        printf("loss %.6e %.6e %d %.3f\n", l0, l1, pre.Samples(), now() - t1);
    }

    if (newton > 0) {
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
        }
        return 0;
    }
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
        printf("hist %d", i);
        for (double r : nk.gmresResidualHistory) printf(" %.3e", r);
        printf("\n");
        nk.enabled = true;
        m0 = nk.matVecs;
        t = now();
        nk.Run(xl);
        const double tl = now() - t;
        const long ml = nk.matVecs - m0;
        const double rl = nk.lastGmresResidual;
        StateVector dp = xp - x0, dl = xl - xp;
        printf("test %d %ld %.4f %ld %.4f %.3e %.3e %.3e %d\n", i, mp, tp, ml, tl, rp, rl,
               sfkrylov::norm(dl) / sfkrylov::norm(dp), nk.fallbacks);
        printf("learned %d %d %.3e\n", i, nk.lastLearnedIterations, nk.lastLearnedResidual);
    }
    return 0;
}
