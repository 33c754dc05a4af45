// Newton-GMRES to a relative residual of 1e-8 on the learned-preconditioner
// problem class: 2D N1 x 1 x N3 stratified shear layers (U = tanh z, background
// buoyancy tanh z), Re drawn uniformly from [1000, 3000] and Ri from
// [0.08, 0.2] per problem, a seeded perturbation of energy `amp`, and
// F(x) = Phi_T(x) - x with Phi_T = T/dt fixed steps after FixedEvolve's
// start-up (StateVector.cpp:85-105).  FP64 time stepping when built with
// USE_DOUBLE (FP32 stepping leaves |Phi(x) - x| a floor near 1e-7 |x|), FP64
// Krylov algebra either way.
//
// Training problems are solved with plain GMRES while every operator
// application (x, A x) is harvested; the network (4 x Conv1D, 64 channels,
// kernel 5, dilations 1, 2, 4, 8, BF16 on the matrix cores) is trained on the
// GPU as a correction of the analytic advected-diffusion preconditioner.
// Each held-out problem is then solved three times from the same start:
//   plain    GMRES without preconditioner (the reference's method),
//   base     the analytic preconditioner alone,
//   learned  the network on the analytic base, judged by plain GMRES (a
//            preconditioned solve that misses the tolerance is redone plain).
// usage: newton_bench key=value ...  (N1 N3 T dt K tol newton ntrain ntest
//        batches delta amp seed methods first; methods is a mask: 1 plain,
//        2 learned, 4 base; the held-out problems are first .. first+ntest-1,
//        so a large set can be measured in parts)
// Prints one line each:
//   train <i> <Re> <Ri> <newton steps> <matvecs> <seconds> <samples so far> <|F|/|F0|>
//   epoch <e> <loss> <seconds>
//   loss <first> <last> <samples> <seconds>
//   test <i> <Re> <Ri> <method> <newton steps> <matvecs> <seconds> <|F|/|F0|> <fallbacks> <|x - x_plain|/|x0|>
//   hist <i> <method> <|F|/|F0|>@<seconds since the solve began>, at every Newton step
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <string>
#include <vector>

#include "stratiflow/DiffusionPreconditioner.h"
#include "stratiflow/LearnedPreconditioner.h"
#include "stratiflow/NewtonKrylov.h"

static double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Phi_T by direct stepping (no trajectory kept)
static StateVector Evolve(const StateVector& at, stratifloat dt, int nsteps) {
    IMEXRK& sv = StateVector::solver();
    sv.u1 = at.u1;
    sv.u2.Zero();
    sv.u3 = at.u3;
    sv.b = at.b;
    sv.p = at.p;
    sv.FilterAll();
    sv.PopulateNodalVariables();
    sv.RemoveDivergence(0.0f);
    sv.deltaT = dt;
    sv.UpdateForTimestep();
    sv.PrepareRun("", false);
    sv.TimeStep();  // synchronises dt and the state with the context
    sfacade::check(sf_step(sfacade::ctx(), nsteps - 1), sfacade::ctx());
    StateVector r;
    r.u1 = sv.u1;
    r.u3 = sv.u3;
    r.b = sv.b;
    r.EnforceBCs();
    return r;
}

class PhiNewton : public NewtonKrylov<StateVector> {
public:
    stratifloat dt = stratifloat(0.01);
    int nsteps = 200;

protected:
    StateVector EvalFunction(const StateVector& at) override {
        StateVector r = Evolve(at, dt, nsteps);
        r -= at;
        return r;
    }
};

struct Problem {
    double Re, Ri;
    StateVector x0;
};

static Problem make_problem(unsigned seed, double amp) {
    std::mt19937_64 g(seed);
    std::uniform_real_distribution<double> re(1000.0, 3000.0), ri(0.08, 0.2);
    Problem p;
    p.Re = re(g);
    p.Ri = ri(g);
    flowParams.Re = stratifloat(p.Re);
    flowParams.Ri = stratifloat(p.Ri);
    StateVector::ResetForParams();
    SeedRandom(seed);
    p.x0.ExciteLowWavenumbers(stratifloat(amp));
    return p;
}

static void use(const Problem& p) {
    flowParams.Re = stratifloat(p.Re);
    flowParams.Ri = stratifloat(p.Ri);
    StateVector::ResetForParams();
}

int main(int argc, char** argv) {
    setvbuf(stdout, nullptr, _IOLBF, 0);
    std::map<std::string, double> o = {{"N1", 256}, {"N3", 256}, {"T", 2}, {"dt", 0.01}, {"K", 61}, {"tol", 1e-2},
                                       {"newton", 12}, {"ntrain", 3}, {"ntest", 3}, {"batches", 4000},
                                       {"delta", 0.3}, {"amp", 1e-6}, {"seed", 7000}, {"methods", 7},
                                       {"first", 0}};
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        const size_t e = a.find('=');
        if (e == std::string::npos || !o.count(a.substr(0, e))) {
            fprintf(stderr, "unknown argument %s\n", argv[i]);
            return 2;
        }
        o[a.substr(0, e)] = std::atof(a.c_str() + e + 1);
    }
    gridParams = {int(o["N1"]), 1, int(o["N3"]), Dimensionality::TwoDimensional};
    flowParams.L1 = stratifloat(2 * 3.14159265358979 / 0.44);
    flowParams.L2 = 1;
    flowParams.L3 = 10;
    flowParams.Re = 2000;
    flowParams.Ri = stratifloat(0.12);
    flowParams.Pr = 1;
    StateVector::ResetForParams();
    StateVector::solver().SetBackground(InitialU);
    StateVector::solver().SetBackgroundStratification([](stratifloat z) {
        const stratifloat c = std::cosh(z);
        return 1 / (c * c);
    });
    const double T = o["T"];
    const int nsteps = int(std::lround(T / o["dt"]));
    const unsigned seed = unsigned(o["seed"]);

    const DiffusionPreconditioner base(T, o["delta"], 1e9, true);
    LearnedPreconditioner net(64, 4, 1);
    net.SetDilations({1, 2, 4, 8});
    net.SetBase(&base);
    typedef WithLearnedPreconditioner<PhiNewton> Learned;
    typedef WithDiffusionPreconditioner<PhiNewton> Based;
    auto configure = [&](PhiNewton& nk) {
        nk.verbose = false;
        nk.dt = stratifloat(o["dt"]);
        nk.nsteps = nsteps;
        nk.SetKrylovDimension(int(o["K"]));
        nk.gmresTolerance = o["tol"];
        nk.maxNewtonSteps = int(o["newton"]);
    };
    Learned nl(net);
    configure(nl);
    Based nb(base);
    configure(nb);

    // residual of x for the current problem
    auto residual = [&](const StateVector& x) {
        PhiNewton probe;
        configure(probe);
        probe.SetKrylovDimension(2);
        probe.maxNewtonSteps = 0;
        StateVector y = x;
        probe.Run(y);
        return probe.residualHistory.front();
    };

    // training: plain Newton runs, every operator application harvested
    nl.harvest = true;
    nl.enabled = false;
    for (int i = 0; i < int(o["ntrain"]); ++i) {
        Problem p = make_problem(seed + i, o["amp"]);
        use(p);
        const double r0 = residual(p.x0);
        nl.targetResidual = 1e-8 * r0;
        StateVector x = p.x0;
        const long m0 = nl.matVecs;
        const double t0 = now();
        nl.Run(x);
        printf("train %d %.2f %.4f %d %ld %.3f %d %.3e\n", i, p.Re, p.Ri, int(nl.residualHistory.size()) - 1,
               nl.matVecs - m0, now() - t0, net.Samples(), nl.residualHistory.back() / r0);
    }
    nl.harvest = false;
    if (net.Samples() > 0) {
        // about `batches` minibatch updates of 64 profiles, one epoch (a pass
        // over every harvested sample) at a time
        const int epochs = std::max(1, int(std::ceil(o["batches"] * 64.0 / std::max(1, net.Samples()))));
        double l0 = -1, l1 = 0;
        const double t0 = now();
        for (int e = 0; e < epochs; ++e) {
            double a = 0, b = 0;
            net.Train(1, 64, 1e-3, &a, &b);
            if (l0 < 0) l0 = a;
            l1 = b;
            printf("epoch %d %.6e %.3f\n", e, b, now() - t0);
        }
        printf("loss %.6e %.6e %d %.3f\n", l0, l1, net.Samples(), now() - t0);
    }

    // held-out problems, three methods from the same start
    for (int i = int(o["first"]); i < int(o["first"]) + int(o["ntest"]); ++i) {
        Problem p = make_problem(seed + 1000 + i, o["amp"]);
        use(p);
        const double r0 = residual(p.x0);
        const double x0n = sfkrylov::norm(p.x0);
        StateVector xplain;
        for (int m = 0; m < 3; ++m) {
            if (!((int(o["methods"]) >> m) & 1)) continue;  // 1 plain, 2 learned, 4 base
            if (m == 1 && !net.Trained()) continue;
            PhiNewton& nk = m == 2 ? static_cast<PhiNewton&>(nb) : static_cast<PhiNewton&>(nl);
            nl.enabled = (m == 1);
            const char* name = m == 0 ? "plain" : (m == 1 ? "learned" : "base");
            nk.targetResidual = 1e-8 * r0;
            StateVector x = p.x0;
            const long m0 = nk.matVecs;
            const int f0 = nl.fallbacks;
            const double t0 = now();
            nk.Run(x);
            const double sec = now() - t0;
            if (m == 0) xplain = x;
            StateVector d = x - xplain;
            printf("test %d %.2f %.4f %s %d %ld %.3f %.3e %d %.3e\n", i, p.Re, p.Ri, name,
                   int(nk.residualHistory.size()) - 1, nk.matVecs - m0, sec, nk.residualHistory.back() / r0,
                   m == 1 ? nl.fallbacks - f0 : 0, sfkrylov::norm(d) / x0n);
            printf("hist %d %s", i, name);
            for (size_t k = 0; k < nk.residualHistory.size(); ++k)
                printf(" %.3e@%.3f", nk.residualHistory[k] / r0, nk.residualSeconds[k]);
            printf("\n");
        }
    }
    return 0;
}
