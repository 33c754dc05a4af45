// Newton-Krylov search for a steady state of the 2D stratified shear layer
// (Newton-GMRES with FP64 Krylov algebra over FP32 or FP64 time stepping),
// then the leading Arnoldi eigenvalues of the linearised map about it.
//   newton_c3 <N1> <N3> <T> <plain steps> <amplitude> <K> <max Newton> <rel tol> <out dir> [Delta] [pdelta]
// pdelta > 0: GMRES right-preconditioned by the advected diffusion preconditioner
// (include/stratiflow/DiffusionPreconditioner.h) with that delta, on every mode.
// Re=2000, Pr=1, Ri=0.12, U = tanh(z), B = tanh(z) (dB/dz = sech^2), dt = 0.01.
// Prints:
//   newton <step> <residual> <GMRES iterations> <seconds> <GMRES relative residual>
//   done <Newton steps> <final/initial residual> <seconds> <matvecs>
//   ev <k> <re> <im>
// and writes <out>/guess.fields and <out>/solution.fields.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "stratiflow/Arnoldi.h"
#include "stratiflow/DiffusionPreconditioner.h"
#include "stratiflow/NewtonKrylov.h"

static double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Phi_T: T/dt fixed steps with the start-up of FixedEvolve (StateVector.cpp:85-105),
// without keeping the trajectory
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
    if (nsteps > 0) {
        sv.TimeStep();  // synchronises dt and the state with the context
        sfacade::check(sf_step(sfacade::ctx(), nsteps - 1), sfacade::ctx());
    }
    StateVector r;
    r.u1 = sv.u1;
    r.u3 = sv.u3;
    r.b = sv.b;
    r.EnforceBCs();
    return r;
}

// F(x) = Phi_T(x) - x
class SteadyNewton : public NewtonKrylov<StateVector> {
public:
    stratifloat dt = stratifloat(0.01);
    int nsteps = 1100;

protected:
    StateVector EvalFunction(const StateVector& at) override {
        StateVector r = Evolve(at, dt, nsteps);
        r -= at;
        return r;
    }
};

class LinearArnoldi : public Arnoldi<StateVector> {
public:
    explicit LinearArnoldi(stratifloat t) { T = t; }

protected:
    StateVector EvalFunction(const StateVector& at) override {
        linearAboutStart = at;
        return at;
    }
    StateVector EvalLinearised(const StateVector& at) override {
        StateVector result;
        at.LinearEvolve(T, linearAboutStart, result);
        return result;
    }
};

int main(int argc, char** argv) {
    setvbuf(stdout, nullptr, _IOLBF, 0);  // one line at a time, also into files
    if (argc < 10) return 2;
    const int N1 = std::atoi(argv[1]), N3 = std::atoi(argv[2]);
    const double T = std::atof(argv[3]);
    const int nplain = std::atoi(argv[4]);
    const double amplitude = std::atof(argv[5]);
    const int K = std::atoi(argv[6]);
    const int maxNewton = std::atoi(argv[7]);
    const double reltol = std::atof(argv[8]);
    const std::string dir = argv[9];
    const double Delta = argc > 10 ? std::atof(argv[10]) : 100.0;  // trust region (NewtonKrylov.h:29)
    const double pdelta = argc > 11 ? std::atof(argv[11]) : 0.0;
    gridParams = {N1, 1, N3, Dimensionality::TwoDimensional};
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

    // seeded perturbation, then plain steps to a Kelvin-Helmholtz-like state
    SeedRandom(2026);
    StateVector x;
    x.ExciteLowWavenumbers(stratifloat(amplitude));
    x.SaveToFile(dir + "/seed");
    x.LoadFromFile(dir + "/seed.fields");
    if (nplain > 0) x = Evolve(x, stratifloat(0.01), nplain);
    x.SaveToFile(dir + "/guess");
    x.LoadFromFile(dir + "/guess.fields");

    WithDiffusionPreconditioner<SteadyNewton> nk(DiffusionPreconditioner(T, pdelta > 0 ? pdelta : 1.0, 1e9, true));
    nk.enabled = pdelta > 0;
    nk.verbose = false;
    nk.targetResidual = 0;  // the relative tolerance below decides
    nk.Delta = Delta;
    nk.nsteps = int(std::lround(T / 0.01));
    nk.SetKrylovDimension(K);
    nk.gmresTolerance = 1e-3;
    nk.maxNewtonSteps = 1;
    const double t0 = now();
    double r0 = -1, r = 0;
    int steps = 0;
    for (int it = 0; it < maxNewton; ++it) {
        const double t = now();
        nk.Run(x);  // one Newton step (evaluates the residual first)
        r = nk.residualHistory.front();
        if (r0 < 0) r0 = r;
        printf("newton %d %.6e %d %.3f %.3e\n", it, r, nk.lastKrylovDimension, now() - t, nk.lastGmresResidual);
        fflush(stdout);
        steps = it;
        if (r <= reltol * r0) break;
    }
    // the residual of the final state
    {
        SteadyNewton check;
        check.nsteps = nk.nsteps;
        check.verbose = false;
        check.SetKrylovDimension(2);
        check.maxNewtonSteps = 0;
        StateVector y = x;
        check.Run(y);
        r = check.residualHistory.front();
    }
    printf("done %d %.6e %.3f %ld\n", steps, r / r0, now() - t0, nk.matVecs);
    x.SaveToFile(dir + "/solution");

    LinearArnoldi ar{stratifloat(T)};
    ar.K = 40;
    ar.verbose = false;
    SeedRandom(11);
    StateVector q0;
    q0.ExciteLowWavenumbers(stratifloat(0.1));
    q0.SaveToFile(dir + "/q0");
    q0.LoadFromFile(dir + "/q0.fields");
    ar.startVector = &q0;
    StateVector lead;
    const double ta = now();
    ar.Run(x, lead, true, false);
    printf("arnoldi_seconds %.3f\n", now() - ta);
    for (size_t k = 0; k < ar.eigenvalues.size() && k < 10; ++k)
        printf("ev %d %.9e %.9e\n", int(k), ar.eigenvalues[k].real(), ar.eigenvalues[k].imag());
    return 0;
}
