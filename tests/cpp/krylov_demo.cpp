// Drives the Newton-GMRES and Arnoldi drivers (include/stratiflow/NewtonKrylov.h,
// Arnoldi.h) through the C++14 facade on a small 2D problem and prints what the
// harness compares against the CPU restatement over the oracle:
//   res <k> <value>       Newton residual history
//   gres <k> <value>      GMRES residual history of the solve
//   H <i> <j> <value>     Hessenberg matrix of the GMRES solve
//   ev <k> <re> <im>      Arnoldi eigenvalues, by decreasing modulus
// and writes x0.fields, x1.fields (after one Newton step) and q0.fields (the
// Arnoldi start vector) into the directory given as argv[1].
#include <cstdio>
#include <string>
#include <vector>

#include "stratiflow/Arnoldi.h"
#include "stratiflow/NewtonKrylov.h"

// F(x) = Phi(x) - x with Phi = nsteps fixed-dt steps (FixedEvolve start-up)
class FixedNewton : public NewtonKrylov<StateVector> {
public:
    stratifloat dt = stratifloat(0.02);
    int nsteps = 25;

protected:
    StateVector EvalFunction(const StateVector& at) override {
        StateVector s = at;  // copies do not carry p (StateVector.h:16-23)
        std::vector<StateVector> traj;
        s.FixedEvolve(dt, nsteps + 1, traj);
        StateVector r = traj[nsteps];
        r -= at;
        return r;
    }
};

// Arnoldi of LinearEvolve over T about the given state
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
    const std::string dir = argc > 1 ? argv[1] : ".";
    const int K = argc > 2 ? std::atoi(argv[2]) : 9;
    gridParams = {32, 1, 48, Dimensionality::TwoDimensional};
    flowParams.L1 = stratifloat(2 * 3.14159265358979 / 0.44);
    flowParams.L2 = 1;
    flowParams.L3 = 10;
    flowParams.Re = 100;
    flowParams.Ri = stratifloat(0.1);
    flowParams.Pr = 1;
    StateVector::ResetForParams();
    StateVector::solver().SetBackground(InitialU);

    SeedRandom(7);
    StateVector x;
    x.ExciteLowWavenumbers(stratifloat(1e-3));
    // through the nodal .fields file and back, so that the state the harness
    // reads is exactly this one (a .fields file cannot hold the imaginary part
    // of the j1 = 0 coefficients that the random excitation draws)
    x.SaveToFile(dir + "/x0");
    x.LoadFromFile(dir + "/x0.fields");

    FixedNewton nk;
    nk.verbose = false;
    nk.SetKrylovDimension(K);
    nk.gmresTolerance = 1e-14;  // build every Krylov vector
    nk.maxNewtonSteps = 1;
    StateVector x1 = x;
    nk.Run(x1);
    for (size_t k = 0; k < nk.residualHistory.size(); ++k) printf("res %d %.17e\n", int(k), nk.residualHistory[k]);
    for (size_t k = 0; k < nk.gmresResidualHistory.size(); ++k)
        printf("gres %d %.17e\n", int(k), nk.gmresResidualHistory[k]);
    const sfdense::MatrixX& H = nk.Hessenberg();
    for (int j = 0; j < K - 1; ++j)
        for (int i = 0; i <= j + 1 && i < K; ++i) printf("H %d %d %.17e\n", i, j, H(i, j));
    x1.SaveToFile(dir + "/x1");

    SeedRandom(11);
    StateVector q0;
    q0.ExciteLowWavenumbers(stratifloat(0.1));
    q0.SaveToFile(dir + "/q0");
    q0.LoadFromFile(dir + "/q0.fields");
    LinearArnoldi ar(stratifloat(0.5));
    ar.K = K + 4;
    ar.startVector = &q0;
    ar.verbose = false;
    StateVector lead;
    const stratifloat g = ar.Run(x, lead, true, false);
    printf("lead %.17e\n", double(g));
    for (size_t k = 0; k < ar.eigenvalues.size(); ++k)
        printf("ev %d %.17e %.17e\n", int(k), ar.eigenvalues[k].real(), ar.eigenvalues[k].imag());
    printf("lead_norm %.17e\n", double(lead.Norm()));
    return 0;
}
