// Drives the C++14 facade (include/stratiflow/stratiflow.h) the way the
// reference's drivers do: StateVector arithmetic, FixedEvolve, Dot/Norm and
// .fields round trips.  Prints one "key value" pair per line for the test
// harness to compare against the CPU oracle.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "stratiflow/stratiflow.h"

int main(int argc, char** argv) {
    const char* out = argc > 1 ? argv[1] : "facade_demo.fields";
    gridParams = {48, 1, 64, Dimensionality::TwoDimensional};
    flowParams.L1 = stratifloat(2 * 3.14159265358979 / 0.44);
    flowParams.L2 = 1;
    flowParams.L3 = 10;
    flowParams.Re = 1000;
    flowParams.Ri = 0.1f;
    flowParams.Pr = 1;
    StateVector::ResetForParams();

    // a seeded low-wavenumber perturbation written through element access
    StateVector x;
    std::vector<double> z, zf;
    VerticalPointsBoth(z, zf);
    srand(17);
    for (int j1 = 1; j1 < 4; ++j1) {
        const complex a(stratifloat(rand() % 1000) / 1000 - 0.5f, stratifloat(rand() % 1000) / 1000 - 0.5f);
        for (int j3 = 1; j3 < gridParams.N3 - 1; ++j3) {
            const stratifloat g = std::exp(-stratifloat(zf[j3] * zf[j3]) / 4);
            x.u1(j1, 0, j3) = stratifloat(0.01) * a * g;
            x.b(j1, 0, j3) = stratifloat(0.005) * a * g;
        }
    }
    x.EnforceBCs();
    printf("norm0 %.9e\n", double(x.Norm()));
    x.SaveToFile(std::string(out) + "_init");

    StateVector y = x;
    y *= 2.0f;
    y -= x;
    StateVector d = y - x;
    printf("selfdiff %.9e\n", double(d.Norm()));

    StateVector::solver().SetBackground(InitialU);
    std::vector<StateVector> traj;
    x.FixedEvolve(0.01f, 20, traj);
    printf("ke20 %.9e\n", double(StateVector::solver().KE()));
    printf("pe20 %.9e\n", double(StateVector::solver().PE()));
    printf("traj_norm10 %.9e\n", double(traj[10].Norm()));

    // adjoint of the first 10 steps, started from the initial state
    StateVector adj;
    x.AdjointEvolve(0.01f, 10, traj, adj);
    printf("adj_ke %.9e\n", double(StateVector::solver().KE()));
    printf("adj_pe %.9e\n", double(StateVector::solver().PE()));

    StateVector r;
    x.FullEvolve(0.5f, r, false, false);
    printf("full_norm %.9e\n", double(r.Norm()));
    r.SaveToFile(out);
    StateVector back;
    back.LoadFromFile(std::string(out) + ".fields");
    StateVector diff = back - r;
    printf("fields_roundtrip %.9e\n", double(diff.Norm() / r.Norm()));
    return 0;
}
