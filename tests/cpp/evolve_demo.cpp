// Drives the facade's CFL-adaptive FullEvolve (StateVector.cpp:3-83) with
// snapshots and the mixing integral, consumes the snapshots the way the
// reference's Adjoint.cpp:11-61 does (directory listing, FixedEvolve between
// two snapshots, AdjointEvolve back over the interval), and runs the adaptive
// LinearEvolve (StateVector.cpp:107-156) to a time that is not a multiple of
// its step.  Prints "key value" lines for the test harness; the initial state
// is written to <out>_init.fields so the oracle can start from it.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "stratiflow/stratiflow.h"

struct Snap {
    int step;
    stratifloat time;
    std::string file;
};

int main(int argc, char** argv) {
    const std::string out = argc > 1 ? argv[1] : "evolve_demo";
    const stratifloat T = argc > 2 ? stratifloat(atof(argv[2])) : 30.0f;
    const stratifloat TL = argc > 3 ? stratifloat(atof(argv[3])) : 0.537f;
    gridParams = {48, 1, 64, Dimensionality::TwoDimensional};
    flowParams.L1 = stratifloat(2 * 3.14159265358979 / 0.44);
    flowParams.L2 = 1;
    flowParams.L3 = 10;
    flowParams.Re = 1000;
    flowParams.Ri = 0.1f;
    flowParams.Pr = 1;
    StateVector::ResetForParams();

    StateVector x;
    std::vector<double> z, zf;
    VerticalPointsBoth(z, zf);
    srand(23);
    for (int j1 = 1; j1 < 5; ++j1) {
        const complex a(stratifloat(rand() % 1000) / 1000 - 0.5f, stratifloat(rand() % 1000) / 1000 - 0.5f);
        for (int j3 = 1; j3 < gridParams.N3 - 1; ++j3) {
            const stratifloat g = std::exp(-stratifloat(zf[j3] * zf[j3]) / 4);
            x.u1(j1, 0, j3) = stratifloat(0.02) * a * g;
            x.b(j1, 0, j3) = stratifloat(0.01) * a * g;
        }
    }
    x.EnforceBCs();
    x.SaveToFile(out + "_init");

    StateVector r;
    const stratifloat mixing = x.FullEvolve(T, r, true, false, true);
    printf("mixing %.9e\n", double(mixing));
    printf("final_ke %.9e\n", double(StateVector::solver().KE()));
    printf("final_pe %.9e\n", double(StateVector::solver().PE()));
    printf("final_dt %.9e\n", double(StateVector::solver().deltaT));

    // the snapshot directory as Adjoint.cpp:14-35 reads it
    std::vector<Snap> snaps;
    DIR* dir = opendir("snapshots");
    if (!dir) {
        printf("nosnapshots 1\n");
        return 1;
    }
    while (dirent* f = readdir(dir)) {
        const std::string name(f->d_name);
        const int ext = int(name.find(".fields")), hy = int(name.find("-"));
        if (ext != -1 && hy != -1)
            snaps.push_back({std::stoi(name.substr(0, hy)), std::stof(name.substr(hy + 1, ext - hy - 1)),
                             "snapshots/" + name});
    }
    closedir(dir);
    std::sort(snaps.begin(), snaps.end(), [](const Snap& a, const Snap& b) { return a.step < b.step; });
    printf("snapshots %d\n", int(snaps.size()));
    for (const Snap& s : snaps) printf("snap_%d %.9e\n", s.step, double(s.time));

    // the last snapshot is the final state
    StateVector last;
    last.LoadFromFile(snaps.back().file);
    StateVector d = last - r;
    printf("last_snapshot_vs_result %.9e\n", double(d.Norm() / r.Norm()));

    // Adjoint.cpp:39-61 over the last interval
    const Snap& above = snaps[snaps.size() - 1];
    const Snap& below = snaps[snaps.size() - 2];
    const int steps = above.step - below.step;
    const stratifloat dt = (above.time - below.time) / stratifloat(steps);
    std::vector<StateVector> inter;
    StateVector direct;
    direct.LoadFromFile(below.file);
    direct.FixedEvolve(dt, steps, inter);
    direct.LoadFromFile(above.file);
    inter.push_back(direct);
    StateVector adj = r;
    adj.AdjointEvolve(dt, steps, inter, adj);
    printf("adjoint_interval_steps %d\n", steps);
    printf("adjoint_norm %.9e\n", double(adj.Norm()));

    // adaptive linearised evolution to TL about the final state
    StateVector lin;
    x.LinearEvolve(TL, r, lin);
    printf("linear_norm %.9e\n", double(lin.Norm()));
    printf("linear_ke %.9e\n", double(StateVector::solver().KE()));
    printf("linear_dt %.9e\n", double(StateVector::solver().deltaT));
    return 0;
}
