// Facade diagnostics and LoadAndInterpolate (StateVector.h:151-197, 275-466;
// IMEXRK.h:445-459) on a 48 x 1 x 96 2D grid.  Usage:
//   diag_demo <input 48x1x64 .fields> <interpolated output> <state output>
// Prints one "name value" line per diagnostic of a seeded state.
#include <cstdio>
#include <string>

#include "stratiflow/stratiflow.h"

int main(int argc, char** argv) {
    setvbuf(stdout, nullptr, _IOLBF, 0);  // one line at a time, also into files
    if (argc < 4) return 2;
    gridParams = {48, 1, 96, Dimensionality::TwoDimensional};
    flowParams.L1 = stratifloat(2 * 3.14159265358979 / 0.44);
    flowParams.L2 = 1;
    flowParams.L3 = 10;
    flowParams.Re = 1000;
    flowParams.Ri = stratifloat(0.1);
    flowParams.Pr = 1;
    StateVector::ResetForParams();

    // a 48 x 1 x 64 file interpolated onto 96 vertical points
    StateVector x;
    x.LoadAndInterpolate<48, 1, 64>(argv[1]);
    x.SaveToFile(argv[2]);

    // diagnostics of a seeded state
    SeedRandom(3);
    StateVector y;
    y.ExciteLowWavenumbers(stratifloat(1e-2));
    // the imaginary parts that ExciteLowWavenumbers gives the (0,0) modes do
    // not survive the nodal file; reloading makes y what the file holds
    y.SaveToFile(argv[3]);
    y.LoadFromFile(std::string(argv[3]) + ".fields");  // the solver now holds y
    printf("jover_k %.9e\n", double(StateVector::solver().JoverK()));
    printf("energy %.9e\n", double(y.Energy()));
    printf("enstrophy %.9e\n", double(y.Enstrophy()));
    printf("other_axis %.9e\n", double(y.OtherAxis()));
    printf("minimum_ri %.9e\n", double(y.MinimumRi()));
    printf("integrate_b %.9e\n", double(IntegrateAllSpace(y.b, flowParams.L1, flowParams.L2, flowParams.L3)));
    StateVector z = y;
    z.AddBackground();
    Neumann1D zave, yave;
    HorizontalAverage(z.u1, zave);
    HorizontalAverage(y.u1, yave);
    printf("background_mid %.9e\n", double(zave.Get()[gridParams.N3 / 2] - yave.Get()[gridParams.N3 / 2]));
    z.RemoveBackground();
    StateVector d = z - y;
    printf("background_roundtrip %.9e\n", double(d.Norm() / y.Norm()));
    y.PlotAll(std::string(argv[3]) + "_plots");
    return 0;
}
