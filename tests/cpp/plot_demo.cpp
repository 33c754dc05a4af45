// HeatPlot and the plot members of the facade (Graph.h:7-109; IMEXRK.h:164-303;
// StateVector.h:476-500) on a 48 x 1 x 64 2D grid.  Usage:
//   plot_demo <out dir>
// Writes into <out dir>: linear.bin / wave.bin (the resampled images of two
// nodal fields, float32, rows top first), zf.bin / z.bin (vertical points,
// float64), linear.png, state/ (StateVector::PlotAll) and images/ (IMEXRK's
// PlotAll after PrepareRun), and prints "size <width> <height>".
#include <cmath>
#include <cstdio>
#include <fstream>
#include <string>

#include "stratiflow/stratiflow.h"

static void dump(const std::string& path, const void* p, size_t bytes) {
    std::ofstream f(path, std::ios::binary);
    f.write(static_cast<const char*>(p), std::streamsize(bytes));
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string out = argv[1];
    gridParams = {48, 1, 64, Dimensionality::TwoDimensional};
    flowParams.L1 = stratifloat(2 * 3.14159265358979 / 0.44);
    flowParams.L2 = 1;
    flowParams.L3 = 10;
    flowParams.Re = 1000;
    flowParams.Ri = stratifloat(0.1);
    flowParams.Pr = 1;
    StateVector::ResetForParams();
    const GridParams& g = gridParams;
    std::vector<double> z, zf;
    VerticalPointsBoth(z, zf);
    dump(out + "/z.bin", z.data(), z.size() * sizeof(double));
    dump(out + "/zf.bin", zf.data(), zf.size() * sizeof(double));

    // a field linear in z (Neumann points) and one that varies in x and z
    NeumannNodal lin, wave;
    for (int n1 = 0; n1 < g.N1; ++n1)
        for (int n3 = 0; n3 < g.N3; ++n3) {
            const double x = n1 * flowParams.L1 / g.N1;
            lin(n1, 0, n3) = 0.25 + 0.5 * zf[n3];
            wave(n1, 0, n3) = std::cos(2 * M_PI * x / flowParams.L1) * std::exp(-0.1 * zf[n3] * zf[n3]) + 0.01 * n3;
        }
    int w = 0, h = 0;
    std::vector<stratifloat> a = HeatPlotData(lin, flowParams.L1, flowParams.L3, 0, w, h);
    std::vector<stratifloat> b = HeatPlotData(wave, flowParams.L1, flowParams.L3, 0, w, h);
    std::vector<float> af(a.begin(), a.end()), bf(b.begin(), b.end());
    dump(out + "/linear.bin", af.data(), af.size() * sizeof(float));
    dump(out + "/wave.bin", bf.data(), bf.size() * sizeof(float));
    HeatPlot(lin, flowParams.L1, flowParams.L3, 0, out + "/linear.png");
    printf("size %d %d\n", w, h);

    // a seeded state: StateVector::PlotAll and IMEXRK's PlotAll
    SeedRandom(5);
    StateVector x;
    x.ExciteLowWavenumbers(stratifloat(1e-2));
    x.PlotAll(out + "/state");
    IMEXRK& sv = StateVector::solver();
    sv.u1 = x.u1;
    sv.u3 = x.u3;
    sv.b = x.b;
    sv.SetBackground(InitialU);
    sv.PrepareRun(out + "/images", true);
    sv.PlotAll("0-0.png", true);
    return 0;
}
