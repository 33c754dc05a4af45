// params.dat and .params files of the facade (Parameters.cpp:55-111,
// OSUtils.cpp:84-95), run in the current directory without a GPU.  Writes
// params.dat and v.params, then prints one line per check:
//   loaded <value read back from v.params>
//   L3 <flowParams.L3 after LoadParameters restored it>
//   mismatch <message thrown for a grid that differs from the file>
// followed by PrintParameters()'s lines.
#include <cstdio>
#include <iostream>
#include <string>

#include "stratiflow/stratiflow.h"

int main() {
    gridParams = {384, 192, 384, Dimensionality::ThreeDimensional};
    flowParams.L1 = stratifloat(2 * 3.14159265358979 / 0.44);
    flowParams.L2 = flowParams.L1 / 2;
    flowParams.L3 = 10;
    flowParams.Re = 4000;
    flowParams.Ri = stratifloat(0.16);
    flowParams.Pr = 1;
    flowParams.EvolveBackground = false;
    DumpParameters();

    SaveValueToFile(stratifloat(0.123456789), "v.params");
    stratifloat v = 0;
    LoadValueFromFile(v, "v.params");
    std::cout << std::setprecision(9) << "loaded " << v << std::endl;

    flowParams.L3 = 5;
    LoadParameters("params.dat");
    std::cout << "L3 " << flowParams.L3 << std::endl;

    gridParams.N1 = 192;
    try {
        LoadParameters("params.dat");
        std::cout << "mismatch none" << std::endl;
    } catch (const std::string& e) {
        std::cout << "mismatch " << e << std::endl;
    }
    gridParams.N1 = 384;
    PrintParameters();
    return 0;
}
