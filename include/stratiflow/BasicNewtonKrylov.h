// BasicNewtonKrylov.h -- fixed points of the time-T flow map (BasicNewtonKrylov.h:1-16).
#pragma once

#include "NewtonKrylov.h"
#include "stratiflow.h"

class BasicNewtonKrylov : public NewtonKrylov<StateVector> {
protected:
    // F(x) = Phi_T(x) - x
    StateVector EvalFunction(const StateVector& at) override {
        StateVector result;
        at.FullEvolve(T, result, false, false);
        result -= at;
        return result;
    }
};
