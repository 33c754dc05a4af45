// DiffusionPreconditioner.h -- an analytic right preconditioner for the
// Newton-GMRES driver on F(x) = Phi_T(x) - x (no counterpart in the reference,
// whose GMRES is unpreconditioned: NewtonKrylov.h:138-235).
//
// The slow directions of A = Phi_T' - I are modes that the background barely
// advects and that mostly diffuse (near the critical level U = 0): for them
// A ~ exp(T L) - I with L = nu (Dzz - k^2) - i kx U(z), eigenvalues near 0.
// Apply(v) is, per field with its own boundary condition and diffusivity
// (1/Re for the velocity, 1/(Re Pr) for b),
//   (T L - delta)^-1 v - v/2   on modes with |k| <= kmax,
//   -v                         on the others,
// where (T L)^-1 - 1/2 is the Cayley approximation of (exp(T L) - I)^-1 and
// delta > 0 keeps the conserved Neumann means finite.  Without `advect` L is
// the diffusion alone.  One library call per field
// (sf_field_diffusion_precondition).  DiffusionPreconditioner(T, 0.3, 1e9,
// true) is the measured setting for the 64x1x64 Newton problems of
// tests/cpp/precond_demo.cpp.
#pragma once

#include <stdexcept>
#include <string>

#include "NewtonKrylov.h"
#include "stratiflow.h"

class DiffusionPreconditioner {
public:
    // T: the integration time of Phi_T; advect: include -i kx U(z) in L
    explicit DiffusionPreconditioner(double T, double delta = 1e-3, double kmax = 0, bool advect = false)
        : T_(T), delta_(delta), kmax_(kmax), advect_(advect) {}

    StateVector Apply(const StateVector& v) const {
        StateVector out;
        const double nu = 1.0 / double(flowParams.Re);
        const double nub = nu / double(flowParams.Pr);
        run(v.u1.dev(false), out.u1.dev(true), T_ * nu);
        if (gridParams.ThirdDimension()) run(v.u2.dev(false), out.u2.dev(true), T_ * nu);
        run(v.u3.dev(false), out.u3.dev(true), T_ * nu);
        run(v.b.dev(false), out.b.dev(true), T_ * nub);
        out.EnforceBCs();
        return out;
    }

    double T() const { return T_; }
    double delta() const { return delta_; }
    double kmax() const { return kmax_; }

private:
    void run(const sf_field* in, sf_field* out, double tnu) const {
        if (sf_field_diffusion_precondition(in, out, tnu, delta_, kmax_, advect_ ? T_ : 0.0) != SF_OK)
            throw std::runtime_error(std::string("sf_field_diffusion_precondition: ") +
                                     sf_last_error(sfacade::ctx()));
    }
    double T_, delta_, kmax_;
    bool advect_;
};

// Adds the diffusion preconditioner to any NewtonKrylov<StateVector> driver:
// while `enabled`, GMRES runs in its flexible right-preconditioned form.
template <class Solver>
class WithDiffusionPreconditioner : public Solver {
public:
    explicit WithDiffusionPreconditioner(const DiffusionPreconditioner& p) : pre_(p) {}
    bool enabled = true;

protected:
    bool HasPreconditioner() const override { return enabled; }
    StateVector Precondition(const StateVector& v) override { return pre_.Apply(v); }

private:
    DiffusionPreconditioner pre_;
};
