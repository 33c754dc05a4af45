// LearnedPreconditioner.h -- the learned right preconditioner for the
// Newton-GMRES driver (no counterpart in the reference, whose GMRES is
// unpreconditioned: NewtonKrylov.h:138-235).
//
// LearnedPreconditioner wraps the library's network (sf_net_*): a few
// Conv1D layers over the vertical profile of each horizontal mode, BF16 on the
// matrix cores, trained on the GPU.  Apply(v) = v + s * net(v / s), s = RMS(v),
// is the identity until the network is trained.  With an analytic base B
// (SetBase, a DiffusionPreconditioner) the network learns a correction to it:
// Apply(v) = B v + s * net(v / s), trained on the pairs (w, q - B w + w).
//
// WithLearnedPreconditioner<Solver> adds it to any NewtonKrylov<StateVector>
// driver: while `harvest` is set, every GMRES operator application (x, A x)
// becomes a training pair; once trained and `enabled`, GMRES runs in its
// flexible right-preconditioned form.  Plain GMRES remains the fallback
// (enabled = false) and the judge of the converged solution.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "DiffusionPreconditioner.h"
#include "NewtonKrylov.h"
#include "stratiflow.h"

class LearnedPreconditioner {
public:
    explicit LearnedPreconditioner(int hidden = 64, int layers = 4, unsigned long long seed = 0) {
        if (sf_net_create(sfacade::ctx(), hidden, layers, seed, &net_) != SF_OK)
            throw std::runtime_error(std::string("sf_net_create: ") + sf_net_last_error(nullptr));
    }
    ~LearnedPreconditioner() {
        if (net_) sf_net_destroy(net_);
    }
    LearnedPreconditioner(const LearnedPreconditioner&) = delete;
    LearnedPreconditioner& operator=(const LearnedPreconditioner&) = delete;

    StateVector Apply(const StateVector& v) {
        StateVector out;
        sf_field* in[4] = {v.u1.dev(false), v.u2.dev(false), v.u3.dev(false), v.b.dev(false)};
        sf_field* o[4] = {out.u1.dev(true), out.u2.dev(true), out.u3.dev(true), out.b.dev(true)};
        check(sf_net_apply(net_, in, o));
        if (base_) {
            out -= v;
            out += base_->Apply(v);
        }
        out.EnforceBCs();
        return out;
    }
    // analytic base preconditioner that the network corrects (nullptr: none);
    // it must outlive this object
    void SetBase(const DiffusionPreconditioner* base) { base_ = base; }
    // dilation of each layer's taps (1 = neighbouring points); dilations 1, 2,
    // 4, 8 let four layers see 61 vertical points instead of 17
    void SetDilations(const std::vector<int>& d) { check(sf_net_set_dilations(net_, d.data())); }
    // a pair with A q = w
    void AddPair(const StateVector& w, const StateVector& q) {
        StateVector t;
        const StateVector* target = &q;
        if (base_) {  // the network's own output should be q - B w + w
            t = q;
            t -= base_->Apply(w);
            t += w;
            target = &t;
        }
        sf_field* a[4] = {w.u1.dev(false), w.u2.dev(false), w.u3.dev(false), w.b.dev(false)};
        sf_field* b[4] = {target->u1.dev(false), target->u2.dev(false), target->u3.dev(false),
                          target->b.dev(false)};
        check(sf_net_add_pair(net_, a, b));
    }
    int Samples() const {
        int n = 0;
        sf_net_samples(net_, &n);
        return n;
    }
    void Train(int epochs, int batch, double lr, double* first = nullptr, double* last = nullptr) {
        double a = 0, b = 0;
        check(sf_net_train(net_, epochs, batch, lr, &a, &b));
        if (first) *first = a;
        if (last) *last = b;
        trained_ = true;
    }
    bool Trained() const { return trained_; }

private:
    void check(int rc) const {
        if (rc != SF_OK) throw std::runtime_error(std::string("sf_net: ") + sf_net_last_error(net_));
    }
    sf_net* net_ = nullptr;
    bool trained_ = false;
    const DiffusionPreconditioner* base_ = nullptr;
};

template <class Solver>
class WithLearnedPreconditioner : public Solver {
public:
    explicit WithLearnedPreconditioner(LearnedPreconditioner& p) : pre_(p) {}
    bool harvest = false;  // record (A x, x) pairs of every GMRES operator application
    bool enabled = true;   // use the preconditioner once it is trained
    bool fallback = true;  // redo a solve with plain GMRES when the learned one misses the tolerance
    int fallbacks = 0;     // how many solves fell back
    double lastLearnedResidual = 0;  // GMRES residual of the last preconditioned attempt
    int lastLearnedIterations = 0;   // and its Krylov dimension

    // GMRES with the learned preconditioner, judged by the plain method: a
    // preconditioned solve that ends above the tolerance is repeated without it
    void SolveLinear(const StateVector& rhs, StateVector& x, double epsilon, double Delta_) override {
        Solver::SolveLinear(rhs, x, epsilon, Delta_);
        if (enabled && pre_.Trained()) {
            lastLearnedResidual = this->lastGmresResidual;
            lastLearnedIterations = this->lastKrylovDimension;
        }
        if (fallback && enabled && pre_.Trained() && !(this->lastGmresResidual < epsilon)) {
            ++fallbacks;
            const bool e = enabled;
            enabled = false;
            Solver::SolveLinear(rhs, x, epsilon, Delta_);
            enabled = e;
        }
    }

protected:
    bool HasPreconditioner() const override { return enabled && pre_.Trained(); }
    StateVector Precondition(const StateVector& v) override { return pre_.Apply(v); }
    void OnMatVec(const StateVector& x, const StateVector& Ax) override {
        if (harvest) pre_.AddPair(Ax, x);
    }

private:
    LearnedPreconditioner& pre_;
};
