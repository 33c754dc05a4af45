// Trust region of the flexible (right-preconditioned) GMRES in NewtonKrylov.h
// on a host vector type: F(x) = A x - b with a dense A whose eigenvalues spread
// over [0.5, 5], so that 30 Krylov vectors do not solve it.  One
// Newton step from x0 = 0.1 (all entries) with a trust region smaller than the full Newton
// step must move x by |dx| <= Delta whatever the preconditioner's scale
// (M^-1 = s I), and with the region inactive the flexible and plain solves
// give the same dx.  Prints "key value" lines.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "stratiflow/NewtonKrylov.h"

static const int N = 200;  // Krylov dimension 30 stays far from exhausting the space

struct Vec {
    std::vector<double> v;
    Vec() : v(N, 0.0) {}
    double Dot(const Vec& o) const {
        double s = 0;
        for (int i = 0; i < N; ++i) s += v[i] * o.v[i];
        return s;
    }
    double Norm() const { return std::sqrt(Dot(*this)); }
    void MulAdd(double a, const Vec& o) {
        for (int i = 0; i < N; ++i) v[i] += a * o.v[i];
    }
    Vec& operator*=(double a) {
        for (double& x : v) x *= a;
        return *this;
    }
    Vec& operator+=(const Vec& o) {
        MulAdd(1, o);
        return *this;
    }
    Vec& operator-=(const Vec& o) {
        MulAdd(-1, o);
        return *this;
    }
    void Zero() { std::fill(v.begin(), v.end(), 0.0); }
    void EnforceBCs() {}
    void SaveToFile(const std::string&) const {}
};

struct Problem : NewtonKrylov<Vec> {
    std::vector<double> A;
    Vec b;
    double scale = 0;  // preconditioner M^-1 = scale * I (0: none)
    Problem() : A(N * N) {
        srand(5);
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j)
                A[i * N + j] = (i == j ? 0.5 + 4.5 * i / (N - 1) : 0.0) + 0.2 * (rand() / double(RAND_MAX) - 0.5) / 14.0;
        for (int i = 0; i < N; ++i) b.v[i] = rand() / double(RAND_MAX) - 0.5;
        verbose = false;
        fdEpsilon = 1e-7;
    }
    Vec EvalFunction(const Vec& x) override {
        Vec r;
        for (int i = 0; i < N; ++i) {
            double s = -b.v[i];
            for (int j = 0; j < N; ++j) s += A[i * N + j] * x.v[j];
            r.v[i] = s;
        }
        return r;
    }
    bool HasPreconditioner() const override { return scale != 0; }
    Vec Precondition(const Vec& v) override {
        Vec r = v;
        r *= scale;
        return r;
    }
};

static Vec one_step(double scale, double Delta) {
    Problem p;
    p.scale = scale;
    p.Delta = Delta;
    p.gmresTolerance = 1e-10;
    p.maxNewtonSteps = 1;
    p.SetKrylovDimension(31);
    Vec x;
    for (int i = 0; i < N; ++i) x.v[i] = 0.1;  // the derivative step scales with |x|
    const Vec x0 = x;
    p.Run(x);
    x -= x0;
    return x;
}

int main() {
    const double Delta = 0.05;
    for (double s : {0.0, 1.0, 10.0, 0.1}) {
        const Vec dx = one_step(s, Delta);
        printf("dx_norm_s%g %.12e\n", s, dx.Norm());
    }
    // region inactive: Delta above the (30-vector) unconstrained step
    Vec plain = one_step(0.0, 1e3), flex = one_step(10.0, 1e3);
    Vec d = flex;
    d -= plain;
    printf("full_step_norm %.12e\n", plain.Norm());
    printf("flex_vs_plain %.12e\n", d.Norm() / plain.Norm());
    return 0;
}
