// NewtonKrylov.h -- Newton-GMRES with a hookstep trust region (NewtonKrylov.h:1-241).
//
// Same class template, member names and algorithm as the reference:
//   Run(x):   Newton iteration x += dx until |F(x)| < targetResidual, with
//             F = EvalFunction (NewtonKrylov.h:16-106);
//   GMRES:    modified Gram-Schmidt Arnoldi on A = -dF/dx (finite-difference
//             directional derivative, NewtonKrylov.h:119-132), least squares in
//             the singular basis of the Hessenberg matrix with the trust region
//             |z| <= Delta (Chandler & Kerswell 2013; NewtonKrylov.h:138-235).
// Differences, all deliberate:
//   * Krylov inner products are accumulated in FP64 on the GPU (DotF64) when
//     the vector type has one, and the Hessenberg matrix, SVD and least squares
//     are FP64 on the host (the reference uses stratifloat throughout);
//   * the SVD uses the thin U of H ((k+1) x k); the reference divides the k+1
//     entries of U^T beta by the k singular values, a size mismatch that Eigen
//     leaves undefined in release builds;
//   * Krylov vectors are allocated as GMRES reaches them (the reference
//     allocates all K = 2048 up front), and Run stops after maxNewtonSteps;
//   * the finite-difference step is fdEpsilon*|x|/|v| with fdEpsilon = 1e-7 in
//     double precision (the reference's constant) and sqrt(FLT_EPSILON) in
//     single precision, where a 1e-7 relative step is below FP32 rounding;
//   * an optional right preconditioner (Precondition / HasPreconditioner)
//     turns GMRES into its flexible form: the preconditioned directions are
//     kept beside the Krylov basis and the update is built from them; the
//     trust region then bounds the update itself, |dx| = |Z y| <= Delta, in
//     the metric of the directions' FP64 Gram matrix (the directions are not
//     orthonormal), instead of |y|;
//   * writing ICs/<step>.fields every Newton step is optional (saveICs), and
//     there is no plotting.
#pragma once

#include <cfloat>
#include <chrono>
#include <cmath>
#include <iostream>
#include <string>
#include <vector>

#include "dense.h"
#include "stratiflow.h"

namespace sfkrylov {
// FP64 inner product when the vector type provides one
template <typename V>
auto dot64(const V& a, const V& b, int) -> decltype(a.DotF64(b)) {
    return a.DotF64(b);
}
template <typename V>
double dot64(const V& a, const V& b, long) {
    return double(a.Dot(b));
}
template <typename V>
double dot(const V& a, const V& b) {
    return dot64(a, b, 0);
}
template <typename V>
double norm(const V& a) {
    return std::sqrt(dot(a, a));
}
}  // namespace sfkrylov

template <typename VectorType>
class NewtonKrylov {
public:
    NewtonKrylov() : H(K, K - 1) { H.setZero(); }
    virtual ~NewtonKrylov() {}

    // Newton-Raphson with GMRES (NewtonKrylov.h:16-106); x is the initial
    // guess and the result
    void Run(VectorType& x) {
        if (saveICs) MakeCleanDir("ICs");
        VectorType dx;
        int step = 0;
        residualHistory.clear();
        residualSeconds.clear();
        const auto start = std::chrono::steady_clock::now();
        while (true) {
            step++;
            if (saveICs) x.SaveToFile("ICs/" + std::to_string(step));
            VectorType rhs = EvalFunction(x);
            linearAboutStart = x;
            linearAboutEnd = rhs;
            const double residual = sfkrylov::norm(rhs);
            residualHistory.push_back(residual);
            residualSeconds.push_back(
                std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count());
            if (verbose) std::cout << "NEWTON STEP " << step << ", RESIDUAL: " << residual << std::endl;
            newtonSteps = step;
            if (residual < targetResidual || step > maxNewtonSteps) return;
            // the Krylov space is rebuilt for every Newton step (NewtonKrylov.h:88-93)
            vectorsToReuse = 0;
            H.setZero();
            SolveLinear(rhs, dx, gmresTolerance, Delta);
            x += dx;
            // repeated, as relative precision can make it better (NewtonKrylov.h:101-104)
            EnforceConstraints(x);
            EnforceConstraints(x);
            EnforceConstraints(x);
        }
    }

    virtual void EnforceConstraints(VectorType& at) { (void)at; }

    // configuration; the reference's constants are the defaults
    double targetResidual = 1e-7;   // NewtonKrylov.h:31
    double gmresTolerance = 0.01;   // NewtonKrylov.h:96
    double Delta = 100;             // NewtonKrylov.h:29
    int maxNewtonSteps = 1 << 30;
    double fdEpsilon = sizeof(stratifloat) == 4 ? 3.4526698e-4 : 1e-7;
    bool saveICs = false;
    bool verbose = true;

    // what the last solve did
    int newtonSteps = 0;
    long matVecs = 0;  // operator applications (time integrations) over all solves
    int lastKrylovDimension = 0;
    double lastGmresResidual = 0;
    std::vector<double> residualHistory;
    std::vector<double> residualSeconds;  // wall time since Run began at each residual
    std::vector<double> gmresResidualHistory;  // every GMRES step of the last solve

    // solves A x = rhs with A = -dF/dx (NewtonKrylov.h:134-235); public so
    // that a caller can run one linear solve about a chosen state
    void GMRES(const VectorType& rhs, VectorType& x, double epsilon, double Delta_ = 0) {
        if (int(q.size()) < 1) q.resize(1);
        q[0] = rhs;
        q[0].EnforceBCs();
        const double beta = sfkrylov::norm(q[0]);
        if (verbose) std::cout << beta << std::endl;
        q[0] *= stratifloat(1 / beta);
        sfdense::VectorX y;
        int Kused = K;
        gmresResidualHistory.clear();
        const bool flexible = HasPreconditioner();
        for (int k = 1; k < K; k++) {
            if (k > vectorsToReuse) {
                vectorsToReuse = k;
                if (int(q.size()) < k + 1) q.resize(k + 1);
                // q_k = A q_{k-1}, or A M^-1 q_{k-1} with a preconditioner
                if (flexible) {
                    if (int(Z.size()) < k) Z.resize(k);
                    Z[k - 1] = Precondition(q[k - 1]);
                    // Gram matrix of the directions, for the trust region on Z y
                    if (G.rows < K) G = sfdense::MatrixX(K, K);
                    for (int j = 0; j < k; ++j) G(j, k - 1) = G(k - 1, j) = sfkrylov::dot(Z[j], Z[k - 1]);
                    q[k] = EvalDerivative(Z[k - 1]);
                } else {
                    q[k] = EvalDerivative(q[k - 1]);
                }
                q[k] *= stratifloat(-1.0);  // factor of -1 for Newton iteration
                ++matVecs;
                OnMatVec(flexible ? Z[k - 1] : q[k - 1], q[k]);
                // modified Gram-Schmidt, FP64 coefficients
                for (int j = 0; j < k; j++) {
                    H(j, k - 1) = sfkrylov::dot(q[j], q[k]);
                    q[k].MulAdd(stratifloat(-H(j, k - 1)), q[j]);
                }
                H(k, k - 1) = sfkrylov::norm(q[k]);
                q[k] *= stratifloat(1 / H(k, k - 1));
                q[k].EnforceBCs();
            }
            const sfdense::MatrixX subH = H.block(k + 1, k);
            double residual = 0, mu = 0;
            sfdense::MatrixX R;
            if (flexible && Delta_ > 0 && sfdense::cholesky_upper(G, k, R)) {
                y = sfdense::hookstep_metric(subH, R, beta, Delta_, &residual, &mu);
            } else {
                y = sfdense::hookstep(subH, beta, Delta_, &residual, &mu);
                if (flexible && Delta_ > 0) {
                    // dependent directions: scale the update back into the region
                    double zz = 0;
                    for (int i = 0; i < k; ++i)
                        for (int j = 0; j < k; ++j) zz += y[i] * G(i, j) * y[j];
                    const double len = std::sqrt(std::max(zz, 0.0));
                    if (len > Delta_)
                        for (double& v : y) v *= Delta_ / len;
                }
            }
            if (verbose) {
                std::cout << "For |z|<Delta, mu=" << mu << std::endl;
                std::cout << "GMRES STEP " << k << ", RESIDUAL: " << residual << std::endl;
            }
            gmresResidualHistory.push_back(residual);
            lastGmresResidual = residual;
            if (residual < epsilon) {
                Kused = k + 1;
                break;
            }
        }
        lastKrylovDimension = Kused - 1;
        x.Zero();
        for (int k = 0; k < Kused - 1; k++) x.MulAdd(stratifloat(y[k]), flexible ? Z[k] : q[k]);
    }

    // the Hessenberg matrix of the last GMRES solve (rows x cols = K x K-1)
    const sfdense::MatrixX& Hessenberg() const { return H; }

protected:
    virtual VectorType EvalFunction(const VectorType& at) = 0;
    // right preconditioner M^-1; return true from HasPreconditioner to use it
    virtual bool HasPreconditioner() const { return false; }
    virtual VectorType Precondition(const VectorType& v) { return v; }
    // the linear solve of each Newton step (GMRES, NewtonKrylov.h:96)
    virtual void SolveLinear(const VectorType& rhs, VectorType& dx, double epsilon, double Delta_) {
        vectorsToReuse = 0;
        GMRES(rhs, dx, epsilon, Delta_);
    }
    // every operator application of GMRES: Ax = A x with A = -dF/dx
    virtual void OnMatVec(const VectorType& x, const VectorType& Ax) {
        (void)x;
        (void)Ax;
    }

    stratifloat T = 11;  // time interval for integration (NewtonKrylov.h:113)

public:
    int K = 2048;  // max Krylov vectors (NewtonKrylov.h:237); set before Run
    void SetKrylovDimension(int k) {
        K = k;
        H = sfdense::MatrixX(K, K - 1);
        vectorsToReuse = 0;
    }

private:
    VectorType linearAboutStart;
    VectorType linearAboutEnd;

    // (F(x + eps v) - F(x)) / eps, NewtonKrylov.h:119-132
    VectorType EvalDerivative(const VectorType& at) {
        const double eps = fdEpsilon * sfkrylov::norm(linearAboutStart) / sfkrylov::norm(at);
        VectorType temp = linearAboutStart;
        temp.MulAdd(stratifloat(eps), at);
        temp = EvalFunction(temp);
        temp -= linearAboutEnd;
        temp *= stratifloat(1 / eps);
        return temp;
    }

    int vectorsToReuse = 0;
    std::vector<VectorType> q;
    std::vector<VectorType> Z;  // preconditioned directions (flexible GMRES)
    sfdense::MatrixX G;         // their FP64 Gram matrix Z^T Z
    sfdense::MatrixX H;         // upper Hessenberg matrix
};
