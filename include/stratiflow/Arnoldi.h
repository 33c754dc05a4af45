// Arnoldi.h -- leading eigenvalues of the linearised flow map (Arnoldi.h:1-227).
//
// Same class template and algorithm as the reference: starting from a
// low-wavenumber excitation, build a Krylov basis of EvalLinearised with
// modified Gram-Schmidt, take the eigenvalues of the Hessenberg matrix, and
// return the largest modulus; the leading three eigenvectors are mapped back
// to state space with any phase-shift component (ddx of the base state)
// removed.  Differences, all deliberate:
//   * Gram-Schmidt coefficients are FP64 GPU reductions (DotF64) and the
//     Hessenberg matrix and its eigen-decomposition are FP64 on the host
//     (sfdense::hessenberg_eigenvalues/eigenvector replace Eigen's EigenSolver);
//   * the per-step progress report is optional (verbose), the eigenvalue
//     decomposition of every intermediate k can be skipped (reportEvery), and
//     writing eigReal/eigImag/... .fields files is optional (saveFiles);
//   * all eigenvalues of the final Hessenberg matrix are kept in
//     `eigenvalues` (complex, sorted by decreasing modulus) for the caller.
#pragma once

#include <algorithm>
#include <complex>
#include <iostream>
#include <vector>

#include "NewtonKrylov.h"
#include "dense.h"
#include "stratiflow.h"

template <typename VectorType>
class Arnoldi {
public:
    Arnoldi() {}
    virtual ~Arnoldi() {}

    stratifloat Run(const VectorType& at, VectorType& result, bool removePhaseShift = true, bool getSecond = false) {
        EvalFunction(at);

        VectorType phaseShift;
        phaseShift.u1 = ddx(at.u1);
        phaseShift.u2 = ddx(at.u2);
        phaseShift.u3 = ddx(at.u3);
        phaseShift.b = ddx(at.b);

        q.assign(1, VectorType());
        if (startVector) {
            q[0] = *startVector;
        } else {
            q[0].ExciteLowWavenumbers(0.1);
        }
        q[0].EnforceBCs();
        q[0] *= stratifloat(1 / sfkrylov::norm(q[0]));

        H = sfdense::MatrixX(K, K - 1);
        for (int k = 1; k < K; k++) {
            q.push_back(EvalLinearised(q[k - 1]));
            for (int j = 0; j < k; j++) {
                H(j, k - 1) = sfkrylov::dot(q[j], q[k]);
                q[k].MulAdd(stratifloat(-H(j, k - 1)), q[j]);
            }
            H(k, k - 1) = sfkrylov::norm(q[k]);
            q[k] *= stratifloat(1 / H(k, k - 1));
            q[k].EnforceBCs();

            if (verbose && reportEvery > 0 && (k % reportEvery == 0)) {
                const sfdense::VectorXc ev = sfdense::hessenberg_eigenvalues(H.block(k, k));
                std::vector<double> m(ev.size());
                for (size_t i = 0; i < ev.size(); ++i) m[i] = std::abs(ev[i]);
                std::sort(m.begin(), m.end(), std::greater<double>());
                std::cout << "At step " << k << ", maximum growth rates: " << (m.size() > 0 ? m[0] : 0) << " & "
                          << (m.size() > 1 ? m[1] : 0) << " & " << (m.size() > 2 ? m[2] : 0) << std::endl;
            }
        }

        const int n = K - 1;
        const sfdense::MatrixX subH = H.block(n, n);
        eigenvalues = sfdense::hessenberg_eigenvalues(subH);
        // decreasing modulus; of a conjugate pair the one with positive imaginary part first
        std::stable_sort(eigenvalues.begin(), eigenvalues.end(), [](const sfdense::cplx& a, const sfdense::cplx& b) {
            const double ma = std::abs(a), mb = std::abs(b);
            if (std::fabs(ma - mb) > 1e-12 * std::max(ma, mb)) return ma > mb;
            return a.imag() > b.imag();
        });

        std::vector<sfdense::VectorXc> vecs;
        for (int e = 0; e < std::min(3, n); ++e) vecs.push_back(sfdense::hessenberg_eigenvector(subH, eigenvalues[e]));

        // phase shift in Arnoldi space (Arnoldi.h:120-133)
        sfdense::VectorX ps(n);
        double ps2 = 0;
        for (int j = 0; j < n; j++) {
            ps[j] = sfkrylov::dot(phaseShift, q[j]);
            ps2 += ps[j] * ps[j];
        }
        for (auto& v : vecs) {
            if (removePhaseShift && ps2 > 0) {
                // Eigen's dot conjugates its first argument: proj = sum conj(v_j) ps_j
                sfdense::cplx proj = 0;
                for (int j = 0; j < n; ++j) proj += std::conj(v[j]) * ps[j];
                for (int j = 0; j < n; ++j) v[j] -= (proj / ps2) * ps[j];
            }
            double nv = 0;
            for (auto& c : v) nv += std::norm(c);
            nv = std::sqrt(nv);
            for (auto& c : v) c /= nv;
        }

        // back to state space (Arnoldi.h:138-156)
        std::vector<VectorType> re(vecs.size()), im(vecs.size());
        for (auto& r : re) r.Zero();
        for (auto& r : im) r.Zero();
        for (int k = 0; k < n; k++) {
            for (size_t e = 0; e < vecs.size(); ++e) {
                re[e].MulAdd(stratifloat(vecs[e][k].real()), q[k]);
                im[e].MulAdd(stratifloat(vecs[e][k].imag()), q[k]);
            }
        }
        if (saveFiles) {
            const char* rn[3] = {"eigReal", "eig2Real", "eig3Real"};
            const char* in[3] = {"eigImag", "eig2Imag", "eig3Imag"};
            for (size_t e = 0; e < vecs.size(); ++e) {
                re[e].SaveToFile(rn[e]);
                im[e].SaveToFile(in[e]);
            }
        }
        if (verbose) {
            for (size_t e = 0; e < vecs.size(); ++e)
                std::cout << "Final eigenvalue: " << eigenvalues[e] << std::endl;
        }
        eigenvectorsReal = re;
        eigenvectorsImag = im;
        if (getSecond && re.size() > 1) {
            result = re[1];
            return stratifloat(std::abs(eigenvalues[1]));
        }
        if (!re.empty()) result = re[0];
        return stratifloat(eigenvalues.empty() ? 0.0 : std::abs(eigenvalues[0]));
    }

protected:
    virtual VectorType EvalFunction(const VectorType& at) = 0;
    virtual VectorType EvalLinearised(const VectorType& at) = 0;

    stratifloat T = 11;  // time interval for integration (Arnoldi.h:199)

    VectorType linearAboutStart;

public:
    int K = 1024;                        // max iterations (Arnoldi.h:204)
    std::vector<VectorType> q;           // Krylov basis
    sfdense::MatrixX H;                  // upper Hessenberg matrix
    sfdense::VectorXc eigenvalues;       // of the final Hessenberg matrix, by decreasing modulus
    std::vector<VectorType> eigenvectorsReal, eigenvectorsImag;  // leading three
    const VectorType* startVector = nullptr;  // replaces the random excitation when set
    bool verbose = true;
    int reportEvery = 1;
    bool saveFiles = false;
};

class BasicArnoldi : public Arnoldi<StateVector> {  // Arnoldi.h:209-227
protected:
    StateVector EvalFunction(const StateVector& at) override {
        StateVector result;
        at.FullEvolve(T, result, false);
        linearAboutStart = at;
        return result;
    }
    StateVector EvalLinearised(const StateVector& at) override {
        StateVector result;
        at.LinearEvolve(T, linearAboutStart, result);
        return result;
    }
};
