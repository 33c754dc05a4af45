// dense.h -- small FP64 dense linear algebra for the Krylov drivers.
//
// The reference solves the GMRES least-squares problem with Eigen's JacobiSVD
// (NewtonKrylov.h:191-215) and finds Arnoldi eigenvalues with Eigen's
// EigenSolver (Arnoldi.h:57-60, 98-117).  Eigen is not a dependency here, so
// the same decompositions are written here from their textbook definitions:
//   * jacobi_svd: one-sided (Hestenes) Jacobi SVD, A = U diag(d) V^T, thin U,
//     singular values in decreasing order;
//   * hessenberg_eigen: eigenvalues of a real upper Hessenberg matrix by the
//     shifted QR iteration (complex Wilkinson shifts, Givens rotations, with
//     deflation), eigenvectors by inverse iteration on H - lambda I.
// Everything is double precision on the host: these matrices are at most the
// Krylov dimension on a side.
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <vector>

namespace sfdense {

typedef std::complex<double> cplx;

// column-major dense matrix
template <typename S>
struct Mat {
    int rows = 0, cols = 0;
    std::vector<S> a;
    Mat() {}
    Mat(int r, int c) : rows(r), cols(c), a(size_t(r) * size_t(c), S(0)) {}
    S& operator()(int i, int j) { return a[size_t(j) * rows + i]; }
    const S& operator()(int i, int j) const { return a[size_t(j) * rows + i]; }
    void setZero() { std::fill(a.begin(), a.end(), S(0)); }
    // top-left r x c block
    Mat block(int r, int c) const {
        Mat b(r, c);
        for (int j = 0; j < c; ++j)
            for (int i = 0; i < r; ++i) b(i, j) = (*this)(i, j);
        return b;
    }
};
typedef Mat<double> MatrixX;
typedef Mat<cplx> MatrixXc;
typedef std::vector<double> VectorX;
typedef std::vector<cplx> VectorXc;

inline double norm(const VectorX& v) {
    double s = 0;
    for (double x : v) s += x * x;
    return std::sqrt(s);
}

inline VectorX matvec(const MatrixX& A, const VectorX& x) {
    VectorX y(A.rows, 0.0);
    for (int j = 0; j < A.cols; ++j)
        for (int i = 0; i < A.rows; ++i) y[i] += A(i, j) * x[j];
    return y;
}

// One-sided Jacobi SVD of A (m x n, m >= n): A = U diag(d) V^T with U m x n
// (orthonormal columns where d > 0), V n x n orthogonal, d descending.
inline void jacobi_svd(const MatrixX& A, MatrixX& U, VectorX& d, MatrixX& V) {
    const int m = A.rows, n = A.cols;
    if (m < n) throw std::invalid_argument("jacobi_svd: rows < cols");
    U = A;
    V = MatrixX(n, n);
    for (int i = 0; i < n; ++i) V(i, i) = 1.0;
    const double eps = std::numeric_limits<double>::epsilon();
    for (int sweep = 0; sweep < 100; ++sweep) {
        bool rotated = false;
        for (int p = 0; p < n - 1; ++p) {
            for (int q = p + 1; q < n; ++q) {
                double alpha = 0, beta = 0, gamma = 0;
                for (int i = 0; i < m; ++i) {
                    alpha += U(i, p) * U(i, p);
                    beta += U(i, q) * U(i, q);
                    gamma += U(i, p) * U(i, q);
                }
                if (gamma == 0.0 || std::fabs(gamma) <= eps * std::sqrt(alpha * beta)) continue;
                rotated = true;
                // rotation that zeroes the (p, q) entry of U^T U
                const double zeta = (beta - alpha) / (2.0 * gamma);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / std::sqrt(1.0 + t * t), s = c * t;
                for (int i = 0; i < m; ++i) {
                    const double up = U(i, p), uq = U(i, q);
                    U(i, p) = c * up - s * uq;
                    U(i, q) = s * up + c * uq;
                }
                for (int i = 0; i < n; ++i) {
                    const double vp = V(i, p), vq = V(i, q);
                    V(i, p) = c * vp - s * vq;
                    V(i, q) = s * vp + c * vq;
                }
            }
        }
        if (!rotated) break;
    }
    d.assign(n, 0.0);
    for (int j = 0; j < n; ++j) {
        double s = 0;
        for (int i = 0; i < m; ++i) s += U(i, j) * U(i, j);
        d[j] = std::sqrt(s);
        if (d[j] > 0)
            for (int i = 0; i < m; ++i) U(i, j) /= d[j];
    }
    // sort by decreasing singular value
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return d[a] > d[b]; });
    MatrixX U2(m, n), V2(n, n);
    VectorX d2(n);
    for (int k = 0; k < n; ++k) {
        d2[k] = d[order[k]];
        for (int i = 0; i < m; ++i) U2(i, k) = U(i, order[k]);
        for (int i = 0; i < n; ++i) V2(i, k) = V(i, order[k]);
    }
    U = U2;
    V = V2;
    d = d2;
}

// Least squares min |H y - beta e1| in the singular basis of H, with the
// trust-region limit |z| <= Delta applied as in NewtonKrylov.h:196-215
// (z_j = p_j d_j / (d_j^2 + mu), mu the first multiple of 1e-5 that meets the
// bound).  Returns y; the
// relative residual |H y - beta e1| / beta goes to *residual and the final mu
// to *mu_out.
inline VectorX hookstep(const MatrixX& H, double beta, double Delta, double* residual, double* mu_out) {
    MatrixX U, V;
    VectorX d;
    jacobi_svd(H, U, d, V);
    const int n = H.cols;
    VectorX p(n), z(n);
    for (int j = 0; j < n; ++j) p[j] = U(0, j) * beta;  // p = U^T (beta e1)
    for (int j = 0; j < n; ++j) z[j] = d[j] > 0 ? p[j] / d[j] : std::numeric_limits<double>::infinity();
    double mu = 0;
    if (Delta > 0 && norm(z) > Delta) {
        // The reference raises mu in steps of 1e-5 until |z| <= Delta.  |z(mu)|
        // decreases monotonically in mu, so the first grid point m*1e-5 that
        // satisfies the bound is found by doubling and bisection over m.
        auto znorm = [&](double m) {
            double s2 = 0;
            for (int j = 0; j < n; ++j) {
                const double zj = p[j] * d[j] / (d[j] * d[j] + m * 1e-5);
                s2 += zj * zj;
            }
            return std::sqrt(s2);
        };
        double lo = 0, hi = 1;  // |z(lo)| > Delta >= |z(hi)| once hi is found
        while (znorm(hi) > Delta) {
            lo = hi;
            hi *= 2;
        }
        while (hi - lo > 1) {
            const double mid = std::floor(0.5 * (lo + hi));
            if (znorm(mid) > Delta) lo = mid;
            else hi = mid;
        }
        mu = hi * 1e-5;
        for (int j = 0; j < n; ++j) z[j] = p[j] * d[j] / (d[j] * d[j] + mu);
    }
    VectorX y(n, 0.0);
    for (int k = 0; k < n; ++k)
        for (int i = 0; i < n; ++i) y[i] += V(i, k) * z[k];
    VectorX r = matvec(H, y);
    r[0] -= beta;
    if (residual) *residual = norm(r) / beta;
    if (mu_out) *mu_out = mu;
    return y;
}

// Upper Cholesky factor R (n x n) of a symmetric positive definite G (the
// top-left n x n block of G is read): G = R^T R.  Returns false when a pivot
// is not positive (G singular to working precision).
inline bool cholesky_upper(const MatrixX& G, int n, MatrixX& R) {
    R = MatrixX(n, n);
    for (int j = 0; j < n; ++j) {
        double s = G(j, j);
        for (int k = 0; k < j; ++k) s -= R(k, j) * R(k, j);
        if (!(s > 0)) return false;
        R(j, j) = std::sqrt(s);
        for (int i = j + 1; i < n; ++i) {
            double t = G(j, i);
            for (int k = 0; k < j; ++k) t -= R(k, j) * R(k, i);
            R(j, i) = t / R(j, j);
        }
    }
    return true;
}

// Hookstep in a metric: min |H y - beta e1| subject to sqrt(y^T G y) <= Delta,
// with G = R^T R.  With w = R y the problem is the plain hookstep on H R^-1,
// so the trust region bounds |Z y| for the directions Z whose Gram matrix is
// G (the flexible GMRES update dx = Z y).  Same outputs as hookstep.
inline VectorX hookstep_metric(const MatrixX& H, const MatrixX& R, double beta, double Delta, double* residual,
                               double* mu_out) {
    const int m = H.rows, n = H.cols;
    // HR = H R^-1, column by column (R upper triangular)
    MatrixX HR(m, n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) {
            double t = H(i, j);
            for (int k = 0; k < j; ++k) t -= HR(i, k) * R(k, j);
            HR(i, j) = t / R(j, j);
        }
    VectorX w = hookstep(HR, beta, Delta, residual, mu_out);
    // y = R^-1 w
    VectorX y(n, 0.0);
    for (int i = n - 1; i >= 0; --i) {
        double t = w[i];
        for (int k = i + 1; k < n; ++k) t -= R(i, k) * y[k];
        y[i] = t / R(i, i);
    }
    return y;
}

// ---- eigenvalues of a real upper Hessenberg matrix -------------------------

// Wilkinson shift: eigenvalue of the trailing 2x2 block closer to its (1,1)
// corner entry
inline cplx wilkinson(cplx a, cplx b, cplx c, cplx dd) {
    const cplx tr = a + dd, det = a * dd - b * c;
    const cplx disc = std::sqrt(tr * tr / 4.0 - det);
    const cplx l1 = tr / 2.0 + disc, l2 = tr / 2.0 - disc;
    return std::abs(l1 - dd) < std::abs(l2 - dd) ? l1 : l2;
}

// Eigenvalues of the upper Hessenberg matrix H (n x n) by shifted QR.
inline VectorXc hessenberg_eigenvalues(const MatrixX& Hr) {
    const int n = Hr.rows;
    MatrixXc H(n, n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) H(i, j) = (i <= j + 1) ? cplx(Hr(i, j), 0) : cplx(0, 0);
    VectorXc ev(n);
    const double eps = std::numeric_limits<double>::epsilon();
    int hi = n - 1;
    int iter = 0;
    while (hi >= 0) {
        if (hi == 0) {
            ev[0] = H(0, 0);
            break;
        }
        // find the start of the active unreduced block
        int lo = hi;
        while (lo > 0) {
            const double s = std::abs(H(lo - 1, lo - 1)) + std::abs(H(lo, lo));
            if (std::abs(H(lo, lo - 1)) <= eps * (s > 0 ? s : 1.0)) {
                H(lo, lo - 1) = 0;
                break;
            }
            --lo;
        }
        if (lo == hi) {  // 1x1 block deflated
            ev[hi] = H(hi, hi);
            --hi;
            iter = 0;
            continue;
        }
        if (++iter > 100 * n) throw std::runtime_error("hessenberg_eigenvalues: no convergence");
        cplx mu = wilkinson(H(hi - 1, hi - 1), H(hi - 1, hi), H(hi, hi - 1), H(hi, hi));
        if (iter % 11 == 10) mu += std::abs(H(hi, hi - 1));  // exceptional shift
        // QR step on rows/cols lo..hi of H - mu I with Givens rotations
        std::vector<cplx> cs(hi - lo), sn(hi - lo);
        for (int k = lo; k <= hi; ++k) H(k, k) -= mu;
        for (int k = lo; k < hi; ++k) {
            const cplx x = H(k, k), y = H(k + 1, k);
            const double r = std::sqrt(std::norm(x) + std::norm(y));
            cplx c = 1, s = 0;
            if (r > 0) {
                c = x / r;
                s = y / r;
            }
            cs[k - lo] = c;
            sn[k - lo] = s;
            // rows k, k+1 <- G^H rows
            for (int j = k; j < n; ++j) {
                const cplx a = H(k, j), b = H(k + 1, j);
                H(k, j) = std::conj(c) * a + std::conj(s) * b;
                H(k + 1, j) = -s * a + c * b;
            }
        }
        for (int k = lo; k < hi; ++k) {
            const cplx c = cs[k - lo], s = sn[k - lo];
            // cols k, k+1 <- cols G
            for (int i = 0; i <= std::min(k + 2, hi); ++i) {
                const cplx a = H(i, k), b = H(i, k + 1);
                H(i, k) = a * c + b * s;
                H(i, k + 1) = -a * std::conj(s) + b * std::conj(c);
            }
        }
        for (int k = lo; k <= hi; ++k) H(k, k) += mu;
    }
    // real input: snap near-real values and order conjugate pairs (+imag first)
    for (auto& e : ev)
        if (std::fabs(e.imag()) <= 1e-13 * std::max(1.0, std::abs(e))) e = cplx(e.real(), 0);
    return ev;
}

// Solve (H - lambda I) x = b for Hessenberg H by Gaussian elimination with
// partial pivoting between adjacent rows.
inline VectorXc hessenberg_shift_solve(const MatrixX& Hr, cplx lambda, VectorXc b) {
    const int n = Hr.rows;
    MatrixXc A(n, n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) A(i, j) = cplx(Hr(i, j), 0) - (i == j ? lambda : cplx(0, 0));
    const double tiny = 1e-14 * (1.0 + std::abs(lambda));
    for (int k = 0; k < n - 1; ++k) {
        if (std::abs(A(k + 1, k)) > std::abs(A(k, k))) {
            for (int j = k; j < n; ++j) std::swap(A(k, j), A(k + 1, j));
            std::swap(b[k], b[k + 1]);
        }
        if (std::abs(A(k, k)) < tiny) A(k, k) = tiny;
        const cplx f = A(k + 1, k) / A(k, k);
        for (int j = k; j < n; ++j) A(k + 1, j) -= f * A(k, j);
        b[k + 1] -= f * b[k];
    }
    if (std::abs(A(n - 1, n - 1)) < tiny) A(n - 1, n - 1) = tiny;
    VectorXc x(n);
    for (int i = n - 1; i >= 0; --i) {
        cplx s = b[i];
        for (int j = i + 1; j < n; ++j) s -= A(i, j) * x[j];
        x[i] = s / A(i, i);
    }
    return x;
}

// unit eigenvector of H for the eigenvalue lambda (inverse iteration)
inline VectorXc hessenberg_eigenvector(const MatrixX& H, cplx lambda) {
    const int n = H.rows;
    VectorXc x(n, cplx(1.0, 0.0));
    for (int it = 0; it < 3; ++it) {
        x = hessenberg_shift_solve(H, lambda, x);
        double s = 0;
        for (auto& v : x) s += std::norm(v);
        s = std::sqrt(s);
        for (auto& v : x) v /= s;
    }
    // fix the phase: largest component real and positive
    int im = 0;
    for (int i = 1; i < n; ++i)
        if (std::abs(x[i]) > std::abs(x[im])) im = i;
    const cplx ph = std::abs(x[im]) > 0 ? std::conj(x[im]) / std::abs(x[im]) : cplx(1, 0);
    for (auto& v : x) v *= ph;
    return x;
}

}  // namespace sfdense
