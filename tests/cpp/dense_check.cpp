// Runs the FP64 dense helpers of include/stratiflow/dense.h on seeded random
// matrices and prints inputs and outputs for the harness to compare with
// numpy's LAPACK routines:
//   M <case> <rows> <cols> <values, column-major>
//   SV <case> <singular values>
//   REC <case> <|A - U diag(d) V^T|_max>   ORTH <case> <|V^T V - I|_max>
//   LS <case> <Delta> <residual> <mu> <y...>
//   EV <case> <re im pairs>
//   EVEC <case> <k> <|H v - lambda v| / |v|>
#include <cstdio>
#include <random>

#include "stratiflow/dense.h"

using namespace sfdense;

static void print_mat(int c, const MatrixX& A) {
    printf("M %d %d %d", c, A.rows, A.cols);
    for (double v : A.a) printf(" %.17e", v);
    printf("\n");
}

int main() {
    std::mt19937 gen(1234);
    std::normal_distribution<double> nd(0.0, 1.0);
    int c = 0;
    // least-squares cases: (k+1) x k upper Hessenberg, as GMRES builds them
    for (int k : {1, 2, 5, 12, 30}) {
        MatrixX H(k + 1, k);
        for (int j = 0; j < k; ++j)
            for (int i = 0; i <= j + 1; ++i) H(i, j) = nd(gen);
        print_mat(c, H);
        MatrixX U, V;
        VectorX d;
        jacobi_svd(H, U, d, V);
        printf("SV %d", c);
        for (double s : d) printf(" %.17e", s);
        printf("\n");
        double rec = 0, orth = 0;
        for (int i = 0; i < H.rows; ++i)
            for (int j = 0; j < H.cols; ++j) {
                double s = 0;
                for (int m = 0; m < k; ++m) s += U(i, m) * d[m] * V(j, m);
                rec = std::max(rec, std::fabs(s - H(i, j)));
            }
        for (int i = 0; i < k; ++i)
            for (int j = 0; j < k; ++j) {
                double s = 0;
                for (int m = 0; m < k; ++m) s += V(m, i) * V(m, j);
                orth = std::max(orth, std::fabs(s - (i == j ? 1.0 : 0.0)));
            }
        printf("REC %d %.3e\nORTH %d %.3e\n", c, rec, c, orth);
        for (double Delta : {100.0, 0.5}) {
            double res = 0, mu = 0;
            const VectorX y = hookstep(H, 2.0, Delta, &res, &mu);
            printf("LS %d %.17e %.17e %.17e", c, Delta, res, mu);
            for (double v : y) printf(" %.17e", v);
            printf("\n");
        }
        ++c;
    }
    // eigenvalue cases: square upper Hessenberg
    for (int n : {1, 2, 3, 6, 15, 40}) {
        MatrixX H(n, n);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i <= j + 1 && i < n; ++i) H(i, j) = nd(gen);
        print_mat(c, H);
        const VectorXc ev = hessenberg_eigenvalues(H);
        printf("EV %d", c);
        for (const cplx& e : ev) printf(" %.17e %.17e", e.real(), e.imag());
        printf("\n");
        for (int k = 0; k < std::min(n, 3); ++k) {
            const VectorXc v = hessenberg_eigenvector(H, ev[k]);
            double r = 0, nv = 0;
            for (int i = 0; i < n; ++i) {
                cplx s = 0;
                for (int j = 0; j < n; ++j) s += H(i, j) * v[j];
                r += std::norm(s - ev[k] * v[i]);
                nv += std::norm(v[i]);
            }
            printf("EVEC %d %d %.3e\n", c, k, std::sqrt(r / nv));
        }
        ++c;
    }
    return 0;
}
