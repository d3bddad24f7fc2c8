// Small dense and Krylov eigenproblems on the host, with the Hamiltonian
// applied on the device:
//  * subspace_eigen: the Hamiltonian in the span of a few states, Loewdin
//    symmetrised, S^-1/2 H S^-1/2 with eigenvalues of S below 1e-14
//    dropped (TUPSQuantities::doSubspaceDiagonalisation,
//    src/tupsquantities.cpp:1600-1768);
//  * lowest_eigenpairs: the k lowest eigenpairs of H by Lanczos with full
//    reorthogonalisation (the ground-state overlaps of writeProperties,
//    src/tupsquantities.cpp:226-270, which diagonalise densely or by power
//    iteration instead).
#include <algorithm>
#include <cmath>

#include "newton.hpp"
#include "spectral.hpp"

namespace ae {

void subspace_eigen(int n, const std::vector<double> &H, const std::vector<double> &S, std::vector<double> &evals,
                    int &nullRank)
{
    std::vector<double> w, V;
    symmetric_eigen(n, S, w, V);
    nullRank = 0;
    for (double &x : w) {
        if (x < 1e-14) {  // negative (numerically) or null directions
            x = 0;
            nullRank++;
        } else {
            x = 1.0 / std::sqrt(x);
        }
    }
    // X = V diag(w) V^T, then X H X
    std::vector<double> X((size_t)n * n, 0.0), T((size_t)n * n, 0.0), A((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double s = 0;
            for (int k = 0; k < n; k++) s += V[(size_t)i * n + k] * w[k] * V[(size_t)j * n + k];
            X[(size_t)i * n + j] = s;
        }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double s = 0;
            for (int k = 0; k < n; k++) s += X[(size_t)i * n + k] * H[(size_t)k * n + j];
            T[(size_t)i * n + j] = s;
        }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double s = 0;
            for (int k = 0; k < n; k++) s += T[(size_t)i * n + k] * X[(size_t)k * n + j];
            A[(size_t)i * n + j] = s;
        }
    for (int i = 0; i < n; i++)  // exact symmetry for the Jacobi solver
        for (int j = 0; j < i; j++) A[(size_t)i * n + j] = A[(size_t)j * n + i] = 0.5 * (A[(size_t)i * n + j] + A[(size_t)j * n + i]);
    std::vector<double> U;
    symmetric_eigen(n, A, evals, U);
}

bool lowest_eigenpairs(int k, long long N, const std::function<bool(const double *, double *)> &applyH,
                       const std::vector<double> &start, std::vector<double> &evals, std::vector<double> &evecs,
                       int maxIter, double tol)
{
    if (N < 1 || k < 1) return false;
    k = (int)std::min<long long>(k, N);
    const int mmax = (int)std::min<long long>(N, std::max(maxIter, 2 * k + 10));
    std::vector<std::vector<double>> Q;  // orthonormal Krylov basis
    std::vector<double> alpha, beta;
    std::vector<double> q(start);
    if ((long long)q.size() != N) q.assign(N, 0.0);
    // a fixed pseudo-random admixture, so that no eigenvector is missed
    // because the start state happens to be orthogonal to it
    unsigned long long r = 0x9e3779b97f4a7c15ull;
    for (long long i = 0; i < N; i++) {
        r ^= r << 13;
        r ^= r >> 7;
        r ^= r << 17;
        q[i] += 1e-3 * ((double)(r >> 11) * (1.0 / 9007199254740992.0) - 0.5);
    }
    auto nrm = [&](const std::vector<double> &v) {
        double s = 0;
        for (double x : v) s += x * x;
        return std::sqrt(s);
    };
    double b = nrm(q);
    for (auto &x : q) x /= b;
    std::vector<double> z(N);
    std::vector<double> prevVals;
    for (int j = 0; j < mmax; j++) {
        Q.push_back(q);
        if (!applyH(Q.back().data(), z.data())) return false;
        double a = 0;
        for (long long i = 0; i < N; i++) a += Q.back()[i] * z[i];
        alpha.push_back(a);
        // full reorthogonalisation (twice) against the whole basis
        for (int pass = 0; pass < 2; pass++)
            for (const auto &v : Q) {
                double c = 0;
                for (long long i = 0; i < N; i++) c += v[i] * z[i];
                for (long long i = 0; i < N; i++) z[i] -= c * v[i];
            }
        b = nrm(z);
        // Ritz values of the tridiagonal matrix so far
        const int m = (int)alpha.size();
        std::vector<double> Tm((size_t)m * m, 0.0), w, Y;
        for (int i = 0; i < m; i++) {
            Tm[(size_t)i * m + i] = alpha[i];
            if (i + 1 < m) Tm[(size_t)i * m + i + 1] = Tm[(size_t)(i + 1) * m + i] = beta[i];
        }
        symmetric_eigen(m, Tm, w, Y);
        bool done = b < 1e-12 || m == mmax;  // invariant subspace or basis exhausted
        if (!done && m >= k) {
            // residual of Ritz pair t: |beta_m * y_t[m-1]|
            done = true;
            for (int t = 0; t < k && done; t++) {
                const double res = std::fabs(b * Y[(size_t)(m - 1) * m + t]);
                done = res <= tol * std::max(1.0, std::fabs(w[t]));
            }
        }
        if (done) {
            const int kk = std::min(k, m);
            evals.assign(w.begin(), w.begin() + kk);
            evecs.assign((size_t)kk * N, 0.0);
            for (int t = 0; t < kk; t++)
                for (int i = 0; i < m; i++) {
                    const double c = Y[(size_t)i * m + t];
                    for (long long x = 0; x < N; x++) evecs[(size_t)t * N + x] += c * Q[i][x];
                }
            return kk == k;
        }
        beta.push_back(b);
        for (long long i = 0; i < N; i++) q[i] = z[i] / b;
    }
    return false;
}

} // namespace ae
