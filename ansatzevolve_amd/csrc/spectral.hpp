// Host eigenproblems with the Hamiltonian applied on the device (spectral.cpp).
#pragma once
#include <functional>
#include <vector>

namespace ae {

// Eigenvalues (ascending) of S^-1/2 H S^-1/2 for n states with Gram matrices
// H and S (row-major); eigenvalues of S below 1e-14 are dropped (their
// directions give zero eigenvalues) and counted in nullRank.
void subspace_eigen(int n, const std::vector<double> &H, const std::vector<double> &S, std::vector<double> &evals,
                    int &nullRank);

// The k lowest eigenpairs of the symmetric operator applyH (y = H x over N
// entries) by Lanczos with full reorthogonalisation from `start`; evecs
// holds the k eigenvectors one after the other (k x N).  Converged when
// every residual is below tol * max(1, |eigenvalue|); false if not within
// maxIter steps.
bool lowest_eigenpairs(int k, long long N, const std::function<bool(const double *, double *)> &applyH,
                       const std::vector<double> &start, std::vector<double> &evals, std::vector<double> &evecs,
                       int maxIter, double tol);

} // namespace ae
