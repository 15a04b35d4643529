// Symmetric tridiagonal eigen tools for the Lanczos matrix T (diagonal a[0..m), off-diagonal
// b[1..m), T(i,i+1) = b[i+1]), the matrix TridiagonalMatrix::buildDenseMatrix forms
// (PsimagLite/TridiagonalMatrix.h:149-163). The reference diagonalizes it with LAPACK dsyev
// (Matrix.cpp:15-45); here the lowest eigenvalue comes from Sturm-sequence bisection and its
// eigenvector from inverse iteration, which is all the ground-state solver needs.
#pragma once

#include <vector>

namespace dmk {

// Number of eigenvalues of T strictly below x.
int sturm_count(const double* a, const double* b, int m, double x);

// k-th smallest eigenvalue (k = 0 lowest) of T to full double precision.
double tridiag_eigenvalue(const double* a, const double* b, int m, int k);

// Unit eigenvector of T for the (accurate) eigenvalue lambda, by inverse iteration with a
// partially pivoted LU of T - lambda I.
std::vector<double> tridiag_eigenvector(const double* a, const double* b, int m, double lambda);

// All eigenpairs of T, ascending: evals[k], and evecs column-major m x m (column k is the unit
// eigenvector of evals[k]). Implicit QL iteration with Wilkinson shifts (the EISPACK tql2
// method), rotations accumulated into the eigenvector matrix; the Krylov propagator needs the
// whole spectrum of T (PsimagLite::diag(T, eigs, 'V') in TimeVectorsKrylov.h:224-225).
void tridiag_eigen_all(const double* a, const double* b, int m, double* evals, double* evecs);

} // namespace dmk
