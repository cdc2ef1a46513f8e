// Host FP64 dense kernels for the lead problems (no host LAPACK is available): LU with partial
// pivoting, Cholesky, complex Schur decomposition with eigenvalue reordering, singular value and
// Hermitian eigen decompositions, orthonormal bases.
//
// They provide what crates/caroli-dense/src/lib.rs (Lu :172-319, schur :94-101, partition /
// reorder in schur.rs:199-251, svd :347-355, hermitian_eigen :359-367, orthonormal_columns) gives
// the mode solver of crates/caroli-green/src/modes.rs, written here from the textbook algorithms
// (Golub & Van Loan: Householder Hessenberg reduction and the implicit single-shift QR iteration,
// Givens swaps of adjacent eigenvalues; one-sided Jacobi SVD; Householder tridiagonalisation and
// the implicit QL iteration with Wilkinson shifts for Hermitian matrices).
#pragma once
#include <vector>

#include "cmat.hpp"

namespace caroli::dense {

/// LU factorisation with partial pivoting (row interchanges) of a square matrix.
struct Lu {
    CMat lu;
    std::vector<int> piv;  // row exchanged with row k at step k
    explicit Lu(const CMat& a);
    /// Smallest and largest |U_ii| ((inf, 0) when empty).
    std::pair<double, double> pivot_range() const;
    /// b <- a^{-1} b.
    void solve_in_place(CMat& b) const;
    CMat solve(const CMat& b) const {
        CMat x = b;
        solve_in_place(x);
        return x;
    }
    /// Reciprocal 1-norm condition number from the exact inverse (0 when singular).
    double inverse_condition_1(double norm_1) const;
};

double norm_1(const CMat& a);
/// x solving x a = b.
CMat solve_right(const CMat& b, const CMat& a);
CMat inverse(const CMat& a);

/// Lower Cholesky factor of a Hermitian positive definite matrix; false if not definite.
bool cholesky(const CMat& a, CMat& l);

/// Complex Schur form a = q t q^H, t upper triangular. False if the QR iteration fails.
bool schur(const CMat& a, CMat& q, CMat& t);
/// Swaps the adjacent diagonal entries k, k+1 of the Schur form (t, q).
void swap_adjacent(CMat& t, CMat& q, int k);
/// Stable partition of the Schur form: diagonal entries with flags first. Returns their count.
int partition(CMat& t, CMat& q, std::vector<char> flags);

/// Singular values (descending) and vectors: a = u diag(s) v^H (thin, a is m x n).
bool svd(const CMat& a, CMat& u, std::vector<double>& s, CMat& v);

/// Eigenvalues (ascending) and orthonormal eigenvectors of a Hermitian matrix.
bool hermitian_eigen(const CMat& a, std::vector<double>& w, CMat& v);

/// Orthonormal basis of the columns of a full-rank tall matrix (Gram-Schmidt, twice).
CMat orthonormal_columns(const CMat& a);

}  // namespace caroli::dense
