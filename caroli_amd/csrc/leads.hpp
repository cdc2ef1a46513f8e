// Lead self-energies on the GPU (leads.hip).
#pragma once
#include <complex>
#include <cstdint>
#include <vector>

#include "cmat.hpp"

namespace caroli::gpu {

/// A HIP device is present for the lead kernels.
bool lead_device_available();

/// out[z] = Y diag(diags[z]) Y^H for z < count (n x n complex, column-major), in FP64 on the GPU.
/// lead_id names the eigenvector matrix Y (its device copy is kept for the next calls).
void cell_sigmas(uint64_t lead_id, const std::complex<double>* y, int n, const std::complex<double>* diags, int count,
                 std::complex<double>* out);

/// The eigendecomposition of the Hermitian part of `a` in FP64 on the GPU (rocSOLVER's Jacobi
/// zheevj; the host's solver when that does not converge or gives non-finite values): eigenvalues
/// ascending in w, unitary eigenvectors in the columns of v; false when neither converges.
bool hermitian_eigen(const CMat& a, std::vector<double>& w, CMat& v);

}  // namespace caroli::gpu
