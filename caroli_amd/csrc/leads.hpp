// Lead self-energies on the GPU (leads.hip).
#pragma once
#include <complex>
#include <cstdint>

namespace caroli::gpu {

/// A HIP device is present for the lead kernels.
bool lead_device_available();

/// out[z] = Y diag(diags[z]) Y^H for z < count (n x n complex, column-major), in FP64 on the GPU.
/// lead_id names the eigenvector matrix Y (its device copy is kept for the next calls).
void cell_sigmas(uint64_t lead_id, const std::complex<double>* y, int n, const std::complex<double>* diags, int count,
                 std::complex<double>* out);

}  // namespace caroli::gpu
