// Lead modes and mode-matching self-energies, FP64 on the host.
//
// Follows crates/caroli-green/src/modes.rs: the stabilised 2r pencil of the compact SVD
// `H1 = u v^H` (LeadModes::solve_with, :159-486), the Hermitian definite fast path
// (DefiniteCoupling, :558-711), the energy shifts off band edges (LeadModes::shifted, :130-145)
// and `Sigma_R = u X_g X_a^-1 u^H`, `Sigma_L = v Y_a Y_g^-1 v^H` (:494-511).
#pragma once
#include <optional>
#include <vector>

#include "cmat.hpp"

namespace caroli {

/// Error kinds of a Green's function computation (crates/caroli-green/src/lib.rs:35-47).
enum Status : int {
    kOk = 0,
    kSingular = 1,       // GreenError::Singular(E)
    kEigen = 2,          // GreenError::Eigen
    kModeCount = 3,      // GreenError::ModeCount
    kNoConvergence = 4,  // GreenError::NoConvergence
    kInvalid = 6,        // bad argument at the boundary
    kDevice = 7,         // device error
};

struct LeadModes {
    CMat u, v;            // H1 = u v^H (n x r)
    CMat right, left;     // orthonormal bases of right- and left-going modes (2r x r)
    int channels = 0;
    cplx energy;
    // Scalar coupling (H1 + H1^H = 2 s c I, the Hermitian cell fast path): the bases are
    // Y diag(.) with Y the unitary eigenvectors of the cell, so Sigma_R = c Y diag(s lambda_R) Y^H
    // and Sigma_L = c Y diag(s / lambda_L) Y^H, one product each instead of a solve and two.
    const CMat* cell_y = nullptr;
    double cell_c = 0.0, cell_s = 1.0;
    std::vector<cplx> cell_right, cell_left;  // lambda of the right- and left-going modes
    CMat sigma_right() const;
    CMat sigma_left() const;
};

/// The energy-independent part of a lead's mode problem (modes.rs:519-538).
class ModeSolver {
   public:
    ModeSolver(const CMat& h0, const CMat& h1);
    /// Modes at the real energy `e`, shifted off band edges when needed.
    Status modes_at(double e, LeadModes& out) const;

   private:
    CMat h0_, h1_;
    struct Definite {
        double s;
        CMat l;
        bool cell = false;
        double c = 0;
        std::vector<double> theta;
        CMat y;
    };
    std::optional<Definite> definite_;
};

/// Lead self-energies Sigma_R / Sigma_L of one lead at energy `e` (modes.rs:494-511, :715-724).
Status lead_self_energies(const ModeSolver& solver, double e, CMat* sigma_left, CMat* sigma_right, int* channels);

}  // namespace caroli
