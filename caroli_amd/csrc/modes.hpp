// Lead modes and mode-matching self-energies, FP64 on the host.
//
// Follows crates/caroli-green/src/modes.rs: the stabilised 2r pencil of the compact SVD
// `H1 = u v^H` (LeadModes::solve_with, :159-486), the Hermitian definite fast path
// (DefiniteCoupling, :558-711), the energy shifts off band edges (LeadModes::shifted, :130-145)
// and `Sigma_R = u X_g X_a^-1 u^H`, `Sigma_L = v Y_a Y_g^-1 v^H` (:494-511).
#pragma once
#include <cstdint>
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
    /// The propagating modes psi_n = phi lambda^n on a cell (phi in columns), normalised to unit
    /// current J(a, b) = i(a_n^H H1 b_{n+1} - a_{n+1}^H H1^H b_n), and their Bloch factors
    /// (modes.rs:64-72): right-going (J = +1) and left-going (J = -1).
    CMat right_modes, left_modes;
    std::vector<cplx> right_lambdas, left_lambdas;
    // Scalar coupling (H1 + H1^H = 2 s c I, the Hermitian cell fast path): the bases are
    // Y diag(.) with Y the unitary eigenvectors of the cell, so Sigma_R = c Y diag(s lambda_R) Y^H
    // and Sigma_L = c Y diag(s / lambda_L) Y^H, one product each instead of a solve and two.
    const CMat* cell_y = nullptr;
    double cell_c = 0.0, cell_s = 1.0;
    std::vector<cplx> cell_right, cell_left;  // lambda of the right- and left-going modes
    CMat sigma_right() const;
    CMat sigma_left() const;
};

/// The modes of a lead with a scalar coupling (H1 + H1^H = 2 s c I) at one energy, per eigenvector
/// of the cell (modes.rs:621-710 with P = c 1): lambda of its right- and left-going mode, and which
/// ones propagate. Sigma_R = Y diag(c s lambda_R) Y^H and Sigma_L = Y diag(c s / lambda_L) Y^H.
struct CellModes {
    const CMat* y = nullptr;  // unitary eigenvectors of the cell (owned by the solver)
    uint64_t lead_id = 0;     // the solver's id (names y)
    double c = 0.0, s = 1.0, energy = 0.0;
    std::vector<cplx> right, left;
    std::vector<int> propagating;
    std::vector<cplx> sigma_right_diag() const;
    std::vector<cplx> sigma_left_diag() const;
};

/// The energy-independent part of a lead's mode problem (modes.rs:519-538).
/// A Hermitian eigensolver (eigenvalues ascending, unitary eigenvectors; false: no convergence).
using HermitianEigen = bool (*)(const CMat& a, std::vector<double>& w, CMat& v);

class ModeSolver {
   public:
    /// eigen: the solver of the scalar-coupling cell's eigendecomposition (the host's when null).
    ModeSolver(const CMat& h0, const CMat& h1, HermitianEigen eigen = nullptr);
    /// Modes at the real energy `e`, shifted off band edges when needed.
    Status modes_at(double e, LeadModes& out) const;
    /// Modes at a complex energy z (modes.rs:531-537): the definite fast path only on the real
    /// axis, else the general pencil, with the same shifts.
    Status modes_at(cplx z, LeadModes& out) const;
    /// The scalar-coupling case with the cell's eigenvectors cached (cell_modes_at applies).
    bool cell() const { return definite_ && definite_->cell; }
    /// The cell modes at `e`, shifted off band edges as modes_at does (no bases built).
    Status cell_modes_at(double e, CellModes& out) const;
    uint64_t id() const { return id_; }

   private:
    uint64_t id_;
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
