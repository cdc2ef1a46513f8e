// The objects behind the C ABI handles: a system (NestedTwoTerminal) and its lead
// self-energies (LeadSelfEnergies).
#pragma once
#include <memory>
#include <optional>
#include <mutex>

#include "modes.hpp"
#include "symbolic.hpp"

namespace caroli {

namespace gpu {
struct Structure;  // device copy of the plans, shared by systems of one structure
}

/// Algorithmic work of one transmission point: 8 flops per complex multiply-add; 16-byte complex
/// words for one assembly of each front plus one write and one read of its Schur complement.
struct Work {
    double flops = 0, bytes = 0;
};
Work algorithmic_work(const Symbolic& sym, size_t nnz);

}  // namespace caroli

struct cg_system {
    caroli::SystemSpec spec;
    caroli::Symbolic sym;
    /// Lead mode solvers per block, built on the first self-energy request (a system used only
    /// with self-energies computed elsewhere never builds them); the right one only when the
    /// block's leads differ.
    mutable std::vector<std::optional<caroli::ModeSolver>> left_modes, right_modes;
    /// Key of the symbolic structure: systems with equal keys share device plans.
    uint64_t key = 0;
    mutable std::mutex mu;
    mutable std::shared_ptr<caroli::gpu::Structure> structure;
    /// Device copy of the Hamiltonian values (complex f64), uploaded on first use.
    mutable void* d_vals = nullptr;
    /// Multi-terminal device (NestedMultiTerminal): its leads as attached, the orbital kind of each
    /// cell orbital (sector labels), and per block and lead the mode solver, built on the first
    /// self-energy request (leads of a block with equal cells share one, nested_multi.rs:130-145).
    bool multi = false;
    std::vector<caroli::AttachedLead> multi_leads;
    std::vector<std::vector<int>> orbital_kinds;
    mutable std::vector<std::vector<std::shared_ptr<caroli::ModeSolver>>> lead_modes;
};

struct cg_sigma {
    struct Block {
        caroli::CMat sl, sr, gl, gr;  // Sigma_L, Sigma_R and their broadenings
        caroli::CMat wl, wr;          // Gamma = W W^H (one column per channel)
        int channels_l = 0, channels_r = 0;
        caroli::CMat left_h00, left_h01, right_h00, right_h01;
        std::vector<size_t> left_orbitals, right_orbitals;
    };
    double energy = 0;
    std::vector<Block> blocks;
    /// Multi-terminal devices: T is a matrix over labels, one per lead or per (lead, sector).
    /// Each block holds Sigma of all its leads on its kept orbitals in sl (sr is zero, both
    /// kept x kept), and in wab the channels of every label (W W^H = Gamma of that label on the
    /// kept orbitals), label l in columns lab_off[l] .. lab_off[l + 1].
    struct Label {
        int32_t lead, sector;
    };
    bool multi = false;
    std::vector<Label> labels;
    struct MultiBlock {
        caroli::CMat wab;
        std::vector<int32_t> lab_off;
        std::vector<caroli::AttachedLead> leads;  // the block's leads they were computed for
        std::vector<caroli::CMat> sigma;          // Sigma_p on each lead's orbitals in the block
    };
    std::vector<MultiBlock> mblocks;
};

namespace caroli {
/// Gamma = i (Sigma - Sigma^H) (crates/caroli-green/src/lib.rs:68-71).
CMat broadening(const CMat& s);
/// W with W W^H = Gamma for a positive semidefinite Gamma: diagonally pivoted Cholesky, stopped
/// when the largest remaining diagonal falls below 1e-12 of the largest diagonal (the role of
/// gamma_root, crates/caroli-green/src/nested.rs:1038-1047; W differs from it by a unitary,
/// which leaves |W_L^H G W_R|^2 = Tr[Gamma_L G Gamma_R G^H] unchanged).
CMat gamma_factor(const CMat& g);
/// True when `g` holds self-energies of the leads and blocks of `s` (nested.rs:693-707).
bool sigma_matches(const cg_system& s, const cg_sigma& g);
}  // namespace caroli
