// Symbolic phase of nested dissection: the geometric elimination tree, the front boundaries,
// the assembly plans and the independent blocks of orbitals.
//
// Follows crates/caroli-green/src/nested.rs: `SiteTree::new`/`split` (:61-71, :1051-1093),
// `SiteTree::tree` (:74-96), `symbolic` (:1097-1122), `Plan::new`/`front_vars`/
// `separator_entries` (:127-179), `Components` (:378-420) and the block selection of
// `NestedTwoTerminal::from_model` (:559-609).
#pragma once
#include <vector>

#include "model.hpp"

namespace caroli {

/// Domains with at most this many orbitals are not split further (nested.rs:34).
constexpr int kLeafOrbitals = 16;
/// Fronts with fewer variables than this take the unblocked elimination (nested.rs:37).
constexpr int kSmallFront = 64;

struct Node {
    std::vector<size_t> eliminate;  // sorted orbitals eliminated here
    std::vector<size_t> boundary;   // sorted orbitals eliminated later or never
    std::vector<int> children;
};

/// An H entry of a front: F[a, b] -= (conj ? conj(vals[k]) : vals[k]).
struct PlanEntry {
    int a, b;
    int64_t k;
    bool conj;
};

/// Assembly of a front when every child hands over exactly its boundary.
struct Plan {
    std::vector<size_t> vars;  // eliminated here (sorted), then the boundary
    int ne = 0;
    std::vector<int> diagonal;  // positions receiving z
    std::vector<PlanEntry> entries;
    /// For each child, the front position of each of its boundary variables.
    std::vector<std::vector<int>> child_pos;
};

struct Tree {
    std::vector<Node> nodes;  // post-order
    std::vector<int> top;     // children of the root (0 or 1 node)
    std::vector<Plan> plans;
    /// Height of each node (leaves 0): nodes of one height are independent.
    std::vector<int> height;
    int levels = 0;
};

/// `(sites eliminated, children)` of each node in post-order, and the children of the root.
struct SiteTree {
    std::vector<std::pair<std::vector<size_t>, std::vector<int>>> nodes;
    std::vector<int> top;
};

SiteTree site_tree(const SystemSpec& s, const std::vector<size_t>& interior);

/// Position of each variable of a front (sorted lookup).
struct Positions {
    std::vector<std::pair<size_t, int>> v;
    explicit Positions(const std::vector<size_t>& vars);
    int get(size_t x) const;  // -1 if absent
};

/// Front variables of `node` given what its children hand over: (vars, ne) (nested.rs:147-159).
std::pair<std::vector<size_t>, int> front_vars(const Node& node, const std::vector<const std::vector<size_t>*>& handed);
/// H entries the separator of `node` brings to a front with positions `pos` (nested.rs:164-179).
std::vector<PlanEntry> separator_entries(const SparseMatrix& h, const Node& node, const Positions& pos);

Tree build_tree(const SiteTree& st, const SystemSpec& s, const std::vector<char>& mine);

/// One independent block of orbitals with its tree and its leads (nested.rs:425-436).
struct Block {
    Tree tree;
    CMat left_h00, left_h01, right_h00, right_h01;
    std::vector<size_t> left_orbitals, right_orbitals;  // device orbitals of the end slices
    std::vector<size_t> kept;                           // left, then right unless the same
    int right = 0;                                      // position of the right slice in kept
    bool same_lead = false;                             // right lead equal to the left one
    /// Multi-terminal blocks (nested_multi.rs:25-34): each lead restricted to the block (no
    /// orbitals where it does not reach it) and its cell orbitals in the block. Such a block has
    /// left = right = kept, so the root's border block holds -G on all kept orbitals.
    bool multi = false;
    std::vector<AttachedLead> leads;
    std::vector<std::vector<size_t>> lead_cells;
};

struct Symbolic {
    std::vector<Block> blocks;
};

Symbolic symbolic_phase(const SystemSpec& s);
/// NestedMultiTerminal::from_model (nested_multi.rs:60-161): the tree over the sites that no lead
/// couples to; every orbital of the other sites stays at the root; the independent blocks of
/// orbitals (device, lead cells and couplings joined) that some lead reaches.
Symbolic multi_symbolic(const SystemSpec& s, const std::vector<AttachedLead>& leads);

}  // namespace caroli
