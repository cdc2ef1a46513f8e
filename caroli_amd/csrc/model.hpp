// Host-side input builders: lattices, regions, tight-binding models, the sparse Hamiltonian of a
// region, the lead cells, reproducible random streams and disorder.
//
// Behaviour follows caroli-lattice (crates/caroli-lattice/src/lib.rs), caroli-model
// (crates/caroli-model/src/{lib.rs,sparse.rs,presets.rs,gamma.rs}) and caroli-disorder
// (crates/caroli-disorder/src/{lib.rs,rng.rs}): same site order, same orbital layout, same
// sparse entries and the same random draws, so a seed gives the same Hamiltonian.
#pragma once
#include <array>
#include <functional>
#include <map>
#include <string>
#include <vector>

#include "cmat.hpp"

namespace caroli {

using Cell = std::array<int64_t, 3>;

struct Sublattice {
    std::array<double, 3> offset;
    int orbitals;
};

/// Bravais lattice with a basis (crates/caroli-lattice/src/lib.rs:27-62).
struct Lattice {
    std::vector<std::array<double, 3>> vectors;
    std::vector<Sublattice> subs;
    int dim() const { return int(vectors.size()); }
    static Lattice square(double a, int orbitals);
    static Lattice cubic(double a, int orbitals);
    static Lattice honeycomb(double a, int orbitals);
};

struct Site {
    Cell cell;
    int sub;
    bool operator<(const Site& o) const {
        return cell != o.cell ? cell < o.cell : sub < o.sub;
    }
    bool operator==(const Site& o) const { return cell == o.cell && sub == o.sub; }
};

/// Ordered set of sites; site i holds orbitals offset(i) .. offset(i) + orbitals(i)
/// (crates/caroli-lattice/src/lib.rs:98-194).
struct Region {
    Lattice lattice;
    std::vector<Site> sites;
    std::vector<size_t> offsets;
    size_t orbitals = 0;
    std::map<Site, size_t> index;

    static Region from_sites(const Lattice& lat, std::vector<Site> sites);
    /// First axis slowest, sublattice fastest (lib.rs:126-140).
    static Region cuboid(const Lattice& lat, std::array<size_t, 3> extent);
    int site_orbitals(size_t i) const { return lattice.subs[sites[i].sub].orbitals; }
    long find(const Site& s) const {
        auto it = index.find(s);
        return it == index.end() ? -1 : long(it->second);
    }
    /// Slices along `axis`, by increasing coordinate, each in region order (lib.rs:188-194).
    std::vector<Region> slices(int axis) const;
    Region translated(Cell d) const;
};

struct Hopping {
    int from, to;
    Cell offset;
    CMat matrix;
};

/// Square complex CSR matrix with sorted columns (crates/caroli-model/src/sparse.rs:13-44).
struct SparseMatrix {
    size_t n = 0;
    std::vector<int64_t> row_ptr;
    std::vector<uint32_t> cols;
    std::vector<cplx> vals;
    /// Duplicates summed, columns sorted in each row (sparse.rs:23-44).
    static SparseMatrix from_triplets(size_t n, std::vector<std::tuple<size_t, size_t, cplx>> t);
};

/// Tight-binding model: on-site matrix per sublattice, hoppings, extra site-dependent on-site
/// terms (crates/caroli-model/src/lib.rs:188-470).
struct Model {
    Lattice lattice;
    std::vector<CMat> onsite;
    std::vector<Hopping> hoppings;
    std::vector<std::function<bool(const Site&, CMat&)>> extra;

    explicit Model(Lattice lat);
    Model& set_onsite(int sub, const CMat& m);
    Model& hopping(int from, int to, Cell offset, const CMat& m);
    CMat onsite_matrix(const Site& s) const;
    SparseMatrix sparse_hamiltonian(const Region& r) const;
    CMat hamiltonian(const Region& r) const;
    CMat coupling(const Region& a, const Region& b) const;

   private:
    template <class Sink>
    void build(Sink& sink, const Region& r) const;
    template <class Sink>
    void couplings_one_way(Sink& sink, const Region& a, const Region& b) const;
};

/// BHZ parameters (crates/caroli-model/src/presets.rs:11-27).
struct BhzParams {
    double a = 1.0, b = -0.376406, c = 0.0, d = -0.280933, mass = -0.137, lambda = 0.0;
};
Model bhz(const BhzParams& p);
Model square_model();
Model cubic_model();
Model graphene_model();
/// 4x4 Gamma matrices of the BHZ basis (crates/caroli-model/src/gamma.rs).
CMat gamma_matrix(const std::string& name);

/// xoshiro256** with splitmix64 and gfortran seeding (crates/caroli-disorder/src/rng.rs:10-90).
struct Xoshiro256 {
    uint64_t s[4];
    static Xoshiro256 seed_from_u64(uint64_t seed);
    static Xoshiro256 gfortran_put(const int32_t put[8]);
    static Xoshiro256 gfortran(int32_t seed);
    uint64_t next_u64();
    double uniform();
    size_t integer(size_t lo, size_t hi);
};

/// Partial Fisher-Yates (crates/caroli-disorder/src/lib.rs:18-26).
std::vector<Site> choose_sites(Xoshiro256& rng, const std::vector<Site>& cand, size_t n);

/// A system as the transport engine receives it: the sparse Hamiltonian of the region, the
/// geometry the elimination tree is cut from, and the two lead cells.
struct SystemSpec {
    SparseMatrix h;
    std::vector<Cell> cells;      // per site
    std::vector<int> site_orbitals;
    std::vector<size_t> site_offsets;
    int dim = 2;
    std::array<int64_t, 3> range{1, 1, 1};  // hopping range per axis
    int axis = 0;
    CMat left_h00, left_h01, right_h00, right_h01;
};

/// What `NestedTwoTerminal::from_model(device, lead, region, axis)` reads from its models
/// (crates/caroli-green/src/nested.rs:530-557, gauge.rs:81-98 without a field).
SystemSpec system_from_models(const Model& device, const Model& lead, const Region& region, int axis);

/// A semi-infinite lead attached to a device (crates/caroli-green/src/multi.rs:15-25): its cell,
/// the coupling from a cell to the next one further from the device, the device orbitals coupled
/// to its first cell (increasing) and that coupling (orbitals x cell orbitals).
struct AttachedLead {
    CMat h00, h01;
    std::vector<size_t> orbitals;
    CMat coupling;
};

/// A multi-terminal device (NestedMultiTerminal::from_model, nested_multi.rs:60-161): the device
/// Hamiltonian and geometry in `base` (its two-terminal lead fields unused), the leads, and for
/// each lead the index within its site of every orbital of its cell (sector labels).
struct MultiSpec {
    SystemSpec base;
    std::vector<AttachedLead> leads;
    std::vector<std::vector<int>> orbital_kinds;
};

/// The leads of `lead` continuing each interface slice outward (attach_leads, multi.rs:112-140,
/// without a magnetic field): cell = the slice, h01 = its coupling to the slice moved one step
/// outward, coupling = every device orbital the first lead cell reaches.
MultiSpec multi_from_models(const Model& device, const Model& lead, const Region& region,
                            const std::vector<Region>& slices, const std::vector<Cell>& outward);

/// Preset model 0 BHZ, 1 square, 2 zigzag graphene, 3 cubic (the clean model of preset_system).
Model preset_model(int model);
/// `base` with an extra on-site potential on every orbital of each listed site.
Model with_site_potential(const Model& base, const Region& region, const double* potential);

/// Named workloads: the model, region and seeded disorder of a preset system.
///  model 0: BHZ (BhzParams::default());  1: square lattice, t = 1;
///        2: zigzag graphene, t = 1 (crates/caroli-bench/src/main.rs:86-91);  3: cubic, t = 1.
///  disorder 0: none;  1: Anderson `w (u - 0.5)` on every orbital of every site;
///           2: magnetic impurities with polarized moments (m_z = 1), opposite g, exchange J.
/// Disorder follows the `caroli run` seeding (crates/caroli-cli/src/build.rs:251-306):
/// `Xoshiro256::seed_from_u64(seed)`, draws over all region sites in region order.
struct PresetArgs {
    int model = 0;
    size_t extent[3] = {16, 10, 0};
    int axis = 0;
    int disorder = 0;
    double strength = 0.0;       // Anderson width w, or exchange J
    double concentration = 0.0;  // magnetic impurities per site
    uint64_t seed = 1;
    /// Optional extra potential per site (added to every orbital of the site), or null.
    const double* site_potential = nullptr;
};
SystemSpec preset_system(const PresetArgs& a);

/// Classical magnetic impurities and their exchange (crates/caroli-disorder/src/lib.rs:18-100,
/// caroli-model presets.rs:60-80): moments 0 isotropic (n_z uniform in [-1, 1]), 1 in a cone of
/// half-angle moment_param (radians) around z, 2 polarized with mean m_z = moment_param; coupling
/// 0 opposite g (s tau_z), 1 same g (s); exchange J (n . S) per impurity site.
struct MagneticArgs {
    int moments = 2;
    double moment_param = 1.0;
    int coupling = 0;
    double exchange = 0.0;
};
/// n distinct sites among `candidates` (partial Fisher-Yates), then all n_z draws and all
/// azimuths (MagneticImpurities::generate), and each site's 4 x 4 exchange matrix.
std::map<Site, CMat> magnetic_exchange(Xoshiro256& rng, const std::vector<Site>& candidates, size_t n,
                                       const MagneticArgs& m);

/// One disorder of the CLI input (crates/caroli-cli/src/input.rs:131-170): on the region's sites
/// inside `cells` (per direction [first, end) when has_cells[d]) and of the sublattices in the mask
/// (0: all), in region order, drawn from Xoshiro256::seed_from_u64(seed).
struct DisorderSpec {
    int kind = 0;  // 0 Anderson (strength W), 1 magnetic
    double strength = 0.0;
    double concentration = 0.0;
    MagneticArgs magnetic;  // coupling 2: the Pauli matrices (two-orbital sites)
    int64_t cells[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    bool has_cells[3] = {false, false, false};
    uint64_t sublattices = 0;
    uint64_t seed = 1;
};
/// The CLI's disordered model (crates/caroli-cli/src/build.rs:251-310): the preset `model` (0 BHZ,
/// 1 square, 2 zigzag graphene, 3 cubic) on Region::cuboid(extent) with every disorder added, as a
/// system along `axis` with the clean preset as leads.
SystemSpec disordered_system(int model, const size_t extent[3], int axis, const std::vector<DisorderSpec>& disorders);
/// The same model with leads on the given faces (sign, axis) of the cuboid, a multi-terminal device
/// (build.rs:371-413, Leads::Faces).
MultiSpec disordered_multi(int model, const size_t extent[3], const std::vector<std::pair<int, int>>& faces,
                           const std::vector<DisorderSpec>& disorders);

/// BHZ ribbon extent[0] x extent[1] (axis 0, clean BHZ leads) with int(conc nx_imp ny) magnetic
/// impurities among the sites x0 <= x < x0 + nx_imp, all y, listed with x fastest (the thesis
/// impurity region, crates/caroli-validation/src/thesis.rs:11-16, :66-78), drawn from
/// Xoshiro256::seed_from_u64(seed) (seed_mode 0) or Xoshiro256::gfortran(seed) (1).
SystemSpec bhz_magnetic_system(const size_t extent[2], int seed_mode, int64_t seed, int64_t x0, int64_t nx_imp,
                               double concentration, const MagneticArgs& m);

}  // namespace caroli
