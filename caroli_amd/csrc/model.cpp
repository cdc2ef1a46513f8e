// Host-side input builders (see model.hpp for the reference files they follow).
#include "model.hpp"

#include <algorithm>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <tuple>

namespace caroli {

Lattice Lattice::square(double a, int orbitals) {
    return Lattice{{{a, 0, 0}, {0, a, 0}}, {{{0, 0, 0}, orbitals}}};
}

Lattice Lattice::cubic(double a, int orbitals) {
    return Lattice{{{a, 0, 0}, {0, a, 0}, {0, 0, a}}, {{{0, 0, 0}, orbitals}}};
}

// crates/caroli-lattice/src/lib.rs:57-62
Lattice Lattice::honeycomb(double a, int orbitals) {
    double s3 = std::sqrt(3.0);
    return Lattice{{{s3 * a, 0, 0}, {0.5 * s3 * a, 1.5 * a, 0}}, {{{0, 0, 0}, orbitals}, {{0, a, 0}, orbitals}}};
}

Region Region::from_sites(const Lattice& lat, std::vector<Site> sites) {
    Region r;
    r.lattice = lat;
    r.sites = std::move(sites);
    r.offsets.reserve(r.sites.size());
    for (size_t i = 0; i < r.sites.size(); ++i) {
        const Site& s = r.sites[i];
        if (s.sub < 0 || s.sub >= int(lat.subs.size())) throw std::invalid_argument("unknown sublattice");
        if (!r.index.emplace(s, i).second) throw std::invalid_argument("duplicate site");
        r.offsets.push_back(r.orbitals);
        r.orbitals += lat.subs[s.sub].orbitals;
    }
    return r;
}

// crates/caroli-lattice/src/lib.rs:126-140
Region Region::cuboid(const Lattice& lat, std::array<size_t, 3> extent) {
    int dim = lat.dim();
    auto n = [&](int a) -> int64_t { return a < dim ? int64_t(extent[a]) : 1; };
    std::vector<Site> sites;
    for (int64_t i = 0; i < n(0); ++i)
        for (int64_t j = 0; j < n(1); ++j)
            for (int64_t k = 0; k < n(2); ++k)
                for (int sub = 0; sub < int(lat.subs.size()); ++sub) sites.push_back(Site{{i, j, k}, sub});
    return from_sites(lat, std::move(sites));
}

// crates/caroli-lattice/src/lib.rs:188-194
std::vector<Region> Region::slices(int axis) const {
    std::map<int64_t, std::vector<Site>> by;
    for (auto& s : sites) by[s.cell[axis]].push_back(s);
    std::vector<Region> out;
    for (auto& kv : by) out.push_back(from_sites(lattice, kv.second));
    return out;
}

Region Region::translated(Cell d) const {
    std::vector<Site> t;
    t.reserve(sites.size());
    for (auto& s : sites) t.push_back(Site{{s.cell[0] + d[0], s.cell[1] + d[1], s.cell[2] + d[2]}, s.sub});
    return from_sites(lattice, std::move(t));
}

// crates/caroli-model/src/sparse.rs:23-44
SparseMatrix SparseMatrix::from_triplets(size_t n, std::vector<std::tuple<size_t, size_t, cplx>> t) {
    std::stable_sort(t.begin(), t.end(), [](auto& x, auto& y) {
        return std::get<0>(x) != std::get<0>(y) ? std::get<0>(x) < std::get<0>(y) : std::get<1>(x) < std::get<1>(y);
    });
    SparseMatrix m;
    m.n = n;
    m.row_ptr.assign(n + 1, 0);
    long lr = -1, lc = -1;
    for (auto& [r, c, v] : t) {
        if (r >= n || c >= n) throw std::invalid_argument("entry outside the matrix");
        if (long(r) == lr && long(c) == lc) {
            m.vals.back() += v;
        } else {
            m.cols.push_back(uint32_t(c));
            m.vals.push_back(v);
            m.row_ptr[r + 1] += 1;
            lr = long(r);
            lc = long(c);
        }
    }
    for (size_t r = 0; r < n; ++r) m.row_ptr[r + 1] += m.row_ptr[r];
    return m;
}

Model::Model(Lattice lat) : lattice(std::move(lat)) {
    for (auto& s : lattice.subs) onsite.emplace_back(s.orbitals, s.orbitals);
}

Model& Model::set_onsite(int sub, const CMat& m) {
    onsite.at(sub) = m;
    return *this;
}

Model& Model::hopping(int from, int to, Cell offset, const CMat& m) {
    if (offset == Cell{0, 0, 0} && from == to) throw std::invalid_argument("an on-site term is not a hopping");
    hoppings.push_back(Hopping{from, to, offset, m});
    return *this;
}

// crates/caroli-model/src/lib.rs:383-391: sublattice matrix plus the extra terms in order.
CMat Model::onsite_matrix(const Site& s) const {
    CMat m = onsite[s.sub];
    for (auto& term : extra) {
        CMat v(m.rows, m.cols);
        if (term(s, v))
            for (size_t k = 0; k < m.a.size(); ++k) m.a[k] += v.a[k];
    }
    return m;
}

namespace {

struct DenseSink {
    CMat& h;
    void block(size_t r0, size_t c0, const CMat& m, bool adjoint) {
        int rows = adjoint ? m.cols : m.rows, cols = adjoint ? m.rows : m.cols;
        for (int j = 0; j < cols; ++j)
            for (int i = 0; i < rows; ++i) h(int(r0 + i), int(c0 + j)) += adjoint ? std::conj(m(j, i)) : m(i, j);
    }
};

// crates/caroli-model/src/lib.rs:683-703: zero entries are skipped.
struct TripletSink {
    std::vector<std::tuple<size_t, size_t, cplx>> t;
    void block(size_t r0, size_t c0, const CMat& m, bool adjoint) {
        int rows = adjoint ? m.cols : m.rows, cols = adjoint ? m.rows : m.cols;
        for (int j = 0; j < cols; ++j)
            for (int i = 0; i < rows; ++i) {
                cplx v = adjoint ? std::conj(m(j, i)) : m(i, j);
                if (v != cplx(0.0)) t.emplace_back(r0 + i, c0 + j, v);
            }
    }
};

Site shifted(const Site& s, const Cell& off, int64_t sign, int sub) {
    return Site{{s.cell[0] + sign * off[0], s.cell[1] + sign * off[1], s.cell[2] + sign * off[2]}, sub};
}

}  // namespace

// crates/caroli-model/src/lib.rs:542-566
template <class Sink>
void Model::couplings_one_way(Sink& sink, const Region& a, const Region& b) const {
    for (size_t i = 0; i < a.sites.size(); ++i) {
        const Site& site = a.sites[i];
        for (auto& hop : hoppings) {
            if (hop.from == site.sub) {
                long j = b.find(shifted(site, hop.offset, 1, hop.to));
                if (j >= 0) sink.block(a.offsets[i], b.offsets[j], hop.matrix, false);
            }
            if (hop.to == site.sub) {
                long j = b.find(shifted(site, hop.offset, -1, hop.from));
                if (j >= 0) sink.block(a.offsets[i], b.offsets[j], hop.matrix, true);
            }
        }
    }
}

// crates/caroli-model/src/lib.rs:427-437 (no periods)
template <class Sink>
void Model::build(Sink& sink, const Region& r) const {
    for (size_t i = 0; i < r.sites.size(); ++i) sink.block(r.offsets[i], r.offsets[i], onsite_matrix(r.sites[i]), false);
    couplings_one_way(sink, r, r);
}

SparseMatrix Model::sparse_hamiltonian(const Region& r) const {
    TripletSink t;
    build(t, r);
    return SparseMatrix::from_triplets(r.orbitals, std::move(t.t));
}

CMat Model::hamiltonian(const Region& r) const {
    CMat h(int(r.orbitals), int(r.orbitals));
    DenseSink s{h};
    build(s, r);
    return h;
}

CMat Model::coupling(const Region& a, const Region& b) const {
    CMat h(int(a.orbitals), int(b.orbitals));
    DenseSink s{h};
    couplings_one_way(s, a, b);
    return h;
}

// crates/caroli-model/src/gamma.rs: the 4x4 matrices by their thesis names.
CMat gamma_matrix(const std::string& name) {
    const cplx O(0, 0), L(1, 0), I(0, 1);
    auto rows = [](std::initializer_list<std::initializer_list<cplx>> r) {
        CMat m(4, 4);
        int i = 0;
        for (auto& row : r) {
            int j = 0;
            for (auto& v : row) m(i, j++) = v;
            ++i;
        }
        return m;
    };
    if (name == "0") return rows({{L, O, O, O}, {O, L, O, O}, {O, O, L, O}, {O, O, O, L}});
    if (name == "12") return rows({{L, O, O, O}, {O, L, O, O}, {O, O, -L, O}, {O, O, O, -L}});
    if (name == "12p") return rows({{L, O, O, O}, {O, -L, O, O}, {O, O, -L, O}, {O, O, O, L}});
    if (name == "13") return rows({{O, O, I, O}, {O, O, O, I}, {-I, O, O, O}, {O, -I, O, O}});
    if (name == "14") return rows({{O, O, L, O}, {O, O, O, -L}, {L, O, O, O}, {O, -L, O, O}});
    if (name == "23") return rows({{O, O, L, O}, {O, O, O, L}, {L, O, O, O}, {O, L, O, O}});
    if (name == "23p") return rows({{O, O, L, O}, {O, O, O, -L}, {L, O, O, O}, {O, -L, O, O}});
    if (name == "31") return rows({{O, O, -I, O}, {O, O, O, -I}, {I, O, O, O}, {O, I, O, O}});
    if (name == "31p") return rows({{O, O, -I, O}, {O, O, O, I}, {I, O, O, O}, {O, -I, O, O}});
    if (name == "33") return rows({{O, L, O, O}, {L, O, O, O}, {O, O, O, -L}, {O, O, -L, O}});
    if (name == "44") return rows({{O, -I, O, O}, {I, O, O, O}, {O, O, O, -I}, {O, O, I, O}});
    if (name == "55") return rows({{L, O, O, O}, {O, -L, O, O}, {O, O, L, O}, {O, O, O, -L}});
    throw std::invalid_argument("unknown gamma matrix " + name);
}

namespace {
// crates/caroli-model/src/lib.rs:705-708: sum of coefficient times matrix.
CMat combine(std::initializer_list<std::pair<cplx, CMat>> terms) {
    CMat out(terms.begin()->second.rows, terms.begin()->second.cols);
    for (size_t k = 0; k < out.a.size(); ++k) {
        cplx s = 0.0;
        for (auto& t : terms) s += t.first * t.second.a[k];
        out.a[k] = s;
    }
    return out;
}
CMat scalar(double x) {
    CMat m(1, 1);
    m(0, 0) = x;
    return m;
}
}  // namespace

// crates/caroli-model/src/presets.rs:34-56
Model bhz(const BhzParams& p) {
    auto r = [](double x) { return cplx(x, 0); };
    auto i = [](double x) { return cplx(0, x); };
    CMat on = combine({{r(p.c - 4.0 * p.d), gamma_matrix("0")}, {r(p.mass - 4.0 * p.b), gamma_matrix("55")}});
    CMat tx = combine({{r(p.d), gamma_matrix("0")},
                       {r(p.b), gamma_matrix("55")},
                       {i(-0.5 * p.a), gamma_matrix("33")},
                       {i(0.5 * p.lambda), gamma_matrix("13")}});
    CMat ty = combine({{r(p.d), gamma_matrix("0")},
                       {r(p.b), gamma_matrix("55")},
                       {i(0.5 * p.a), gamma_matrix("44")},
                       {i(-0.5 * p.lambda), gamma_matrix("14")}});
    Model m(Lattice::square(1.0, 4));
    m.set_onsite(0, on).hopping(0, 0, {1, 0, 0}, tx).hopping(0, 0, {0, 1, 0}, ty);
    return m;
}

Model square_model() {
    Model m(Lattice::square(1.0, 1));
    m.hopping(0, 0, {1, 0, 0}, scalar(-1.0)).hopping(0, 0, {0, 1, 0}, scalar(-1.0));
    return m;
}

Model cubic_model() {
    Model m(Lattice::cubic(1.0, 1));
    m.hopping(0, 0, {1, 0, 0}, scalar(-1.0)).hopping(0, 0, {0, 1, 0}, scalar(-1.0)).hopping(0, 0, {0, 0, 1}, scalar(-1.0));
    return m;
}

// crates/caroli-bench/src/main.rs:86-91
Model graphene_model() {
    Model m(Lattice::honeycomb(1.0, 1));
    m.hopping(0, 1, {0, 0, 0}, scalar(-1.0)).hopping(0, 1, {0, -1, 0}, scalar(-1.0)).hopping(0, 1, {1, -1, 0}, scalar(-1.0));
    return m;
}

// crates/caroli-disorder/src/rng.rs:19-25
static uint64_t splitmix64(uint64_t& state) {
    state += 0x9e3779b97f4a7c15ULL;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

Xoshiro256 Xoshiro256::seed_from_u64(uint64_t seed) {
    Xoshiro256 r;
    uint64_t sm = seed;
    for (auto& w : r.s) w = splitmix64(sm);
    return r;
}

// crates/caroli-disorder/src/rng.rs:15-16, 37-43
Xoshiro256 Xoshiro256::gfortran_put(const int32_t put[8]) {
    static const uint64_t keys[4] = {0xbd0c5b6e50c2df49ULL, 0xd46061cd46e1df38ULL, 0xbb4f4d4ed6103544ULL,
                                     0x114a583d0756ad39ULL};
    auto w = [&](int k) { return uint64_t(uint32_t(put[k])); };
    uint64_t raw[4] = {w(7) | w(6) << 32, w(5) | w(4) << 32, w(3) | w(2) << 32, w(1) | w(0) << 32};
    Xoshiro256 r;
    for (int i = 0; i < 4; ++i) r.s[i] = raw[i] ^ keys[i];
    return r;
}

Xoshiro256 Xoshiro256::gfortran(int32_t seed) {
    int32_t put[8];
    for (auto& p : put) p = seed;
    return gfortran_put(put);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

// crates/caroli-disorder/src/rng.rs:53-64
uint64_t Xoshiro256::next_u64() {
    uint64_t result = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}

double Xoshiro256::uniform() { return double(next_u64() >> 11) * (1.0 / double(uint64_t(1) << 53)); }

size_t Xoshiro256::integer(size_t lo, size_t hi) { return lo + size_t(double(hi - lo + 1) * uniform()); }

// crates/caroli-disorder/src/lib.rs:18-26
std::vector<Site> choose_sites(Xoshiro256& rng, const std::vector<Site>& cand, size_t n) {
    if (n > cand.size()) throw std::invalid_argument("more impurities than sites");
    std::vector<size_t> tab(cand.size());
    for (size_t i = 0; i < tab.size(); ++i) tab[i] = i;
    for (size_t i = 0; i < n; ++i) {
        size_t r = rng.integer(i + 1, cand.size()) - 1;
        std::swap(tab[i], tab[r]);
    }
    std::vector<Site> out;
    for (size_t i = 0; i < n; ++i) out.push_back(cand[tab[i]]);
    return out;
}

// crates/caroli-green/src/nested.rs:530-557 and gauge.rs:81-98 (no field: the gauge is trivial)
SystemSpec system_from_models(const Model& device, const Model& lead, const Region& region, int axis) {
    for (auto& hop : lead.hoppings)
        if (std::llabs(hop.offset[axis]) > 1) throw std::invalid_argument("a lead hopping skips a cell along the axis");
    SystemSpec s;
    s.h = device.sparse_hamiltonian(region);
    s.dim = region.lattice.dim();
    s.axis = axis;
    s.range = {1, 1, 1};
    for (auto& hop : device.hoppings)
        for (int d = 0; d < 3; ++d) s.range[d] = std::max<int64_t>(s.range[d], std::llabs(hop.offset[d]));
    for (size_t i = 0; i < region.sites.size(); ++i) {
        s.cells.push_back(region.sites[i].cell);
        s.site_orbitals.push_back(region.site_orbitals(i));
        s.site_offsets.push_back(region.offsets[i]);
    }
    auto slices = region.slices(axis);
    Cell step{0, 0, 0};
    step[axis] = 1;
    Cell back{-step[0], -step[1], -step[2]};
    auto cell_pair = [&](const Region& c, Cell along) {
        CMat h00 = lead.hamiltonian(c), h01 = lead.coupling(c, c.translated(step));
        Region n = c.translated(along);
        CMat h00n = lead.hamiltonian(n), h01n = lead.coupling(n, n.translated(step));
        double tol = 1e-12 * std::max(h00.norm_max(), h01.norm_max());
        for (size_t k = 0; k < h00.a.size(); ++k)
            if (std::abs(h00.a[k] - h00n.a[k]) > tol) throw std::invalid_argument("the lead model changes along the lead");
        for (size_t k = 0; k < h01.a.size(); ++k)
            if (std::abs(h01.a[k] - h01n.a[k]) > tol) throw std::invalid_argument("the lead model changes along the lead");
        return std::make_pair(h00, h01);
    };
    std::tie(s.left_h00, s.left_h01) = cell_pair(slices.front(), back);
    std::tie(s.right_h00, s.right_h01) = cell_pair(slices.back(), step);
    return s;
}

// crates/caroli-disorder/src/lib.rs:18-100 (choose_sites, Moments::draw, exchange)
std::map<Site, CMat> magnetic_exchange(Xoshiro256& rng, const std::vector<Site>& candidates, size_t n,
                                       const MagneticArgs& m) {
    auto sites = choose_sites(rng, candidates, n);
    // Moments::nz (lib.rs:44-52): all n_z draws first, then all azimuths.
    auto nz_of = [&](double u) {
        switch (m.moments) {
            case 0:
                return 2.0 * u - 1.0;
            case 1:
                return -(1.0 - std::cos(m.moment_param)) * u + 1.0;
            case 2:
                if (!(m.moment_param >= 0.0 && m.moment_param <= 1.0))
                    throw std::invalid_argument("polarisation outside [0, 1]");
                return 2.0 * (1.0 - m.moment_param) * u + 2.0 * m.moment_param - 1.0;
            default:
                throw std::invalid_argument("unknown moment distribution");
        }
    };
    std::vector<double> nz(n), phi(n);
    for (auto& z : nz) z = nz_of(rng.uniform());
    for (auto& p : phi) p = 2.0 * M_PI * rng.uniform();
    // bhz_spin (presets.rs:69-74): (s_x, s_y, s_z) for same g, times tau_z for opposite g.
    // Pauli matrices for two-orbital sites (crates/caroli-cli/src/build.rs:322-333).
    const bool same = m.coupling == 1;
    if (m.coupling < 0 || m.coupling > 2) throw std::invalid_argument("unknown spin matrices");
    CMat sx, sy, sz;
    if (m.coupling == 2) {
        sx = CMat(2, 2);
        sy = CMat(2, 2);
        sz = CMat(2, 2);
        sx(0, 1) = sx(1, 0) = 1.0;
        sy(0, 1) = cplx(0.0, -1.0);
        sy(1, 0) = cplx(0.0, 1.0);
        sz(0, 0) = 1.0;
        sz(1, 1) = -1.0;
    } else {
        sx = gamma_matrix(same ? "23" : "23p");
        sy = gamma_matrix(same ? "31" : "31p");
        sz = gamma_matrix(same ? "12" : "12p");
    }
    std::map<Site, CMat> map;
    for (size_t k = 0; k < n; ++k) {
        double perp = std::sqrt(1.0 - nz[k] * nz[k]);
        double v3[3] = {perp * std::cos(phi[k]), perp * std::sin(phi[k]), nz[k]};
        // MagneticImpurities::exchange (lib.rs:90-100): J (n_x S_x + n_y S_y + n_z S_z)
        CMat v = combine({{cplx(m.exchange * v3[0]), sx}, {cplx(m.exchange * v3[1]), sy}, {cplx(m.exchange * v3[2]), sz}});
        auto it = map.find(sites[k]);
        if (it == map.end())
            map.emplace(sites[k], v);
        else
            for (size_t e = 0; e < v.a.size(); ++e) it->second.a[e] += v.a[e];
    }
    return map;
}

SystemSpec bhz_magnetic_system(const size_t extent[2], int seed_mode, int64_t seed, int64_t x0, int64_t nx_imp,
                               double concentration, const MagneticArgs& m) {
    Model lead = bhz(BhzParams{});
    Region region = Region::cuboid(lead.lattice, {extent[0], extent[1], 0});
    if (x0 < 0 || nx_imp < 0 || size_t(x0 + nx_imp) > extent[0]) throw std::invalid_argument("impurity region outside");
    std::vector<Site> cand;
    for (size_t y = 0; y < extent[1]; ++y)
        for (int64_t x = 0; x < nx_imp; ++x) cand.push_back(Site{{x0 + x, int64_t(y), 0}, 0});
    const size_t nimp = size_t(concentration * double(nx_imp) * double(extent[1]));
    if (seed_mode != 0 && seed_mode != 1) throw std::invalid_argument("unknown seeding");
    auto rng = seed_mode == 0 ? Xoshiro256::seed_from_u64(uint64_t(seed)) : Xoshiro256::gfortran(int32_t(seed));
    auto map = std::make_shared<std::map<Site, CMat>>(magnetic_exchange(rng, cand, nimp, m));
    Model device = lead;
    device.extra.push_back([map](const Site& s, CMat& v) {
        auto it = map->find(s);
        if (it == map->end()) return false;
        v = it->second;
        return true;
    });
    return system_from_models(device, lead, region, 0);
}

namespace {
// The CLI's device model on its cuboid region (crates/caroli-cli/src/build.rs:251-310): every
// disorder added in order.
Model disordered_model(const Model& lead, const Region& region, const std::vector<DisorderSpec>& disorders) {
    Model device = lead;
    for (const DisorderSpec& d : disorders) {
        std::vector<Site> sites;
        for (const Site& s : region.sites) {
            bool in = true;
            for (int k = 0; k < 3; ++k)
                if (d.has_cells[k]) in = in && s.cell[k] >= d.cells[k][0] && s.cell[k] < d.cells[k][1];
            if (d.sublattices) in = in && s.sub < 64 && ((d.sublattices >> s.sub) & 1u);
            if (in) sites.push_back(s);
        }
        auto rng = Xoshiro256::seed_from_u64(d.seed);
        if (d.kind == 0) {
            // Anderson::generate (crates/caroli-disorder/src/lib.rs:112-115): w (u - 1/2) per site.
            auto map = std::make_shared<std::map<Site, double>>();
            for (auto& s : sites) (*map)[s] = d.strength * (rng.uniform() - 0.5);
            device.extra.push_back([map](const Site& s, CMat& v) {
                auto it = map->find(s);
                if (it == map->end()) return false;
                for (int i = 0; i < v.rows; ++i) v(i, i) = it->second;
                return true;
            });
        } else if (d.kind == 1) {
            const int orbitals = lead.lattice.subs.empty() ? 0 : lead.lattice.subs[0].orbitals;
            if ((d.magnetic.coupling == 2) != (orbitals == 2) || (d.magnetic.coupling != 2 && orbitals != 4))
                throw std::invalid_argument("spin matrices of another size than the sites' orbitals");
            const size_t n = size_t(d.concentration * double(sites.size()));
            auto map = std::make_shared<std::map<Site, CMat>>(magnetic_exchange(rng, sites, n, d.magnetic));
            device.extra.push_back([map](const Site& s, CMat& v) {
                auto it = map->find(s);
                if (it == map->end()) return false;
                v = it->second;
                return true;
            });
        } else {
            throw std::invalid_argument("unknown disorder");
        }
    }
    return device;
}
}  // namespace

SystemSpec disordered_system(int model, const size_t extent[3], int axis, const std::vector<DisorderSpec>& disorders) {
    Model lead = preset_model(model);
    Region region = Region::cuboid(lead.lattice, {extent[0], extent[1], extent[2]});
    return system_from_models(disordered_model(lead, region, disorders), lead, region, axis);
}

// build.rs:371-413 (Leads::Faces): one interface per face, the region's sites on it, stepping out.
MultiSpec disordered_multi(int model, const size_t extent[3], const std::vector<std::pair<int, int>>& faces,
                           const std::vector<DisorderSpec>& disorders) {
    Model lead = preset_model(model);
    Region region = Region::cuboid(lead.lattice, {extent[0], extent[1], extent[2]});
    std::vector<Region> slices;
    std::vector<Cell> outward;
    for (auto [sign, axis] : faces) {
        if ((sign != 1 && sign != -1) || axis < 0 || axis >= region.lattice.dim())
            throw std::invalid_argument("a face such as -x or +y");
        int64_t edge = sign < 0 ? INT64_MAX : INT64_MIN;
        for (auto& st : region.sites) edge = sign < 0 ? std::min(edge, st.cell[axis]) : std::max(edge, st.cell[axis]);
        std::vector<Site> on;
        for (auto& st : region.sites)
            if (st.cell[axis] == edge) on.push_back(st);
        slices.push_back(Region::from_sites(region.lattice, on));
        Cell d{0, 0, 0};
        d[axis] = sign;
        outward.push_back(d);
    }
    return multi_from_models(disordered_model(lead, region, disorders), lead, region, slices, outward);
}

Model preset_model(int model) {
    if (model < 0 || model > 3) throw std::invalid_argument("unknown preset model");
    return model == 0 ? bhz(BhzParams{}) : model == 1 ? square_model() : model == 2 ? graphene_model() : cubic_model();
}

Model with_site_potential(const Model& base, const Region& region, const double* potential) {
    Model m = base;
    auto pot = std::make_shared<std::map<Site, double>>();
    for (size_t i = 0; i < region.sites.size(); ++i) (*pot)[region.sites[i]] = potential[i];
    m.extra.push_back([pot](const Site& s, CMat& v) {
        auto it = pot->find(s);
        if (it == pot->end()) return false;
        for (int i = 0; i < v.rows; ++i) v(i, i) = it->second;
        return true;
    });
    return m;
}

// attach_leads (crates/caroli-green/src/multi.rs:112-140) and the device part of
// NestedMultiTerminal::from_model (nested_multi.rs:60-64, :155-160).
MultiSpec multi_from_models(const Model& device, const Model& lead, const Region& region,
                            const std::vector<Region>& slices, const std::vector<Cell>& outward) {
    if (slices.size() != outward.size()) throw std::invalid_argument("one outward step per interface");
    MultiSpec m;
    SystemSpec& s = m.base;
    s.h = device.sparse_hamiltonian(region);
    s.dim = region.lattice.dim();
    s.axis = 0;
    s.range = {1, 1, 1};
    for (auto& hop : device.hoppings)
        for (int d = 0; d < 3; ++d) s.range[d] = std::max<int64_t>(s.range[d], std::llabs(hop.offset[d]));
    for (size_t i = 0; i < region.sites.size(); ++i) {
        s.cells.push_back(region.sites[i].cell);
        s.site_orbitals.push_back(region.site_orbitals(i));
        s.site_offsets.push_back(region.offsets[i]);
    }
    for (size_t p = 0; p < slices.size(); ++p) {
        const Region& slice = slices[p];
        const Cell d = outward[p];
        if (slice.sites.empty()) throw std::invalid_argument("empty interface slice");
        if (d == Cell{0, 0, 0}) throw std::invalid_argument("interface without an outward step");
        const Region first = slice.translated(d);
        for (auto& site : first.sites)
            if (region.find(site) >= 0) throw std::invalid_argument("the first lead cell overlaps the device");
        const CMat skip = lead.coupling(slice, slice.translated(Cell{2 * d[0], 2 * d[1], 2 * d[2]}));
        if (skip.norm_max() != 0.0) throw std::invalid_argument("a lead hopping reaches beyond the next cell");
        AttachedLead l;
        l.h00 = lead.hamiltonian(slice);
        l.h01 = lead.coupling(slice, first);
        const CMat v = lead.coupling(region, first);
        for (int i = 0; i < v.rows; ++i)
            for (int j = 0; j < v.cols; ++j)
                if (v(i, j) != cplx(0.0)) {
                    l.orbitals.push_back(size_t(i));
                    break;
                }
        l.coupling = CMat(int(l.orbitals.size()), v.cols);
        for (int j = 0; j < v.cols; ++j)
            for (size_t a = 0; a < l.orbitals.size(); ++a) l.coupling(int(a), j) = v(int(l.orbitals[a]), j);
        std::vector<int> kinds;
        for (size_t i = 0; i < slice.sites.size(); ++i)
            for (int o = 0; o < slice.site_orbitals(i); ++o) kinds.push_back(o);
        m.leads.push_back(std::move(l));
        m.orbital_kinds.push_back(std::move(kinds));
    }
    return m;
}

SystemSpec preset_system(const PresetArgs& a) {
    Model lead = preset_model(a.model);
    Region region = Region::cuboid(lead.lattice, {a.extent[0], a.extent[1], a.extent[2]});
    Model device = lead;
    // crates/caroli-cli/src/build.rs:260-306: one draw stream over the region sites.
    auto rng = Xoshiro256::seed_from_u64(a.seed);
    if (a.disorder == 1) {
        // crates/caroli-disorder/src/lib.rs:112-115 (Anderson::generate)
        auto map = std::make_shared<std::map<Site, double>>();
        for (auto& s : region.sites) (*map)[s] = a.strength * (rng.uniform() - 0.5);
        device.extra.push_back([map](const Site& s, CMat& v) {
            auto it = map->find(s);
            if (it == map->end()) return false;
            for (int i = 0; i < v.rows; ++i) v(i, i) = it->second;
            return true;
        });
    } else if (a.disorder == 2) {
        if (a.model != 0) throw std::invalid_argument("magnetic impurities need the BHZ model");
        // crates/caroli-cli/src/build.rs:296-306: floor(c #sites) impurities over the region
        // sites in region order, Polarized { mz: 1 }, opposite g.
        size_t nimp = size_t(a.concentration * double(region.sites.size()));
        MagneticArgs m;
        m.moments = 2;
        m.moment_param = 1.0;
        m.coupling = 0;
        m.exchange = a.strength;
        auto map = std::make_shared<std::map<Site, CMat>>(magnetic_exchange(rng, region.sites, nimp, m));
        device.extra.push_back([map](const Site& s, CMat& v) {
            auto it = map->find(s);
            if (it == map->end()) return false;
            v = it->second;
            return true;
        });
    }
    if (a.site_potential) device = with_site_potential(device, region, a.site_potential);
    return system_from_models(device, lead, region, a.axis);
}

}  // namespace caroli
