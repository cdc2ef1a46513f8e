// The C ABI (include/caroli_amd.h): system construction, host self-energies, and the calls into
// the GPU engine.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <exception>
#include <map>
#include <string>

#include "../../include/caroli_amd.h"
#include "engine.hpp"
#include "checked.hpp"
#include "dense.hpp"
#include "leads.hpp"
#include "system.hpp"

using namespace caroli;

namespace caroli {

CMat broadening(const CMat& s) {
    CMat g(s.rows, s.cols);
    for (int j = 0; j < s.cols; ++j)
        for (int i = 0; i < s.rows; ++i) g(i, j) = cplx(0, 1) * (s(i, j) - std::conj(s(j, i)));
    return g;
}

CMat gamma_factor(const CMat& g) {
    const int n = g.rows;
    std::vector<double> d(n);
    double top = 0.0;
    for (int i = 0; i < n; ++i) {
        d[i] = g(i, i).real();
        top = std::max(top, d[i]);
    }
    std::vector<std::vector<cplx>> cols;
    std::vector<char> used(n, 0);
    while (int(cols.size()) < n) {
        int p = -1;
        for (int i = 0; i < n; ++i)
            if (!used[i] && (p < 0 || d[i] > d[p])) p = i;
        if (p < 0 || !(d[p] > 1e-12 * top)) break;
        used[p] = 1;
        std::vector<cplx> w(n);
        double r = std::sqrt(d[p]);
        for (int i = 0; i < n; ++i) {
            cplx v = g(i, p);
            for (auto& c : cols) v -= c[i] * std::conj(c[p]);
            w[i] = v / r;
        }
        for (int i = 0; i < n; ++i) d[i] -= std::norm(w[i]);
        cols.push_back(std::move(w));
    }
    CMat w(n, int(cols.size()));
    for (int c = 0; c < w.cols; ++c)
        for (int i = 0; i < n; ++i) w(i, c) = cols[c][i];
    return w;
}

bool same_leads(const std::vector<AttachedLead>& a, const std::vector<AttachedLead>& b) {
    if (a.size() != b.size()) return false;
    for (size_t p = 0; p < a.size(); ++p)
        if (!(a[p].h00 == b[p].h00 && a[p].h01 == b[p].h01 && a[p].orbitals == b[p].orbitals &&
              a[p].coupling == b[p].coupling))
            return false;
    return true;
}

bool sigma_matches(const cg_system& s, const cg_sigma& g) {
    if (g.blocks.size() != s.sym.blocks.size() || g.multi != s.multi) return false;
    if (s.multi) {
        if (g.mblocks.size() != g.blocks.size()) return false;
        for (size_t b = 0; b < g.blocks.size(); ++b)
            if (!same_leads(s.sym.blocks[b].leads, g.mblocks[b].leads) ||
                g.blocks[b].left_orbitals != s.sym.blocks[b].kept)
                return false;
        return true;
    }
    for (size_t b = 0; b < g.blocks.size(); ++b) {
        auto& x = s.sym.blocks[b];
        auto& y = g.blocks[b];
        if (!(x.left_h00 == y.left_h00 && x.left_h01 == y.left_h01 && x.right_h00 == y.right_h00 &&
              x.right_h01 == y.right_h01 && x.left_orbitals == y.left_orbitals && x.right_orbitals == y.right_orbitals))
            return false;
    }
    return true;
}

Work algorithmic_work(const Symbolic& sym, size_t nnz) {
    Work w;
    const double es = 16.0;
    for (auto& b : sym.blocks) {
        auto& t = b.tree;
        for (size_t k = 0; k < t.nodes.size(); ++k) {
            double ne = double(t.plans[k].ne), n = double(t.plans[k].vars.size()), nb = n - ne;
            w.flops += 8.0 / 3.0 * ne * ne * ne + 8.0 * ne * ne * nb + 8.0 * ne * nb * nb;
            double kids = 0.0;
            for (int c : t.nodes[k].children) {
                double m = double(t.nodes[c].boundary.size());
                kids += m * m;
            }
            w.bytes += es * (n * n + kids + nb * nb);
        }
        double k = double(b.kept.size()), nl = double(b.left_orbitals.size()), nr = double(b.right_orbitals.size());
        w.flops += 8.0 / 3.0 * k * k * k + 8.0 * k * k * nr + 8.0 * nl * nr * (nl + nr);
        w.bytes += es * k * k;
    }
    w.bytes += double(nnz) * (es + 4.0);
    return w;
}

}  // namespace caroli

namespace {

// The message of the last exception a C ABI call of this thread turned into a status code.
thread_local std::string last_error;
void note_error(const std::exception& e) { last_error = e.what(); }

uint64_t fnv(uint64_t h, const void* p, size_t n) {
    auto* c = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) h = (h ^ c[i]) * 1099511628211ULL;
    return h;
}

template <class V>
uint64_t fnv_vec(uint64_t h, const V& v) {
    return fnv(h, v.data(), v.size() * sizeof(v[0]));
}

uint64_t pattern_hash(uint64_t h, const CMat& m) {
    for (auto& z : m.a) {
        char nz = z != cplx(0.0);
        h = fnv(h, &nz, 1);
    }
    return h;
}

CMat read_mat(const double* p, size_t n) {
    CMat m(static_cast<int>(n), static_cast<int>(n));
    for (size_t k = 0; k < n * n; ++k) m.a[k] = cplx(p[2 * k], p[2 * k + 1]);
    return m;
}

void write_mat(const CMat& m, double* p) {
    for (size_t k = 0; k < m.a.size(); ++k) {
        p[2 * k] = m.a[k].real();
        p[2 * k + 1] = m.a[k].imag();
    }
}

int32_t finish(cg_system* s, cg_system** out) {
    std::unique_ptr<cg_system> hold(s);  // freed if the symbolic phase refuses the system
    s->sym = symbolic_phase(s->spec);
    uint64_t h = 1469598103934665603ULL;
    h = fnv(h, &s->spec.h.n, sizeof(s->spec.h.n));
    h = fnv_vec(h, s->spec.h.row_ptr);
    h = fnv_vec(h, s->spec.h.cols);
    h = fnv_vec(h, s->spec.cells);
    h = fnv_vec(h, s->spec.site_orbitals);
    h = fnv(h, &s->spec.dim, sizeof(int));
    h = fnv(h, s->spec.range.data(), sizeof(int64_t) * 3);
    h = fnv(h, &s->spec.axis, sizeof(int));
    for (auto* m : {&s->spec.left_h00, &s->spec.left_h01, &s->spec.right_h00, &s->spec.right_h01}) h = pattern_hash(h, *m);
    s->key = h;
    *out = hold.release();
    return CG_OK;
}

// The device arrays of cg_system_create and cg_multi_create. The CSR must start at 0 and be
// non-decreasing, its columns in range; every site must have a positive number of orbitals and
// the sites must cover the n orbitals exactly once (the reference builds these arrays itself and
// never sees others: crates/caroli-model/src/lib.rs:427-470). False when they are malformed.
bool read_device(uint64_t n, const int64_t* row_ptr, const uint32_t* cols, const double* vals, uint64_t nsites,
                 const int64_t* cells, const int32_t* site_orbitals, const uint64_t* site_offsets, int32_t dim,
                 const int64_t* range, SystemSpec& sp) {
    if (!row_ptr || !cols || !vals || !cells || !site_orbitals || !site_offsets || !range || dim < 1 || dim > 3 ||
        nsites == 0 || n == 0)
        return false;
    if (row_ptr[0] != 0) return false;
    for (uint64_t r = 0; r < n; ++r)
        if (row_ptr[r] > row_ptr[r + 1]) return false;
    for (int d = 0; d < dim; ++d)
        if (range[d] < 0) return false;
    std::vector<char> covered(n, 0);
    for (uint64_t i = 0; i < nsites; ++i) {
        if (site_orbitals[i] <= 0 || site_offsets[i] >= n || uint64_t(site_orbitals[i]) > n - site_offsets[i])
            return false;
        for (int32_t o = 0; o < site_orbitals[i]; ++o) {
            if (covered[site_offsets[i] + uint64_t(o)]) return false;  // overlapping sites
            covered[site_offsets[i] + uint64_t(o)] = 1;
        }
    }
    for (char c : covered)
        if (!c) return false;
    sp.h.n = n;
    sp.h.row_ptr.assign(row_ptr, row_ptr + n + 1);
    size_t nnz = size_t(row_ptr[n]);
    sp.h.cols.assign(cols, cols + nnz);
    for (auto c : sp.h.cols)
        if (c >= n) return false;
    sp.h.vals.resize(nnz);
    for (size_t e = 0; e < nnz; ++e) sp.h.vals[e] = cplx(vals[2 * e], vals[2 * e + 1]);
    for (size_t i = 0; i < nsites; ++i) {
        sp.cells.push_back({cells[3 * i], cells[3 * i + 1], cells[3 * i + 2]});
        sp.site_orbitals.push_back(site_orbitals[i]);
        sp.site_offsets.push_back(site_offsets[i]);
    }
    sp.dim = dim;
    sp.range = {range[0], range[1], range[2]};
    return true;
}

// A multi-terminal system from its device and leads: the symbolic phase of nested_multi.rs:60-161
// and the structure key (the device pattern and the leads' patterns and orbitals).
int32_t finish_multi(cg_system* s, std::vector<AttachedLead> leads, std::vector<std::vector<int>> kinds, cg_system** out) {
    std::unique_ptr<cg_system> hold(s);
    s->multi = true;
    s->sym = multi_symbolic(s->spec, leads);
    s->multi_leads = std::move(leads);
    s->orbital_kinds = std::move(kinds);
    uint64_t h = 7809847782465536322ULL;  // differs from every two-terminal key's start
    h = fnv(h, &s->spec.h.n, sizeof(s->spec.h.n));
    h = fnv_vec(h, s->spec.h.row_ptr);
    h = fnv_vec(h, s->spec.h.cols);
    h = fnv_vec(h, s->spec.cells);
    h = fnv_vec(h, s->spec.site_orbitals);
    h = fnv(h, &s->spec.dim, sizeof(int));
    h = fnv(h, s->spec.range.data(), sizeof(int64_t) * 3);
    for (auto& l : s->multi_leads) {
        h = fnv_vec(h, l.orbitals);
        for (auto* m : {&l.h00, &l.h01, &l.coupling}) h = pattern_hash(h, *m);
    }
    s->key = h;
    *out = hold.release();
    return CG_OK;
}
}  // namespace

extern "C" {

int32_t cg_system_create(uint64_t n, const int64_t* row_ptr, const uint32_t* cols, const double* vals, uint64_t nsites,
                         const int64_t* cells, const int32_t* site_orbitals, const uint64_t* site_offsets, int32_t dim,
                         const int64_t* range, int32_t axis, const double* left_h00, const double* left_h01, uint64_t nl,
                         const double* right_h00, const double* right_h01, uint64_t nr, cg_system** out) {
    if (!out || !left_h00 || !left_h01 || !right_h00 || !right_h01 || axis < 0 || axis >= dim || nl == 0 || nr == 0)
        return CG_ERR_INVALID;
    try {
        auto s = std::make_unique<cg_system>();
        auto& sp = s->spec;
        if (!read_device(n, row_ptr, cols, vals, nsites, cells, site_orbitals, site_offsets, dim, range, sp))
            return CG_ERR_INVALID;
        sp.axis = axis;
        sp.left_h00 = read_mat(left_h00, nl);
        sp.left_h01 = read_mat(left_h01, nl);
        sp.right_h00 = read_mat(right_h00, nr);
        sp.right_h01 = read_mat(right_h01, nr);
        return finish(s.release(), out);
    } catch (const std::exception& e_) {
        note_error(e_);
        return CG_ERR_INVALID;
    }
}

int32_t cg_bhz_magnetic_create(const uint64_t* extent, int32_t seed_mode, int64_t seed, int64_t x0, int64_t nx_imp,
                               double concentration, int32_t moments, double moment_param, int32_t coupling,
                               double exchange, cg_system** out) {
    if (!out || !extent || extent[0] < 3 || extent[1] < 1 || !(concentration >= 0.0 && concentration <= 1.0))
        return CG_ERR_INVALID;
    try {
        MagneticArgs m;
        m.moments = moments;
        m.moment_param = moment_param;
        m.coupling = coupling;
        m.exchange = exchange;
        const size_t ext[2] = {size_t(extent[0]), size_t(extent[1])};
        auto* s = new cg_system;
        s->spec = bhz_magnetic_system(ext, seed_mode, seed, x0, nx_imp, concentration, m);
        return finish(s, out);
    } catch (const std::exception& e_) {
        note_error(e_);
        return CG_ERR_INVALID;
    }
}

}  // extern "C"

namespace {
bool disorder_specs(int64_t count, const cg_disorder* disorders, uint64_t config, std::vector<DisorderSpec>& ds) {
    for (int64_t k = 0; k < count; ++k) {
        const cg_disorder& c = disorders[k];
        DisorderSpec d;
        d.kind = c.kind;
        d.strength = c.strength;
        d.concentration = c.concentration;
        d.magnetic.moments = c.moments;
        d.magnetic.moment_param = c.moment_param;
        d.magnetic.coupling = c.spin;
        d.magnetic.exchange = c.exchange;
        for (int a = 0; a < 3; ++a) {
            d.has_cells[a] = c.has_cells[a] != 0;
            d.cells[a][0] = c.cells[a][0];
            d.cells[a][1] = c.cells[a][1];
        }
        d.sublattices = c.sublattices;
        d.seed = c.seed + config;
        if (!(d.concentration >= 0.0 && d.concentration <= 1.0)) return false;
        ds.push_back(d);
    }
    return true;
}
}  // namespace

extern "C" {

int32_t cg_multi_disordered_create(int32_t model, const uint64_t* extent, int32_t faces, const int32_t* face_sign,
                                   const int32_t* face_axis, int64_t count, const cg_disorder* disorders,
                                   uint64_t config, cg_system** out) {
    if (!out || !extent || faces < 2 || !face_sign || !face_axis || count < 0 || (count > 0 && !disorders))
        return CG_ERR_INVALID;
    try {
        std::vector<DisorderSpec> ds;
        if (!disorder_specs(count, disorders, config, ds)) return CG_ERR_INVALID;
        std::vector<std::pair<int, int>> fs;
        for (int32_t k = 0; k < faces; ++k) fs.emplace_back(face_sign[k], face_axis[k]);
        const size_t ext[3] = {size_t(extent[0]), size_t(extent[1]), size_t(extent[2])};
        MultiSpec m = disordered_multi(model, ext, fs, ds);
        auto s = std::make_unique<cg_system>();
        s->spec = std::move(m.base);
        return finish_multi(s.release(), std::move(m.leads), std::move(m.orbital_kinds), out);
    } catch (const std::exception& e_) {
        note_error(e_);
        return CG_ERR_INVALID;
    }
}

int32_t cg_disordered_create(int32_t model, const uint64_t* extent, int32_t axis, int64_t count,
                             const cg_disorder* disorders, uint64_t config, cg_system** out) {
    if (!out || !extent || count < 0 || (count > 0 && !disorders)) return CG_ERR_INVALID;
    try {
        std::vector<DisorderSpec> ds;
        if (!disorder_specs(count, disorders, config, ds)) return CG_ERR_INVALID;
        const size_t ext[3] = {size_t(extent[0]), size_t(extent[1]), size_t(extent[2])};
        auto* s = new cg_system;
        s->spec = disordered_system(model, ext, axis, ds);
        return finish(s, out);
    } catch (const std::exception& e_) {
        note_error(e_);
        return CG_ERR_INVALID;
    }
}

int32_t cg_preset_create(int32_t model, const uint64_t* extent, int32_t axis, int32_t disorder, double strength,
                         double concentration, uint64_t seed, const double* site_potential, cg_system** out) {
    if (!out || !extent) return CG_ERR_INVALID;
    try {
        PresetArgs a;
        a.model = model;
        for (int d = 0; d < 3; ++d) a.extent[d] = extent[d];
        a.axis = axis;
        a.disorder = disorder;
        a.strength = strength;
        a.concentration = concentration;
        a.seed = seed;
        a.site_potential = site_potential;
        auto* s = new cg_system;
        s->spec = preset_system(a);
        return finish(s, out);
    } catch (const std::exception& e_) {
        note_error(e_);
        return CG_ERR_INVALID;
    }
}

void cg_system_destroy(cg_system* s) {
    if (!s) return;
    gpu::release(s);
    delete s;
}

int32_t cg_uniforms(int32_t mode, uint64_t seed, const int32_t* put, int64_t count, double* out) {
    if (count < 0 || (count > 0 && !out) || mode < 0 || mode > 2 || (mode == 2 && !put)) return CG_ERR_INVALID;
    Xoshiro256 r = mode == 0 ? Xoshiro256::seed_from_u64(seed)
                   : mode == 1 ? Xoshiro256::gfortran(int32_t(seed))
                               : Xoshiro256::gfortran_put(put);
    for (int64_t i = 0; i < count; ++i) out[i] = r.uniform();
    return CG_OK;
}

int32_t cg_block_count(const cg_system* s) { return s ? int32_t(s->sym.blocks.size()) : -1; }

int32_t cg_system_sizes(const cg_system* s, uint64_t* n, uint64_t* nnz, uint64_t* nsites) {
    if (!s) return CG_ERR_INVALID;
    if (n) *n = s->spec.h.n;
    if (nnz) *nnz = s->spec.h.vals.size();
    if (nsites) *nsites = s->spec.cells.size();
    return CG_OK;
}

int32_t cg_block_sizes(const cg_system* s, int32_t b, uint64_t* nl, uint64_t* nr, uint64_t* nodes, uint64_t* levels) {
    if (!s || b < 0 || b >= int32_t(s->sym.blocks.size())) return CG_ERR_INVALID;
    auto& blk = s->sym.blocks[b];
    if (nl) *nl = blk.left_orbitals.size();
    if (nr) *nr = blk.right_orbitals.size();
    if (nodes) *nodes = blk.tree.nodes.size();
    if (levels) *levels = blk.tree.levels;
    return CG_OK;
}

int32_t cg_system_arrays(const cg_system* s, int64_t* row_ptr, uint32_t* cols, double* vals, int64_t* cells,
                         int32_t* site_orbitals, uint64_t* site_offsets, int32_t* dim, int64_t* range, int32_t* axis,
                         double* left_h00, double* left_h01, double* right_h00, double* right_h01) {
    if (!s) return CG_ERR_INVALID;
    auto& sp = s->spec;
    if (row_ptr) std::memcpy(row_ptr, sp.h.row_ptr.data(), sp.h.row_ptr.size() * sizeof(int64_t));
    if (cols) std::memcpy(cols, sp.h.cols.data(), sp.h.cols.size() * sizeof(uint32_t));
    if (vals) std::memcpy(vals, sp.h.vals.data(), sp.h.vals.size() * sizeof(cplx));
    if (cells)
        for (size_t i = 0; i < sp.cells.size(); ++i)
            for (int d = 0; d < 3; ++d) cells[3 * i + d] = sp.cells[i][d];
    if (site_orbitals) std::memcpy(site_orbitals, sp.site_orbitals.data(), sp.site_orbitals.size() * sizeof(int32_t));
    if (site_offsets)
        for (size_t i = 0; i < sp.site_offsets.size(); ++i) site_offsets[i] = sp.site_offsets[i];
    if (dim) *dim = sp.dim;
    if (range)
        for (int d = 0; d < 3; ++d) range[d] = sp.range[d];
    if (axis) *axis = sp.axis;
    if (left_h00) write_mat(sp.left_h00, left_h00);
    if (left_h01) write_mat(sp.left_h01, left_h01);
    if (right_h00) write_mat(sp.right_h00, right_h00);
    if (right_h01) write_mat(sp.right_h01, right_h01);
    return CG_OK;
}

int32_t cg_tree_node(const cg_system* s, int32_t b, int32_t k, uint64_t* ne, uint64_t* nb, uint64_t* eliminate,
                     uint64_t* boundary) {
    if (!s || b < 0 || b >= int32_t(s->sym.blocks.size())) return CG_ERR_INVALID;
    auto& t = s->sym.blocks[b].tree;
    if (k < 0 || k >= int32_t(t.nodes.size())) return CG_ERR_INVALID;
    auto& nd = t.nodes[k];
    if (ne) *ne = nd.eliminate.size();
    if (nb) *nb = nd.boundary.size();
    if (eliminate) std::copy(nd.eliminate.begin(), nd.eliminate.end(), eliminate);
    if (boundary) std::copy(nd.boundary.begin(), nd.boundary.end(), boundary);
    return CG_OK;
}

int32_t cg_work(const cg_system* s, double* flops, double* bytes) {
    if (!s) return CG_ERR_INVALID;
    Work w = algorithmic_work(s->sym, s->spec.h.vals.size());
    if (flops) *flops = w.flops;
    if (bytes) *bytes = w.bytes;
    return CG_OK;
}

// NestedTwoTerminal::self_energies (nested.rs:652-676): per block, the left lead's Sigma_L and
// the right lead's Sigma_R (the left modes serve both when the leads are equal).
// The scalar-coupling leads go to the GPU (leads.hip) when a device is present, unless
// CAROLI_SIGMA_HOST=1 keeps them on the host.
static bool sigma_host_forced() {
    const char* v = std::getenv("CAROLI_SIGMA_HOST");
    return v && std::strcmp(v, "1") == 0;
}

static bool cell_on_gpu(const ModeSolver& m) {
    if (!m.cell()) return false;
    return !sigma_host_forced() && gpu::lead_device_available();
}

// Cells of at least this many orbitals take the GPU eigensolver when the leads go to the GPU.
constexpr int kGpuEigenMin = 128;

// The cell eigensolver of a lead: rocSOLVER on the GPU for large cells, the host's otherwise.
static HermitianEigen cell_eigen(const CMat& h0) {
    return h0.rows >= kGpuEigenMin && !sigma_host_forced() && gpu::lead_device_available() ? gpu::hermitian_eigen
                                                                                           : nullptr;
}

void ensure_modes(const cg_system* s) {
    std::lock_guard<std::mutex> g(s->mu);
    if (!s->left_modes.empty() || s->sym.blocks.empty()) return;
    s->left_modes.resize(s->sym.blocks.size());
    s->right_modes.resize(s->sym.blocks.size());
    for (size_t b = 0; b < s->sym.blocks.size(); ++b) {
        const auto& blk = s->sym.blocks[b];
        s->left_modes[b].emplace(blk.left_h00, blk.left_h01, cell_eigen(blk.left_h00));
        if (!blk.same_lead) s->right_modes[b].emplace(blk.right_h00, blk.right_h01, cell_eigen(blk.right_h00));
    }
}

// Gamma = W W^H of a cell self-energy Sigma = Y diag(d) Y^H: Gamma = Y diag(-2 Im d) Y^H, so W is
// the propagating eigenvectors scaled by sqrt(-2 Im d) (the evanescent modes have real d).
static CMat cell_gamma_factor(const CellModes& m, const std::vector<cplx>& d) {
    const CMat& y = *m.y;
    CMat w(y.rows, int(m.propagating.size()));
    for (int q = 0; q < w.cols; ++q) {
        const int k = m.propagating[q];
        const double g = std::sqrt(std::max(0.0, -2.0 * d[k].imag()));
        for (int i = 0; i < y.rows; ++i) w(i, q) = y(i, k) * g;
    }
    return w;
}

// Block b's self-energies from cell modes on the GPU: Sigma_L from the left lead's modes, Sigma_R
// from the right lead's (one launch when the leads are equal).
static Status cell_block_sigmas(const ModeSolver& left, const ModeSolver& right, bool same, double energy, cg_sigma::Block& sb) {
    CellModes cl, cr;
    if (Status st = left.cell_modes_at(energy, cl); st != kOk) return st;
    if (same) {
        cr = cl;
    } else if (Status st = right.cell_modes_at(energy, cr); st != kOk) {
        return st;
    }
    const std::vector<cplx> dl = cl.sigma_left_diag(), dr = cr.sigma_right_diag();
    const int nl = cl.y->rows, nr = cr.y->rows;
    sb.sl = CMat(nl, nl);
    sb.sr = CMat(nr, nr);
    if (same) {
        std::vector<cplx> d(dl);
        d.insert(d.end(), dr.begin(), dr.end());
        std::vector<cplx> both(size_t(2) * nl * nl);
        gpu::cell_sigmas(cl.lead_id, cl.y->a.data(), nl, d.data(), 2, both.data());
        std::copy(both.begin(), both.begin() + sb.sl.a.size(), sb.sl.a.begin());
        std::copy(both.begin() + sb.sl.a.size(), both.end(), sb.sr.a.begin());
    } else {
        gpu::cell_sigmas(cl.lead_id, cl.y->a.data(), nl, dl.data(), 1, sb.sl.a.data());
        gpu::cell_sigmas(cr.lead_id, cr.y->a.data(), nr, dr.data(), 1, sb.sr.a.data());
    }
    sb.gl = broadening(sb.sl);
    sb.gr = broadening(sb.sr);
    sb.wl = cell_gamma_factor(cl, dl);
    sb.wr = cell_gamma_factor(cr, dr);
    sb.channels_l = int(cl.propagating.size());
    sb.channels_r = int(cr.propagating.size());
    return kOk;
}

// Self-energies of every block at E + i eta (eta = 0: the real axis, nested.rs:652-676).
static int32_t self_energies_at(const cg_system* s, double energy, double eta, cg_sigma** out) {
    if (s->multi) return CG_ERR_INVALID;  // cg_multi_self_energies
    try {
        ensure_modes(s);
        auto* g = new cg_sigma;
        g->energy = energy;
        for (size_t b = 0; b < s->sym.blocks.size(); ++b) {
            auto& blk = s->sym.blocks[b];
            cg_sigma::Block sb;
            const ModeSolver& left = *s->left_modes[b];
            const ModeSolver& right = blk.same_lead ? left : *s->right_modes[b];
            if (eta == 0.0 && cell_on_gpu(left) && cell_on_gpu(right)) {
                if (Status st = cell_block_sigmas(left, right, blk.same_lead, energy, sb); st != kOk) {
                    delete g;
                    return st;
                }
                sb.left_h00 = blk.left_h00;
                sb.left_h01 = blk.left_h01;
                sb.right_h00 = blk.right_h00;
                sb.right_h01 = blk.right_h01;
                sb.left_orbitals = blk.left_orbitals;
                sb.right_orbitals = blk.right_orbitals;
                g->blocks.push_back(std::move(sb));
                continue;
            }
            LeadModes ml, mr;
            Status st = eta == 0.0 ? s->left_modes[b]->modes_at(energy, ml)
                                   : s->left_modes[b]->modes_at(cplx(energy, eta), ml);
            if (st == kOk) {
                if (blk.same_lead)
                    mr = ml;
                else
                    st = eta == 0.0 ? s->right_modes[b]->modes_at(energy, mr)
                                    : s->right_modes[b]->modes_at(cplx(energy, eta), mr);
            }
            if (st != kOk) {
                delete g;
                return st;
            }
            sb.sl = ml.sigma_left();
            sb.sr = mr.sigma_right();
            sb.gl = broadening(sb.sl);
            sb.gr = broadening(sb.sr);
            sb.wl = gamma_factor(sb.gl);
            sb.wr = gamma_factor(sb.gr);
            sb.channels_l = ml.channels;
            sb.channels_r = mr.channels;
            sb.left_h00 = blk.left_h00;
            sb.left_h01 = blk.left_h01;
            sb.right_h00 = blk.right_h00;
            sb.right_h01 = blk.right_h01;
            sb.left_orbitals = blk.left_orbitals;
            sb.right_orbitals = blk.right_orbitals;
            g->blocks.push_back(std::move(sb));
        }
        *out = g;
        return CG_OK;
    } catch (const std::exception& e_) {
        note_error(e_);
        return CG_ERR_EIGEN;
    }
}

int32_t cg_self_energies(const cg_system* s, double energy, cg_sigma** out) {
    if (!s || !out) return CG_ERR_INVALID;
    return self_energies_at(s, energy, 0.0, out);
}

int32_t cg_self_energies_broadened(const cg_system* s, double energy, double eta, cg_sigma** out) {
    if (!s || !out || !(eta >= 0.0) || !std::isfinite(eta)) return CG_ERR_INVALID;
    return self_energies_at(s, energy, eta, out);
}

int32_t cg_sigma_from_arrays(const cg_system* s, double energy, const double* sigmas, cg_sigma** out) {
    if (!s || !sigmas || !out || s->multi) return CG_ERR_INVALID;
    auto* g = new cg_sigma;
    g->energy = energy;
    const double* p = sigmas;
    for (auto& blk : s->sym.blocks) {
        cg_sigma::Block sb;
        size_t nl = blk.left_orbitals.size(), nr = blk.right_orbitals.size();
        sb.sl = read_mat(p, nl);
        p += 2 * nl * nl;
        sb.sr = read_mat(p, nr);
        p += 2 * nr * nr;
        sb.gl = broadening(sb.sl);
        sb.gr = broadening(sb.sr);
        sb.wl = gamma_factor(sb.gl);
        sb.wr = gamma_factor(sb.gr);
        sb.left_h00 = blk.left_h00;
        sb.left_h01 = blk.left_h01;
        sb.right_h00 = blk.right_h00;
        sb.right_h01 = blk.right_h01;
        sb.left_orbitals = blk.left_orbitals;
        sb.right_orbitals = blk.right_orbitals;
        g->blocks.push_back(std::move(sb));
    }
    *out = g;
    return CG_OK;
}

int32_t cg_sigma_block(const cg_sigma* g, int32_t b, double* sigma_l, double* sigma_r, int32_t* channels_l,
                       int32_t* channels_r) {
    if (!g || b < 0 || b >= int32_t(g->blocks.size())) return CG_ERR_INVALID;
    auto& sb = g->blocks[b];
    if (sigma_l) write_mat(sb.sl, sigma_l);
    if (sigma_r) write_mat(sb.sr, sigma_r);
    if (channels_l) *channels_l = sb.channels_l;
    if (channels_r) *channels_r = sb.channels_r;
    return CG_OK;
}

double cg_sigma_energy(const cg_sigma* g) { return g ? g->energy : NAN; }

void cg_sigma_destroy(cg_sigma* g) { delete g; }

int32_t cg_batch_transmission(const cg_system* const* systems, const cg_sigma* const* sigmas, int64_t count,
                              int32_t precision, double* out_t, int32_t* out_status, cg_batch_stats* stats) {
    if (count < 0 || (count > 0 && (!systems || !sigmas || !out_t))) return CG_ERR_INVALID;
    if (precision < CG_DOUBLE || precision > CG_MIXED) return CG_ERR_INVALID;
    for (int64_t p = 0; p < count; ++p)
        if (!systems[p] || !sigmas[p] || !sigma_matches(*systems[p], *sigmas[p])) return CG_ERR_INVALID;
    try {
        return gpu::batch_transmission(systems, sigmas, count, precision, out_t, out_status, stats);
    } catch (const std::exception& e_) {
        note_error(e_);
        return CG_ERR_DEVICE;
    }
}

int32_t cg_batch_create(const cg_system* const* systems, const cg_sigma* const* sigmas, int64_t count, int32_t precision,
                        cg_batch** out) {
    if (!out || count < 0 || (count > 0 && (!systems || !sigmas))) return CG_ERR_INVALID;
    if (precision < CG_DOUBLE || precision > CG_MIXED) return CG_ERR_INVALID;
    for (int64_t p = 0; p < count; ++p)
        if (!systems[p] || !sigmas[p] || !sigma_matches(*systems[p], *sigmas[p])) return CG_ERR_INVALID;
    try {
        return gpu::create_batch(systems, sigmas, count, precision, out);
    } catch (const std::exception& e_) {
        note_error(e_);
        return CG_ERR_DEVICE;
    }
}

int32_t cg_batch_run(cg_batch* batch, double* out_t, int32_t* out_status, cg_batch_stats* stats) {
    if (!batch || !out_t) return CG_ERR_INVALID;
    try {
        return gpu::run_batch(batch, out_t, out_status, stats);
    } catch (const std::exception& e_) {
        note_error(e_);
        return CG_ERR_DEVICE;
    }
}

void cg_batch_destroy(cg_batch* batch) { gpu::destroy_batch(batch); }

int32_t cg_get_tuning(cg_tuning* out) {
    if (!out) return CG_ERR_INVALID;
    return gpu::get_tuning(out);
}

int32_t cg_set_tuning(const cg_tuning* t) { return gpu::set_tuning(t); }

int32_t cg_batch_autotune(cg_batch* batch, cg_tuning* out, double* out_seconds) {
    if (!batch) return CG_ERR_INVALID;
    try {
        return gpu::autotune_batch(batch, out, out_seconds);
    } catch (const std::exception& e_) {
        note_error(e_);
        return CG_ERR_DEVICE;
    }
}

int32_t cg_front_schur(const double* f, int64_t n, int64_t ne, int64_t ne1, int32_t precision, double* out_s,
                       int32_t* out_status) {
    if (!f || !out_status || n <= 0 || ne < 0 || ne > n || ne1 < 0 || ne1 > ne || (ne < n && !out_s) ||
        (precision != CG_DOUBLE && precision != CG_SINGLE) || n > 8192)
        return CG_ERR_INVALID;
    try {
        return gpu::front_schur(f, n, ne, ne1, precision, out_s, out_status);
    } catch (const std::exception& e_) {
        note_error(e_);
        return CG_ERR_DEVICE;
    }
}

int32_t cg_front_schur_noalloc(const double* f, int64_t n, int64_t ne, int64_t ne1, int32_t precision, double* out_s,
                               int32_t* out_status) {
    if (!f || !out_status || n <= 0 || ne < 0 || ne > n || ne1 < 0 || ne1 > ne || (ne < n && !out_s) ||
        (precision != CG_DOUBLE && precision != CG_SINGLE))
        return CG_ERR_INVALID;
    if (n > CG_FRONT_NOALLOC_MAX_N) return CG_ERR_CAPACITY;
    try {
        return gpu::front_schur_noalloc(f, n, ne, ne1, precision, out_s, out_status);
    } catch (const std::exception& e_) {
        note_error(e_);
        return CG_ERR_DEVICE;
    }
}

int32_t cg_kernel_timing(int32_t enable) {
    try {
        return gpu::kernel_timing(enable);
    } catch (const std::exception& e_) {
        note_error(e_);
        return CG_ERR_DEVICE;
    }
}

int32_t cg_kernel_stats(cg_kernel_stat* out, int32_t max_records) {
    if (!out || max_records < 0) return CG_ERR_INVALID;
    return gpu::kernel_stats(out, max_records);
}

int32_t cg_batch_transmission_checked(const cg_system* const* systems, const cg_sigma* const* sigmas, int64_t count,
                                      const cg_precision_options* options, double* out_t, int32_t* out_status,
                                      double* out_dt, cg_precision_report* report, cg_batch_stats* stats) {
    if (!options || !report || count < 0 || (count > 0 && (!systems || !sigmas || !out_t || !out_status || !out_dt)))
        return CG_ERR_INVALID;
    if (options->precision < CG_DOUBLE || options->precision > CG_MIXED || !(options->tolerance >= 0.0))
        return CG_ERR_INVALID;
    bool first = true;
    auto run = [&](const std::vector<int64_t>& idx, int32_t precision, double* t, int32_t* st) {
        std::vector<const cg_system*> ss;
        std::vector<const cg_sigma*> gg;
        for (int64_t p : idx) {
            ss.push_back(systems[p]);
            gg.push_back(sigmas[p]);
        }
        cg_batch_stats local{};
        const int32_t rc = cg_batch_transmission(ss.data(), gg.data(), int64_t(idx.size()), precision, t, st, &local);
        if (first && stats) *stats = local;
        first = false;
        return rc;
    };
    CheckedReport rep;
    const int32_t rc = checked_batch(count, options->precision, options->tolerance, options->every, run, out_t,
                                     out_status, out_dt, rep);
    *report = cg_precision_report{};
    report->precision = rep.precision;
    report->checked = rep.checked;
    report->median = rep.median;
    report->max = rep.max;
    report->failed = rep.failed;
    report->fell_back = rep.fell_back ? 1 : 0;
    std::snprintf(report->fallback, sizeof report->fallback, "%s", rep.fallback.c_str());
    return rc;
}

int32_t cg_transmission_with(const cg_system* s, const cg_sigma* g, int32_t precision, double* out_t) {
    if (!s || !g || !out_t) return CG_ERR_INVALID;
    int32_t status = CG_OK;
    int32_t rc = cg_batch_transmission(&s, &g, 1, precision, out_t, &status, nullptr);
    return rc != CG_OK ? rc : status;
}

int32_t cg_transmission(const cg_system* s, double energy, int32_t precision, double* out_t) {
    if (!s || !out_t) return CG_ERR_INVALID;
    cg_sigma* g = nullptr;
    int32_t rc = cg_self_energies(s, energy, &g);
    if (rc != CG_OK) {
        *out_t = NAN;
        return rc;
    }
    rc = cg_transmission_with(s, g, precision, out_t);
    cg_sigma_destroy(g);
    return rc;
}

int32_t cg_transmissions(const cg_system* s, const double* energies, int64_t count, int32_t precision, double* out_t,
                         int32_t* out_status) {
    if (!s || (count > 0 && (!energies || !out_t))) return CG_ERR_INVALID;
    std::vector<cg_sigma*> sig(count, nullptr);
    std::vector<int64_t> ok;
    for (int64_t p = 0; p < count; ++p) {
        int32_t rc = cg_self_energies(s, energies[p], &sig[p]);
        out_t[p] = NAN;
        if (out_status) out_status[p] = rc;
        if (rc == CG_OK) ok.push_back(p);
    }
    std::vector<const cg_system*> ss(ok.size(), s);
    std::vector<const cg_sigma*> gg;
    for (auto p : ok) gg.push_back(sig[p]);
    std::vector<double> t(ok.size());
    std::vector<int32_t> st(ok.size());
    int32_t rc = ok.empty() ? CG_OK : cg_batch_transmission(ss.data(), gg.data(), int64_t(ok.size()), precision, t.data(), st.data(), nullptr);
    for (size_t q = 0; q < ok.size(); ++q) {
        out_t[ok[q]] = t[q];
        if (out_status) out_status[ok[q]] = st[q];
    }
    for (auto* g : sig) cg_sigma_destroy(g);
    return rc;
}

int32_t cg_backward(const cg_system* s, const cg_sigma* g, double* out_t, double* out_grad) {
    if (!s || !g || !out_t || !out_grad || s->multi || !sigma_matches(*s, *g)) return CG_ERR_INVALID;
    try {
        return gpu::backward(s, g, out_t, out_grad);
    } catch (const std::exception& e_) {
        note_error(e_);
        return CG_ERR_DEVICE;
    }
}

int32_t cg_backward_batch(const cg_system* const* systems, const cg_sigma* const* sigmas, int64_t count, double* out_t,
                          double* out_grad, int32_t* out_status) {
    if (count < 0 || (count > 0 && (!systems || !sigmas || !out_t || !out_grad))) return CG_ERR_INVALID;
    for (int64_t p = 0; p < count; ++p)
        if (!systems[p] || !sigmas[p] || systems[p]->multi || !sigma_matches(*systems[p], *sigmas[p]))
            return CG_ERR_INVALID;
    try {
        return gpu::backward_batch(systems, sigmas, count, out_t, out_grad, out_status);
    } catch (const std::exception& e_) {
        note_error(e_);
        return CG_ERR_DEVICE;
    }
}

int32_t cg_scattering_states(const cg_system* s, const cg_sigma* g, int64_t* cols, double* psi_left,
                             double* psi_right) {
    if (!s || !g || !cols || s->multi || !sigma_matches(*s, *g)) return CG_ERR_INVALID;
    try {
        return gpu::scattering_states(s, g, cols, psi_left, psi_right);
    } catch (const std::exception& e_) {
        note_error(e_);
        return CG_ERR_DEVICE;
    }
}

namespace {
/// (psi_p)_{a k}: states of lead p, n orbitals x cols[p] channels, interleaved complex.
struct States {
    std::vector<const double*> m;  // one per lead (or label)
    const int64_t* cols;
    int64_t n;
    cplx at(size_t p, int64_t a, int64_t k) const {
        const double* v = m[p] + 2 * (a + k * n);
        return cplx(v[0], v[1]);
    }
    /// G^<_{ba} = i sum_p f_p sum_k psi_bk conj(psi_ak) (nested.rs:936-943).
    cplx lesser(int64_t b, int64_t a, const double* f) const {
        cplx acc = 0.0;
        for (size_t p = 0; p < m.size(); ++p) {
            cplx sp = 0.0;
            for (int64_t k = 0; k < cols[p]; ++k) sp += at(p, b, k) * std::conj(at(p, a, k));
            acc += sp * f[p];
        }
        return cplx(0.0, 1.0) * acc;
    }
};
}  // namespace

namespace {
bool states_ok(int32_t leads, const double* const* psi, const int64_t* cols) {
    if (leads < 1 || !psi || !cols) return false;
    for (int32_t p = 0; p < leads; ++p)
        if (cols[p] < 0 || (cols[p] > 0 && !psi[p])) return false;
    return true;
}

int32_t site_densities(const cg_system* s, const States& st, const double* f, const double* op, int32_t op_dim,
                       double* out) {
    const auto& sp = s->spec;
    const size_t sites = sp.site_orbitals.size();
    if (op)
        for (size_t i = 0; i < sites; ++i)
            if (sp.site_orbitals[i] != op_dim) return CG_ERR_INVALID;  // the operator acts on a site's orbitals
    const int64_t* cols = st.cols;
    for (size_t i = 0; i < sites; ++i) {
        const int64_t o0 = int64_t(sp.site_offsets[i]), ni = sp.site_orbitals[i];
        cplx acc = 0.0;
        for (size_t p = 0; p < st.m.size(); ++p)
            for (int64_t k = 0; k < cols[p]; ++k)
                for (int64_t a = 0; a < ni; ++a) {
                    const cplx left = std::conj(st.at(p, o0 + a, k));
                    if (!op) {
                        acc += left * st.at(p, o0 + a, k) * f[p];
                    } else {
                        cplx ob = 0.0;
                        for (int64_t b = 0; b < ni; ++b)
                            ob += cplx(op[2 * (a + b * ni)], op[2 * (a + b * ni) + 1]) * st.at(p, o0 + b, k);
                        acc += left * ob * f[p];
                    }
                }
        out[i] = acc.real() / (2.0 * M_PI);
    }
    return CG_OK;
}

int32_t bond_currents(const cg_system* s, const States& st, const double* f, int64_t* count, int64_t* from,
                      int64_t* to, double* current) {
    const auto& sp = s->spec;
    const auto& h = sp.h;
    std::vector<int64_t> owner(h.n, 0);
    for (size_t i = 0; i < sp.site_orbitals.size(); ++i)
        for (int o = 0; o < sp.site_orbitals[i]; ++o) owner[sp.site_offsets[i] + size_t(o)] = int64_t(i);
    // Each bond (i < j) once, ordered by (i, j): 2 sum_{a in i, b in j} H_ab G^<_ba (nested.rs:949-956).
    std::map<std::pair<int64_t, int64_t>, cplx> bonds;
    const bool fill = from && to && current;
    for (size_t a = 0; a < h.n; ++a)
        for (int64_t e = h.row_ptr[a]; e < h.row_ptr[a + 1]; ++e) {
            const size_t b = h.cols[e];
            const int64_t i = owner[a], j = owner[b];
            if (i < j) {
                cplx& acc = bonds[{i, j}];
                if (fill) acc += h.vals[e] * st.lesser(int64_t(b), int64_t(a), f) * 2.0;
            }
        }
    if (!fill) {
        *count = int64_t(bonds.size());
        return CG_OK;
    }
    if (*count < int64_t(bonds.size())) return CG_ERR_INVALID;
    int64_t q = 0;
    for (const auto& [ij, acc] : bonds) {
        from[q] = ij.first;
        to[q] = ij.second;
        current[q] = acc.real();
        ++q;
    }
    *count = q;
    return CG_OK;
}

int32_t operator_currents(const cg_system* s, const States& st, const double* f, const double* op, int32_t op_dim,
                          int64_t* count, int64_t* from, int64_t* to, double* current) {
    const auto& sp = s->spec;
    const auto& h = sp.h;
    std::vector<int64_t> owner(h.n, 0);
    for (size_t i = 0; i < sp.site_orbitals.size(); ++i)
        for (int o = 0; o < sp.site_orbitals[i]; ++o) owner[sp.site_offsets[i] + size_t(o)] = int64_t(i);
    // The blocks H_ij of every bond i < j (nested.rs:958-975); the operator acts on a site's orbitals.
    const int m = op_dim;
    std::map<std::pair<int64_t, int64_t>, std::vector<cplx>> blocks;
    for (size_t a = 0; a < h.n; ++a)
        for (int64_t e = h.row_ptr[a]; e < h.row_ptr[a + 1]; ++e) {
            const size_t b = h.cols[e];
            const int64_t i = owner[a], j = owner[b];
            if (i < j) {
                if (sp.site_orbitals[size_t(i)] != m || sp.site_orbitals[size_t(j)] != m) return CG_ERR_INVALID;
                auto& blk = blocks[{i, j}];
                if (blk.empty()) blk.assign(size_t(m) * m, cplx(0.0, 0.0));
                blk[(a - sp.site_offsets[size_t(i)]) + size_t(m) * (b - sp.site_offsets[size_t(j)])] += h.vals[e];
            }
        }
    if (!(from && to && current)) {
        *count = int64_t(blocks.size());
        return CG_OK;
    }
    if (*count < int64_t(blocks.size())) return CG_ERR_INVALID;
    auto O = [&](int a, int b) { return cplx(op[2 * (a + b * m)], op[2 * (a + b * m) + 1]); };
    int64_t q = 0;
    for (const auto& [ij, hij] : blocks) {
        const int64_t oi = int64_t(sp.site_offsets[size_t(ij.first)]), oj = int64_t(sp.site_offsets[size_t(ij.second)]);
        // Tr[(O H_ij + H_ij O) G^<_ji] = sum_{a, b} (O H + H O)_ab G^<_{oj + b, oi + a} (nested.rs:976-985)
        cplx acc = 0.0;
        for (int a = 0; a < m; ++a)
            for (int b = 0; b < m; ++b) {
                cplx w = 0.0;
                for (int c = 0; c < m; ++c) w += O(a, c) * hij[c + size_t(m) * b] + hij[a + size_t(m) * c] * O(c, b);
                if (w != cplx(0.0, 0.0)) acc += w * st.lesser(oj + b, oi + a, f);
            }
        from[q] = ij.first;
        to[q] = ij.second;
        current[q] = acc.real();
        ++q;
    }
    *count = q;
    return CG_OK;
}
}  // namespace

int32_t cg_site_densities(const cg_system* s, const double* psi_left, const double* psi_right, const int64_t* cols,
                          const double* f, const double* op, int32_t op_dim, double* out) {
    const double* psi[2] = {psi_left, psi_right};
    if (!s || !psi_left || !psi_right || !states_ok(2, psi, cols) || !f || !out) return CG_ERR_INVALID;
    return site_densities(s, States{{psi_left, psi_right}, cols, int64_t(s->spec.h.n)}, f, op, op_dim, out);
}

int32_t cg_bond_currents(const cg_system* s, const double* psi_left, const double* psi_right, const int64_t* cols,
                         const double* f, int64_t* count, int64_t* from, int64_t* to, double* current) {
    const double* psi[2] = {psi_left, psi_right};
    if (!s || !psi_left || !psi_right || !states_ok(2, psi, cols) || !f || !count) return CG_ERR_INVALID;
    return bond_currents(s, States{{psi_left, psi_right}, cols, int64_t(s->spec.h.n)}, f, count, from, to, current);
}

int32_t cg_operator_currents(const cg_system* s, const double* psi_left, const double* psi_right, const int64_t* cols,
                             const double* f, const double* op, int32_t op_dim, int64_t* count, int64_t* from,
                             int64_t* to, double* current) {
    const double* psi[2] = {psi_left, psi_right};
    if (!s || !psi_left || !psi_right || !states_ok(2, psi, cols) || !f || !op || op_dim <= 0 || !count)
        return CG_ERR_INVALID;
    return operator_currents(s, States{{psi_left, psi_right}, cols, int64_t(s->spec.h.n)}, f, op, op_dim, count, from,
                             to, current);
}

int32_t cg_multi_site_densities(const cg_system* s, int32_t leads, const double* const* psi, const int64_t* cols,
                                const double* f, const double* op, int32_t op_dim, double* out) {
    if (!s || !states_ok(leads, psi, cols) || !f || !out) return CG_ERR_INVALID;
    return site_densities(s, States{{psi, psi + leads}, cols, int64_t(s->spec.h.n)}, f, op, op_dim, out);
}

int32_t cg_multi_bond_currents(const cg_system* s, int32_t leads, const double* const* psi, const int64_t* cols,
                               const double* f, const double* op, int32_t op_dim, int64_t* count, int64_t* from,
                               int64_t* to, double* current) {
    if (!s || !states_ok(leads, psi, cols) || !f || !count || (op && op_dim <= 0)) return CG_ERR_INVALID;
    const States st{{psi, psi + leads}, cols, int64_t(s->spec.h.n)};
    return op ? operator_currents(s, st, f, op, op_dim, count, from, to, current)
              : bond_currents(s, st, f, count, from, to, current);
}

int32_t cg_device_available(void) { return gpu::device_available() ? 1 : 0; }

const char* cg_last_error(void) { return last_error.c_str(); }

// ---- NestedMultiTerminal (crates/caroli-green/src/nested_multi.rs) ----

int32_t cg_multi_create(uint64_t n, const int64_t* row_ptr, const uint32_t* cols, const double* vals, uint64_t nsites,
                        const int64_t* cells, const int32_t* site_orbitals, const uint64_t* site_offsets, int32_t dim,
                        const int64_t* range, int32_t leads, const cg_lead* lead, cg_system** out) {
    if (!out || leads < 1 || !lead) return CG_ERR_INVALID;
    try {
        auto s = std::make_unique<cg_system>();
        if (!read_device(n, row_ptr, cols, vals, nsites, cells, site_orbitals, site_offsets, dim, range, s->spec))
            return CG_ERR_INVALID;
        std::vector<AttachedLead> ls;
        std::vector<std::vector<int>> kinds;
        for (int32_t p = 0; p < leads; ++p) {
            const cg_lead& l = lead[p];
            if (l.cell == 0 || l.coupled == 0 || !l.h00 || !l.h01 || !l.orbitals || !l.coupling) return CG_ERR_INVALID;
            AttachedLead a;
            a.h00 = read_mat(l.h00, l.cell);
            a.h01 = read_mat(l.h01, l.cell);
            a.orbitals.assign(l.orbitals, l.orbitals + l.coupled);
            for (size_t k = 0; k < a.orbitals.size(); ++k)
                if (a.orbitals[k] >= n || (k > 0 && a.orbitals[k] <= a.orbitals[k - 1])) return CG_ERR_INVALID;
            a.coupling = CMat(int(l.coupled), int(l.cell));
            for (size_t e = 0; e < a.coupling.a.size(); ++e) a.coupling.a[e] = cplx(l.coupling[2 * e], l.coupling[2 * e + 1]);
            // Sector labels by cell orbital: without the site layout of the cell, orbital k is kind k.
            std::vector<int> kind(l.cell);
            for (size_t k = 0; k < kind.size(); ++k) kind[k] = int(k);
            if (l.kinds)
                for (size_t k = 0; k < kind.size(); ++k) {
                    if (l.kinds[k] < 0) return CG_ERR_INVALID;
                    kind[k] = l.kinds[k];
                }
            ls.push_back(std::move(a));
            kinds.push_back(std::move(kind));
        }
        return finish_multi(s.release(), std::move(ls), std::move(kinds), out);
    } catch (const std::exception& e_) {
        note_error(e_);
        return CG_ERR_INVALID;
    }
}

int32_t cg_multi_model_create(int32_t model, uint64_t sites, const int64_t* cells, const int32_t* subs,
                              const double* site_potential, int32_t interfaces, const cg_interface* itf, cg_system** out) {
    if (!out || sites == 0 || !cells || !subs || interfaces < 1 || !itf || model < 0 || model > 3) return CG_ERR_INVALID;
    try {
        Model lead = preset_model(model);
        auto to_sites = [](uint64_t count, const int64_t* c, const int32_t* sb) {
            std::vector<Site> v;
            for (uint64_t i = 0; i < count; ++i) v.push_back(Site{{c[3 * i], c[3 * i + 1], c[3 * i + 2]}, sb[i]});
            return v;
        };
        Region region = Region::from_sites(lead.lattice, to_sites(sites, cells, subs));
        Model device = site_potential ? with_site_potential(lead, region, site_potential) : lead;
        std::vector<Region> slices;
        std::vector<Cell> outward;
        for (int32_t p = 0; p < interfaces; ++p) {
            if (itf[p].sites == 0 || !itf[p].cells || !itf[p].subs) return CG_ERR_INVALID;
            Region slice = Region::from_sites(lead.lattice, to_sites(itf[p].sites, itf[p].cells, itf[p].subs));
            for (auto& site : slice.sites)
                if (region.find(site) < 0) return CG_ERR_INVALID;  // an interface of device sites
            slices.push_back(std::move(slice));
            outward.push_back(Cell{itf[p].outward[0], itf[p].outward[1], itf[p].outward[2]});
        }
        MultiSpec m = multi_from_models(device, lead, region, slices, outward);
        auto s = std::make_unique<cg_system>();
        s->spec = std::move(m.base);
        return finish_multi(s.release(), std::move(m.leads), std::move(m.orbital_kinds), out);
    } catch (const std::exception& e_) {
        note_error(e_);
        return CG_ERR_INVALID;
    }
}

int32_t cg_multi_lead_count(const cg_system* s) { return s && s->multi ? int32_t(s->multi_leads.size()) : -1; }

int32_t cg_multi_lead(const cg_system* s, int32_t p, uint64_t* cell, uint64_t* coupled, double* h00, double* h01,
                      uint64_t* orbitals, double* coupling) {
    if (!s || !s->multi || p < 0 || p >= int32_t(s->multi_leads.size())) return CG_ERR_INVALID;
    const AttachedLead& l = s->multi_leads[size_t(p)];
    if (cell) *cell = uint64_t(l.h00.rows);
    if (coupled) *coupled = l.orbitals.size();
    if (h00) write_mat(l.h00, h00);
    if (h01) write_mat(l.h01, h01);
    if (orbitals) std::copy(l.orbitals.begin(), l.orbitals.end(), orbitals);
    if (coupling) write_mat(l.coupling, coupling);
    return CG_OK;
}

}  // extern "C"

namespace {
// Self-energies and broadenings of a multi-terminal system at E + i eta (nested_multi.rs:196-232,
// :246-313), per block:
//  - each lead's surface Green's function g = (z - h00 - Sigma_rest)^-1 with Sigma_rest the
//    right-going self-energy of its own modes (AttachedLead::surface_green_with, multi.rs:43-50),
//    once per distinct lead;
//  - Sigma_p = V g V^H of every lead on the block's kept orbitals, summed into one kept x kept matrix;
//  - per label (p, a) the channels W = V R with R R^H = P_a i(g - g^H) P_a, the lead's broadening
//    projected on sector a (the whole cell without sectors): W W^H = Gamma_{p,a}.
int32_t multi_sigmas(const cg_system* s, double energy, double eta, const std::vector<std::vector<int32_t>>& sectors,
                     cg_sigma** out) {
    const size_t np = s->multi_leads.size();
    auto tag = [&](size_t p, int cell_orbital) -> int32_t {
        if (sectors[p].empty()) return 0;
        return sectors[p][size_t(s->orbital_kinds[p][size_t(cell_orbital)])];
    };
    // Labels (lead, sector) in increasing order (nested_multi.rs:263-270); a lead's Hamiltonian
    // must not couple different sectors (nested_multi.rs:281-287).
    auto g = std::make_unique<cg_sigma>();
    g->energy = energy;
    g->multi = true;
    for (size_t p = 0; p < np; ++p) {
        const AttachedLead& l = s->multi_leads[p];
        for (int k = 0; k < l.h00.rows; ++k)
            if (!sectors[p].empty() && size_t(s->orbital_kinds[p][size_t(k)]) >= sectors[p].size()) return CG_ERR_INVALID;
        for (const CMat* m : {&l.h00, &l.h01})
            for (int b = 0; b < m->cols; ++b)
                for (int a = 0; a < m->rows; ++a)
                    if (tag(p, a) != tag(p, b) && (*m)(a, b) != cplx(0.0)) return CG_ERR_INVALID;
        std::vector<int32_t> own;
        for (int k = 0; k < l.h00.rows; ++k) own.push_back(tag(p, k));
        std::sort(own.begin(), own.end());
        own.erase(std::unique(own.begin(), own.end()), own.end());
        for (int32_t a : own) g->labels.push_back({int32_t(p), a});
    }
    const size_t L = g->labels.size();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->lead_modes.empty()) {
            s->lead_modes.resize(s->sym.blocks.size());
            for (size_t b = 0; b < s->sym.blocks.size(); ++b) {
                const Block& blk = s->sym.blocks[b];
                auto& v = s->lead_modes[b];
                v.resize(np);
                for (size_t p = 0; p < np; ++p) {
                    const AttachedLead& l = blk.leads[p];
                    if (l.orbitals.empty()) continue;
                    for (size_t q = 0; q < p && !v[p]; ++q)
                        if (v[q] && blk.leads[q].h00 == l.h00 && blk.leads[q].h01 == l.h01) v[p] = v[q];
                    if (!v[p]) v[p] = std::make_shared<ModeSolver>(l.h00, l.h01, cell_eigen(l.h00));
                }
            }
        }
    }
    const cplx z(energy, eta);
    for (size_t b = 0; b < s->sym.blocks.size(); ++b) {
        const Block& blk = s->sym.blocks[b];
        const int k = int(blk.kept.size());
        Positions kept(blk.kept);
        cg_sigma::Block sb;
        sb.sl = CMat(k, k);
        sb.sr = CMat(k, k);
        sb.gl = CMat(k, k);
        sb.gr = CMat(k, k);
        sb.wl = CMat(k, 0);
        sb.wr = CMat(k, 0);
        sb.left_orbitals = blk.kept;
        sb.right_orbitals = blk.kept;
        cg_sigma::MultiBlock mb;
        mb.leads = blk.leads;
        mb.sigma.resize(np);
        std::vector<std::vector<cplx>> cols;  // W columns by label
        std::vector<int32_t> count(L, 0);
        std::map<const ModeSolver*, CMat> surfaces;  // one g per distinct lead
        for (size_t p = 0; p < np; ++p) {
            const AttachedLead& l = blk.leads[p];
            if (l.orbitals.empty()) continue;
            const ModeSolver* solver = s->lead_modes[b][p].get();
            auto it = surfaces.find(solver);
            if (it == surfaces.end()) {
                LeadModes md;
                Status st = eta == 0.0 ? solver->modes_at(energy, md) : solver->modes_at(z, md);
                if (st != kOk) return st;
                const CMat rest = md.sigma_right();
                const int nc = l.h00.rows;
                CMat a(nc, nc);
                for (int j = 0; j < nc; ++j)
                    for (int i = 0; i < nc; ++i) a(i, j) = (i == j ? z : cplx(0.0)) - l.h00(i, j) - rest(i, j);
                it = surfaces.emplace(solver, dense::inverse(a)).first;
            }
            const CMat& gs = it->second;
            const CMat sig = matmul(matmul(l.coupling, gs), l.coupling.adjoint());
            mb.sigma[p] = sig;
            std::vector<int> at;
            for (size_t o : l.orbitals) at.push_back(kept.get(o));
            for (size_t y = 0; y < at.size(); ++y)
                for (size_t x = 0; x < at.size(); ++x) sb.sl(at[x], at[y]) += sig(int(x), int(y));
            // Broadening of the cell, projected on each sector of the lead.
            const int nc = gs.rows;
            const std::vector<size_t>& cell = blk.lead_cells[p];
            CMat gam(nc, nc);
            for (int j = 0; j < nc; ++j)
                for (int i = 0; i < nc; ++i) gam(i, j) = cplx(0.0, 1.0) * (gs(i, j) - std::conj(gs(j, i)));
            for (size_t lab = 0; lab < L; ++lab) {
                if (g->labels[lab].lead != int32_t(p)) continue;
                CMat proj(nc, nc);
                for (int j = 0; j < nc; ++j)
                    for (int i = 0; i < nc; ++i)
                        if (tag(p, int(cell[size_t(i)])) == g->labels[lab].sector &&
                            tag(p, int(cell[size_t(j)])) == g->labels[lab].sector)
                            proj(i, j) = gam(i, j);
                const CMat w = matmul(l.coupling, gamma_factor(proj));
                for (int c = 0; c < w.cols; ++c) {
                    std::vector<cplx> col(size_t(k), cplx(0.0));
                    for (size_t x = 0; x < at.size(); ++x) col[size_t(at[x])] = w(int(x), c);
                    cols.push_back(std::move(col));
                }
                count[lab] += w.cols;
            }
        }
        // The columns went in lead by lead and, within a lead, label by label: the label order.
        mb.lab_off.assign(L + 1, 0);
        for (size_t lab = 0; lab < L; ++lab) mb.lab_off[lab + 1] = mb.lab_off[lab] + count[lab];
        mb.wab = CMat(k, int(cols.size()));
        for (size_t c = 0; c < cols.size(); ++c) std::copy(cols[c].begin(), cols[c].end(), mb.wab.col(int(c)));
        g->blocks.push_back(std::move(sb));
        g->mblocks.push_back(std::move(mb));
    }
    *out = g.release();
    return CG_OK;
}
}  // namespace

extern "C" {

int32_t cg_multi_self_energies(const cg_system* s, double energy, double eta, const int32_t* sector_len,
                               const int32_t* sector_tags, cg_sigma** out) {
    if (!s || !s->multi || !out || !(eta >= 0.0) || !std::isfinite(eta) || !std::isfinite(energy)) return CG_ERR_INVALID;
    std::vector<std::vector<int32_t>> sectors(s->multi_leads.size());
    if (sector_len) {
        const int32_t* t = sector_tags;
        for (size_t p = 0; p < sectors.size(); ++p) {
            if (sector_len[p] < 0 || (sector_len[p] > 0 && !t)) return CG_ERR_INVALID;
            sectors[p].assign(t, t + sector_len[p]);
            t += sector_len[p];
        }
    }
    try {
        return multi_sigmas(s, energy, eta, sectors, out);
    } catch (const std::exception& e_) {
        note_error(e_);
        return CG_ERR_EIGEN;
    }
}

int32_t cg_multi_sigma(const cg_system* s, const cg_sigma* g, int32_t p, double* out) {
    if (!s || !g || !out || !s->multi || !sigma_matches(*s, *g) || p < 0 || p >= int32_t(s->multi_leads.size()))
        return CG_ERR_INVALID;
    const std::vector<size_t>& all = s->multi_leads[size_t(p)].orbitals;
    CMat full(int(all.size()), int(all.size()));
    for (size_t b = 0; b < g->mblocks.size(); ++b) {
        const auto& orbs = g->mblocks[b].leads[size_t(p)].orbitals;
        const CMat& sig = g->mblocks[b].sigma[size_t(p)];
        std::vector<int> at;
        for (size_t o : orbs) at.push_back(int(std::lower_bound(all.begin(), all.end(), o) - all.begin()));
        for (size_t y = 0; y < at.size(); ++y)
            for (size_t x = 0; x < at.size(); ++x) full(at[x], at[y]) = sig(int(x), int(y));
    }
    write_mat(full, out);
    return CG_OK;
}

int32_t cg_multi_labels(const cg_sigma* g, int32_t* count, int32_t* lead, int32_t* sector) {
    if (!g || !g->multi || !count) return CG_ERR_INVALID;
    *count = int32_t(g->labels.size());
    for (size_t q = 0; q < g->labels.size(); ++q) {
        if (lead) lead[q] = g->labels[q].lead;
        if (sector) sector[q] = g->labels[q].sector;
    }
    return CG_OK;
}

int32_t cg_multi_batch_transmissions(const cg_system* const* systems, const cg_sigma* const* sigmas, int64_t count,
                                     int32_t precision, double* out, int32_t* out_status, cg_batch_stats* stats) {
    if (count < 0 || (count > 0 && (!systems || !sigmas || !out))) return CG_ERR_INVALID;
    if (precision != CG_DOUBLE && precision != CG_SINGLE) return CG_ERR_INVALID;
    for (int64_t p = 0; p < count; ++p)
        if (!systems[p] || !sigmas[p] || !systems[p]->multi || !sigma_matches(*systems[p], *sigmas[p]) ||
            sigmas[p]->labels.size() != sigmas[0]->labels.size())
            return CG_ERR_INVALID;
    if (count == 0) return CG_OK;
    try {
        cg_batch* bt = nullptr;
        int32_t rc = gpu::create_batch(systems, sigmas, count, precision, &bt);
        if (rc != CG_OK) return rc;
        std::unique_ptr<cg_batch, void (*)(cg_batch*)> hold(bt, gpu::destroy_batch);
        return gpu::run_batch_labels(bt, out, out_status, stats);
    } catch (const std::exception& e_) {
        note_error(e_);
        return CG_ERR_DEVICE;
    }
}

int32_t cg_multi_scattering_states(const cg_system* s, const cg_sigma* g, int64_t* cols, double* const* psi) {
    if (!s || !g || !cols || !s->multi || !sigma_matches(*s, *g)) return CG_ERR_INVALID;
    const size_t L = g->labels.size();
    for (size_t l = 0; l < L; ++l) {
        cols[l] = 0;
        for (auto& mb : g->mblocks) cols[l] += mb.lab_off[l + 1] - mb.lab_off[l];
    }
    if (!psi) return CG_OK;
    const int64_t n = int64_t(s->spec.h.n);
    for (size_t l = 0; l < L; ++l)
        if (cols[l] > 0) {
            if (!psi[l]) return CG_ERR_INVALID;
            std::fill(psi[l], psi[l] + 2 * n * cols[l], 0.0);
        }
    try {
        // nested_multi.rs:354-413: per block the columns of every label, the blocks' columns of a
        // label one after another.
        std::vector<int64_t> c0(L, 0);
        for (size_t b = 0; b < g->mblocks.size(); ++b) {
            const auto& off = g->mblocks[b].lab_off;
            std::vector<cplx> x;
            const int32_t rc = gpu::multi_block_solve(s, g, int(b), nullptr, &x);
            if (rc != CG_OK) return rc;
            for (size_t l = 0; l < L; ++l) {
                const int64_t w = off[l + 1] - off[l];
                if (w > 0)
                    std::memcpy(psi[l] + 2 * n * c0[l], x.data() + size_t(n) * size_t(off[l]),
                                size_t(n) * size_t(w) * sizeof(cplx));
                c0[l] += w;
            }
        }
        return CG_OK;
    } catch (const std::exception& e_) {
        note_error(e_);
        return CG_ERR_DEVICE;
    }
}

}  // extern "C"

namespace {
// The scattering matrix in the basis of the propagating lead modes (crates/caroli-green/src/
// smatrix.rs, NestedMultiTerminal::smatrix, nested_multi.rs:440-470): per block, G on the kept
// orbitals from the GPU root, and for every lead q and incoming mode phi (Bloch factor lambda) the
// source V_q (phi lambda - g H1^H phi) on the device; its wave G V_q s reaches the first cell of
// lead p as psi_1 = g V_p^H G_pq V_q s (minus g H1^H phi for p = q), whose amplitude on the
// outgoing mode (lambda') is i conj(lambda') [phi'^H (H1 F) psi_1 - conj(lambda') phi'^H H1^H psi_1].
struct SBlock {
    std::vector<std::vector<cplx>> in, out;  // per lead: Bloch factors
    std::vector<std::vector<CMat>> s;        // [p][q]
};

int32_t multi_smatrix(const cg_system* s, double energy, bool counts_only, std::vector<std::vector<cplx>>& in,
                      std::vector<std::vector<cplx>>& out, std::vector<std::vector<CMat>>& full) {
    cg_sigma* graw = nullptr;
    const size_t P = s->multi_leads.size();
    int32_t rc = multi_sigmas(s, energy, 0.0, std::vector<std::vector<int32_t>>(P), &graw);
    if (rc != CG_OK) return rc;
    std::unique_ptr<cg_sigma> g(graw);
    const cplx z(energy, 0.0), iu(0.0, 1.0);
    std::vector<SBlock> parts;
    for (size_t b = 0; b < s->sym.blocks.size(); ++b) {
        const Block& blk = s->sym.blocks[b];
        struct Solution {
            LeadModes modes;
            CMat rest, g;
        };
        std::map<const ModeSolver*, std::shared_ptr<Solution>> solved;
        std::vector<std::shared_ptr<Solution>> sol(P);
        for (size_t p = 0; p < P; ++p) {
            const AttachedLead& l = blk.leads[p];
            if (l.orbitals.empty()) continue;
            const ModeSolver* solver = s->lead_modes[b][p].get();
            auto it = solved.find(solver);
            if (it == solved.end()) {
                auto so = std::make_shared<Solution>();
                if (Status st = solver->modes_at(energy, so->modes); st != kOk) return st;
                so->rest = so->modes.sigma_right();
                const int nc = l.h00.rows;
                CMat a(nc, nc);
                for (int j = 0; j < nc; ++j)
                    for (int i = 0; i < nc; ++i) a(i, j) = (i == j ? z : cplx(0.0)) - l.h00(i, j) - so->rest(i, j);
                so->g = dense::inverse(a);
                it = solved.emplace(solver, so).first;
            }
            sol[p] = it->second;
        }
        SBlock part;
        part.in.resize(P);
        part.out.resize(P);
        part.s.assign(P, std::vector<CMat>(P));
        for (size_t p = 0; p < P; ++p)
            if (sol[p]) {
                part.in[p] = sol[p]->modes.left_lambdas;
                part.out[p] = sol[p]->modes.right_lambdas;
            }
        if (counts_only) {  // the mode counts need the leads only
            parts.push_back(std::move(part));
            continue;
        }
        CMat green;
        rc = gpu::multi_block_solve(s, g.get(), int(b), &green, nullptr);
        if (rc != CG_OK) return rc;
        if (!std::isfinite(green.norm_max())) return CG_ERR_SINGULAR;
        Positions kept(blk.kept);
        std::vector<std::vector<int>> at(P);
        for (size_t p = 0; p < P; ++p)
            for (size_t o : blk.leads[p].orbitals) at[p].push_back(kept.get(o));
        for (size_t q = 0; q < P; ++q) {
            for (size_t p = 0; p < P; ++p) part.s[p][q] = CMat(int(part.out[p].size()), int(part.in[q].size()));
            if (!sol[q]) continue;
            const AttachedLead& lq = blk.leads[q];
            const CMat& phi_in = sol[q]->modes.left_modes;
            const std::vector<cplx>& lam_in = sol[q]->modes.left_lambdas;
            const CMat back = matmul(sol[q]->g, matmul(lq.h01.adjoint(), phi_in));
            CMat source(phi_in.rows, phi_in.cols);
            for (int c = 0; c < phi_in.cols; ++c)
                for (int a = 0; a < phi_in.rows; ++a) source(a, c) = phi_in(a, c) * lam_in[size_t(c)] - back(a, c);
            const CMat device_source = matmul(lq.coupling, source);
            for (size_t p = 0; p < P; ++p) {
                if (!sol[p]) continue;
                const AttachedLead& lp = blk.leads[p];
                CMat gpq(int(at[p].size()), int(at[q].size()));
                for (size_t y = 0; y < at[q].size(); ++y)
                    for (size_t x = 0; x < at[p].size(); ++x) gpq(int(x), int(y)) = green(at[p][x], at[q][y]);
                CMat psi1 = matmul(sol[p]->g, matmul(lp.coupling.adjoint(), matmul(gpq, device_source)));
                if (p == q)
                    for (size_t e = 0; e < psi1.a.size(); ++e) psi1.a[e] -= back.a[e];
                const CMat& phi_out = sol[p]->modes.right_modes;
                const std::vector<cplx>& lam_out = sol[p]->modes.right_lambdas;
                const CMat a = matmul(phi_out.adjoint(), matmul(sol[p]->rest, psi1));
                const CMat bm = matmul(phi_out.adjoint(), matmul(lp.h01.adjoint(), psi1));
                CMat& out_pq = part.s[p][q];
                for (int c = 0; c < out_pq.cols; ++c)
                    for (int m = 0; m < out_pq.rows; ++m) {
                        const cplx cj = std::conj(lam_out[size_t(m)]);
                        out_pq(m, c) = iu * cj * (a(m, c) - cj * bm(m, c));
                    }
            }
        }
        parts.push_back(std::move(part));
    }
    // SMatrix::assemble (smatrix.rs:62-90): a lead's modes are those of every block in turn, and
    // the blocks do not mix.
    in.assign(P, {});
    out.assign(P, {});
    for (auto& part : parts)
        for (size_t p = 0; p < P; ++p) {
            in[p].insert(in[p].end(), part.in[p].begin(), part.in[p].end());
            out[p].insert(out[p].end(), part.out[p].begin(), part.out[p].end());
        }
    full.assign(P, std::vector<CMat>(P));
    if (counts_only) return CG_OK;
    for (size_t p = 0; p < P; ++p)
        for (size_t q = 0; q < P; ++q) {
            CMat m(int(out[p].size()), int(in[q].size()));
            int r0 = 0, c0 = 0;
            for (auto& part : parts) {
                const CMat& bq = part.s[p][q];
                for (int c = 0; c < bq.cols; ++c)
                    for (int r = 0; r < bq.rows; ++r) m(r0 + r, c0 + c) = bq(r, c);
                r0 += int(part.out[p].size());
                c0 += int(part.in[q].size());
            }
            full[p][q] = std::move(m);
        }
    return CG_OK;
}
}  // namespace

extern "C" {

int32_t cg_multi_smatrix(const cg_system* s, double energy, int32_t* modes, double* full, double* lambda_out,
                         double* lambda_in) {
    if (!s || !s->multi || !modes || !std::isfinite(energy)) return CG_ERR_INVALID;
    try {
        std::vector<std::vector<cplx>> in, out;
        std::vector<std::vector<CMat>> blocks;
        const int32_t rc = multi_smatrix(s, energy, !full && !lambda_out && !lambda_in, in, out, blocks);
        if (rc != CG_OK) return rc;
        const size_t P = in.size();
        size_t rows = 0;
        for (size_t p = 0; p < P; ++p) {
            modes[p] = int32_t(out[p].size());
            modes[P + p] = int32_t(in[p].size());
            rows += out[p].size();
        }
        if (full) {
            size_t r0 = 0;
            for (size_t p = 0; p < P; ++p) {
                size_t c0 = 0;
                for (size_t q = 0; q < P; ++q) {
                    const CMat& b = blocks[p][q];
                    for (int c = 0; c < b.cols; ++c)
                        for (int r = 0; r < b.rows; ++r) {
                            const size_t e = (r0 + size_t(r)) + (c0 + size_t(c)) * rows;
                            full[2 * e] = b(r, c).real();
                            full[2 * e + 1] = b(r, c).imag();
                        }
                    c0 += in[q].size();
                }
                r0 += out[p].size();
            }
        }
        size_t k = 0;
        for (size_t p = 0; p < P; ++p)
            for (cplx v : out[p]) {
                if (lambda_out) {
                    lambda_out[2 * k] = v.real();
                    lambda_out[2 * k + 1] = v.imag();
                }
                ++k;
            }
        k = 0;
        for (size_t p = 0; p < P; ++p)
            for (cplx v : in[p]) {
                if (lambda_in) {
                    lambda_in[2 * k] = v.real();
                    lambda_in[2 * k + 1] = v.imag();
                }
                ++k;
            }
        return CG_OK;
    } catch (const std::exception& e_) {
        note_error(e_);
        return CG_ERR_DEVICE;
    }
}

int32_t cg_multi_transmissions(const cg_system* s, double energy, int32_t precision, double* out) {
    if (!s || !s->multi || !out) return CG_ERR_INVALID;
    cg_sigma* g = nullptr;
    int32_t rc = cg_multi_self_energies(s, energy, 0.0, nullptr, nullptr, &g);
    if (rc != CG_OK) return rc;
    std::unique_ptr<cg_sigma, void (*)(cg_sigma*)> hold(g, cg_sigma_destroy);
    int32_t st = CG_OK;
    rc = cg_multi_batch_transmissions(&s, const_cast<const cg_sigma* const*>(&g), 1, precision, out, &st, nullptr);
    return rc != CG_OK ? rc : st;
}

}  // extern "C"
