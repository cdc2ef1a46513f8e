// ORACLE — test infrastructure only. A CPU FP64 restatement of the nested-dissection transmission
// of Arnwald/caroli, used as the checker of the GPU engine (by the tests and as the CPU baseline of
// the benchmark). Nothing in the product links or calls this file.
//
// It restates crates/caroli-green/src/nested.rs function by function:
//   split                    nested.rs:1051-1093
//   symbolic                 nested.rs:1097-1122
//   front_vars / entries     nested.rs:147-179
//   Plan::new                nested.rs:127-141
//   Tree::root / front       nested.rs:246-374  (planned and delayed assembly, delays)
//   runs / add_runs          nested.rs:1126-1162
//   small_schur              nested.rs:1167-1223
//   schur (Field::schur c64) crates/caroli-dense/src/field.rs:79-105
//   Components / blocks      nested.rs:378-420, :559-609
//   root_with_leads          nested.rs:806-845
//   lead_green               nested.rs:783-802
//   caroli_trace/broadening  crates/caroli-green/src/lib.rs:68-71, :108-113
//   Factor::back_substitute  nested.rs:200-227 (G columns on lead slices, scattering_states)
// The lead self-energies are inputs (as `transmission_with`, nested.rs:682-713).
// Parity pins: tests/test_host.py checks it against dense inversion of z - H - Sigma on small
// systems (nested.rs:1-7 defines the same quantity), clean-ribbon channel counts, the chain
// closed forms and the reference's golden data.
//
// The scalar type is ORACLE_REAL (double by default). Built with ORACLE_REAL=long double
// (oracle/_ref/liboracle_ld.so) it computes the same elimination with a 64-bit mantissa: the
// tests use it to measure the rounding floor of the double-precision results on large systems.
#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

#ifndef ORACLE_REAL
#define ORACLE_REAL double
#endif
using real = ORACLE_REAL;
using cx = std::complex<real>;

struct Mat {
    int r = 0, c = 0;
    std::vector<cx> a;
    Mat() = default;
    Mat(int rr, int cc) : r(rr), c(cc), a(size_t(rr) * cc) {}
    cx& operator()(int i, int j) { return a[size_t(i) + size_t(j) * r]; }
    cx operator()(int i, int j) const { return a[size_t(i) + size_t(j) * r]; }
};

constexpr size_t kLeaf = 16;    // nested.rs:34
constexpr int kSmall = 64;      // nested.rs:37
constexpr double kTol = 1e-11;  // field.rs:61

struct Csr {
    size_t n;
    std::vector<int64_t> rp;
    std::vector<uint32_t> ci;
    std::vector<cx> v;
};

struct ONode {
    std::vector<size_t> elim, bnd;
    std::vector<int> kids;
};

struct OPlan {
    std::vector<size_t> vars;
    int ne;
};

struct OBlock {
    std::vector<ONode> nodes;
    std::vector<int> top;
    std::vector<OPlan> plans;
    std::vector<size_t> left, right, kept;
    int right_at;
};

struct OSystem {
    Csr h;
    std::vector<OBlock> blocks;
};

// Sorted positions of a variable list.
struct Pos {
    std::vector<std::pair<size_t, int>> p;
    explicit Pos(const std::vector<size_t>& vars) {
        for (size_t k = 0; k < vars.size(); ++k) p.push_back({vars[k], int(k)});
        std::sort(p.begin(), p.end());
    }
    int get(size_t x) const {
        auto it = std::lower_bound(p.begin(), p.end(), std::make_pair(x, INT_MIN));
        return (it != p.end() && it->first == x) ? it->second : -1;
    }
};

bool has(const std::vector<size_t>& s, size_t x) { return std::binary_search(s.begin(), s.end(), x); }

struct Geometry {
    const int64_t* cells;
    const int* orb;
    int dim;
    int64_t range[3];
};

// nested.rs:1051-1093
int split(const Geometry& g, std::vector<size_t> sites, std::vector<std::pair<std::vector<size_t>, std::vector<int>>>& nodes) {
    if (sites.empty()) return -1;
    size_t orbitals = 0;
    for (size_t i : sites) orbitals += g.orb[i];
    int d = 0;
    int64_t lo = 0, hi = -1;
    for (int dd = 0; dd < g.dim; ++dd) {
        int64_t l = LLONG_MAX, h = LLONG_MIN;
        for (size_t i : sites) {
            l = std::min(l, g.cells[3 * i + dd]);
            h = std::max(h, g.cells[3 * i + dd]);
        }
        if (dd == 0 || h - l > hi - lo) {
            d = dd;
            lo = l;
            hi = h;
        }
    }
    if (orbitals <= kLeaf || hi - lo < g.range[d]) {
        nodes.push_back({sites, {}});
        return int(nodes.size()) - 1;
    }
    int64_t mid = lo + (hi - lo + 1 - g.range[d]) / 2;
    std::vector<size_t> a, sep, b;
    for (size_t i : sites) {
        int64_t c = g.cells[3 * i + d];
        (c < mid ? a : c < mid + g.range[d] ? sep : b).push_back(i);
    }
    std::vector<int> kids;
    int x = split(g, a, nodes);
    if (x >= 0) kids.push_back(x);
    int y = split(g, b, nodes);
    if (y >= 0) kids.push_back(y);
    nodes.push_back({sep, kids});
    return int(nodes.size()) - 1;
}

// nested.rs:147-159
std::pair<std::vector<size_t>, int> front_vars(const ONode& node, const std::vector<const std::vector<size_t>*>& handed) {
    std::vector<size_t> e = node.elim;
    for (auto* h : handed)
        for (size_t o : *h)
            if (!has(node.bnd, o)) e.push_back(o);
    std::sort(e.begin(), e.end());
    e.erase(std::unique(e.begin(), e.end()), e.end());
    int ne = int(e.size());
    e.insert(e.end(), node.bnd.begin(), node.bnd.end());
    return {e, ne};
}

// nested.rs:164-179, applied as F[a,b] -= value.
void separator_entries(const Csr& h, const ONode& node, const Pos& pos, Mat& f) {
    for (size_t i : node.elim) {
        int pi = pos.get(i);
        for (int64_t e = h.rp[i]; e < h.rp[i + 1]; ++e) {
            size_t j = h.ci[e];
            cx v = h.v[e];
            if (j == i) {
                f(pi, pi) -= v;
                continue;
            }
            int pj = pos.get(j);
            if (pj < 0) continue;
            bool later = has(node.elim, j) ? j > i : has(node.bnd, j);
            if (later) {
                f(pi, pj) -= v;
                f(pj, pi) -= std::conj(v);
            }
        }
    }
}

struct Factor {
    std::vector<size_t> elim, bnd;
    Mat x;
    std::vector<Factor> kids;
};

struct Front {
    std::vector<size_t> vars;
    Mat s;
    Factor factor;
};

// nested.rs:1167-1223; returns false when a pivot fails the test.
bool small_schur(const Mat& f, int ne, bool want_x, Mat& s, Mat& x) {
    const int n = f.r, nb = n - ne;
    real scale = 0.0;
    for (int j = 0; j < ne; ++j)
        for (int i = 0; i < ne; ++i) scale = std::max(scale, std::abs(f(i, j)));
    Mat w = f;
    for (int k = 0; k < ne; ++k) {
        int p = k;
        real big = -1.0;
        for (int i = k; i < ne; ++i) {
            real m = std::abs(w(i, k));
            if (m > big) {
                big = m;
                p = i;
            }
        }
        if (!(big > kTol * scale)) return false;
        if (p != k)
            for (int j = 0; j < n; ++j) std::swap(w(p, j), w(k, j));
        cx inv = real(1) / w(k, k);
        for (int i = k + 1; i < n; ++i) w(i, k) *= inv;
        for (int j = k + 1; j < n; ++j) {
            cx u = w(k, j);
            if (u == cx(0.0)) continue;
            for (int i = k + 1; i < n; ++i) w(i, j) -= w(i, k) * u;
        }
    }
    s = Mat(nb, nb);
    for (int j = 0; j < nb; ++j)
        for (int i = 0; i < nb; ++i) s(i, j) = w(ne + i, ne + j);
    x = Mat(ne, want_x ? nb : 0);
    if (want_x) {
        for (int j = 0; j < nb; ++j) {
            for (int i = 0; i < ne; ++i) x(i, j) = w(i, ne + j);
            for (int i = ne - 1; i >= 0; --i) {
                cx v = x(i, j) / w(i, i);
                x(i, j) = v;
                for (int l = 0; l < i; ++l) x(l, j) -= w(l, i) * v;
            }
        }
    }
    return true;
}

// LU with partial pivoting (rows), in place; piv[k] = row swapped with k.
void lu(Mat& a, std::vector<int>& piv) {
    const int n = a.r;
    piv.resize(n);
    for (int k = 0; k < n; ++k) {
        int p = k;
        real big = -1.0;
        for (int i = k; i < n; ++i)
            if (std::abs(a(i, k)) > big) {
                big = std::abs(a(i, k));
                p = i;
            }
        piv[k] = p;
        if (p != k)
            for (int j = 0; j < n; ++j) std::swap(a(p, j), a(k, j));
        if (a(k, k) == cx(0.0)) continue;
        cx inv = real(1) / a(k, k);
        for (int i = k + 1; i < n; ++i) a(i, k) *= inv;
        for (int j = k + 1; j < n; ++j) {
            cx u = a(k, j);
            if (u == cx(0.0)) continue;
            for (int i = k + 1; i < n; ++i) a(i, j) -= a(i, k) * u;
        }
    }
}

void lu_solve(const Mat& a, const std::vector<int>& piv, Mat& b) {
    const int n = a.r;
    for (int c = 0; c < b.c; ++c) {
        for (int k = 0; k < n; ++k)
            if (piv[k] != k) std::swap(b(k, c), b(piv[k], c));
        for (int k = 0; k < n; ++k)
            for (int i = k + 1; i < n; ++i) b(i, c) -= a(i, k) * b(k, c);
        for (int k = n - 1; k >= 0; --k) {
            b(k, c) /= a(k, k);
            for (int i = 0; i < k; ++i) b(i, c) -= a(i, k) * b(k, c);
        }
    }
}

// field.rs:79-105: LU of f11, pivot test against max |f11|, X = f11^-1 f12, S = f22 - f21 X.
bool schur(const Mat& f, int ne, Mat& s, Mat& x) {
    const int n = f.r, nb = n - ne;
    Mat f11(ne, ne);
    real scale = 0.0;
    for (int j = 0; j < ne; ++j)
        for (int i = 0; i < ne; ++i) {
            f11(i, j) = f(i, j);
            scale = std::max(scale, std::abs(f(i, j)));
        }
    std::vector<int> piv;
    lu(f11, piv);
    real small = INFINITY;
    for (int i = 0; i < ne; ++i) small = std::min(small, std::abs(f11(i, i)));
    if (!(small > kTol * scale)) return false;
    x = Mat(ne, nb);
    for (int j = 0; j < nb; ++j)
        for (int i = 0; i < ne; ++i) x(i, j) = f(i, ne + j);
    lu_solve(f11, piv, x);
    s = Mat(nb, nb);
    for (int j = 0; j < nb; ++j) {
        for (int i = 0; i < nb; ++i) s(i, j) = f(ne + i, ne + j);
        for (int k = 0; k < ne; ++k) {
            cx xk = x(k, j);
            for (int i = 0; i < nb; ++i) s(i, j) -= f(ne + i, k) * xk;
        }
    }
    return true;
}

// nested.rs:1126-1162: F[p, q] += S[a, b] at the positions of the child's variables.
void extend_add(Mat& f, const Pos& pos, const Front& kid) {
    std::vector<int> at(kid.vars.size());
    for (size_t a = 0; a < kid.vars.size(); ++a) at[a] = pos.get(kid.vars[a]);
    for (size_t b = 0; b < at.size(); ++b)
        for (size_t a = 0; a < at.size(); ++a) f(at[a], at[b]) += kid.s(int(a), int(b));
}

// nested.rs:304-374
Front front(const OSystem& sys, const OBlock& blk, int k, cx z, bool keep) {
    const ONode& node = blk.nodes[k];
    std::vector<Front> kids;
    for (int c : node.kids) kids.push_back(front(sys, blk, c, z, keep));
    bool planned = true;
    for (size_t q = 0; q < kids.size(); ++q) planned &= kids[q].vars == blk.nodes[node.kids[q]].bnd;
    std::vector<size_t> vars;
    int ne;
    if (planned) {
        vars = blk.plans[k].vars;
        ne = blk.plans[k].ne;
    } else {
        std::vector<const std::vector<size_t>*> handed;
        for (auto& kd : kids) handed.push_back(&kd.vars);
        std::tie(vars, ne) = front_vars(node, handed);
    }
    Pos pos(vars);
    Mat f(int(vars.size()), int(vars.size()));
    separator_entries(sys.h, node, pos, f);
    for (size_t o : node.elim) f(pos.get(o), pos.get(o)) += z;
    Front out;
    for (auto& kd : kids) {
        extend_add(f, pos, kd);
        if (keep) out.factor.kids.push_back(std::move(kd.factor));
    }
    Mat s, x;
    bool ok = ne > 0 && (f.r < kSmall ? small_schur(f, ne, keep, s, x) : schur(f, ne, s, x));
    if (!ok) {  // singular separator (or nothing to eliminate): hand the whole front up
        out.vars = vars;
        out.s = std::move(f);
        return out;
    }
    out.vars = node.bnd;
    out.s = std::move(s);
    if (keep) {
        out.factor.elim.assign(vars.begin(), vars.begin() + ne);
        out.factor.bnd = node.bnd;
        out.factor.x = std::move(x);
    }
    return out;
}

// Factor::back_substitute (nested.rs:200-227): psi rows of eliminated variables from their boundary.
void back_substitute(const Factor& fac, Mat& psi) {
    if (!fac.elim.empty()) {
        for (int c = 0; c < psi.c; ++c) {
            for (size_t a = 0; a < fac.elim.size(); ++a) {
                cx acc = 0.0;
                for (size_t b = 0; b < fac.bnd.size(); ++b) acc += fac.x(int(a), int(b)) * psi(int(fac.bnd[b]), c);
                psi(int(fac.elim[a]), c) = -acc;
            }
        }
    }
    for (auto& kd : fac.kids) back_substitute(kd, psi);
}

struct RootOut {
    Mat s;  // Schur complement on kept
    std::vector<size_t> delayed;
    Factor factor;  // delayed elimination with the top factors as children
};

// Tree::root (nested.rs:246-281) and root_with_leads (nested.rs:806-845).
int root_with_leads(const OSystem& sys, const OBlock& blk, cx z, const cx* sl, const cx* sr, bool keep, RootOut& out) {
    std::vector<Front> kids;
    for (int t : blk.top) kids.push_back(front(sys, blk, t, z, keep));
    Pos in_kept(blk.kept);
    std::vector<size_t> delayed;
    for (auto& kd : kids)
        for (size_t o : kd.vars)
            if (in_kept.get(o) < 0) delayed.push_back(o);
    std::sort(delayed.begin(), delayed.end());
    delayed.erase(std::unique(delayed.begin(), delayed.end()), delayed.end());
    std::vector<size_t> vars = delayed;
    vars.insert(vars.end(), blk.kept.begin(), blk.kept.end());
    Pos pos(vars);
    Mat f(int(vars.size()), int(vars.size()));
    for (size_t i : blk.kept) {
        int pi = pos.get(i);
        f(pi, pi) += z;
        for (int64_t e = sys.h.rp[i]; e < sys.h.rp[i + 1]; ++e)
            if (in_kept.get(sys.h.ci[e]) >= 0) f(pi, pos.get(sys.h.ci[e])) -= sys.h.v[e];
    }
    for (auto& kd : kids) {
        extend_add(f, pos, kd);
        if (keep) out.factor.kids.push_back(std::move(kd.factor));
    }
    const int nd = int(delayed.size()), nl = int(blk.left.size()), nr = int(blk.right.size());
    for (int c = 0; c < nl; ++c)
        for (int a = 0; a < nl; ++a) f(nd + a, nd + c) -= sl[a + size_t(c) * nl];
    for (int c = 0; c < nr; ++c)
        for (int a = 0; a < nr; ++a) f(nd + blk.right_at + a, nd + blk.right_at + c) -= sr[a + size_t(c) * nr];
    if (nd > 0) {
        Mat x;
        if (!schur(f, nd, out.s, x)) return 1;  // GreenError::Singular
        out.factor.x = std::move(x);
        out.factor.elim = delayed;
        out.factor.bnd = blk.kept;
    } else {
        out.s = std::move(f);
    }
    out.delayed = delayed;
    return 0;
}

// nested.rs:783-802: G_LR from the LU of S with unit columns on the right slice.
int lead_green(const OSystem& sys, const OBlock& blk, cx z, const cx* sl, const cx* sr, Mat& g) {
    RootOut root;
    if (root_with_leads(sys, blk, z, sl, sr, false, root)) return 1;
    const int k = root.s.r, nl = int(blk.left.size()), nr = int(blk.right.size());
    Mat rhs(k, nr);
    for (int j = 0; j < nr; ++j) rhs(blk.right_at + j, j) = 1.0;
    std::vector<int> piv;
    lu(root.s, piv);
    lu_solve(root.s, piv, rhs);
    for (auto& v : rhs.a)
        if (!std::isfinite(v.real()) || !std::isfinite(v.imag())) return 1;
    g = Mat(nl, nr);
    for (int j = 0; j < nr; ++j)
        for (int i = 0; i < nl; ++i) g(i, j) = rhs(i, j);
    return 0;
}

// lib.rs:68-71
Mat broadening(const cx* s, int n) {
    Mat g(n, n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) g(i, j) = cx(0, 1) * (s[i + size_t(j) * n] - std::conj(s[j + size_t(i) * n]));
    return g;
}

Mat mul(const Mat& a, const Mat& b) {
    Mat c(a.r, b.c);
    for (int j = 0; j < b.c; ++j)
        for (int k = 0; k < a.c; ++k) {
            cx v = b(k, j);
            for (int i = 0; i < a.r; ++i) c(i, j) += a(i, k) * v;
        }
    return c;
}

// lib.rs:108-113
real caroli_trace(const Mat& gp, const Mat& g, const Mat& gq) {
    Mat a = mul(gp, g), b = mul(g, gq);
    real t = 0.0;
    for (size_t e = 0; e < a.a.size(); ++e) t += (a.a[e] * std::conj(b.a[e])).real();
    return t;
}

// nested.rs:378-420
struct UF {
    std::vector<size_t> p;
    explicit UF(size_t n) : p(n) {
        for (size_t i = 0; i < n; ++i) p[i] = i;
    }
    size_t find(size_t i) {
        while (p[i] != i) {
            p[i] = p[p[i]];
            i = p[i];
        }
        return i;
    }
    void join(size_t a, size_t b) {
        size_t x = find(a), y = find(b);
        if (x != y) p[std::max(x, y)] = std::min(x, y);
    }
};

// nested.rs:1097-1122
void symbolic(const Csr& h, std::vector<ONode>& nodes) {
    std::vector<size_t> rank(h.n, SIZE_MAX), last(nodes.size());
    size_t next = 0;
    for (size_t k = 0; k < nodes.size(); ++k) {
        for (size_t o : nodes[k].elim) rank[o] = next++;
        last[k] = next;
    }
    for (size_t k = 0; k < nodes.size(); ++k) {
        std::vector<size_t> b;
        for (size_t o : nodes[k].elim)
            for (int64_t e = h.rp[o]; e < h.rp[o + 1]; ++e) b.push_back(h.ci[e]);
        for (int c : nodes[k].kids) b.insert(b.end(), nodes[c].bnd.begin(), nodes[c].bnd.end());
        std::vector<size_t> keep;
        for (size_t o : b)
            if (rank[o] == SIZE_MAX || rank[o] >= last[k]) keep.push_back(o);
        std::sort(keep.begin(), keep.end());
        keep.erase(std::unique(keep.begin(), keep.end()), keep.end());
        nodes[k].bnd = keep;
    }
}

// The thesis transmission (crates/caroli-validation/src/thesis.rs:89-95): Tr[Re(Gamma_L) G
// Re(Gamma_R) G^H], the broadenings truncated to their real part element by element.
Mat real_part(Mat m) {
    for (auto& v : m.a) v = cx(v.real(), 0.0);
    return m;
}

// Interleaved (re, im) doubles of the C interface -> the oracle's complex scalars.
std::vector<cx> to_cx(const double* v, size_t words) {
    std::vector<cx> out(words);
    for (size_t e = 0; e < words; ++e) out[e] = cx(v[2 * e], v[2 * e + 1]);
    return out;
}

}  // namespace

extern "C" {

// Builds the oracle's copy of a system: the CSR Hamiltonian of the region, the site geometry and
// the lead cells (column-major nl x nl, nr x nr). Returns null on malformed input.
void* oracle_create(uint64_t n, const int64_t* row_ptr, const uint32_t* cols, const double* vals, uint64_t nsites,
                    const int64_t* cells, const int32_t* site_orbitals, const uint64_t* site_offsets, int32_t dim,
                    const int64_t* range, int32_t axis, const double* lh00, const double* lh01, uint64_t nl,
                    const double* rh00, const double* rh01, uint64_t nr) {
    auto* sys = new OSystem;
    sys->h.n = n;
    sys->h.rp.assign(row_ptr, row_ptr + n + 1);
    size_t nnz = size_t(row_ptr[n]);
    sys->h.ci.assign(cols, cols + nnz);
    sys->h.v.resize(nnz);
    for (size_t e = 0; e < nnz; ++e) sys->h.v[e] = cx(vals[2 * e], vals[2 * e + 1]);
    int64_t first = LLONG_MAX, last = LLONG_MIN;
    for (size_t i = 0; i < nsites; ++i) {
        first = std::min(first, cells[3 * i + axis]);
        last = std::max(last, cells[3 * i + axis]);
    }
    std::vector<size_t> left, right, interior;
    for (size_t i = 0; i < nsites; ++i) {
        int64_t c = cells[3 * i + axis];
        for (int o = 0; o < site_orbitals[i]; ++o) {
            if (c == first) left.push_back(site_offsets[i] + o);
            if (c == last) right.push_back(site_offsets[i] + o);
        }
        if (c != first && c != last) interior.push_back(i);
    }
    if (left.size() != nl || right.size() != nr) {
        delete sys;
        return nullptr;
    }
    Geometry g{cells, site_orbitals, dim, {range[0], range[1], range[2]}};
    std::vector<std::pair<std::vector<size_t>, std::vector<int>>> snodes;
    int top = split(g, interior, snodes);
    UF uf(n);
    for (size_t r = 0; r < n; ++r)
        for (int64_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e)
            if (vals[2 * e] != 0.0 || vals[2 * e + 1] != 0.0) uf.join(r, cols[e]);
    auto join_m = [&](const double* m, const std::vector<size_t>& o) {
        size_t k = o.size();
        for (size_t b = 0; b < k; ++b)
            for (size_t a = 0; a < k; ++a)
                if (m[2 * (a + b * k)] != 0.0 || m[2 * (a + b * k) + 1] != 0.0) uf.join(o[a], o[b]);
    };
    join_m(lh00, left);
    join_m(lh01, left);
    join_m(rh00, right);
    join_m(rh01, right);
    std::vector<size_t> roots(n), labels;
    for (size_t i = 0; i < n; ++i) roots[i] = uf.find(i);
    labels = roots;
    std::sort(labels.begin(), labels.end());
    labels.erase(std::unique(labels.begin(), labels.end()), labels.end());
    for (size_t lab : labels) {
        OBlock b;
        for (size_t o : left)
            if (roots[o] == lab) b.left.push_back(o);
        for (size_t o : right)
            if (roots[o] == lab) b.right.push_back(o);
        if (b.left.empty() || b.right.empty()) continue;
        bool same = b.left == b.right;
        b.kept = b.left;
        b.right_at = same ? 0 : int(b.left.size());
        if (!same) b.kept.insert(b.kept.end(), b.right.begin(), b.right.end());
        for (auto& [sites, kids] : snodes) {
            ONode nd;
            for (size_t i : sites)
                for (int o = 0; o < site_orbitals[i]; ++o)
                    if (roots[site_offsets[i] + o] == lab) nd.elim.push_back(site_offsets[i] + o);
            std::sort(nd.elim.begin(), nd.elim.end());
            nd.kids = kids;
            b.nodes.push_back(nd);
        }
        if (top >= 0) b.top.push_back(top);
        symbolic(sys->h, b.nodes);
        for (auto& nd : b.nodes) {
            std::vector<const std::vector<size_t>*> handed;
            for (int c : nd.kids) handed.push_back(&b.nodes[c].bnd);
            OPlan p;
            std::tie(p.vars, p.ne) = front_vars(nd, handed);
            b.plans.push_back(p);
        }
        sys->blocks.push_back(std::move(b));
    }
    return sys;
}

void oracle_destroy(void* h) { delete static_cast<OSystem*>(h); }

int32_t oracle_block_count(void* h) { return int32_t(static_cast<OSystem*>(h)->blocks.size()); }

// Sizes of block b: left and right slice orbitals, tree nodes; fills the orbital lists if given.
int32_t oracle_block_info(void* h, int32_t b, uint64_t* nl, uint64_t* nr, uint64_t* nodes, uint64_t* left, uint64_t* right) {
    auto& blk = static_cast<OSystem*>(h)->blocks.at(b);
    *nl = blk.left.size();
    *nr = blk.right.size();
    *nodes = blk.nodes.size();
    if (left) std::copy(blk.left.begin(), blk.left.end(), left);
    if (right) std::copy(blk.right.begin(), blk.right.end(), right);
    return 0;
}

// Node k of block b: sizes of eliminate/boundary, and their contents when arrays are given.
int32_t oracle_node(void* h, int32_t b, int32_t k, uint64_t* ne, uint64_t* nb, uint64_t* elim, uint64_t* bnd) {
    auto& nd = static_cast<OSystem*>(h)->blocks.at(b).nodes.at(k);
    *ne = nd.elim.size();
    *nb = nd.bnd.size();
    if (elim) std::copy(nd.elim.begin(), nd.elim.end(), elim);
    if (bnd) std::copy(nd.bnd.begin(), nd.bnd.end(), bnd);
    return 0;
}

static int32_t transmission_of(void* h, double energy, const double* sigmas, bool thesis, double* out_t);

// transmission_with (nested.rs:682-713) at `energy`, with the self-energies of each block
// concatenated (Sigma_L then Sigma_R, column-major complex). 0 = ok, 1 = Singular.
int32_t oracle_transmission_with(void* h, double energy, const double* sigmas, double* out_t) {
    return transmission_of(h, energy, sigmas, false, out_t);
}

// The same with the thesis formula (real parts of the broadenings), for the reference's
// ti-cond2t golden data (crates/caroli-validation/tests/ti_conductance_disorder.rs:25-58).
int32_t oracle_transmission_thesis(void* h, double energy, const double* sigmas, double* out_t) {
    return transmission_of(h, energy, sigmas, true, out_t);
}

static int32_t transmission_of(void* h, double energy, const double* sigmas, bool thesis, double* out_t) {
    auto* sys = static_cast<OSystem*>(h);
    size_t words = 0;
    for (auto& blk : sys->blocks) words += blk.left.size() * blk.left.size() + blk.right.size() * blk.right.size();
    const std::vector<cx> sig = to_cx(sigmas, words);
    const cx* s = sig.data();
    cx z(energy, 0.0);
    real total = 0.0;
    for (auto& blk : sys->blocks) {
        int nl = int(blk.left.size()), nr = int(blk.right.size());
        const cx* sl = s;
        const cx* sr = s + size_t(nl) * nl;
        s = sr + size_t(nr) * nr;
        Mat g;
        if (lead_green(*sys, blk, z, sl, sr, g)) {
            *out_t = NAN;
            return 1;
        }
        Mat gl = broadening(sl, nl), gr = broadening(sr, nr);
        if (thesis) {
            gl = real_part(gl);
            gr = real_part(gr);
        }
        total += caroli_trace(gl, g, gr);
    }
    *out_t = double(total);
    return 0;
}

// Columns of G = (z - H - Sigma)^-1 on the right slice of block b (which = 1) or the left slice
// (which = 0), on every device orbital: the root solve with unit sources, then the back-substitution
// through the kept factors (nested.rs:851-899 with W = 1). out is n x n_slice column-major.
int32_t oracle_green_columns(void* h, int32_t b, double energy, const double* sigma_l, const double* sigma_r, int32_t which, double* out) {
    auto* sys = static_cast<OSystem*>(h);
    auto& blk = sys->blocks.at(b);
    RootOut root;
    const std::vector<cx> sl = to_cx(sigma_l, blk.left.size() * blk.left.size());
    const std::vector<cx> sr = to_cx(sigma_r, blk.right.size() * blk.right.size());
    if (root_with_leads(*sys, blk, cx(energy, 0.0), sl.data(), sr.data(), true, root)) return 1;
    const int k = root.s.r;
    const int m = which ? int(blk.right.size()) : int(blk.left.size());
    const int at = which ? blk.right_at : 0;
    Mat x(k, m);
    for (int j = 0; j < m; ++j) x(at + j, j) = 1.0;
    std::vector<int> piv;
    lu(root.s, piv);
    lu_solve(root.s, piv, x);
    Mat psi(int(sys->h.n), m);
    for (size_t a = 0; a < blk.kept.size(); ++a)
        for (int c = 0; c < m; ++c) psi(int(blk.kept[a]), c) = x(int(a), c);
    back_substitute(root.factor, psi);
    for (size_t e = 0; e < psi.a.size(); ++e) {
        out[2 * e] = double(psi.a[e].real());
        out[2 * e + 1] = double(psi.a[e].imag());
    }
    return 0;
}

// CPU baseline: transmission_with for many (system, energy, sigma) points on `threads` threads,
// one point per task (as rayon's par_iter over points). Returns the number of failed points.
int32_t oracle_transmissions_parallel(void** systems, const double* energies, const double** sigmas, int64_t npoints,
                                      int32_t threads, double* out_t) {
    std::atomic<int64_t> next{0};
    std::atomic<int32_t> failed{0};
    auto work = [&]() {
        for (int64_t p; (p = next++) < npoints;)
            if (oracle_transmission_with(systems[p], energies[p], sigmas[p], &out_t[p])) ++failed;
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < std::max(1, threads); ++t) pool.emplace_back(work);
    for (auto& t : pool) t.join();
    return failed;
}

}  // extern "C"
