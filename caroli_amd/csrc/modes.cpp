// Lead modes and mode-matching self-energies (see modes.hpp for the reference lines).
#include "modes.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <limits>

#include "dense.hpp"

namespace caroli {

namespace {

// Tolerances of crates/caroli-green/src/modes.rs:33-50.
constexpr double kRankTol = 1e-12;
constexpr double kUnitTol = 1e-6;
constexpr double kDegenerateTol = 1e-7;
constexpr double kVelocityTol = 1e-8;
constexpr double kDefectiveTol = 1e-7;
constexpr double kSingularTol = 1e-11;
constexpr double kRealStabilisation = 0.6180339887498948;
constexpr double kShifts[5] = {0.0, 1e-9, 1e-8, 1e-7, 1e-6};
constexpr double kHermitianTol = 1e-14;
constexpr double kDefiniteTol = 1e-4;

double row_norm(const CMat& m) {
    double best = 0.0;
    for (int i = 0; i < m.rows; ++i) {
        double s = 0.0;
        for (int j = 0; j < m.cols; ++j) s += std::abs(m(i, j));
        best = std::max(best, s);
    }
    return best;
}

bool is_real(const CMat& m) {
    for (auto& z : m.a)
        if (z.imag() != 0.0) return false;
    return true;
}

CMat rows_of(const CMat& m, int r0, int r1) {
    CMat o(r1 - r0, m.cols);
    for (int j = 0; j < m.cols; ++j)
        for (int i = r0; i < r1; ++i) o(i - r0, j) = m(i, j);
    return o;
}

CMat cols_of(const CMat& m, int c0, int c1) {
    CMat o(m.rows, c1 - c0);
    for (int j = c0; j < c1; ++j)
        for (int i = 0; i < m.rows; ++i) o(i, j - c0) = m(i, j);
    return o;
}

CMat sandwich(const CMat& v, const CMat& m) { return matmul(matmul(v, m), v.adjoint()); }

// Smallest over largest singular value (modes.rs:94-105).
double inverse_condition(const CMat& m) {
    if (m.rows == 0) return 1.0;
    CMat u, v;
    std::vector<double> s;
    if (!dense::svd(m, u, s, v)) return 0.0;
    double lo = std::numeric_limits<double>::infinity(), hi = 0.0;
    for (double x : s) {
        lo = std::min(lo, x);
        hi = std::max(hi, x);
    }
    return hi > 0.0 ? lo / hi : 0.0;
}

// LeadModes::solve_with (modes.rs:159-486): the stabilised 2r pencil, its Schur form, the
// classification of the modes and the orthonormal bases of the right- and left-going subspaces.
Status solve_with(const CMat& h0, const CMat& h1, cplx z, double scale, bool real, LeadModes& out) {
    const int n = h0.rows;
    const cplx one(1.0), zero(0.0), iu(0.0, 1.0);
    const cplx s = real ? cplx(kRealStabilisation) : iu;
    CMat su, sv;
    std::vector<double> ss;
    if (!dense::svd(h1, su, ss, sv)) return kEigen;
    double smax = 0.0;
    for (double x : ss) smax = std::max(smax, x);
    std::vector<int> kept;
    for (int k = 0; k < int(ss.size()); ++k)
        if (ss[k] > kRankTol * smax) kept.push_back(k);
    const int r = int(kept.size());
    out.u = CMat(n, r);
    out.v = CMat(n, r);
    for (int k = 0; k < r; ++k) {
        double sq = std::sqrt(ss[kept[k]]);
        for (int a = 0; a < n; ++a) {
            out.u(a, k) = su(a, kept[k]) * sq;
            out.v(a, k) = sv(a, kept[k]) * sq;
        }
    }
    out.energy = z;
    if (r == 0) {
        out.right = out.left = CMat(0, 0);
        out.right_modes = out.left_modes = CMat(n, 0);
        out.right_lambdas.clear();
        out.left_lambdas.clear();
        out.channels = 0;
        return kOk;
    }
    // Stabilised resolvent (H0 - z + s (uu^H + vv^H))^-1 applied to [u v].
    CMat uv(n, 2 * r);
    for (int k = 0; k < r; ++k)
        for (int a = 0; a < n; ++a) {
            uv(a, k) = out.u(a, k);
            uv(a, r + k) = out.v(a, k);
        }
    CMat stab = matmul(uv, uv.adjoint());
    CMat am(n, n);
    for (int q = 0; q < n; ++q)
        for (int p = 0; p < n; ++p) am(p, q) = h0(p, q) - (p == q ? z : zero) + s * stab(p, q);
    dense::Lu lu_a(am);
    if (real) {
        auto [lo, hi] = lu_a.pivot_range();
        if (!(lo > 1e-10 * hi)) return kEigen;
    }
    CMat g = lu_a.solve(uv);
    for (auto& x : g.a)
        if (std::isnan(x.real()) || std::isnan(x.imag())) return kSingular;
    CMat gu = cols_of(g, 0, r), gv = cols_of(g, r, 2 * r);
    CMat blocks = matmul(uv.adjoint(), g);
    auto B = [&](int i, int j) { return blocks(i, j); };
    const int m2 = 2 * r;
    CMat lmat(m2, m2), rmat(m2, m2);
    for (int q = 0; q < m2; ++q)
        for (int p = 0; p < m2; ++p) {
            bool pt = p < r, qt = q < r;
            int pp = pt ? p : p - r, qq = qt ? q : q - r;
            cplx ugu = B(pp, qq), ugv = B(pp, r + qq), vgu = B(r + pp, qq), vgv = B(r + pp, r + qq);
            if (pt && qt) {
                lmat(p, q) = (p == q ? one : zero) - s * ugu;
                rmat(p, q) = -ugv;
            } else if (pt) {
                lmat(p, q) = ugu;
                rmat(p, q) = s * ugv;
            } else if (qt) {
                lmat(p, q) = -s * vgu;
                rmat(p, q) = -vgv;
            } else {
                lmat(p, q) = vgu;
                rmat(p, q) = s * vgv - (p == q ? one : zero);
            }
        }
    auto pivot_ratio = [](const dense::Lu& lu) {
        auto [lo, hi] = lu.pivot_range();
        return hi > 0.0 ? lo / hi : 0.0;
    };
    enum { Direct, Inverse, Mobius } map;
    cplx sigma = 0.0;
    CMat c;
    dense::Lu lu_l(lmat), lu_r(rmat);
    double pl = pivot_ratio(lu_l), pr = pivot_ratio(lu_r);
    if (pl >= pr && pl > 1e-10) {
        c = lu_l.solve(rmat);
        map = Direct;
    } else if (pr > 1e-10) {
        c = lu_r.solve(lmat);
        map = Inverse;
    } else {
        double best = -1.0;
        static const double real_sigmas[8] = {0.5, 2.0, -0.5, -2.0, 0.37, 2.7, -0.37, -2.7};
        for (int k = 0; k < 8; ++k) {
            cplx sg = real ? cplx(real_sigmas[k]) : std::polar(k % 2 == 0 ? 0.5 : 2.0, 0.3 + 0.9 * k);
            CMat mm(m2, m2);
            for (size_t e = 0; e < mm.a.size(); ++e) mm.a[e] = rmat.a[e] - sg * lmat.a[e];
            dense::Lu lu(mm);
            double ratio = pivot_ratio(lu);
            if (best < 0.0 || ratio > best) {
                best = ratio;
                sigma = sg;
                c = lu.solve(lmat);
            }
        }
        if (std::isnan(best) || best <= 1e-14) return kEigen;
        map = Mobius;
    }
    CMat q0, t0;
    if (!dense::schur(c, q0, t0)) return kEigen;
    double mu_max = 0.0;
    for (int k = 0; k < m2; ++k) mu_max = std::max(mu_max, std::abs(t0(k, k)));
    auto to_lambda = [&](cplx mu, cplx& lam) -> bool {
        if (map == Direct) {
            lam = mu;
            return true;
        }
        if (std::abs(mu) <= 1e-13 * mu_max) return false;
        lam = map == Inverse ? one / mu : sigma + one / mu;
        return true;
    };
    const size_t RIGHT = 0, LEFT = std::numeric_limits<size_t>::max();
    std::vector<size_t> keys(m2);
    std::vector<cplx> lams(m2);
    std::vector<char> finite(m2);
    for (int k = 0; k < m2; ++k) {
        finite[k] = to_lambda(t0(k, k), lams[k]);
        double a = std::abs(lams[k]);
        keys[k] = !finite[k] ? LEFT : a < 1.0 - kUnitTol ? RIGHT : a <= 1.0 + kUnitTol ? 1 : LEFT;
    }
    // Degenerate propagating groups (single linkage), numbered 1, 2, ... in order.
    std::vector<int> prop;
    for (int k = 0; k < m2; ++k)
        if (keys[k] == 1) prop.push_back(k);
    std::vector<size_t> group(prop.size());
    for (size_t a = 0; a < prop.size(); ++a) group[a] = a;
    for (size_t a = 0; a < prop.size(); ++a)
        for (size_t b = 0; b < a; ++b)
            if (std::abs(lams[prop[a]] - lams[prop[b]]) < kDegenerateTol) {
                size_t lo = std::min(group[a], group[b]), hi = std::max(group[a], group[b]);
                for (auto& g2 : group)
                    if (g2 == hi) g2 = lo;
            }
    std::vector<size_t> ids = group;
    std::sort(ids.begin(), ids.end());
    ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
    for (size_t a = 0; a < prop.size(); ++a)
        keys[prop[a]] = 1 + size_t(std::lower_bound(ids.begin(), ids.end(), group[a]) - ids.begin());

    // Ordering A: right-evanescent, then the propagating groups, each contiguous.
    CMat qa = q0, ta = t0;
    std::vector<size_t> keys_a = keys;
    auto partition_by = [&](auto select) {
        std::vector<char> flags(m2);
        for (int k = 0; k < m2; ++k) flags[k] = select(keys_a[k]);
        dense::partition(ta, qa, flags);
        std::stable_partition(keys_a.begin(), keys_a.end(), select);
    };
    partition_by([&](size_t k) { return k == RIGHT; });
    partition_by([&](size_t k) { return k != LEFT; });
    for (int i = 1; i < m2; ++i)  // schur.rs:199-208: insertion sort by adjacent swaps
        for (int j = i; j > 0 && keys_a[j - 1] > keys_a[j]; --j) {
            dense::swap_adjacent(ta, qa, j - 1);
            std::swap(keys_a[j - 1], keys_a[j]);
        }
    int n_re = 0;
    for (auto k : keys_a) n_re += k == RIGHT;
    std::vector<std::vector<cplx>> right, left;
    for (int k = 0; k < n_re; ++k) right.emplace_back(qa.col(k), qa.col(k) + m2);

    double tnorm = ta.norm_max();
    struct Group {
        int g0, g1;
        bool scalar;
    };
    std::vector<Group> groups;
    int g0 = n_re;
    while (g0 < m2 && keys_a[g0] != LEFT) {
        int g1 = g0 + 1;
        while (g1 < m2 && keys_a[g1] == keys_a[g0]) ++g1;
        double off = 0.0;
        for (int a = g0; a < g1; ++a)
            for (int b = a + 1; b < g1; ++b) off = std::max(off, std::abs(ta(a, b)));
        groups.push_back({g0, g1, off <= 1e-8 * tnorm});
        g0 = g1;
    }
    const int np = g0 - n_re;
    CMat y(m2, np);
    std::vector<cplx> plam;
    for (auto& gr : groups)
        for (int k = gr.g0; k < gr.g1; ++k) {
            cplx mu = ta(k, k), lam;
            if (!to_lambda(mu, lam)) return kSingular;
            plam.push_back(lam);
            cplx* col = y.col(k - n_re);
            col[k] = one;
            int start = gr.scalar ? gr.g0 : k;
            std::vector<cplx> acc(start);
            for (int j = 0; j < start; ++j) acc[j] = -ta(j, k);
            for (int j = start - 1; j >= 0; --j) {
                cplx d = ta(j, j) - mu;
                if (d == zero) return kSingular;
                cplx yj = acc[j] / d;
                col[j] = yj;
                for (int a = 0; a < j; ++a) acc[a] -= yj * ta(a, j);
            }
        }
    CMat x = matmul(qa, y);
    for (int p = 0; p < np; ++p) {
        double nn = 0.0;
        for (int i = 0; i < m2; ++i) nn += std::norm(x(i, p));
        nn = std::sqrt(nn);
        if (!(std::isfinite(nn) && nn > 0.0)) return kSingular;
        for (int i = 0; i < m2; ++i) x(i, p) /= nn;
    }
    CMat bb(r, np), cc(r, np);
    for (int p = 0; p < np; ++p)
        for (int q = 0; q < r; ++q) {
            bb(q, p) = -x(q, p) + s * x(r + q, p);
            cc(q, p) = plam[p] * (s * x(q, p) - x(r + q, p));
        }
    CMat phis = matmul(gv, bb), t2 = matmul(gu, cc);
    for (size_t e = 0; e < phis.a.size(); ++e) phis.a[e] += t2.a[e];
    CMat h1phi = matmul(h1, phis), h1dphi = matmul(h1.adjoint(), phis);
    int nright = 0;
    std::vector<std::vector<cplx>> right_phi, left_phi;
    std::vector<cplx> right_lam, left_lam;
    for (auto& gr : groups) {
        int p0 = gr.g0 - n_re, m = gr.g1 - gr.g0;
        CMat xg = cols_of(x, p0, p0 + m), fg = cols_of(phis, p0, p0 + m);
        CMat gram = matmul(xg.adjoint(), xg);
        CMat a = matmul(fg.adjoint(), cols_of(h1phi, p0, p0 + m));
        CMat b = matmul(fg.adjoint(), cols_of(h1dphi, p0, p0 + m));
        CMat cur(m, m);
        for (int q = 0; q < m; ++q)
            for (int p = 0; p < m; ++p) cur(p, q) = iu * (plam[p0 + q] * a(p, q) - std::conj(plam[p0 + p]) * b(p, q));
        CMat lch;
        if (!dense::cholesky(gram, lch)) return kSingular;
        if (inverse_condition(lch) < kDefectiveTol) return kSingular;  // Jordan block: subband edge
        CMat linv = dense::inverse(lch);
        CMat jp = matmul(matmul(linv, cur), linv.adjoint());
        CMat jh(m, m);
        for (int q = 0; q < m; ++q)
            for (int p = 0; p < m; ++p) jh(p, q) = (jp(p, q) + std::conj(jp(q, p))) * 0.5;
        std::vector<double> vel;
        CMat w;
        if (!dense::hermitian_eigen(jh, vel, w)) return kEigen;
        CMat coeffs = matmul(linv.adjoint(), w);
        CMat phi_k = matmul(fg, coeffs), x_k = matmul(xg, coeffs);
        // The Bloch factor of a group: the mean of its members (equal to kDegenerateTol).
        cplx mean = 0.0;
        for (int k = 0; k < m; ++k) mean += plam[p0 + k];
        mean /= double(m);
        const cplx lambda = mean / std::abs(mean);
        for (int k = 0; k < m; ++k) {
            double nn = 0.0;
            for (int i = 0; i < n; ++i) nn += std::norm(phi_k(i, k));
            double vk = vel[k] / nn;
            if (std::isnan(vk) || std::abs(vk) < kVelocityTol * scale) return kSingular;  // subband edge
            // J(phi_k, phi_k) = vel_k: unit current after dividing by its root.
            std::vector<cplx> flux(static_cast<size_t>(n));
            const double root = std::sqrt(std::abs(vel[k]));
            for (int i = 0; i < n; ++i) flux[size_t(i)] = phi_k(i, k) / root;
            if (vk > 0.0) {
                right.emplace_back(x_k.col(k), x_k.col(k) + m2);
                right_phi.push_back(std::move(flux));
                right_lam.push_back(lambda);
                ++nright;
            } else {
                left.emplace_back(x_k.col(k), x_k.col(k) + m2);
                left_phi.push_back(std::move(flux));
                left_lam.push_back(lambda);
            }
        }
    }
    // Ordering B: left-evanescent first.
    {
        CMat qb = q0, tb = t0;
        std::vector<char> flags(m2);
        for (int k = 0; k < m2; ++k) flags[k] = keys[k] == LEFT;
        int n_le = dense::partition(tb, qb, flags);
        for (int k = 0; k < n_le; ++k) left.emplace_back(qb.col(k), qb.col(k) + m2);
    }
    if (int(right.size()) != r || int(left.size()) != r) return kModeCount;
    auto basis = [&](const std::vector<std::vector<cplx>>& cols) {
        CMat m(m2, int(cols.size()));
        for (size_t q = 0; q < cols.size(); ++q) {
            double nn = 0.0;
            for (auto& zz : cols[q]) nn += std::norm(zz);
            nn = std::sqrt(nn);
            for (int p = 0; p < m2; ++p) m(p, int(q)) = cols[q][p] / nn;
        }
        return dense::orthonormal_columns(m);
    };
    out.right = basis(right);
    out.left = basis(left);
    auto invertible = [](const CMat& m) { return dense::Lu(m).inverse_condition_1(dense::norm_1(m)) >= kSingularTol; };
    if (!invertible(rows_of(out.right, 0, r)) || !invertible(rows_of(out.left, r, m2))) return kSingular;
    out.channels = nright;
    auto columns = [&](const std::vector<std::vector<cplx>>& cols) {
        CMat m(n, int(cols.size()));
        for (size_t c = 0; c < cols.size(); ++c) std::copy(cols[c].begin(), cols[c].end(), m.col(int(c)));
        return m;
    };
    out.right_modes = columns(right_phi);
    out.left_modes = columns(left_phi);
    out.right_lambdas = right_lam;
    out.left_lambdas = left_lam;
    return kOk;
}

Status solve_general(const CMat& h0, const CMat& h1, cplx z, double scale, LeadModes& out) {
    bool real = z.imag() == 0.0 && is_real(h0) && is_real(h1);
    if (real && solve_with(h0, h1, z, scale, true, out) == kOk) return kOk;
    return solve_with(h0, h1, z, scale, false, out);
}

// c Y diag(d) Y^H.
CMat cell_sigma(const CMat& y, double c, const std::vector<cplx>& d) {
    CMat w = y;
    for (int k = 0; k < w.cols; ++k)
        for (int i = 0; i < w.rows; ++i) w(i, k) *= c * d[k];
    return matmul(w, y.adjoint());
}

}  // namespace

CMat LeadModes::sigma_right() const {
    const int n = u.rows, r = u.cols;
    if (r == 0) return CMat(n, n);
    if (cell_y) {  // X_g X_a^-1 = Y diag(lambda / s) Y^H, u = s sqrt(c) I
        std::vector<cplx> d(cell_right.size());
        for (size_t k = 0; k < d.size(); ++k) d[k] = cell_s * cell_right[k];
        return cell_sigma(*cell_y, cell_c, d);
    }
    CMat t = dense::solve_right(rows_of(right, r, 2 * r), rows_of(right, 0, r));
    return sandwich(u, t);
}

CMat LeadModes::sigma_left() const {
    const int n = v.rows, r = v.cols;
    if (r == 0) return CMat(n, n);
    if (cell_y) {  // Y_a Y_g^-1 = Y diag(s / lambda) Y^H, v = sqrt(c) I
        std::vector<cplx> d(cell_left.size());
        for (size_t k = 0; k < d.size(); ++k) d[k] = cell_s / cell_left[k];
        return cell_sigma(*cell_y, cell_c, d);
    }
    CMat t = dense::solve_right(rows_of(left, 0, r), rows_of(left, r, 2 * r));
    return sandwich(v, t);
}

std::vector<cplx> CellModes::sigma_right_diag() const {
    std::vector<cplx> d(right.size());
    for (size_t k = 0; k < d.size(); ++k) d[k] = c * s * right[k];
    return d;
}

std::vector<cplx> CellModes::sigma_left_diag() const {
    std::vector<cplx> d(left.size());
    for (size_t k = 0; k < d.size(); ++k) d[k] = c * s / left[k];
    return d;
}

namespace {
std::atomic<uint64_t> next_solver_id{1};
}

// DefiniteCoupling::new (modes.rs:575-618)
ModeSolver::ModeSolver(const CMat& h0, const CMat& h1, HermitianEigen eigen)
    : id_(next_solver_id++), h0_(h0), h1_(h1) {
    const int n = h0.rows;
    auto hermitian = [&](const CMat& m) {
        double asym = 0.0;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) asym = std::max(asym, std::abs(m(i, j) - std::conj(m(j, i))));
        return asym <= kHermitianTol * m.norm_max();
    };
    if (n == 0 || h1.norm_max() == 0.0 || !hermitian(h0) || !hermitian(h1)) return;
    for (double s : {-1.0, 1.0}) {
        CMat p(n, n);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) p(i, j) = (h1(i, j) + std::conj(h1(j, i))) * (0.5 * s);
        CMat l;
        if (!dense::cholesky(p, l)) continue;
        double lo = std::numeric_limits<double>::infinity(), hi = 0.0;
        for (int i = 0; i < n; ++i) {
            lo = std::min(lo, std::abs(l(i, i)));
            hi = std::max(hi, std::abs(l(i, i)));
        }
        if (lo <= kDefiniteTol * hi) continue;
        Definite d;
        d.s = s;
        d.l = l;
        double c = p(0, 0).real();
        bool scalar = true;
        for (int j = 0; j < n && scalar; ++j)
            for (int i = 0; i < n; ++i)
                if (std::abs(p(i, j) - (i == j ? cplx(c) : cplx(0.0))) > kHermitianTol * c) {
                    scalar = false;
                    break;
                }
        if (scalar) {
            CMat h(n, n);
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) h(i, j) = (h0(i, j) + std::conj(h0(j, i))) * 0.5;
            if ((eigen ? eigen : dense::hermitian_eigen)(h, d.theta, d.y)) {
                d.cell = true;
                d.c = c;
                d.l = CMat(n, n);
                for (int i = 0; i < n; ++i) d.l(i, i) = std::sqrt(c);
            }
        }
        definite_ = std::move(d);
        return;
    }
}

namespace {
// The per-mode part of DefiniteCoupling::modes (modes.rs:660-700): for nu_k (the generalized
// eigenvalues of (H0 - e) against P, here -s nu_k = mu_k), the right- and left-going lambda, and
// which modes propagate. kSingular at a subband edge (a degenerate pair) or a unit-modulus
// evanescent root.
Status mode_lambdas(const std::vector<double>& nu, double s, std::vector<cplx>& right, std::vector<cplx>& leftl,
                    std::vector<int>& propagating) {
    const int n = int(nu.size());
    right.assign(n, 0.0);
    leftl.assign(n, 0.0);
    propagating.clear();
    for (int k = 0; k < n; ++k) {
        double mu = -s * nu[k];
        if (std::abs(mu) < 2.0) {
            cplx root(mu / 2.0, std::sqrt(4.0 - mu * mu) / 2.0);
            if (2.0 * root.imag() < kDegenerateTol) return kSingular;
            right[k] = s < 0 ? root : std::conj(root);
            leftl[k] = s < 0 ? std::conj(root) : root;
            propagating.push_back(k);
        } else {
            double big = (mu + (mu > 0 ? 1.0 : mu < 0 ? -1.0 : 0.0) * std::sqrt(mu * mu - 4.0)) / 2.0;
            if (std::abs(big) <= 1.0 + kUnitTol) return kSingular;
            right[k] = 1.0 / big;
            leftl[k] = big;
        }
    }
    return kOk;
}

// DefiniteCoupling::modes (modes.rs:621-710): one Hermitian eigenproblem of size n.
Status definite_modes(const CMat& h0, double s, const CMat& l, bool cell, double c, const std::vector<double>& theta,
                      const CMat& ycell, double e, double scale, LeadModes& out) {
    const int n = h0.rows;
    std::vector<double> nu;
    CMat computed;
    const CMat* y = &ycell;
    if (cell) {
        for (double t : theta) nu.push_back((t - e) / c);
    } else {
        CMat b(n, n);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) b(i, j) = (h0(i, j) + std::conj(h0(j, i))) * 0.5 - (i == j ? cplx(e) : cplx(0.0));
        // m = L^-1 b L^-H
        CMat linv = dense::inverse(l);
        CMat m = matmul(matmul(linv, b), linv.adjoint());
        CMat mh(n, n);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) mh(i, j) = (m(i, j) + std::conj(m(j, i))) * 0.5;
        if (!dense::hermitian_eigen(mh, nu, computed)) return kEigen;
        y = &computed;
    }
    std::vector<cplx> right, leftl;
    std::vector<int> propagating;
    if (Status st = mode_lambdas(nu, s, right, leftl, propagating); st != kOk) return st;
    // Velocities of the propagating modes, psi = L^-H y (L = sqrt(c) I in the cell case); psi is
    // P-normalised, so J(psi, psi) = -2 s Im lambda: unit current after dividing by its root.
    CMat lhinv;
    if (!cell) lhinv = dense::inverse(l.adjoint());
    CMat flux(n, int(propagating.size()));
    for (size_t q = 0; q < propagating.size(); ++q) {
        const int k = propagating[q];
        double nn = 0.0;
        for (int i = 0; i < n; ++i) {
            cplx v = 0.0;
            if (cell) {
                v = (*y)(i, k) / std::sqrt(c);
            } else {
                for (int j = 0; j < n; ++j) v += lhinv(i, j) * (*y)(j, k);
            }
            flux(i, int(q)) = v;
            nn += std::norm(v);
        }
        double v = -2.0 * s * right[k].imag() / nn;
        if (!(std::abs(v) >= kVelocityTol * scale)) return kSingular;
        const double root = std::sqrt(2.0 * std::abs(right[k].imag()));
        for (int i = 0; i < n; ++i) flux(i, int(q)) /= root;
    }
    auto basis = [&](const std::vector<cplx>& lams) {
        CMat m(2 * n, n);
        for (int k = 0; k < n; ++k) {
            double cc = 1.0 / std::sqrt(1.0 + std::norm(lams[k]));
            for (int p = 0; p < n; ++p) {
                m(p, k) = (*y)(p, k) * (s * cc);
                m(n + p, k) = (*y)(p, k) * lams[k] * cc;
            }
        }
        return m;
    };
    out.u = CMat(n, n);
    for (size_t e2 = 0; e2 < l.a.size(); ++e2) out.u.a[e2] = l.a[e2] * s;
    out.v = l;
    out.right = basis(right);
    out.left = basis(leftl);
    out.channels = int(propagating.size());
    out.energy = e;
    out.right_modes = flux;
    out.left_modes = flux;
    out.right_lambdas.clear();
    out.left_lambdas.clear();
    for (int k : propagating) {
        out.right_lambdas.push_back(right[k]);
        out.left_lambdas.push_back(leftl[k]);
    }
    if (cell) {
        out.cell_y = y;
        out.cell_c = c;
        out.cell_s = s;
        out.cell_right = right;
        out.cell_left = leftl;
    }
    return kOk;
}

}  // namespace

// ModeSolver::modes_at with LeadModes::shifted (modes.rs:130-145, :531-537)
Status ModeSolver::modes_at(double e, LeadModes& out) const {
    double scale = std::max(row_norm(h0_) + 2.0 * row_norm(h1_), std::numeric_limits<double>::min());
    Status last = kEigen;
    for (double shift : kShifts) {
        double es = e + shift * scale;
        Status st = definite_ ? definite_modes(h0_, definite_->s, definite_->l, definite_->cell, definite_->c,
                                               definite_->theta, definite_->y, es, scale, out)
                              : solve_general(h0_, h1_, cplx(es, 0.0), scale, out);
        if (st == kOk) return kOk;
        last = st;
    }
    return last;
}

Status ModeSolver::modes_at(cplx z, LeadModes& out) const {
    if (z.imag() == 0.0) return modes_at(z.real(), out);
    double scale = std::max(row_norm(h0_) + 2.0 * row_norm(h1_), std::numeric_limits<double>::min());
    Status last = kEigen;
    for (double shift : kShifts) {
        Status st = solve_general(h0_, h1_, z + shift * scale, scale, out);
        if (st == kOk) return kOk;
        last = st;
    }
    return last;
}

// The cell case of modes_at (modes.rs:130-145, :621-710) without the bases: per energy shift, the
// lambdas of every cell mode and the velocity test of the propagating ones (the cell eigenvectors
// are orthonormal, so |L^-H y_k|^2 = 1 / c).
Status ModeSolver::cell_modes_at(double e, CellModes& out) const {
    if (!cell()) return kInvalid;
    const Definite& d = *definite_;
    double scale = std::max(row_norm(h0_) + 2.0 * row_norm(h1_), std::numeric_limits<double>::min());
    Status last = kEigen;
    for (double shift : kShifts) {
        const double es = e + shift * scale;
        std::vector<double> nu;
        for (double t : d.theta) nu.push_back((t - es) / d.c);
        CellModes m;
        Status st = mode_lambdas(nu, d.s, m.right, m.left, m.propagating);
        if (st == kOk)
            for (int k : m.propagating) {
                double nn = 0.0;
                for (int i = 0; i < d.y.rows; ++i) nn += std::norm(d.y(i, k) / std::sqrt(d.c));
                const double v = -2.0 * d.s * m.right[k].imag() / nn;
                if (!(std::abs(v) >= kVelocityTol * scale)) {
                    st = kSingular;
                    break;
                }
            }
        if (st == kOk) {
            m.y = &d.y;
            m.lead_id = id_;
            m.c = d.c;
            m.s = d.s;
            m.energy = es;
            out = std::move(m);
            return kOk;
        }
        last = st;
    }
    return last;
}

Status lead_self_energies(const ModeSolver& solver, double e, CMat* sigma_left, CMat* sigma_right, int* channels) {
    LeadModes m;
    Status st = solver.modes_at(e, m);
    if (st != kOk) return st;
    if (sigma_left) *sigma_left = m.sigma_left();
    if (sigma_right) *sigma_right = m.sigma_right();
    if (channels) *channels = m.channels;
    return kOk;
}

}  // namespace caroli
