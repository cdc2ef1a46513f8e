// Host FP64 dense kernels for the lead problems (see dense.hpp).
#include "dense.hpp"

#include <algorithm>
#include <cmath>
#include <limits>

namespace caroli::dense {

namespace {
constexpr double kEps = std::numeric_limits<double>::epsilon();
double abs1(cplx z) { return std::abs(z.real()) + std::abs(z.imag()); }
}  // namespace

Lu::Lu(const CMat& a) : lu(a), piv(a.rows) {
    const int n = a.rows;
    for (int k = 0; k < n; ++k) {
        int p = k;
        double big = -1.0;
        for (int i = k; i < n; ++i) {
            double m = std::abs(lu(i, k));
            if (m > big) {
                big = m;
                p = i;
            }
        }
        piv[k] = p;
        if (p != k)
            for (int j = 0; j < n; ++j) std::swap(lu(k, j), lu(p, j));
        cplx d = lu(k, k);
        if (d == cplx(0.0)) continue;
        cplx inv = 1.0 / d;
        for (int i = k + 1; i < n; ++i) lu(i, k) *= inv;
        for (int j = k + 1; j < n; ++j) {
            cplx u = lu(k, j);
            if (u == cplx(0.0)) continue;
            cplx* cj = lu.col(j);
            const cplx* ck = lu.col(k);
            for (int i = k + 1; i < n; ++i) cj[i] -= ck[i] * u;
        }
    }
}

std::pair<double, double> Lu::pivot_range() const {
    double lo = std::numeric_limits<double>::infinity(), hi = 0.0;
    for (int i = 0; i < lu.rows; ++i) {
        double m = std::abs(lu(i, i));
        if (std::isnan(m)) return {m, m};
        lo = std::min(lo, m);
        hi = std::max(hi, m);
    }
    return {lo, hi};
}

void Lu::solve_in_place(CMat& b) const {
    const int n = lu.rows;
    for (int c = 0; c < b.cols; ++c) {
        cplx* x = b.col(c);
        for (int k = 0; k < n; ++k)
            if (piv[k] != k) std::swap(x[k], x[piv[k]]);
        for (int k = 0; k < n; ++k) {
            cplx v = x[k];
            if (v == cplx(0.0)) continue;
            const cplx* ck = lu.col(k);
            for (int i = k + 1; i < n; ++i) x[i] -= ck[i] * v;
        }
        for (int k = n - 1; k >= 0; --k) {
            x[k] /= lu(k, k);
            cplx v = x[k];
            const cplx* ck = lu.col(k);
            for (int i = 0; i < k; ++i) x[i] -= ck[i] * v;
        }
    }
}

double norm_1(const CMat& a) {
    double m = 0.0;
    for (int j = 0; j < a.cols; ++j) {
        double s = 0.0;
        for (int i = 0; i < a.rows; ++i) s += std::abs(a(i, j));
        m = std::max(m, s);
    }
    return m;
}

double Lu::inverse_condition_1(double n1) const {
    const int n = lu.rows;
    if (n == 0) return 1.0;
    auto [lo, hi] = pivot_range();
    if (!(lo > 0.0 && n1 > 0.0)) return 0.0;
    CMat inv = solve(CMat::identity(n));
    double e = norm_1(inv);
    return (std::isfinite(e) && e > 0.0) ? 1.0 / (n1 * e) : 0.0;
}

CMat inverse(const CMat& a) { return Lu(a).solve(CMat::identity(a.rows)); }

// x a = b  <=>  a^T x^T = b^T
CMat solve_right(const CMat& b, const CMat& a) {
    CMat at(a.cols, a.rows), bt(b.cols, b.rows);
    for (int j = 0; j < a.cols; ++j)
        for (int i = 0; i < a.rows; ++i) at(j, i) = a(i, j);
    for (int j = 0; j < b.cols; ++j)
        for (int i = 0; i < b.rows; ++i) bt(j, i) = b(i, j);
    CMat xt = Lu(at).solve(bt);
    CMat x(xt.cols, xt.rows);
    for (int j = 0; j < xt.cols; ++j)
        for (int i = 0; i < xt.rows; ++i) x(j, i) = xt(i, j);
    return x;
}

bool cholesky(const CMat& a, CMat& l) {
    const int n = a.rows;
    l = CMat(n, n);
    for (int j = 0; j < n; ++j) {
        double d = a(j, j).real();
        for (int k = 0; k < j; ++k) d -= std::norm(l(j, k));
        if (!(d > 0.0)) return false;
        double r = std::sqrt(d);
        l(j, j) = r;
        for (int i = j + 1; i < n; ++i) {
            cplx s = a(i, j);
            for (int k = 0; k < j; ++k) s -= l(i, k) * std::conj(l(j, k));
            l(i, j) = s / r;
        }
    }
    return true;
}

namespace {

// Rotation G = [[c, s], [-conj(s), c]] (c real) with G [f; g] = [r; 0].
void givens(cplx f, cplx g, double& c, cplx& s) {
    double af = std::abs(f), ag = std::abs(g);
    if (ag == 0.0) {
        c = 1.0;
        s = 0.0;
        return;
    }
    if (af == 0.0) {
        c = 0.0;
        s = std::conj(g) / ag;
        return;
    }
    double r = std::hypot(af, ag);
    c = af / r;
    s = (f / af) * std::conj(g) / r;
}

// Rows k, k+1 of m over columns [c0, c1): m <- G m.
void rot_rows(CMat& m, int k, int c0, int c1, double c, cplx s) {
    for (int j = c0; j < c1; ++j) {
        cplx x = m(k, j), y = m(k + 1, j);
        m(k, j) = c * x + s * y;
        m(k + 1, j) = -std::conj(s) * x + c * y;
    }
}

// Columns k, k+1 of m over rows [r0, r1): m <- m G^H.
void rot_cols(CMat& m, int k, int r0, int r1, double c, cplx s) {
    for (int i = r0; i < r1; ++i) {
        cplx x = m(i, k), y = m(i, k + 1);
        m(i, k) = c * x + std::conj(s) * y;
        m(i, k + 1) = -s * x + c * y;
    }
}

}  // namespace

bool schur(const CMat& a, CMat& q, CMat& t) {
    const int n = a.rows;
    t = a;
    q = CMat::identity(n);
    for (auto& z : t.a)
        if (!std::isfinite(z.real()) || !std::isfinite(z.imag())) return false;
    // Householder reduction to Hessenberg form.
    std::vector<cplx> v(n);
    for (int k = 0; k + 2 < n; ++k) {
        double xn = 0.0;
        for (int i = k + 1; i < n; ++i) xn += std::norm(t(i, k));
        xn = std::sqrt(xn);
        if (xn == 0.0) continue;
        cplx x0 = t(k + 1, k);
        cplx ph = std::abs(x0) > 0 ? x0 / std::abs(x0) : cplx(1.0);
        cplx alpha = -ph * xn;
        for (int i = 0; i < n; ++i) v[i] = 0.0;
        v[k + 1] = x0 - alpha;
        for (int i = k + 2; i < n; ++i) v[i] = t(i, k);
        double vn = 0.0;
        for (int i = k + 1; i < n; ++i) vn += std::norm(v[i]);
        if (vn == 0.0) continue;
        double beta = 2.0 / vn;
        // t <- (I - beta v v^H) t
        for (int j = k; j < n; ++j) {
            cplx s = 0.0;
            for (int i = k + 1; i < n; ++i) s += std::conj(v[i]) * t(i, j);
            s *= beta;
            for (int i = k + 1; i < n; ++i) t(i, j) -= v[i] * s;
        }
        // t <- t (I - beta v v^H), q <- q (I - beta v v^H)
        for (CMat* m : {&t, &q}) {
            for (int i = 0; i < n; ++i) {
                cplx s = 0.0;
                for (int j = k + 1; j < n; ++j) s += (*m)(i, j) * v[j];
                s *= beta;
                for (int j = k + 1; j < n; ++j) (*m)(i, j) -= s * std::conj(v[j]);
            }
        }
        for (int i = k + 2; i < n; ++i) t(i, k) = 0.0;
    }
    // Implicit single-shift QR iteration on the Hessenberg form.
    int hi = n - 1;
    int iter = 0, total = 0;
    double tnorm = t.norm_max();
    while (hi > 0) {
        int l = hi;
        while (l > 0) {
            double s = abs1(t(l - 1, l - 1)) + abs1(t(l, l));
            if (s == 0.0) s = tnorm;
            if (abs1(t(l, l - 1)) <= kEps * s) {
                t(l, l - 1) = 0.0;
                break;
            }
            --l;
        }
        if (l == hi) {
            --hi;
            iter = 0;
            continue;
        }
        if (++total > 60 * n) return false;
        ++iter;
        cplx mu;
        if (iter % 11 == 10) {
            mu = t(hi, hi) + cplx(abs1(t(hi, hi - 1)), 0.0) * 0.75;  // exceptional shift
        } else {
            // Eigenvalue of the trailing 2x2 block closest to t(hi, hi) (Wilkinson).
            cplx a = t(hi - 1, hi - 1), b = t(hi - 1, hi), c = t(hi, hi - 1), d = t(hi, hi);
            cplx half = (a - d) * 0.5;
            cplx disc = std::sqrt(half * half + b * c);
            cplx m1 = d + half + disc, m2 = d + half - disc;
            mu = std::abs(m1 - d) < std::abs(m2 - d) ? m1 : m2;
        }
        cplx x = t(l, l) - mu, y = t(l + 1, l);
        for (int k = l; k < hi; ++k) {
            if (k > l) {
                x = t(k, k - 1);
                y = t(k + 1, k - 1);
            }
            double c;
            cplx s;
            givens(x, y, c, s);
            rot_rows(t, k, k > l ? k - 1 : l, n, c, s);
            if (k > l) t(k + 1, k - 1) = 0.0;
            rot_cols(t, k, 0, std::min(k + 3, hi + 1), c, s);
            rot_cols(q, k, 0, n, c, s);
        }
    }
    for (int j = 0; j < n; ++j)
        for (int i = j + 1; i < n; ++i) t(i, j) = 0.0;
    return true;
}

void swap_adjacent(CMat& t, CMat& q, int k) {
    const int n = t.rows;
    cplx t11 = t(k, k), t22 = t(k + 1, k + 1);
    double c;
    cplx s;
    givens(t(k, k + 1), t22 - t11, c, s);
    if (c == 1.0 && s == cplx(0.0)) return;
    rot_rows(t, k, k, n, c, s);
    rot_cols(t, k, 0, k + 2, c, s);
    rot_cols(q, k, 0, n, c, s);
    t(k, k) = t22;
    t(k + 1, k + 1) = t11;
    t(k + 1, k) = 0.0;
}

int partition(CMat& t, CMat& q, std::vector<char> flags) {
    const int n = t.rows;
    int placed = 0;
    for (int j = 0; j < n; ++j) {
        if (!flags[j]) continue;
        for (int k = j - 1; k >= placed; --k) {
            swap_adjacent(t, q, k);
            std::swap(flags[k], flags[k + 1]);
        }
        ++placed;
    }
    return placed;
}

bool svd(const CMat& a, CMat& u, std::vector<double>& s, CMat& v) {
    const int m = a.rows, n = a.cols;
    u = a;
    v = CMat::identity(n);
    for (int sweep = 0; sweep < 80; ++sweep) {
        bool rotated = false;
        for (int p = 0; p < n - 1; ++p)
            for (int qq = p + 1; qq < n; ++qq) {
                double alpha = 0, beta = 0;
                cplx gamma = 0;
                for (int i = 0; i < m; ++i) {
                    alpha += std::norm(u(i, p));
                    beta += std::norm(u(i, qq));
                    gamma += std::conj(u(i, p)) * u(i, qq);
                }
                double g = std::abs(gamma);
                if (g == 0.0 || g <= 1e-15 * std::sqrt(alpha * beta)) continue;
                rotated = true;
                cplx ph = gamma / g;
                double zeta = (beta - alpha) / (2.0 * g);
                double tt = (zeta >= 0 ? 1.0 : -1.0) / (std::abs(zeta) + std::sqrt(1.0 + zeta * zeta));
                double c = 1.0 / std::sqrt(1.0 + tt * tt), sn = c * tt;
                for (CMat* mm : {&u, &v}) {
                    for (int i = 0; i < mm->rows; ++i) {
                        cplx x = (*mm)(i, p), y = (*mm)(i, qq);
                        (*mm)(i, p) = c * x - sn * std::conj(ph) * y;
                        (*mm)(i, qq) = sn * ph * x + c * y;
                    }
                }
            }
        if (!rotated) break;
        if (sweep == 79) return false;
    }
    s.assign(n, 0.0);
    for (int j = 0; j < n; ++j) {
        double nn = 0;
        for (int i = 0; i < m; ++i) nn += std::norm(u(i, j));
        s[j] = std::sqrt(nn);
    }
    std::vector<int> order(n);
    for (int j = 0; j < n; ++j) order[j] = j;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return s[x] > s[y]; });
    CMat u2(m, n), v2(n, n);
    std::vector<double> s2(n);
    for (int k = 0; k < n; ++k) {
        int j = order[k];
        s2[k] = s[j];
        for (int i = 0; i < m; ++i) u2(i, k) = s[j] > 0 ? u(i, j) / s[j] : cplx(0.0);
        for (int i = 0; i < n; ++i) v2(i, k) = v(i, j);
    }
    u = u2;
    v = v2;
    s = s2;
    return true;
}

bool hermitian_eigen(const CMat& a0, std::vector<double>& w, CMat& v) {
    // Householder reduction to Hermitian tridiagonal form A = Q T Q^H (reflectors kept), the
    // subdiagonal made real by a diagonal unitary D (T = D T' D^H), the implicit QL iteration with
    // Wilkinson shifts on the real tridiagonal T' = Z diag(w) Z^T, and V = Q D Z.
    const int n = a0.rows;
    CMat a(n, n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) a(i, j) = (a0(i, j) + std::conj(a0(j, i))) * 0.5;
    w.assign(n, 0.0);
    if (n == 0) {
        v = CMat(0, 0);
        return true;
    }
    std::vector<std::vector<cplx>> refl;  // reflector k acts on rows/columns k+1..n-1
    std::vector<cplx> p(n), wv(n);
    for (int k = 0; k + 2 < n; ++k) {
        const int m = n - k - 1;
        std::vector<cplx> x(a.col(k) + k + 1, a.col(k) + n);
        double nx = 0.0;
        for (auto& z : x) nx += std::norm(z);
        nx = std::sqrt(nx);
        if (nx == 0.0) {
            refl.emplace_back();
            continue;
        }
        const cplx ph = std::abs(x[0]) > 0.0 ? x[0] / std::abs(x[0]) : cplx(1.0);
        const cplx alpha = -ph * nx;
        x[0] -= alpha;
        double nv = 0.0;
        for (auto& z : x) nv += std::norm(z);
        nv = std::sqrt(nv);
        for (auto& z : x) z /= nv;
        // A22 <- H A22 H with H = I - 2 x x^H: p = A22 x, g = x^H p, w = 2 p - 2 g x,
        // A22 <- A22 - x w^H - w x^H.
        for (int i = 0; i < m; ++i) p[i] = 0.0;
        for (int j = 0; j < m; ++j) {
            const cplx xj = x[j];
            const cplx* col = a.col(k + 1 + j) + k + 1;
            for (int i = 0; i < m; ++i) p[i] += col[i] * xj;
        }
        cplx g = 0.0;
        for (int i = 0; i < m; ++i) g += std::conj(x[i]) * p[i];
        for (int i = 0; i < m; ++i) wv[i] = 2.0 * p[i] - 2.0 * g.real() * x[i];
        for (int j = 0; j < m; ++j) {
            const cplx cxj = std::conj(x[j]), cwj = std::conj(wv[j]);
            cplx* col = a.col(k + 1 + j) + k + 1;
            for (int i = 0; i < m; ++i) col[i] -= x[i] * cwj + wv[i] * cxj;
        }
        a(k + 1, k) = alpha;
        a(k, k + 1) = std::conj(alpha);
        for (int i = k + 2; i < n; ++i) a(i, k) = a(k, i) = 0.0;
        refl.push_back(std::move(x));
    }
    // Real tridiagonal: d, e (e[i] couples i and i + 1) and the phases of D.
    std::vector<double> d(n), e(n, 0.0);
    std::vector<cplx> phase(n, 1.0);
    for (int i = 0; i < n; ++i) d[i] = a(i, i).real();
    for (int i = 0; i + 1 < n; ++i) {
        const cplx s = a(i + 1, i);
        const double m = std::abs(s);
        e[i] = m;
        phase[i + 1] = m > 0.0 ? phase[i] * s / m : phase[i];
    }
    std::vector<double> z(size_t(n) * n, 0.0);  // column-major Z
    for (int i = 0; i < n; ++i) z[i + size_t(i) * n] = 1.0;
    const double eps = std::numeric_limits<double>::epsilon();
    for (int l = 0; l < n; ++l) {
        int iter = 0, m;
        do {
            for (m = l; m + 1 < n; ++m) {
                const double dd = std::abs(d[m]) + std::abs(d[m + 1]);
                if (std::abs(e[m]) <= eps * dd) break;
            }
            if (m == l) break;
            if (++iter > 60) return false;
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? r : -r));
            double s = 1.0, c = 1.0, pp = 0.0;
            int i;
            bool deflated = false;
            for (i = m - 1; i >= l; --i) {
                const double f = s * e[i], b = c * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] -= pp;
                    e[m] = 0.0;
                    deflated = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - pp;
                r = (d[i] - g) * s + 2.0 * c * b;
                pp = s * r;
                d[i + 1] = g + pp;
                g = c * r - b;
                double* zi = z.data() + size_t(i) * n;
                double* zj = z.data() + size_t(i + 1) * n;
                for (int k = 0; k < n; ++k) {
                    const double t = zj[k];
                    zj[k] = s * zi[k] + c * t;
                    zi[k] = c * zi[k] - s * t;
                }
            }
            if (deflated) continue;
            d[l] -= pp;
            e[l] = g;
            e[m] = 0.0;
        } while (true);
    }
    // V = Q D Z, eigenvalues ascending.
    std::vector<int> order(n);
    for (int j = 0; j < n; ++j) order[j] = j;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return d[x] < d[y]; });
    v = CMat(n, n);
    for (int k = 0; k < n; ++k) {
        w[k] = d[order[k]];
        const double* zc = z.data() + size_t(order[k]) * n;
        for (int i = 0; i < n; ++i) v(i, k) = phase[i] * zc[i];
    }
    for (int k = int(refl.size()) - 1; k >= 0; --k) {
        const auto& x = refl[k];
        if (x.empty()) continue;
        const int m = int(x.size());
        for (int j = 0; j < n; ++j) {
            cplx* col = v.col(j) + k + 1;
            cplx t = 0.0;
            for (int i = 0; i < m; ++i) t += std::conj(x[i]) * col[i];
            t *= 2.0;
            for (int i = 0; i < m; ++i) col[i] -= x[i] * t;
        }
    }
    return true;
}

CMat orthonormal_columns(const CMat& a) {
    CMat q = a;
    for (int j = 0; j < q.cols; ++j) {
        for (int pass = 0; pass < 2; ++pass)
            for (int k = 0; k < j; ++k) {
                cplx d = 0;
                for (int i = 0; i < q.rows; ++i) d += std::conj(q(i, k)) * q(i, j);
                for (int i = 0; i < q.rows; ++i) q(i, j) -= d * q(i, k);
            }
        double nn = 0;
        for (int i = 0; i < q.rows; ++i) nn += std::norm(q(i, j));
        nn = std::sqrt(nn);
        for (int i = 0; i < q.rows; ++i) q(i, j) /= nn;
    }
    return q;
}

}  // namespace caroli::dense
