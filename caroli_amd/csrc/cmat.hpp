// Dense complex matrices of the host side (column-major, complex f64).
#pragma once
#include <complex>
#include <cstddef>
#include <cstdint>
#include <algorithm>
#include <thread>
#include <vector>

namespace caroli {

using cplx = std::complex<double>;

/// Column-major complex matrix: entry (i, j) at i + j * rows.
struct CMat {
    int rows = 0, cols = 0;
    std::vector<cplx> a;
    CMat() = default;
    CMat(int r, int c) : rows(r), cols(c), a(size_t(r) * size_t(c)) {}
    cplx& operator()(int i, int j) { return a[size_t(i) + size_t(j) * size_t(rows)]; }
    const cplx& operator()(int i, int j) const { return a[size_t(i) + size_t(j) * size_t(rows)]; }
    cplx* col(int j) { return a.data() + size_t(j) * size_t(rows); }
    const cplx* col(int j) const { return a.data() + size_t(j) * size_t(rows); }
    static CMat identity(int n) {
        CMat m(n, n);
        for (int i = 0; i < n; ++i) m(i, i) = 1.0;
        return m;
    }
    /// Largest modulus of the entries.
    double norm_max() const {
        double m = 0.0;
        for (auto& z : a) m = std::max(m, std::abs(z));
        return m;
    }
    bool operator==(const CMat& o) const { return rows == o.rows && cols == o.cols && a == o.a; }
    CMat adjoint() const {
        CMat m(cols, rows);
        for (int j = 0; j < cols; ++j)
            for (int i = 0; i < rows; ++i) m(j, i) = std::conj((*this)(i, j));
        return m;
    }
};

/// c = a * b; large products split their columns over up to 16 threads.
inline CMat matmul(const CMat& a, const CMat& b) {
    CMat c(a.rows, b.cols);
    auto columns = [&](int j0, int j1) {
        for (int j = j0; j < j1; ++j)
            for (int k = 0; k < a.cols; ++k) {
                cplx bkj = b(k, j);
                if (bkj == cplx(0.0)) continue;
                const cplx* ak = a.col(k);
                cplx* cj = c.col(j);
                for (int i = 0; i < a.rows; ++i) cj[i] += ak[i] * bkj;
            }
    };
    const double work = double(a.rows) * double(a.cols) * double(b.cols);
    const int threads = work < 6.4e7 ? 1 : int(std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency())));
    if (threads <= 1 || b.cols < 2 * threads) {
        columns(0, b.cols);
        return c;
    }
    std::vector<std::thread> pool;
    const int per = (b.cols + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        const int j0 = t * per, j1 = std::min(b.cols, j0 + per);
        if (j0 < j1) pool.emplace_back(columns, j0, j1);
    }
    for (auto& th : pool) th.join();
    return c;
}

}  // namespace caroli
