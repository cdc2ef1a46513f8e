// Test driver of caroli_amd/csrc/checked.hpp (the checked reduced-precision batch) with the fake
// runner of crates/caroli-green/src/precision.rs:140-148: in single precision the point at index
// 7 is off by 1e-4 and the point at index 3 fails; every other value is exact (its index).
// Prints one line per case: checked failed fell_back precision t[0..9] dt[0..9].
#include <cstdio>
#include <vector>

#include "../../caroli_amd/csrc/checked.hpp"

int main() {
    const int64_t n = 10;
    auto run = [](const std::vector<int64_t>& idx, int32_t precision, double* t, int32_t* st) {
        for (size_t q = 0; q < idx.size(); ++q) {
            const int64_t i = idx[q];
            t[q] = double(i);
            st[q] = 0;
            if (precision != 0 && i == 7) t[q] += 1e-4;
            if (precision != 0 && i == 3) {
                t[q] = -1.0;
                st[q] = 1;
            }
        }
        return int32_t(0);
    };
    struct Case {
        int32_t precision;
        double tolerance;
        int64_t every;
    } cases[] = {{1, 1e-6, 5}, {1, 1e-6, 7}, {0, 0.0, 0}, {1, 1e-3, 1}};
    for (const Case& c : cases) {
        std::vector<double> t(n), dt(n);
        std::vector<int32_t> st(n);
        caroli::CheckedReport rep;
        int32_t rc = caroli::checked_batch(n, c.precision, c.tolerance, c.every, run, t.data(), st.data(), dt.data(), rep);
        std::printf("%d %lld %lld %d %d", rc, (long long)rep.checked, (long long)rep.failed, rep.fell_back ? 1 : 0,
                    rep.precision);
        for (double v : t) std::printf(" %.17g", v);
        for (double v : dt) std::printf(" %.17g", v);
        std::printf(" %.17g %.17g\n", rep.median, rep.max);
    }
    return 0;
}
