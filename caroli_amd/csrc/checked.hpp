// Checked reduced-precision batches: crates/caroli-green/src/precision.rs:91-134 (`checked`),
// over a batched runner instead of a per-point closure.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace caroli {

/// What a checked batch did (PrecisionReport, precision.rs:65-78).
struct CheckedReport {
    int32_t precision = 0;  // precision of the values returned: 0 (double) after a fallback
    int64_t checked = 0;    // points recomputed in double precision and compared
    double median = 0.0, max = 0.0;
    int64_t failed = 0;     // points that failed in the reduced precision, redone in double
    bool fell_back = false;
    std::string fallback;   // why the whole batch was recomputed in double precision
};

/// Runs every point in `precision` through `run(points, precision, t, status)` (a batched
/// runner: t and status have one entry per listed point; it returns a status code for the call).
/// In a reduced precision, the points that failed and one point in `every` (at least one) are
/// recomputed in double precision (precision 0); their differences |T - T_double| go to `dt`
/// (NaN for the points not compared) and the checked points return their double values. When the
/// largest difference exceeds `tolerance` (or is NaN: a point that failed in double only), the
/// whole batch is recomputed in double precision. Returns the first failing call's code, else 0.
template <class Run>
int32_t checked_batch(int64_t count, int32_t precision, double tolerance, int64_t every, Run&& run, double* t,
                      int32_t* status, double* dt, CheckedReport& report) {
    report = CheckedReport{};
    report.precision = precision;
    std::vector<int64_t> all(size_t(std::max<int64_t>(count, 0)));
    for (int64_t p = 0; p < count; ++p) all[size_t(p)] = p;
    for (int64_t p = 0; p < count; ++p) dt[p] = NAN;
    int32_t rc = run(all, precision, t, status);
    if (rc != 0) return rc;
    if (precision == 0 || count == 0) return 0;
    every = std::max<int64_t>(every, 1);
    std::vector<int64_t> redo;
    for (int64_t p = 0; p < count; ++p)
        if (status[p] != 0 || p % every == 0) redo.push_back(p);
    std::vector<double> td(redo.size());
    std::vector<int32_t> sd(redo.size());
    if (!redo.empty()) {
        rc = run(redo, 0, td.data(), sd.data());
        if (rc != 0) return rc;
    }
    std::vector<double> diffs;
    bool nan = false;
    for (size_t q = 0; q < redo.size(); ++q) {
        const int64_t p = redo[q];
        if (status[p] != 0) {
            ++report.failed;
        } else if (sd[q] != 0) {
            diffs.push_back(NAN);
            nan = true;
        } else {
            dt[p] = std::fabs(t[p] - td[q]);
            diffs.push_back(dt[p]);
        }
        t[p] = td[q];
        status[p] = sd[q];
    }
    report.checked = int64_t(diffs.size());
    if (!diffs.empty()) {
        // Sorted with NaN last (f64::total_cmp for these non-negative values), median at len / 2.
        std::sort(diffs.begin(), diffs.end(), [](double a, double b) {
            return std::isnan(b) ? !std::isnan(a) : (!std::isnan(a) && a < b);
        });
        report.median = diffs[diffs.size() / 2];
        report.max = nan ? NAN : diffs.back();
    }
    if (std::isnan(report.max) || report.max > tolerance) {
        char buf[160];
        std::snprintf(buf, sizeof buf,
                      "largest difference %.1e above the tolerance %.1e: all %lld points recomputed in double precision",
                      report.max, tolerance, static_cast<long long>(count));
        report.fallback = buf;
        report.fell_back = true;
        report.precision = 0;
        rc = run(all, 0, t, status);
        if (rc != 0) return rc;
    }
    return 0;
}

}  // namespace caroli
