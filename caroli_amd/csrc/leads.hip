// Lead self-energies on the GPU, FP64: the definite-coupling "cell" case of the mode solver
// (crates/caroli-green/src/modes.rs:621-710 with P = c 1, sigma_left / sigma_right :494-511).
// With the cell Hamiltonian's eigenvectors Y (one eigendecomposition per lead, on the GPU by
// rocSOLVER's Jacobi zheevj), every self-energy of the lead is Sigma = Y diag(d) Y^H,
// d the per-mode factors of that energy: a dense complex GEMM of the cell size (L^2 = 2,304 for a
// cubic lead of cross-section 48), here on FP64 MFMA. The host keeps the O(n) per-mode work.
#include <hip/hip_runtime.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "dense.hpp"
#include "leads.hpp"

namespace caroli::gpu {

namespace {

#define LEAD_CHECK(x)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int kT = 64;   // output tile edge
constexpr int kK = 16;   // K chunk staged in LDS
constexpr int kNT = 256; // four waves, 32 x 32 of the tile each
constexpr int kJacobiSweeps = 100;  // limit of the cell eigensolver's Jacobi sweeps

// out[z] = Y diag(d[z]) Y^H for z = blockIdx.z: n x n complex f64, column-major, leading dimension
// n. Tile (i0, j0) = 64 * (blockIdx.x, blockIdx.y). Per K chunk the workgroup stages Z = Y diag(d)
// rows [i0, i0 + 64) and Y rows [j0, j0 + 64) (16 columns each) in LDS; wave w computes rows
// 32 (w % 2) and columns 32 (w / 2) of the tile as 2 x 2 blocks of v_mfma_f64_16x16x4f64.
// Operand layout of the 16x16x4 MFMA (D[m][q] = sum_k A[m][k] B[k][q]): lane l holds A[l % 16][l / 16]
// and B[l / 16][l % 16]; accumulator element e of lane l is D[l / 16 + 4 e][l % 16]. With
// A[m][k] = conj(Y[j + m, k]) and B[k][q] = Z[i + q, k], D[m][q] = Sigma[i + q, j + m].
__global__ void __launch_bounds__(kNT) cell_sigma_kernel(const double2* __restrict__ y, const double2* __restrict__ d,
                                                         int32_t n, double2* __restrict__ out) {
    __shared__ double2 zs[kK][kT + 1];
    __shared__ double2 ys[kK][kT + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i0 = blockIdx.x * kT, j0 = blockIdx.y * kT;
    const double2* dz = d + int64_t(blockIdx.z) * n;
    double2* o = out + int64_t(blockIdx.z) * n * n;
    const int wi = 32 * (wave & 1), wj = 32 * (wave >> 1);
    f64x4 cr[2][2], ci[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            cr[a][b] = f64x4{0.0, 0.0, 0.0, 0.0};
            ci[a][b] = f64x4{0.0, 0.0, 0.0, 0.0};
        }
    const int lr = tid & 63, lk = tid >> 6;  // staging: row lr of the tile, columns lk + 4 s
    for (int k0 = 0; k0 < n; k0 += kK) {
#pragma unroll
        for (int s = 0; s < kK / 4; ++s) {
            const int kk = lk + 4 * s, k = k0 + kk;
            const int ri = min(i0 + lr, n - 1), rj = min(j0 + lr, n - 1), kc = min(k, n - 1);
            double2 zv = y[ri + int64_t(kc) * n], yv = y[rj + int64_t(kc) * n];
            const double2 dk = dz[kc];
            double2 zz{zv.x * dk.x - zv.y * dk.y, zv.x * dk.y + zv.y * dk.x};
            if (k >= n || i0 + lr >= n) zz = double2{0.0, 0.0};
            if (k >= n || j0 + lr >= n) yv = double2{0.0, 0.0};
            zs[kk][lr] = zz;
            ys[kk][lr] = yv;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < kK; ks += 4) {
            const int kk = ks + lane / 16, r = lane % 16;
            double2 av[2], bv[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) av[a] = ys[kk][wj + 16 * a + r];
#pragma unroll
            for (int b = 0; b < 2; ++b) bv[b] = zs[kk][wi + 16 * b + r];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    // conj(Y) * Z: re = yr zr + yi zi, im = yr zi - yi zr
                    cr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a].x, bv[b].x, cr[a][b], 0, 0, 0);
                    cr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a].y, bv[b].y, cr[a][b], 0, 0, 0);
                    ci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a].x, bv[b].y, ci[a][b], 0, 0, 0);
                    ci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(-av[a].y, bv[b].x, ci[a][b], 0, 0, 0);
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int i = i0 + wi + 16 * b + lane % 16, j = j0 + wj + 16 * a + lane / 16 + 4 * e;
                if (i < n && j < n) o[i + int64_t(j) * n] = double2{cr[a][b][e], ci[a][b][e]};
            }
}

/// Device copies of the eigenvectors of the last few leads (keyed by the mode solver's id).
struct LeadCache {
    struct Entry {
        uint64_t id = 0;
        int n = 0;
        double2* y = nullptr;
        uint64_t used = 0;
    };
    std::mutex mu;
    hipStream_t stream = nullptr;
    std::vector<Entry> entries;
    double2* d = nullptr;
    double2* out = nullptr;
    size_t d_bytes = 0, out_bytes = 0;
    uint64_t clock = 0;
    bool ok = false;
    rocblas_handle blas = nullptr;  // created at the first eigendecomposition
    LeadCache() {
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return;
        if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) return;
        ok = true;
    }
    static LeadCache& get() {
        static LeadCache c;
        return c;
    }
    double2* ensure(double2*& p, size_t& have, size_t bytes) {
        if (bytes <= have) return p;
        if (p) LEAD_CHECK(hipFree(p));
        p = nullptr;
        have = 0;
        LEAD_CHECK(hipMalloc(reinterpret_cast<void**>(&p), bytes));
        have = bytes;
        return p;
    }
    const double2* eigenvectors(uint64_t id, const std::complex<double>* y, int n) {
        for (auto& e : entries)
            if (e.id == id && e.n == n) {
                e.used = ++clock;
                return e.y;
            }
        constexpr size_t kKeep = 4;
        if (entries.size() >= kKeep) {
            auto old = std::min_element(entries.begin(), entries.end(),
                                        [](const Entry& a, const Entry& b) { return a.used < b.used; });
            (void)hipFree(old->y);
            entries.erase(old);
        }
        Entry e;
        e.id = id;
        e.n = n;
        e.used = ++clock;
        const size_t bytes = size_t(n) * size_t(n) * sizeof(double2);
        LEAD_CHECK(hipMalloc(reinterpret_cast<void**>(&e.y), bytes));
        entries.push_back(e);
        LEAD_CHECK(hipMemcpyAsync(e.y, y, bytes, hipMemcpyHostToDevice, stream));
        return e.y;
    }
};

}  // namespace

bool lead_device_available() { return LeadCache::get().ok; }

void cell_sigmas(uint64_t lead_id, const std::complex<double>* y, int n, const std::complex<double>* diags, int count,
                 std::complex<double>* out) {
    LeadCache& c = LeadCache::get();
    if (!c.ok) throw std::runtime_error("no HIP device for the lead self-energies");
    if (n <= 0 || count <= 0) return;
    std::lock_guard<std::mutex> g(c.mu);
    const double2* dy = c.eigenvectors(lead_id, y, n);
    const size_t nn = size_t(n) * size_t(n);
    double2* dd = c.ensure(c.d, c.d_bytes, size_t(count) * n * sizeof(double2));
    double2* dout = c.ensure(c.out, c.out_bytes, size_t(count) * nn * sizeof(double2));
    LEAD_CHECK(hipMemcpyAsync(dd, diags, size_t(count) * n * sizeof(double2), hipMemcpyHostToDevice, c.stream));
    const dim3 grid(unsigned((n + kT - 1) / kT), unsigned((n + kT - 1) / kT), unsigned(count));
    hipLaunchKernelGGL(cell_sigma_kernel, grid, dim3(kNT), 0, c.stream, dy, dd, n, dout);
    LEAD_CHECK(hipGetLastError());
    LEAD_CHECK(hipMemcpyAsync(out, dout, size_t(count) * nn * sizeof(double2), hipMemcpyDeviceToHost, c.stream));
    LEAD_CHECK(hipStreamSynchronize(c.stream));
}

bool hermitian_eigen(const CMat& a, std::vector<double>& w, CMat& v) {
    LeadCache& c = LeadCache::get();
    if (!c.ok) throw std::runtime_error("no HIP device for the lead eigendecomposition");
    const int n = a.rows;
    w.assign(size_t(n), 0.0);
    v = CMat(n, n);
    if (n == 0) return true;
    // The Hermitian part, as the host solver takes it.
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) v(i, j) = (a(i, j) + std::conj(a(j, i))) * 0.5;
    bool good = false;
    {
        std::lock_guard<std::mutex> g(c.mu);
        if (!c.blas) {
            if (rocblas_create_handle(&c.blas) != rocblas_status_success) throw std::runtime_error("rocblas_create_handle");
            if (rocblas_set_stream(c.blas, c.stream) != rocblas_status_success) throw std::runtime_error("rocblas_set_stream");
        }
        const size_t bytes = size_t(n) * size_t(n) * sizeof(double2);
        double2* da = nullptr;
        double *dw = nullptr, *dres = nullptr;
        rocblas_int *dinfo = nullptr, *dsweeps = nullptr;
        LEAD_CHECK(hipMalloc(reinterpret_cast<void**>(&da), bytes));
        LEAD_CHECK(hipMalloc(reinterpret_cast<void**>(&dw), size_t(n) * sizeof(double)));
        LEAD_CHECK(hipMalloc(reinterpret_cast<void**>(&dres), sizeof(double)));
        LEAD_CHECK(hipMalloc(reinterpret_cast<void**>(&dinfo), sizeof(rocblas_int)));
        LEAD_CHECK(hipMalloc(reinterpret_cast<void**>(&dsweeps), sizeof(rocblas_int)));
        rocblas_int info = -1;
        LEAD_CHECK(hipMemcpyAsync(da, v.a.data(), bytes, hipMemcpyHostToDevice, c.stream));
        // Jacobi sweeps to machine precision (abstol 0), eigenvalues ascending. (The
        // divide-and-conquer zheevd gave non-finite eigenvectors for cells with degenerate
        // eigenvalues, as every clean lattice cross-section has.)
        const rocblas_status st =
            rocsolver_zheevj(c.blas, rocblas_esort_ascending, rocblas_evect_original, rocblas_fill_lower, n,
                             reinterpret_cast<rocblas_double_complex*>(da), n, 0.0, dres, kJacobiSweeps, dsweeps, dw,
                             dinfo);
        if (st == rocblas_status_success) {
            LEAD_CHECK(hipMemcpyAsync(v.a.data(), da, bytes, hipMemcpyDeviceToHost, c.stream));
            LEAD_CHECK(hipMemcpyAsync(w.data(), dw, size_t(n) * sizeof(double), hipMemcpyDeviceToHost, c.stream));
            LEAD_CHECK(hipMemcpyAsync(&info, dinfo, sizeof(rocblas_int), hipMemcpyDeviceToHost, c.stream));
        }
        LEAD_CHECK(hipStreamSynchronize(c.stream));
        (void)hipFree(da);
        (void)hipFree(dw);
        (void)hipFree(dres);
        (void)hipFree(dinfo);
        (void)hipFree(dsweeps);
        good = st == rocblas_status_success && info == 0;
        for (size_t k = 0; good && k < w.size(); ++k) good = std::isfinite(w[k]);
        for (size_t k = 0; good && k < v.a.size(); ++k) good = std::isfinite(v.a[k].real()) && std::isfinite(v.a[k].imag());
    }
    // Not converged or not finite: the host's eigensolver.
    return good || dense::hermitian_eigen(a, w, v);
}

}  // namespace caroli::gpu
