// Multifrontal solves with the stored front factors (complex f64 right-hand sides).
//
// A front F = [[F11, F12], [F21, F22]] was factored in place as P F11 = L11 U11 (row pivots within
// the separator rows), L21 = F21 U11^-1, U12 = L11^-1 P F12. For A x = b:
//   forward   y = L11^-1 P b_e,        b_b -= L21 y          (children before parents)
//   backward  x_e = U11^-1 (y - U12 x_b)                     (parents before children)
// and for A^T x = b (used by the adjoint):
//   forward   g = U11^-T b_e,           b_b -= U12^T g
//   backward  x_e = P^T L11^-T (g - L21^T x_b).
// The root is a dense LU of its leading nd + k block. Boundary updates of sibling fronts meet on
// the parent's variables, so they are accumulated with FP64 atomics.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "cx.hpp"
#include "fronts.hpp"

namespace caroli::gpu {

namespace {

constexpr int kThreads = 256;

template <class TF>
__device__ __forceinline__ double2 widen(TF a) {
    return double2{double(a.x), double(a.y)};
}

__device__ __forceinline__ void atomic_sub(double2* p, double2 v) {
    atomicAdd(&p->x, -v.x);
    atomicAdd(&p->y, -v.y);
}

template <class TF, bool kT>
__global__ void __launch_bounds__(kThreads)
    solve_forward_kernel(const NodeDesc* __restrict__ nodes, const int32_t* __restrict__ level, PointsView pv, SolveView sv) {
    extern __shared__ double2 y[];  // ne x c, column-major
    const NodeDesc nd = nodes[level[blockIdx.x]];
    const int64_t p = blockIdx.y;
    const TF* F = reinterpret_cast<const TF*>(pv.arena) + p * pv.arena_stride + nd.off;
    const int32_t* piv = pv.piv + p * pv.piv_stride + nd.piv_off;
    const int32_t* vars = sv.vars + nd.var_off;
    double2* X = sv.x + p * sv.x_stride;
    const int n = nd.n, ne = nd.ne, nb = n - ne, c = sv.c, tid = threadIdx.x;
    auto f = [&](int i, int j) { return widen(F[i + int64_t(j) * n]); };
    for (int idx = tid; idx < ne * c; idx += kThreads) y[idx] = X[vars[idx % ne] + int64_t(idx / ne) * sv.ldx];
    __syncthreads();
    if (!kT) {
        for (int col = tid; col < c; col += kThreads)
            for (int k = 0; k < ne; ++k) {
                int q = piv[k];
                if (q != k) {
                    double2 t = y[k + col * ne];
                    y[k + col * ne] = y[q + col * ne];
                    y[q + col * ne] = t;
                }
            }
        __syncthreads();
        for (int k = 0; k < ne; ++k) {  // unit lower L11
            for (int idx = tid; idx < (ne - k - 1) * c; idx += kThreads) {
                int i = k + 1 + idx % (ne - k - 1), col = idx / (ne - k - 1);
                y[i + col * ne] = csub(y[i + col * ne], cmul(f(i, k), y[k + col * ne]));
            }
            __syncthreads();
        }
        for (int idx = tid; idx < nb * c; idx += kThreads) {  // b_b -= L21 y
            int b = idx % nb, col = idx / nb;
            double2 acc{0.0, 0.0};
            for (int k = 0; k < ne; ++k) acc = cadd(acc, cmul(f(ne + b, k), y[k + col * ne]));
            atomic_sub(&X[vars[ne + b] + int64_t(col) * sv.ldx], acc);
        }
    } else {
        for (int k = 0; k < ne; ++k) {  // U11^T g = b_e (lower, non-unit)
            for (int col = tid; col < c; col += kThreads) y[k + col * ne] = cmul(y[k + col * ne], crecip(f(k, k)));
            __syncthreads();
            for (int idx = tid; idx < (ne - k - 1) * c; idx += kThreads) {
                int i = k + 1 + idx % (ne - k - 1), col = idx / (ne - k - 1);
                y[i + col * ne] = csub(y[i + col * ne], cmul(f(k, i), y[k + col * ne]));
            }
            __syncthreads();
        }
        for (int idx = tid; idx < nb * c; idx += kThreads) {  // b_b -= U12^T g
            int b = idx % nb, col = idx / nb;
            double2 acc{0.0, 0.0};
            for (int k = 0; k < ne; ++k) acc = cadd(acc, cmul(f(k, ne + b), y[k + col * ne]));
            atomic_sub(&X[vars[ne + b] + int64_t(col) * sv.ldx], acc);
        }
    }
    for (int idx = tid; idx < ne * c; idx += kThreads) X[vars[idx % ne] + int64_t(idx / ne) * sv.ldx] = y[idx];
}

template <class TF, bool kT>
__global__ void __launch_bounds__(kThreads)
    solve_backward_kernel(const NodeDesc* __restrict__ nodes, const int32_t* __restrict__ level, PointsView pv, SolveView sv) {
    extern __shared__ double2 y[];  // ne x c
    const NodeDesc nd = nodes[level[blockIdx.x]];
    const int64_t p = blockIdx.y;
    const TF* F = reinterpret_cast<const TF*>(pv.arena) + p * pv.arena_stride + nd.off;
    const int32_t* piv = pv.piv + p * pv.piv_stride + nd.piv_off;
    const int32_t* vars = sv.vars + nd.var_off;
    double2* X = sv.x + p * sv.x_stride;
    const int n = nd.n, ne = nd.ne, nb = n - ne, c = sv.c, tid = threadIdx.x;
    auto f = [&](int i, int j) { return widen(F[i + int64_t(j) * n]); };
    // y = x_e - (U12 or L21^T) x_b
    for (int idx = tid; idx < ne * c; idx += kThreads) {
        int k = idx % ne, col = idx / ne;
        double2 acc = X[vars[k] + int64_t(col) * sv.ldx];
        for (int b = 0; b < nb; ++b) {
            double2 xb = X[vars[ne + b] + int64_t(col) * sv.ldx];
            acc = csub(acc, cmul(kT ? f(ne + b, k) : f(k, ne + b), xb));
        }
        y[idx] = acc;
    }
    __syncthreads();
    if (!kT) {
        for (int k = ne - 1; k >= 0; --k) {  // U11 back substitution
            for (int col = tid; col < c; col += kThreads) y[k + col * ne] = cmul(y[k + col * ne], crecip(f(k, k)));
            __syncthreads();
            for (int idx = tid; idx < k * c; idx += kThreads) {
                int i = idx % k, col = idx / k;
                y[i + col * ne] = csub(y[i + col * ne], cmul(f(i, k), y[k + col * ne]));
            }
            __syncthreads();
        }
    } else {
        for (int k = ne - 1; k >= 0; --k) {  // L11^T (unit upper)
            for (int idx = tid; idx < k * c; idx += kThreads) {
                int i = idx % k, col = idx / k;
                y[i + col * ne] = csub(y[i + col * ne], cmul(f(k, i), y[k + col * ne]));
            }
            __syncthreads();
        }
        for (int col = tid; col < c; col += kThreads)  // P^T: the swaps in reverse order
            for (int k = ne - 1; k >= 0; --k) {
                int q = piv[k];
                if (q != k) {
                    double2 t = y[k + col * ne];
                    y[k + col * ne] = y[q + col * ne];
                    y[q + col * ne] = t;
                }
            }
        __syncthreads();
    }
    for (int idx = tid; idx < ne * c; idx += kThreads) X[vars[idx % ne] + int64_t(idx / ne) * sv.ldx] = y[idx];
}

// Dense solve with the root's LU (its leading ne x ne block, complex f64).
template <bool kT>
__global__ void __launch_bounds__(kThreads)
    solve_root_kernel(const NodeDesc* __restrict__ nodes, const int32_t* __restrict__ level, PointsView pv, SolveView sv) {
    extern __shared__ double2 y[];
    const NodeDesc nd = nodes[level[0]];
    const int64_t p = blockIdx.x;
    const double2* F = reinterpret_cast<const double2*>(pv.root_arena) + p * pv.root_stride + nd.off;
    const int32_t* piv = pv.piv + p * pv.piv_stride + nd.piv_off;
    const int32_t* vars = sv.vars + nd.var_off;
    double2* X = sv.x + p * sv.x_stride;
    const int n = nd.n, ne = nd.ne, c = sv.c, tid = threadIdx.x;
    auto f = [&](int i, int j) { return F[i + int64_t(j) * n]; };
    for (int idx = tid; idx < ne * c; idx += kThreads) y[idx] = X[vars[idx % ne] + int64_t(idx / ne) * sv.ldx];
    __syncthreads();
    auto swaps = [&](bool reverse) {
        for (int col = tid; col < c; col += kThreads)
            for (int s = 0; s < ne; ++s) {
                int k = reverse ? ne - 1 - s : s, q = piv[k];
                if (q != k) {
                    double2 t = y[k + col * ne];
                    y[k + col * ne] = y[q + col * ne];
                    y[q + col * ne] = t;
                }
            }
        __syncthreads();
    };
    if (!kT) {
        swaps(false);
        for (int k = 0; k < ne; ++k) {
            for (int idx = tid; idx < (ne - k - 1) * c; idx += kThreads) {
                int i = k + 1 + idx % (ne - k - 1), col = idx / (ne - k - 1);
                y[i + col * ne] = csub(y[i + col * ne], cmul(f(i, k), y[k + col * ne]));
            }
            __syncthreads();
        }
        for (int k = ne - 1; k >= 0; --k) {
            for (int col = tid; col < c; col += kThreads) y[k + col * ne] = cmul(y[k + col * ne], crecip(f(k, k)));
            __syncthreads();
            for (int idx = tid; idx < k * c; idx += kThreads) {
                int i = idx % k, col = idx / k;
                y[i + col * ne] = csub(y[i + col * ne], cmul(f(i, k), y[k + col * ne]));
            }
            __syncthreads();
        }
    } else {
        for (int k = 0; k < ne; ++k) {  // U^T a = b
            for (int col = tid; col < c; col += kThreads) y[k + col * ne] = cmul(y[k + col * ne], crecip(f(k, k)));
            __syncthreads();
            for (int idx = tid; idx < (ne - k - 1) * c; idx += kThreads) {
                int i = k + 1 + idx % (ne - k - 1), col = idx / (ne - k - 1);
                y[i + col * ne] = csub(y[i + col * ne], cmul(f(k, i), y[k + col * ne]));
            }
            __syncthreads();
        }
        for (int k = ne - 1; k >= 0; --k) {  // L^T (unit upper)
            for (int idx = tid; idx < k * c; idx += kThreads) {
                int i = idx % k, col = idx / k;
                y[i + col * ne] = csub(y[i + col * ne], cmul(f(k, i), y[k + col * ne]));
            }
            __syncthreads();
        }
        swaps(true);
    }
    for (int idx = tid; idx < ne * c; idx += kThreads) X[vars[idx % ne] + int64_t(idx / ne) * sv.ldx] = y[idx];
}

// B[orbs[a], col] = W[a, col] (or conj) for a < m, per point; B must be zero elsewhere.
__global__ void __launch_bounds__(kThreads)
    scatter_kernel(SolveView sv, const int32_t* __restrict__ orbs, int32_t m, const double2* __restrict__ w,
                   int64_t w_stride, int32_t conj) {
    const int64_t p = blockIdx.y;
    double2* X = sv.x + p * sv.x_stride;
    const double2* W = w + p * w_stride;
    for (int idx = blockIdx.x * kThreads + threadIdx.x; idx < m * sv.c; idx += gridDim.x * kThreads) {
        int a = idx % m, col = idx / m;
        double2 v = W[a + int64_t(col) * m];
        if (conj) v.y = -v.y;
        X[orbs[a] + int64_t(col) * sv.ldx] = v;
    }
}

// D = B - (z - H - Sigma_L - Sigma_R) X on the rows of a block, in FP64 (the residual of the
// refinement). Sigma_L acts on the left-slice rows (positions left_pos), Sigma_R on the right.
__global__ void __launch_bounds__(kThreads)
    residual_kernel(PointsView pv, SolveView xv, const double2* __restrict__ b, double2* __restrict__ d,
                    const int32_t* __restrict__ rows, int32_t nrows, const int64_t* __restrict__ row_ptr,
                    const uint32_t* __restrict__ cols, const int32_t* __restrict__ left_pos,
                    const int32_t* __restrict__ right_pos, const int32_t* __restrict__ left_orbs, int32_t nl,
                    const int32_t* __restrict__ right_orbs, int32_t nr) {
    const int64_t p = blockIdx.y;
    const double2* X = xv.x + p * xv.x_stride;
    const double2* B = b + p * xv.x_stride;
    double2* D = d + p * xv.x_stride;
    const double2* vals = pv.vals[p];
    const double2 z = pv.z[p];
    const double2* sl = pv.sigma + p * pv.sigma_stride;
    const double2* sr = sl + int64_t(nl) * nl;
    for (int idx = blockIdx.x * kThreads + threadIdx.x; idx < nrows * xv.c; idx += gridDim.x * kThreads) {
        const int r = rows[idx % nrows], col = idx / nrows;
        const double2* x = X + int64_t(col) * xv.ldx;
        double2 acc = cmul(z, x[r]);
        for (int64_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e) acc = csub(acc, cmul(vals[e], x[cols[e]]));
        const int lp = left_pos[r], rp = right_pos[r];
        if (lp >= 0)
            for (int a = 0; a < nl; ++a) acc = csub(acc, cmul(sl[lp + int64_t(a) * nl], x[left_orbs[a]]));
        if (rp >= 0)
            for (int a = 0; a < nr; ++a) acc = csub(acc, cmul(sr[rp + int64_t(a) * nr], x[right_orbs[a]]));
        const int64_t at = r + int64_t(col) * xv.ldx;
        D[at] = csub(B[at], acc);
    }
}

// X += D on the block rows; ratio[p] = max |D| / max |X| (the relative correction).
__global__ void __launch_bounds__(kThreads)
    update_kernel(SolveView xv, const double2* __restrict__ d, const int32_t* __restrict__ rows, int32_t nrows,
                  double* __restrict__ ratio) {
    __shared__ double sd[kThreads / 64], sx[kThreads / 64];
    const int64_t p = blockIdx.x;
    double2* X = xv.x + p * xv.x_stride;
    const double2* D = d + p * xv.x_stride;
    double md = 0.0, mx = 0.0;
    for (int idx = threadIdx.x; idx < nrows * xv.c; idx += kThreads) {
        const int64_t at = rows[idx % nrows] + int64_t(idx / nrows) * xv.ldx;
        double2 v = cadd(X[at], D[at]);
        X[at] = v;
        md = fmax(md, cabs(D[at]));
        mx = fmax(mx, cabs(v));
    }
    for (int off = 32; off > 0; off >>= 1) {
        md = fmax(md, __shfl_xor(md, off, 64));
        mx = fmax(mx, __shfl_xor(mx, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        sd[threadIdx.x >> 6] = md;
        sx[threadIdx.x >> 6] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kThreads / 64; ++w) {
            md = fmax(md, sd[w]);
            mx = fmax(mx, sx[w]);
        }
        md = fmax(md, sd[0]);
        mx = fmax(mx, sx[0]);
        ratio[p] = mx > 0.0 ? md / mx : (md > 0.0 ? INFINITY : 0.0);
    }
}

// out[a, col] = X[orbs[a], col] per point (m x c, column-major).
__global__ void __launch_bounds__(kThreads)
    gather_kernel(SolveView xv, const int32_t* __restrict__ orbs, int32_t m, double2* __restrict__ out) {
    const int64_t p = blockIdx.y;
    const double2* X = xv.x + p * xv.x_stride;
    double2* O = out + p * int64_t(m) * xv.c;
    for (int idx = blockIdx.x * kThreads + threadIdx.x; idx < m * xv.c; idx += gridDim.x * kThreads)
        O[idx] = X[orbs[idx % m] + int64_t(idx / m) * xv.ldx];
}

}  // namespace

inline dim3 rows_grid(int64_t work, int32_t points) {
    return dim3(unsigned(std::min<int64_t>((work + kThreads - 1) / kThreads, 4096)), unsigned(points));
}

void launch_scatter(const SolveView& sv, const int32_t* orbs, int32_t m, const double2* w, int64_t w_stride, bool conj,
                    int32_t points, hipStream_t st) {
    if (points <= 0 || m <= 0 || sv.c <= 0) return;
    hipLaunchKernelGGL(scatter_kernel, rows_grid(int64_t(m) * sv.c, points), dim3(kThreads), 0, st, sv, orbs, m, w,
                       w_stride, conj ? 1 : 0);
}

void launch_residual(const PointsView& pv, const SolveView& xv, const double2* b, double2* d, const int32_t* rows,
                     int32_t nrows, const int64_t* row_ptr, const uint32_t* cols, const int32_t* left_pos,
                     const int32_t* right_pos, const int32_t* left_orbs, int32_t nl, const int32_t* right_orbs,
                     int32_t nr, int32_t points, hipStream_t st) {
    if (points <= 0 || xv.c <= 0) return;
    hipLaunchKernelGGL(residual_kernel, rows_grid(int64_t(nrows) * xv.c, points), dim3(kThreads), 0, st, pv, xv, b, d,
                       rows, nrows, row_ptr, cols, left_pos, right_pos, left_orbs, nl, right_orbs, nr);
}

void launch_update(const SolveView& xv, const double2* d, const int32_t* rows, int32_t nrows, double* ratio,
                   int32_t points, hipStream_t st) {
    if (points <= 0 || xv.c <= 0) return;
    hipLaunchKernelGGL(update_kernel, dim3(unsigned(points)), dim3(kThreads), 0, st, xv, d, rows, nrows, ratio);
}

void launch_gather(const SolveView& xv, const int32_t* orbs, int32_t m, double2* out, int32_t points, hipStream_t st) {
    if (points <= 0 || m <= 0 || xv.c <= 0) return;
    hipLaunchKernelGGL(gather_kernel, rows_grid(int64_t(m) * xv.c, points), dim3(kThreads), 0, st, xv, orbs, m, out);
}

using SolveKernel = void (*)(const NodeDesc*, const int32_t*, PointsView, SolveView);

void launch_solve_forward(const NodeDesc* nodes, const int32_t* level, int32_t count, const PointsView& pv,
                          const SolveView& sv, int32_t points, bool single, bool transpose, hipStream_t st) {
    if (count <= 0 || points <= 0) return;
    SolveKernel k = single ? (transpose ? solve_forward_kernel<float2, true> : solve_forward_kernel<float2, false>)
                           : (transpose ? solve_forward_kernel<double2, true> : solve_forward_kernel<double2, false>);
    hipLaunchKernelGGL(k, dim3(unsigned(count), unsigned(points)), dim3(kThreads), size_t(kSolveLds) * sizeof(double2),
                       st, nodes, level, pv, sv);
}

void launch_solve_backward(const NodeDesc* nodes, const int32_t* level, int32_t count, const PointsView& pv,
                           const SolveView& sv, int32_t points, bool single, bool transpose, hipStream_t st) {
    if (count <= 0 || points <= 0) return;
    SolveKernel k = single ? (transpose ? solve_backward_kernel<float2, true> : solve_backward_kernel<float2, false>)
                           : (transpose ? solve_backward_kernel<double2, true> : solve_backward_kernel<double2, false>);
    hipLaunchKernelGGL(k, dim3(unsigned(count), unsigned(points)), dim3(kThreads), size_t(kSolveLds) * sizeof(double2),
                       st, nodes, level, pv, sv);
}

void launch_solve_root(const NodeDesc* nodes, const int32_t* level, const PointsView& pv, const SolveView& sv,
                       int32_t points, bool transpose, hipStream_t st) {
    if (points <= 0) return;
    const size_t lds = size_t(kSolveLds) * sizeof(double2);
    if (transpose)
        hipLaunchKernelGGL((solve_root_kernel<true>), dim3(unsigned(points)), dim3(kThreads), lds, st, nodes, level, pv, sv);
    else
        hipLaunchKernelGGL((solve_root_kernel<false>), dim3(unsigned(points)), dim3(kThreads), lds, st, nodes, level, pv, sv);
}

}  // namespace caroli::gpu
