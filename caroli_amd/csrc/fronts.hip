// Front kernels of the multifrontal elimination on CDNA4 (gfx950).
//
// One workgroup (256 lanes = 4 wavefronts of 64) per (front, point): the launch grid is the fronts
// of one tree level times the points of the batch, so thousands of small fronts fill the 256 CUs.
//  - assemble: F = z - H on the separator, the children's Schur complements added at their
//    positions (extend-add, coalesced along the columns), and at the root the lead self-energies
//    and the unit borders that make the Green's function block the Schur complement of the front.
//  - factor: partial LU with row pivoting restricted to the separator rows, the pivot test of
//    crates/caroli-dense/src/field.rs:61 and nested.rs:1167-1223, and the rank-1 updates that leave
//    S = F22 - F21 F11^-1 F12 in the trailing block. Small fronts run in LDS.
//  - trace: the Caroli trace Tr[Gamma_L G Gamma_R G^H] as one fused reduction.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>

#include "cx.hpp"
#include "fronts.hpp"

namespace caroli::gpu {

namespace {

constexpr int kThreads = 256;


// Block-wide max of (value, -index): the largest value, the smallest index on ties.
__device__ void block_argmax(double& v, int& idx, double* sv, int* si) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int off = 32; off > 0; off >>= 1) {
        double ov = __shfl_xor(v, off, 64);
        int oi = __shfl_xor(idx, off, 64);
        if (ov > v || (ov == v && oi < idx)) {
            v = ov;
            idx = oi;
        }
    }
    if (lane == 0) {
        sv[wave] = v;
        si[wave] = idx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < int(blockDim.x) / 64; ++w)
            if (sv[w] > sv[0] || (sv[w] == sv[0] && si[w] < si[0])) {
                sv[0] = sv[w];
                si[0] = si[w];
            }
    }
    __syncthreads();
    v = sv[0];
    idx = si[0];
    __syncthreads();
}

__device__ double block_max(double v, double* sv) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    if ((threadIdx.x & 63) == 0) sv[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = sv[0];
    for (int w = 1; w < int(blockDim.x) / 64; ++w) r = fmax(r, sv[w]);
    __syncthreads();
    return r;
}

__device__ double block_sum(double v, double* sv) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) sv[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.0;
    for (int w = 0; w < int(blockDim.x) / 64; ++w) r += sv[w];
    __syncthreads();
    return r;
}

// Assembly of one front (see the file comment). TF: type of the front, TK: type of the children.
template <class TF, class TK, bool kRoot>
__global__ void __launch_bounds__(kThreads)
    assemble_kernel(const NodeDesc* __restrict__ nodes, const int32_t* __restrict__ level,
                    const Entry* __restrict__ ent, const int32_t* __restrict__ diag, const KidDesc* __restrict__ kids,
                    const int32_t* __restrict__ maps, PointsView pv) {
    const NodeDesc nd = nodes[level[blockIdx.x]];
    const int64_t p = blockIdx.y;
    TF* F = (kRoot ? reinterpret_cast<TF*>(pv.root_arena) + p * pv.root_stride
                   : reinterpret_cast<TF*>(pv.arena) + p * pv.arena_stride) +
            nd.off;
    const TK* kidbase = reinterpret_cast<const TK*>(pv.arena) + p * pv.arena_stride;
    const int n = nd.n, tid = threadIdx.x;
    const int64_t n2 = int64_t(n) * n;
    const TF zero = cx<TF>(0.0, 0.0);
    for (int64_t i = tid; i < n2; i += kThreads) F[i] = zero;
    __syncthreads();
    const double2* vals = pv.vals[p];
    for (int e = nd.ent_beg + tid; e < nd.ent_end; e += kThreads) {
        Entry en = ent[e];
        double2 v = vals[en.k];
        if (en.conj) v.y = -v.y;
        int64_t at = en.a + int64_t(en.b) * n;
        F[at] = csub(F[at], convert<TF>(v));
    }
    __syncthreads();
    const TF z = convert<TF>(pv.z[p]);
    for (int d = nd.diag_beg + tid; d < nd.diag_end; d += kThreads) {
        int64_t q = diag[d];
        F[q + q * n] = cadd(F[q + q * n], z);
    }
    __syncthreads();
    if (kRoot) {
        const double2* sg = pv.sigma + p * pv.sigma_stride;
        const int nl = nd.nl, nr = nd.nr, b0 = nd.nd, r0 = nd.nd + nd.right_at, k0 = nd.ne;
        for (int idx = tid; idx < nl * nl; idx += kThreads) {
            int a = idx % nl, c = idx / nl;
            int64_t at = (b0 + a) + int64_t(b0 + c) * n;
            F[at] = csub(F[at], convert<TF>(sg[idx]));
        }
        __syncthreads();
        for (int idx = tid; idx < nr * nr; idx += kThreads) {
            int a = idx % nr, c = idx / nr;
            int64_t at = (r0 + a) + int64_t(r0 + c) * n;
            F[at] = csub(F[at], convert<TF>(sg[nl * nl + idx]));
        }
        const TF one = cx<TF>(1.0, 0.0);
        for (int j = tid; j < nr; j += kThreads) F[(r0 + j) + int64_t(k0 + j) * n] = one;
        for (int i = tid; i < nl; i += kThreads) F[(k0 + i) + int64_t(b0 + i) * n] = one;
        __syncthreads();
    }
    for (int c = nd.kid_beg; c < nd.kid_end; ++c) {
        const KidDesc kd = kids[c];
        const TK* S = kidbase + kd.src;
        const int32_t* map = maps + kd.map_beg;
        const int m = kd.size;
        const int64_t m2 = int64_t(m) * m;
        for (int64_t idx = tid; idx < m2; idx += kThreads) {
            int a = int(idx % m), b = int(idx / m);
            int64_t at = map[a] + int64_t(map[b]) * n;
            F[at] = cadd(F[at], convert<TF>(S[a + int64_t(b) * kd.ld]));
        }
        __syncthreads();
    }
}

// Partial LU of one front (see NodeDesc). W points at the working copy (LDS or the front itself).
template <class T>
__device__ void eliminate(T* W, const NodeDesc& nd, int32_t* piv, int32_t* status, double tol, bool root) {
    __shared__ double sv[16];
    __shared__ int si[16];
    const int n = nd.n, ne = nd.ne, ne1 = nd.ne1, tid = threadIdx.x;
    double m = 0.0;
    for (int idx = tid; idx < ne1 * ne1; idx += kThreads) m = fmax(m, cabs(W[(idx % ne1) + int64_t(idx / ne1) * n]));
    const double scale = block_max(m, sv);
    for (int k = 0; k < ne; ++k) {
        const int lim = k < ne1 ? ne1 : ne;
        double big = -1.0;
        int at = k;
        for (int i = k + tid; i < lim; i += kThreads) {
            double v = cabs(W[i + int64_t(k) * n]);
            if (v > big) {
                big = v;
                at = i;
            }
        }
        block_argmax(big, at, sv, si);
        const bool ok = k < ne1 ? (big > tol * scale) : (big > 0.0 && big < INFINITY);
        if (!ok) {
            if (tid == 0) *status = (k < ne1 && !root) ? kFrontDelayed : kFrontSingular;
            return;
        }
        if (at != k)
            for (int j = tid; j < n; j += kThreads) {
                T x = W[k + int64_t(j) * n];
                W[k + int64_t(j) * n] = W[at + int64_t(j) * n];
                W[at + int64_t(j) * n] = x;
            }
        if (tid == 0) piv[k] = at;
        __syncthreads();
        const T inv = crecip(W[k + int64_t(k) * n]);
        for (int i = k + 1 + tid; i < n; i += kThreads) W[i + int64_t(k) * n] = cmul(W[i + int64_t(k) * n], inv);
        __syncthreads();
        const int r = n - k - 1;
        const int64_t r2 = int64_t(r) * r;
        const T* lk = W + int64_t(k) * n;
        for (int64_t idx = tid; idx < r2; idx += kThreads) {
            int i = k + 1 + int(idx % r), j = k + 1 + int(idx / r);
            T u = W[k + int64_t(j) * n];
            W[i + int64_t(j) * n] = csub(W[i + int64_t(j) * n], cmul(lk[i], u));
        }
        __syncthreads();
    }
    if (tid == 0) *status = kFrontOk;
}

// MFMA 16x16x4 on one wavefront: lane l holds A[l % 16][l / 16] and B[l / 16][l % 16]. The
// accumulator element j of lane l holds D[row(l, j)][l % 16], with row = 4 (l / 16) + j for
// f32 and row = l / 16 + 4 j for f64 (measured on gfx950).
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
template <class T>
struct Mfma;
template <>
struct Mfma<double2> {
    using acc = f64x4;
    static __device__ __forceinline__ acc mac(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int j) { return lane / 16 + 4 * j; }
};
template <>
struct Mfma<float2> {
    using acc = f32x4;
    static __device__ __forceinline__ acc mac(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int j) { return 4 * (lane / 16) + j; }
};

constexpr int kPanel = 16;  // panel width of the blocked elimination

// Blocked right-looking partial LU of a front in global memory (the same factorisation as
// `eliminate`, panel by panel): unblocked pivoting inside a panel of kPanel columns, the
// triangular solve of the panel's row block, then the trailing update on MFMA, computed as
// C^T -= U12^T L21^T so that the front's columns are read and written contiguously.
template <class T>
__device__ void eliminate_blocked(T* F, const NodeDesc& nd, int32_t* piv, int32_t* status, double tol, bool root) {
    using R = typename Real<T>::type;
    using M = Mfma<T>;
    __shared__ double sv[16];
    __shared__ int si[16];
    __shared__ T l11[kPanel * kPanel];
    const int n = nd.n, ne = nd.ne, ne1 = nd.ne1, tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    auto at = [&](int i, int j) -> T& { return F[i + int64_t(j) * n]; };
    double m = 0.0;
    for (int idx = tid; idx < ne1 * ne1; idx += kThreads) m = fmax(m, cabs(at(idx % ne1, idx / ne1)));
    const double scale = block_max(m, sv);
    for (int k0 = 0; k0 < ne; k0 += kPanel) {
        const int w = min(kPanel, ne - k0);
        // 1. Panel: columns [k0, k0 + w), rows [k0, n).
        for (int k = k0; k < k0 + w; ++k) {
            const int lim = k < ne1 ? ne1 : ne;
            double big = -1.0;
            int p = k;
            for (int i = k + tid; i < lim; i += kThreads) {
                double v = cabs(at(i, k));
                if (v > big) {
                    big = v;
                    p = i;
                }
            }
            block_argmax(big, p, sv, si);
            const bool ok = k < ne1 ? (big > tol * scale) : (big > 0.0 && big < INFINITY);
            if (!ok) {
                if (tid == 0) *status = (k < ne1 && !root) ? kFrontDelayed : kFrontSingular;
                return;
            }
            if (p != k)
                for (int j = tid; j < n; j += kThreads) {
                    T x = at(k, j);
                    at(k, j) = at(p, j);
                    at(p, j) = x;
                }
            if (tid == 0) piv[k] = p;
            __syncthreads();
            const T inv = crecip(at(k, k));
            for (int i = k + 1 + tid; i < n; i += kThreads) at(i, k) = cmul(at(i, k), inv);
            __syncthreads();
            const int cols = k0 + w - k - 1, rows = n - k - 1;
            for (int idx = tid; idx < cols * rows; idx += kThreads) {
                int i = k + 1 + idx % rows, j = k + 1 + idx / rows;
                at(i, j) = csub(at(i, j), cmul(at(i, k), at(k, j)));
            }
            __syncthreads();
        }
        const int j1 = k0 + w;  // first trailing row and column
        if (j1 >= n) break;
        // 2. U12 = L11^-1 A12 on the panel's rows, one column per lane.
        for (int idx = tid; idx < kPanel * kPanel; idx += kThreads) {
            int r = idx % kPanel, c = idx / kPanel;
            l11[idx] = (r < w && c < w) ? at(k0 + r, k0 + c) : cx<T>(0.0, 0.0);
        }
        __syncthreads();
        for (int j = j1 + tid; j < n; j += kThreads) {
            T u[kPanel];
#pragma unroll
            for (int r = 0; r < kPanel; ++r) u[r] = r < w ? at(k0 + r, j) : cx<T>(0.0, 0.0);
#pragma unroll
            for (int r = 1; r < kPanel; ++r)
#pragma unroll
                for (int c = 0; c < r; ++c) u[r] = csub(u[r], cmul(l11[r + c * kPanel], u[c]));
#pragma unroll
            for (int r = 0; r < kPanel; ++r)
                if (r < w) at(k0 + r, j) = u[r];
        }
        __syncthreads();
        // 3. Trailing update on MFMA: each wave owns columns of 16x16 tiles.
        const int rem = n - j1, tiles = (rem + 15) / 16;
        for (int tj = wave; tj < tiles; tj += kThreads / 64) {
            const int j0 = j1 + 16 * tj;
            // A' = U12^T: lane holds U12[k0 + 4 s + lane / 16][j0 + lane % 16] for the 4 slices.
            R ar[4], ai[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                int r = 4 * s + lane / 16, j = j0 + lane % 16;
                T v = (r < w && j < n) ? at(k0 + r, j) : cx<T>(0.0, 0.0);
                ar[s] = v.x;
                ai[s] = v.y;
            }
            for (int ti = 0; ti < tiles; ++ti) {
                const int i0 = j1 + 16 * ti;
                typename M::acc cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    // B' = L21^T: lane holds L21[i0 + lane % 16][k0 + 4 s + lane / 16].
                    int c = 4 * s + lane / 16, i = i0 + lane % 16;
                    T b = (c < w && i < n) ? at(i, k0 + c) : cx<T>(0.0, 0.0);
                    cr = M::mac(ar[s], b.x, cr);
                    cr = M::mac(-ai[s], b.y, cr);
                    ci = M::mac(ar[s], b.y, ci);
                    ci = M::mac(ai[s], b.x, ci);
                }
                // D'[row(lane, jj)][lane % 16] = C[i0 + lane % 16][j0 + row(lane, jj)].
                const int i = i0 + lane % 16;
                if (i < n)
#pragma unroll
                    for (int jj = 0; jj < 4; ++jj) {
                        int j = j0 + M::row(lane, jj);
                        if (j < n) {
                            T& d = at(i, j);
                            d.x -= cr[jj];
                            d.y -= ci[jj];
                        }
                    }
            }
        }
        __syncthreads();
    }
    if (tid == 0) *status = kFrontOk;
}

template <class T, bool kRoot>
__global__ void __launch_bounds__(kThreads)
    factor_blocked_kernel(const NodeDesc* __restrict__ nodes, const int32_t* __restrict__ level, PointsView pv, double tol) {
    const NodeDesc nd = nodes[level[blockIdx.x]];
    const int64_t p = blockIdx.y;
    T* F = (kRoot ? reinterpret_cast<T*>(pv.root_arena) + p * pv.root_stride
                  : reinterpret_cast<T*>(pv.arena) + p * pv.arena_stride) +
           nd.off;
    eliminate_blocked(F, nd, pv.piv + p * pv.piv_stride + nd.piv_off, pv.status + p * pv.status_stride + nd.slot, tol,
                      kRoot);
}

// One 16x16 tile of C -= A B on MFMA, computed as C^T -= B^T A^T so that the front's columns are
// read and written contiguously: A = F[i0:i0+16, k0:k1] (rows of L), B = F[k0:k1, j0:j0+16]
// (rows of U), K = k1 - k0 in slices of 4. Every lane runs every MFMA; entries outside [0, n)
// and those `skip` selects are not stored.
template <class T, class Skip>
__device__ __forceinline__ void tile_update(T* F, int n, int i0, int j0, int k0, int k1, int lane, Skip skip) {
    using M = Mfma<T>;
    auto at = [&](int i, int j) -> T& { return F[i + int64_t(j) * n]; };
    typename M::acc cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
    const int li = lane % 16, lk = lane / 16;
    for (int kk = k0; kk < k1; kk += 4) {
        const int k = kk + lk;
        const int j = j0 + li, i = i0 + li;
        T a = (k < k1 && j < n) ? at(k, j) : cx<T>(0.0, 0.0);
        T b = (k < k1 && i < n) ? at(i, k) : cx<T>(0.0, 0.0);
        cr = M::mac(a.x, b.x, cr);
        cr = M::mac(-a.y, b.y, cr);
        ci = M::mac(a.x, b.y, ci);
        ci = M::mac(a.y, b.x, ci);
    }
    const int i = i0 + li;
    for (int jj = 0; jj < 4; ++jj) {
        const int j = j0 + M::row(lane, jj);
        const bool store = i < n && j < n && !skip(i, j);
        if (store) {
            T& d = at(i, j);
            d.x -= cr[jj];
            d.y -= ci[jj];
        }
    }
}
// Blocked partial LU with a deferred Schur update: pivoting inside panels of kPanel columns,
// the triangular solve of each panel's row block, and MFMA updates of the separator rows and
// columns only; S = F22 - L21 U12 is formed once at the end with K = ne, so the Schur block is
// read and written once instead of once per panel.
template <class T>
__device__ void eliminate_deferred(T* F, const NodeDesc& nd, int32_t* piv, int32_t* status, double tol, bool root) {
    __shared__ double sv[16];
    __shared__ int si[16];
    __shared__ T l11[kPanel * kPanel];
    const int n = nd.n, ne = nd.ne, ne1 = nd.ne1, tid = threadIdx.x, nt = blockDim.x;
    const int lane = tid & 63, wave = tid >> 6, waves = nt >> 6;
    auto at = [&](int i, int j) -> T& { return F[i + int64_t(j) * n]; };
    double m = 0.0;
    for (int idx = tid; idx < ne1 * ne1; idx += nt) m = fmax(m, cabs(at(idx % ne1, idx / ne1)));
    const double scale = block_max(m, sv);
    for (int k0 = 0; k0 < ne; k0 += kPanel) {
        const int w = min(kPanel, ne - k0);
        for (int k = k0; k < k0 + w; ++k) {
            const int lim = k < ne1 ? ne1 : ne;
            double big = -1.0;
            int p = k;
            for (int i = k + tid; i < lim; i += nt) {
                double v = cabs(at(i, k));
                if (v > big) { big = v; p = i; }
            }
            block_argmax(big, p, sv, si);
            const bool ok = k < ne1 ? (big > tol * scale) : (big > 0.0 && big < INFINITY);
            if (!ok) { if (tid == 0) *status = (k < ne1 && !root) ? kFrontDelayed : kFrontSingular; return; }
            if (p != k)
                for (int j = tid; j < n; j += nt) { T x = at(k, j); at(k, j) = at(p, j); at(p, j) = x; }
            if (tid == 0) piv[k] = p;
            __syncthreads();
            const T inv = crecip(at(k, k));
            for (int i = k + 1 + tid; i < n; i += nt) at(i, k) = cmul(at(i, k), inv);
            __syncthreads();
            const int cols = k0 + w - k - 1, rows = n - k - 1;
            for (int idx = tid; idx < cols * rows; idx += nt) {
                int i = k + 1 + idx % rows, j = k + 1 + idx / rows;
                at(i, j) = csub(at(i, j), cmul(at(i, k), at(k, j)));
            }
            __syncthreads();
        }
        const int j1 = k0 + w;
        if (j1 >= n) break;
        for (int idx = tid; idx < kPanel * kPanel; idx += nt) {
            int r = idx % kPanel, c = idx / kPanel;
            l11[idx] = (r < w && c < w) ? at(k0 + r, k0 + c) : cx<T>(0.0, 0.0);
        }
        __syncthreads();
        for (int j = j1 + tid; j < n; j += nt)
            for (int r = 1; r < w; ++r) {
                T acc = at(k0 + r, j);
                for (int c = 0; c < r; ++c) acc = csub(acc, cmul(l11[r + c * kPanel], at(k0 + c, j)));
                at(k0 + r, j) = acc;
            }
        __syncthreads();
        if (j1 < ne) {
            const int tiles_c = (n - j1 + 15) / 16;
            for (int tj = wave; tj < tiles_c; tj += waves) {
                const int j0 = j1 + 16 * tj;
                const int rows_end = j0 < ne ? n : ne;
                for (int i0 = j1; i0 < rows_end; i0 += 16)
                    tile_update(F, n, i0, j0, k0, j1, lane, [&](int i, int j) { return i >= ne && j >= ne; });
            }
            __syncthreads();
        }
    }
    if (ne < n) {
        const int tiles = (n - ne + 15) / 16;
        for (int t = wave; t < tiles * tiles; t += waves) {
            const int tj = t / tiles, ti = t % tiles;
            tile_update(F, n, ne + 16 * ti, ne + 16 * tj, 0, ne, lane, [](int, int) { return false; });
        }
    }
    __syncthreads();
    if (tid == 0) *status = kFrontOk;
}

template <class T, bool kRoot>
__global__ void __launch_bounds__(kThreads)
    factor_deferred_kernel(const NodeDesc* __restrict__ nodes, const int32_t* __restrict__ level, PointsView pv, double tol) {
    const NodeDesc nd = nodes[level[blockIdx.x]];
    const int64_t p = blockIdx.y;
    T* F = (kRoot ? reinterpret_cast<T*>(pv.root_arena) + p * pv.root_stride
                  : reinterpret_cast<T*>(pv.arena) + p * pv.arena_stride) +
           nd.off;
    eliminate_deferred(F, nd, pv.piv + p * pv.piv_stride + nd.piv_off, pv.status + p * pv.status_stride + nd.slot, tol,
                       kRoot);
}

template <class T, int NMAX>
__global__ void __launch_bounds__(kThreads)
    factor_small_kernel(const NodeDesc* __restrict__ nodes, const int32_t* __restrict__ level, PointsView pv, double tol) {
    __shared__ T w[NMAX * NMAX];
    const NodeDesc nd = nodes[level[blockIdx.x]];
    const int64_t p = blockIdx.y;
    T* F = reinterpret_cast<T*>(pv.arena) + p * pv.arena_stride + nd.off;
    const int n2 = nd.n * nd.n;
    for (int i = threadIdx.x; i < n2; i += kThreads) w[i] = F[i];
    __syncthreads();
    eliminate(w, nd, pv.piv + p * pv.piv_stride + nd.piv_off, pv.status + p * pv.status_stride + nd.slot, tol, false);
    __syncthreads();
    for (int i = threadIdx.x; i < n2; i += kThreads) F[i] = w[i];
}

template <class T, bool kRoot>
__global__ void __launch_bounds__(kThreads)
    factor_global_kernel(const NodeDesc* __restrict__ nodes, const int32_t* __restrict__ level, PointsView pv, double tol) {
    const NodeDesc nd = nodes[level[blockIdx.x]];
    const int64_t p = blockIdx.y;
    T* F = (kRoot ? reinterpret_cast<T*>(pv.root_arena) + p * pv.root_stride
                  : reinterpret_cast<T*>(pv.arena) + p * pv.arena_stride) +
           nd.off;
    eliminate(F, nd, pv.piv + p * pv.piv_stride + nd.piv_off, pv.status + p * pv.status_stride + nd.slot, tol, kRoot);
}

__global__ void __launch_bounds__(kThreads) trace_kernel(const NodeDesc* __restrict__ rootp, PointsView pv) {
    __shared__ double sv[16];
    const NodeDesc nd = *rootp;
    const int64_t p = blockIdx.x;
    const double2* F = reinterpret_cast<const double2*>(pv.root_arena) + p * pv.root_stride + nd.off;
    const int n = nd.n, k0 = nd.ne, nl = nd.nl, nr = nd.nr;
    const double2* gl = pv.gamma + p * pv.sigma_stride;
    const double2* gr = gl + int64_t(nl) * nl;
    // The border block holds -G_LR; the sign cancels in A conj(B).
    auto g = [&](int i, int j) { return F[(k0 + i) + int64_t(k0 + j) * n]; };
    double acc = 0.0;
    for (int idx = threadIdx.x; idx < nl * nr; idx += kThreads) {
        int i = idx % nl, j = idx / nl;
        double2 a = cx<double2>(0, 0), b = cx<double2>(0, 0);
        for (int k = 0; k < nl; ++k) a = cadd(a, cmul(gl[i + int64_t(k) * nl], g(k, j)));
        for (int k = 0; k < nr; ++k) b = cadd(b, cmul(g(i, k), gr[k + int64_t(j) * nr]));
        acc += a.x * b.x + a.y * b.y;
    }
    double t = block_sum(acc, sv);
    if (threadIdx.x == 0) {
        int32_t st = pv.status[p * pv.status_stride + nd.slot];
        pv.t_out[p] = st == kFrontOk ? t : NAN;
    }
}

inline dim3 grid2(int32_t count, int32_t points) { return dim3(unsigned(count), unsigned(points)); }

}  // namespace

void launch_assemble_c128(const NodeDesc* nodes, const int32_t* level, int32_t count, const Entry* ent,
                          const int32_t* diag, const KidDesc* kids, const int32_t* maps, const PointsView& pv,
                          int32_t points, hipStream_t st) {
    if (count <= 0 || points <= 0) return;
    hipLaunchKernelGGL((assemble_kernel<double2, double2, false>), grid2(count, points), dim3(kThreads), 0, st, nodes,
                       level, ent, diag, kids, maps, pv);
}

void launch_assemble_c64(const NodeDesc* nodes, const int32_t* level, int32_t count, const Entry* ent,
                         const int32_t* diag, const KidDesc* kids, const int32_t* maps, const PointsView& pv,
                         int32_t points, hipStream_t st) {
    if (count <= 0 || points <= 0) return;
    hipLaunchKernelGGL((assemble_kernel<float2, float2, false>), grid2(count, points), dim3(kThreads), 0, st, nodes,
                       level, ent, diag, kids, maps, pv);
}

void launch_assemble_root(const NodeDesc* nodes, const int32_t* level, int32_t count, const Entry* ent,
                          const int32_t* diag, const KidDesc* kids, const int32_t* maps, const PointsView& pv,
                          int32_t points, int single, hipStream_t st) {
    if (count <= 0 || points <= 0) return;
    if (single)
        hipLaunchKernelGGL((assemble_kernel<double2, float2, true>), grid2(count, points), dim3(kThreads), 0, st, nodes,
                           level, ent, diag, kids, maps, pv);
    else
        hipLaunchKernelGGL((assemble_kernel<double2, double2, true>), grid2(count, points), dim3(kThreads), 0, st, nodes,
                           level, ent, diag, kids, maps, pv);
}

bool blocked_fronts() {
    const char* v = std::getenv("CAROLI_FRONT_KERNEL");
    return !(v && std::strcmp(v, "unblocked") == 0);
}

bool deferred_fronts() {
    const char* v = std::getenv("CAROLI_FRONT_KERNEL");
    return v && std::strcmp(v, "deferred") == 0;
}

void launch_factor_c128(const NodeDesc* nodes, const int32_t* level, int32_t count, const PointsView& pv,
                        int32_t points, double tol, bool small, bool root, hipStream_t st) {
    if (count <= 0 || points <= 0) return;
    const bool blk = blocked_fronts();
    if (deferred_fronts() && !small) {
        if (root)
            hipLaunchKernelGGL((factor_deferred_kernel<double2, true>), grid2(count, points), dim3(kThreads), 0, st,
                               nodes, level, pv, tol);
        else
            hipLaunchKernelGGL((factor_deferred_kernel<double2, false>), grid2(count, points), dim3(kThreads), 0, st,
                               nodes, level, pv, tol);
        return;
    }
    if (root && blk)
        hipLaunchKernelGGL((factor_blocked_kernel<double2, true>), grid2(count, points), dim3(kThreads), 0, st, nodes,
                           level, pv, tol);
    else if (root)
        hipLaunchKernelGGL((factor_global_kernel<double2, true>), grid2(count, points), dim3(kThreads), 0, st, nodes,
                           level, pv, tol);
    else if (small)
        hipLaunchKernelGGL((factor_small_kernel<double2, kLdsFrontC128>), grid2(count, points), dim3(kThreads), 0, st,
                           nodes, level, pv, tol);
    else if (blk)
        hipLaunchKernelGGL((factor_blocked_kernel<double2, false>), grid2(count, points), dim3(kThreads), 0, st, nodes,
                           level, pv, tol);
    else
        hipLaunchKernelGGL((factor_global_kernel<double2, false>), grid2(count, points), dim3(kThreads), 0, st, nodes,
                           level, pv, tol);
}

void launch_factor_c64(const NodeDesc* nodes, const int32_t* level, int32_t count, const PointsView& pv,
                       int32_t points, double tol, bool small, hipStream_t st) {
    if (count <= 0 || points <= 0) return;
    if (small)
        hipLaunchKernelGGL((factor_small_kernel<float2, kLdsFrontC64>), grid2(count, points), dim3(kThreads), 0, st,
                           nodes, level, pv, tol);
    else
        hipLaunchKernelGGL((factor_global_kernel<float2, false>), grid2(count, points), dim3(kThreads), 0, st, nodes,
                           level, pv, tol);
}

void launch_trace(const NodeDesc* root, const PointsView& pv, int32_t points, hipStream_t st) {
    if (points <= 0) return;
    hipLaunchKernelGGL(trace_kernel, dim3(unsigned(points)), dim3(kThreads), 0, st, root, pv);
}

}  // namespace caroli::gpu
