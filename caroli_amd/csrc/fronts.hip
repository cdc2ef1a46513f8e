// Front kernels of the multifrontal elimination on CDNA4 (gfx950).
//
// Every launch covers the fronts of one tree level (or one size class of them) times the points
// of the batch, so thousands of fronts fill the 256 CUs.
//  - front_lds_kernel: small fronts (n <= 64 in complex f64, <= 90 in f32) are assembled straight
//    into LDS (F = z - H on the separator plus the children's Schur complements) and eliminated
//    there, one wave per front: partial LU with row pivoting restricted to the separator rows and
//    the pivot test of crates/caroli-dense/src/field.rs:61 / nested.rs:1167-1223.
//  - large fronts: assemble_kernel writes F to HBM; border_kernel factors the separator block and
//    the border strips (F11 = P^T L11 U11, L21, U12) panel by panel; schur_kernel forms
//    S = F22 - L21 U12 on MFMA over 64 x 64 tiles, many workgroups per front.
//  - the root is a large front bordered by the lead slices (assemble_kernel<root>), whose Schur
//    block is -G_LR; trace_kernel evaluates Tr[Gamma_L G Gamma_R G^H] from it.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <mutex>
#include <set>
#include <stdexcept>

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

// Largest value and smallest index on ties over one wavefront, with a payload; every lane gets
// the result.
__device__ __forceinline__ void wave_argmax(double& v, int& idx, int& pay) {
    for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(v, off, 64);
        const int oi = __shfl_xor(idx, off, 64), op = __shfl_xor(pay, off, 64);
        if (ov > v || (ov == v && oi < idx)) {
            v = ov;
            idx = oi;
            pay = op;
        }
    }
}
__device__ __forceinline__ void wave_argmax(double& v, int& idx) {
    int pay = 0;
    wave_argmax(v, idx, pay);
}

// ---------------------------------------------------------------------------------------------
// Small fronts: assembly and elimination in LDS, one workgroup of NT lanes per (front, point).
//
// The front is assembled straight into LDS (no HBM copy of the assembled front), eliminated with
// one lane per row, and written back once: the whole factored front when the factors are kept
// for the solves (pv.keep), else only the Schur block the parent reads. Lane t owns physical row
// t of the LDS copy for the whole elimination; a row interchange only swaps the logical positions
// two lanes track, and the rows are written back in logical order, which leaves the factors of
// interchanging whole rows (LAPACK's convention, the one the solves use).
template <class T, int NT, bool kAssemble>
__global__ void __launch_bounds__(NT)
    front_lds_kernel(const NodeDesc* __restrict__ nodes, const int32_t* __restrict__ level,
                     const Entry* __restrict__ ent, const int32_t* __restrict__ diag,
                     const KidDesc* __restrict__ kids, const int32_t* __restrict__ maps, PointsView pv, double tol) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ double sv[NT / 64];
    __shared__ int32_t perm[kLdsMaxN];
    __shared__ double rv[2][NT / 64];
    __shared__ int32_t ri[2][NT / 64], rp[2][NT / 64];
    T* W = reinterpret_cast<T*>(lds_raw);
    const NodeDesc nd = nodes[level[blockIdx.x]];
    const int64_t p = blockIdx.y;
    T* F = reinterpret_cast<T*>(pv.arena) + p * pv.arena_stride + nd.off;
    int32_t* piv = pv.piv + p * pv.piv_stride + nd.piv_off;
    int32_t* status = pv.status + p * pv.status_stride + nd.slot;
    const int n = nd.n, ne = nd.ne, ne1 = nd.ne1, tid = threadIdx.x, lane = tid & 63;
    const int n2 = n * n;
    if (kAssemble) {
        const T zero = cx<T>(0.0, 0.0);
        for (int i = tid; i < n2; i += NT) W[i] = zero;
        __syncthreads();
        const double2* vals = pv.vals[p];
        for (int e = nd.ent_beg + tid; e < nd.ent_end; e += NT) {
            Entry en = ent[e];
            double2 v = vals[en.k];
            if (en.conj) v.y = -v.y;
            const int at = en.a + en.b * n;
            W[at] = csub(W[at], convert<T>(v));
        }
        __syncthreads();
        const T z = convert<T>(pv.z[p]);
        for (int d = nd.diag_beg + tid; d < nd.diag_end; d += NT) {
            const int q = diag[d];
            W[q + q * n] = cadd(W[q + q * n], z);
        }
        __syncthreads();
        const T* kidbase = reinterpret_cast<const T*>(pv.arena) + p * pv.arena_stride;
        for (int c = nd.kid_beg; c < nd.kid_end; ++c) {
            const KidDesc kd = kids[c];
            const T* S = kidbase + kd.src;
            const int32_t* map = maps + kd.map_beg;
            const int m = kd.size, m2 = m * m;
            for (int idx = tid; idx < m2; idx += NT) {
                const int a = idx % m, b = idx / m;
                const int at = map[a] + map[b] * n;
                W[at] = cadd(W[at], S[a + int64_t(b) * kd.ld]);
            }
            __syncthreads();
        }
    } else {
        for (int i = tid; i < n2; i += NT) W[i] = F[i];
        __syncthreads();
    }
    double m = 0.0;
    for (int idx = tid; idx < ne1 * ne1; idx += NT) m = fmax(m, cabs(W[(idx % ne1) + (idx / ne1) * n]));
    const double scale = block_max(m, sv);
    // Lane t owns physical row t and tracks its logical position lg.
    int lg = tid;
    for (int k = 0; k < ne; ++k) {
        const int lim = k < ne1 ? ne1 : ne;
        double big = -1.0;
        int at = kLdsMaxN, ph = 0;  // pivot: value, logical row, physical row
        if (tid < n && lg >= k && lg < lim) {
            big = cabs(W[tid + k * n]);
            at = lg;
            ph = tid;
        }
        wave_argmax(big, at, ph);
        if constexpr (NT > 64) {  // combine the waves (double-buffered by step parity)
            const int par = k & 1;
            if (lane == 0) {
                rv[par][tid >> 6] = big;
                ri[par][tid >> 6] = at;
                rp[par][tid >> 6] = ph;
            }
            __syncthreads();
            for (int w = 0; w < NT / 64; ++w) {
                const double ov = rv[par][w];
                const int oi = ri[par][w];
                if (ov > big || (ov == big && oi < at)) {
                    big = ov;
                    at = oi;
                    ph = rp[par][w];
                }
            }
        }
        const bool ok = k < ne1 ? (big > tol * scale) : (big > 0.0 && big < INFINITY);
        if (!ok) {
            if (tid == 0) *status = k < ne1 ? kFrontDelayed : kFrontSingular;
            return;
        }
        if (tid == 0) piv[k] = at;
        if (tid == ph)
            lg = k;
        else if (lg == k)
            lg = at;
        if (tid < n && lg > k) {
            const T l = cmul(W[tid + k * n], crecip(W[ph + k * n]));
            W[tid + k * n] = l;
            for (int j = k + 1; j < n; ++j) W[tid + j * n] = csub(W[tid + j * n], cmul(l, W[ph + j * n]));
        }
        __syncthreads();
    }
    if (tid == 0) *status = kFrontOk;
    if (tid < n) perm[lg] = tid;
    __syncthreads();
    if (pv.keep) {
        for (int idx = tid; idx < n2; idx += NT) {
            const int i = idx % n, j = idx / n;
            F[idx] = W[perm[i] + j * n];
        }
    } else {
        const int nb = n - ne, nb2 = nb * nb;
        for (int idx = tid; idx < nb2; idx += NT) {
            const int i = ne + idx % nb, j = ne + idx / nb;
            F[i + j * n] = W[perm[i] + j * n];
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Large fronts, step 1: the border. Partial LU of F11 (row pivoting within the separator rows,
// the pivot test of field.rs:61), L21 = F21 U11^-1 and U12 = L11^-1 P F12, in place, panel by
// panel (kPanel columns), without touching F22: the Schur update S = F22 - L21 U12 is left to
// schur_kernel, which spreads it over many workgroups. Per panel:
//   1. the panel's separator rows [k0, ne) are factored in LDS (when they fit, else in place);
//   2. its interchanges are applied to the other columns, the L21 rows are solved against the
//      panel's U11 block (one lane per row), and the panel goes back to the front;
//   3. the panel's U rows are solved with its unit-lower L11 block (one lane per column);
//   4. the rest of the border (rows or columns below ne) gets the rank-kPanel update on MFMA.
constexpr int kBorderThreads = 256;

template <class T>
__device__ __forceinline__ void mfma_tile_update(T* F, int n, int ne, int i0, int j0, int k0, int w, int lane) {
    // F[i, j] -= sum_c L[i, k0 + c] U[k0 + c, j] over c < w for the 16 x 16 tile at (i0, j0), as
    // C^T -= U^T L^T so that lanes read and write the front's columns contiguously. Entries in
    // F22 (i, j >= ne) are not stored.
    using M = Mfma<T>;
    typename M::acc cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
    const int li = lane % 16, lk = lane / 16;
#pragma unroll
    for (int kk = 0; kk < kPanel; kk += 4) {
        const int c = kk + lk;
        const int j = j0 + li, i = i0 + li;
        // Clamped loads and a select of the value (a select of addresses would go through flat).
        T a = F[(k0 + min(c, w - 1)) + int64_t(min(j, n - 1)) * n];
        T b = F[min(i, n - 1) + int64_t(k0 + min(c, w - 1)) * n];
        if (!(c < w && j < n)) a = cx<T>(0.0, 0.0);
        if (!(c < w && i < n)) b = cx<T>(0.0, 0.0);
        cr = M::mac(a.x, b.x, cr);
        cr = M::mac(-a.y, b.y, cr);
        ci = M::mac(a.x, b.y, ci);
        ci = M::mac(a.y, b.x, ci);
    }
    const int i = i0 + li;
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
        const int j = j0 + M::row(lane, jj);
        if (i < n && j < n && (i < ne || j < ne)) {
            T& d = F[i + int64_t(j) * n];
            d.x -= cr[jj];
            d.y -= ci[jj];
        }
    }
}

// The border factorisation of one front; kLds: the panel's separator rows are held in LDS (P).
template <class T, bool kRoot, bool kLds>
__device__ __forceinline__ void border_factor(T* F, T* P, int32_t* spiv, int n, int ne, int ne1, int32_t* piv,
                                              int32_t* status, double scale, double tol) {
    constexpr int NT = kBorderThreads;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    auto at = [&](int i, int j) -> T& { return F[i + int64_t(j) * n]; };
    for (int k0 = 0; k0 < ne; k0 += kPanel) {
        const int w = min(kPanel, ne - k0), R = ne - k0, k1 = k0 + w;
        // Panel element (row i in [k0, ne), panel column c).
        auto pp = [&](int i, int c) -> T& { return kLds ? P[(i - k0) + c * R] : at(i, k0 + c); };
        if (kLds)
            for (int e = tid; e < R * w; e += NT) P[e] = at(k0 + e % R, k0 + e / R);
        __syncthreads();
        // 1. Panel factorisation on the separator rows.
        for (int c = 0; c < w; ++c) {
            const int k = k0 + c, lim = k < ne1 ? ne1 : ne;
            double big = -1.0;
            int q = k;
            for (int i = k + lane; i < lim; i += 64) {
                const double v = cabs(pp(i, c));
                if (v > big) {
                    big = v;
                    q = i;
                }
            }
            wave_argmax(big, q);  // every wave finds the same pivot
            const bool ok = k < ne1 ? (big > tol * scale) : (big > 0.0 && big < INFINITY);
            if (!ok) {
                if (tid == 0) *status = (k < ne1 && !kRoot) ? kFrontDelayed : kFrontSingular;
                return;
            }
            if (tid == 0) {
                piv[k] = q;
                spiv[c] = q;
            }
            if (q != k && tid < w) {
                T x = pp(k, tid);
                pp(k, tid) = pp(q, tid);
                pp(q, tid) = x;
            }
            __syncthreads();
            const T inv = crecip(pp(k, c));
            for (int i = k + 1 + tid; i < ne; i += NT) {
                const T l = cmul(pp(i, c), inv);
                pp(i, c) = l;
                for (int c2 = c + 1; c2 < w; ++c2) pp(i, c2) = csub(pp(i, c2), cmul(l, pp(k, c2)));
            }
            __syncthreads();
        }
        // 2. Interchanges on the other columns (separator rows only), the panel back in place, and
        //    the L21 rows: l U11 = f with U11 the panel's upper triangle.
        for (int j = tid; j < n; j += NT) {
            if (j >= k0 && j < k1) continue;
            for (int c = 0; c < w; ++c) {
                const int q = spiv[c], k = k0 + c;
                if (q != k) {
                    T x = at(k, j);
                    at(k, j) = at(q, j);
                    at(q, j) = x;
                }
            }
        }
        if (kLds)
            for (int e = tid; e < R * w; e += NT) at(k0 + e % R, k0 + e / R) = P[e];
        for (int i = ne + tid; i < n; i += NT) {
            T l[kPanel];
#pragma unroll
            for (int c = 0; c < kPanel; ++c) {
                l[c] = at(i, k0 + min(c, w - 1));
                if (c >= w) l[c] = cx<T>(0.0, 0.0);
            }
#pragma unroll
            for (int c = 0; c < kPanel; ++c) {
                if (c < w) {
                    T acc = l[c];
#pragma unroll
                    for (int c2 = 0; c2 < c; ++c2) acc = csub(acc, cmul(l[c2], pp(k0 + c2, c)));
                    l[c] = cmul(acc, crecip(pp(k0 + c, c)));
                }
            }
#pragma unroll
            for (int c = 0; c < kPanel; ++c)
                if (c < w) at(i, k0 + c) = l[c];
        }
        __syncthreads();
        if (k1 >= n) break;
        // 3. U rows of the panel: unit-lower solve, one lane per column.
        for (int j = k1 + tid; j < n; j += NT) {
            T u[kPanel];
#pragma unroll
            for (int c = 0; c < kPanel; ++c) {
                u[c] = at(k0 + min(c, w - 1), j);
                if (c >= w) u[c] = cx<T>(0.0, 0.0);
            }
#pragma unroll
            for (int c = 1; c < kPanel; ++c) {
                if (c < w) {
                    T acc = u[c];
#pragma unroll
                    for (int c2 = 0; c2 < c; ++c2) acc = csub(acc, cmul(pp(k0 + c, c2), u[c2]));
                    u[c] = acc;
                }
            }
#pragma unroll
            for (int c = 1; c < kPanel; ++c)
                if (c < w) at(k0 + c, j) = u[c];
        }
        __syncthreads();
        // 4. Rank-w update of the border below and right of the panel, 16 x 16 tiles on MFMA; the
        //    tiles inside F22 are skipped.
        if (k1 < ne) {
            const int tiles = (n - k1 + 15) / 16;
            for (int t = wave; t < tiles * tiles; t += NT / 64) {
                const int i0 = k1 + 16 * (t % tiles), j0 = k1 + 16 * (t / tiles);
                if (i0 >= ne && j0 >= ne) continue;
                mfma_tile_update(F, n, ne, i0, j0, k0, w, lane);
            }
        }
        __syncthreads();
    }
    if (tid == 0) *status = kFrontOk;
}

template <class T, bool kRoot>
__global__ void __launch_bounds__(kBorderThreads)
    border_kernel(const NodeDesc* __restrict__ nodes, const int32_t* __restrict__ level, PointsView pv, double tol,
                  int lds_rows) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ double sv[kBorderThreads / 64];
    __shared__ int32_t spiv[kPanel];
    const NodeDesc nd = nodes[level[blockIdx.x]];
    const int64_t p = blockIdx.y;
    T* F = (kRoot ? reinterpret_cast<T*>(pv.root_arena) + p * pv.root_stride
                  : reinterpret_cast<T*>(pv.arena) + p * pv.arena_stride) +
           nd.off;
    int32_t* piv = pv.piv + p * pv.piv_stride + nd.piv_off;
    int32_t* status = pv.status + p * pv.status_stride + nd.slot;
    const int n = nd.n, ne = nd.ne, ne1 = nd.ne1;
    double m = 0.0;
    for (int idx = threadIdx.x; idx < ne1 * ne1; idx += kBorderThreads)
        m = fmax(m, cabs(F[(idx % ne1) + int64_t(idx / ne1) * n]));
    const double scale = block_max(m, sv);
    if (ne <= lds_rows)
        border_factor<T, kRoot, true>(F, reinterpret_cast<T*>(lds_raw), spiv, n, ne, ne1, piv, status, scale, tol);
    else
        border_factor<T, kRoot, false>(F, nullptr, spiv, n, ne, ne1, piv, status, scale, tol);
}

// ---------------------------------------------------------------------------------------------
// Large fronts, step 2: S = F22 - L21 U12 (K = ne) over 64 x 64 tiles of the Schur block, one
// workgroup per (tile, point), on MFMA. Each wave owns a 32 x 32 quarter (2 x 2 blocks of 16 x 16,
// complex = 4 real MFMAs per block and K-step of 4); L21 and U12 are staged through LDS in
// chunks of kChunk columns / rows. Fronts whose border failed are skipped.
constexpr int kChunk = 16, kTileLd = kTile + 1;

template <class T, bool kRoot>
__global__ void __launch_bounds__(256)
    schur_kernel(const NodeDesc* __restrict__ nodes, const Tile* __restrict__ tiles, PointsView pv) {
    using M = Mfma<T>;
    __shared__ T As[kChunk * kTileLd];  // As[k * kTileLd + i] = L21[i0 + i, kc + k]
    __shared__ T Bs[kChunk * kTileLd];  // Bs[k * kTileLd + j] = U12[kc + k, j0 + j]
    const Tile tl = tiles[blockIdx.x];
    const NodeDesc nd = nodes[tl.node];
    const int64_t p = blockIdx.y;
    if (pv.status[p * pv.status_stride + nd.slot] != kFrontOk) return;
    T* F = (kRoot ? reinterpret_cast<T*>(pv.root_arena) + p * pv.root_stride
                  : reinterpret_cast<T*>(pv.arena) + p * pv.arena_stride) +
           nd.off;
    const int n = nd.n, ne = nd.ne, nb = n - ne, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wi = (wave & 1) * 32, wj = (wave >> 1) * 32;
    const T zero = cx<T>(0.0, 0.0);
    typename M::acc cr[2][2], ci[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            cr[x][y] = typename M::acc{0, 0, 0, 0};
            ci[x][y] = typename M::acc{0, 0, 0, 0};
        }
    for (int kc = 0; kc < ne; kc += kChunk) {
        // Clamped loads and a select of the value (a select of addresses would go through flat).
        for (int e = tid; e < kTile * kChunk; e += 256) {
            const int r = e % kTile, c = e / kTile, i = tl.i0 + r, k = kc + c;
            T v = F[(ne + min(i, nb - 1)) + int64_t(min(k, ne - 1)) * n];
            if (!(i < nb && k < ne)) v = zero;
            As[c * kTileLd + r] = v;
        }
        for (int e = tid; e < kTile * kChunk; e += 256) {
            const int c = e % kChunk, r = e / kChunk, j = tl.j0 + r, k = kc + c;
            T v = F[min(k, ne - 1) + int64_t(ne + min(j, nb - 1)) * n];
            if (!(j < nb && k < ne)) v = zero;
            Bs[c * kTileLd + r] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kChunk; kk += 4) {
            const int k = kk + lane / 16;
            T a[2], b[2];
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                a[s] = Bs[k * kTileLd + wj + 16 * s + lane % 16];  // U12^T operand
                b[s] = As[k * kTileLd + wi + 16 * s + lane % 16];  // L21^T operand
            }
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y) {
                    cr[x][y] = M::mac(a[x].x, b[y].x, cr[x][y]);
                    cr[x][y] = M::mac(-a[x].y, b[y].y, cr[x][y]);
                    ci[x][y] = M::mac(a[x].x, b[y].y, ci[x][y]);
                    ci[x][y] = M::mac(a[x].y, b[y].x, ci[x][y]);
                }
        }
        __syncthreads();
    }
    // Block (x, y) of the wave: lane l, element jj holds (L21 U12)[i, j] with
    // i = wi + 16 y + l % 16 and j = wj + 16 x + row(l, jj).
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            const int i = tl.i0 + wi + 16 * y + lane % 16;
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                const int j = tl.j0 + wj + 16 * x + M::row(lane, jj);
                if (i < nb && j < nb) {
                    T& d = F[(ne + i) + int64_t(ne + j) * n];
                    d.x -= cr[x][y][jj];
                    d.y -= ci[x][y][jj];
                }
            }
        }
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

// Dynamic LDS above the 64 KB default (gfx950 has 160 KB per workgroup), once per kernel.
void allow_lds(const void* kernel) {
    static std::mutex mu;
    static std::set<const void*> done;
    std::lock_guard<std::mutex> g(mu);
    if (done.insert(kernel).second)
        (void)hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, 96 * 1024);
}

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

bool unblocked_fronts() {
    const char* v = std::getenv("CAROLI_FRONT_KERNEL");
    return v && std::strcmp(v, "unblocked") == 0;
}

void launch_front_lds(const NodeDesc* nodes, const int32_t* level, int32_t count, int32_t nmax, const Entry* ent,
                      const int32_t* diag, const KidDesc* kids, const int32_t* maps, const PointsView& pv,
                      int32_t points, double tol, bool single, bool assemble, hipStream_t st) {
    if (count <= 0 || points <= 0) return;
    if (nmax > kLdsMaxN) throw std::invalid_argument("front too large for the LDS kernel");
    const size_t lds = size_t(nmax) * size_t(nmax) * (single ? sizeof(float2) : sizeof(double2));
    const bool wide = nmax > 64;
    auto go = [&](auto kernel, int nt) {
        allow_lds(reinterpret_cast<const void*>(kernel));
        hipLaunchKernelGGL(kernel, grid2(count, points), dim3(nt), lds, st, nodes, level, ent, diag, kids, maps, pv, tol);
    };
    if (single) {
        if (assemble)
            wide ? go(front_lds_kernel<float2, 128, true>, 128) : go(front_lds_kernel<float2, 64, true>, 64);
        else
            wide ? go(front_lds_kernel<float2, 128, false>, 128) : go(front_lds_kernel<float2, 64, false>, 64);
    } else {
        if (assemble)
            wide ? go(front_lds_kernel<double2, 128, true>, 128) : go(front_lds_kernel<double2, 64, true>, 64);
        else
            wide ? go(front_lds_kernel<double2, 128, false>, 128) : go(front_lds_kernel<double2, 64, false>, 64);
    }
}

int32_t border_lds_rows(int32_t max_ne, bool single) {
    const int32_t cap = kBorderLdsBytes / (kPanel * int32_t(single ? sizeof(float2) : sizeof(double2)));
    return max_ne <= cap ? max_ne : 0;
}

void launch_border(const NodeDesc* nodes, const int32_t* level, int32_t count, int32_t max_ne, const PointsView& pv,
                   int32_t points, double tol, bool single, bool root, hipStream_t st) {
    if (count <= 0 || points <= 0) return;
    const int32_t rows = border_lds_rows(max_ne, single);
    const size_t lds = size_t(rows) * kPanel * (single ? sizeof(float2) : sizeof(double2));
    const dim3 g = grid2(count, points), b(kBorderThreads);
    auto go = [&](auto kernel) {
        allow_lds(reinterpret_cast<const void*>(kernel));
        hipLaunchKernelGGL(kernel, g, b, lds, st, nodes, level, pv, tol, rows);
    };
    if (single)
        go(border_kernel<float2, false>);
    else if (root)
        go(border_kernel<double2, true>);
    else
        go(border_kernel<double2, false>);
}

void launch_schur(const NodeDesc* nodes, const Tile* tiles, int32_t count, const PointsView& pv, int32_t points,
                  bool single, bool root, hipStream_t st) {
    if (count <= 0 || points <= 0) return;
    const dim3 g = grid2(count, points), b(256);
    if (single)
        hipLaunchKernelGGL((schur_kernel<float2, false>), g, b, 0, st, nodes, tiles, pv);
    else if (root)
        hipLaunchKernelGGL((schur_kernel<double2, true>), g, b, 0, st, nodes, tiles, pv);
    else
        hipLaunchKernelGGL((schur_kernel<double2, false>), g, b, 0, st, nodes, tiles, pv);
}

void launch_factor_unblocked(const NodeDesc* nodes, const int32_t* level, int32_t count, const PointsView& pv,
                             int32_t points, double tol, bool single, bool root, hipStream_t st) {
    if (count <= 0 || points <= 0) return;
    if (single)
        hipLaunchKernelGGL((factor_global_kernel<float2, false>), grid2(count, points), dim3(kThreads), 0, st, nodes,
                           level, pv, tol);
    else if (root)
        hipLaunchKernelGGL((factor_global_kernel<double2, true>), grid2(count, points), dim3(kThreads), 0, st, nodes,
                           level, pv, tol);
    else
        hipLaunchKernelGGL((factor_global_kernel<double2, false>), grid2(count, points), dim3(kThreads), 0, st, nodes,
                           level, pv, tol);
}

void launch_trace(const NodeDesc* root, const PointsView& pv, int32_t points, hipStream_t st) {
    if (points <= 0) return;
    hipLaunchKernelGGL(trace_kernel, dim3(unsigned(points)), dim3(kThreads), 0, st, root, pv);
}

}  // namespace caroli::gpu
