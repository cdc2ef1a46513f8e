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

#include "fronts.hpp"

namespace caroli::gpu {

namespace {

constexpr int kThreads = 256;

template <class T>
struct Real;
template <>
struct Real<double2> {
    using type = double;
};
template <>
struct Real<float2> {
    using type = float;
};

template <class T>
__device__ __forceinline__ T cx(double re, double im) {
    using R = typename Real<T>::type;
    T t;
    t.x = R(re);
    t.y = R(im);
    return t;
}
template <class T>
__device__ __forceinline__ T cmul(T a, T b) {
    T c;
    c.x = a.x * b.x - a.y * b.y;
    c.y = a.x * b.y + a.y * b.x;
    return c;
}
template <class T>
__device__ __forceinline__ T cadd(T a, T b) {
    a.x += b.x;
    a.y += b.y;
    return a;
}
template <class T>
__device__ __forceinline__ T csub(T a, T b) {
    a.x -= b.x;
    a.y -= b.y;
    return a;
}
template <class T>
__device__ __forceinline__ double cabs(T a) {
    return hypot(double(a.x), double(a.y));
}
template <class T>
__device__ __forceinline__ T crecip(T a) {
    // 1 / a with one division of the squared modulus, scaled against overflow.
    using R = typename Real<T>::type;
    R s = fmax(fabs(a.x), fabs(a.y));
    R ar = a.x / s, ai = a.y / s;
    R d = (ar * ar + ai * ai) * s;
    T r;
    r.x = ar / d;
    r.y = -ai / d;
    return r;
}
template <class TO, class TI>
__device__ __forceinline__ TO convert(TI a) {
    using R = typename Real<TO>::type;
    TO t;
    t.x = R(a.x);
    t.y = R(a.y);
    return t;
}

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
        for (int w = 1; w < kThreads / 64; ++w)
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
    for (int w = 1; w < kThreads / 64; ++w) r = fmax(r, sv[w]);
    __syncthreads();
    return r;
}

__device__ double block_sum(double v, double* sv) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) sv[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.0;
    for (int w = 0; w < kThreads / 64; ++w) r += sv[w];
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
    __shared__ double sv[kThreads / 64];
    __shared__ int si[kThreads / 64];
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
    __shared__ double sv[kThreads / 64];
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

void launch_factor_c128(const NodeDesc* nodes, const int32_t* level, int32_t count, const PointsView& pv,
                        int32_t points, double tol, bool small, bool root, hipStream_t st) {
    if (count <= 0 || points <= 0) return;
    if (root)
        hipLaunchKernelGGL((factor_global_kernel<double2, true>), grid2(count, points), dim3(kThreads), 0, st, nodes,
                           level, pv, tol);
    else if (small)
        hipLaunchKernelGGL((factor_small_kernel<double2, kLdsFrontC128>), grid2(count, points), dim3(kThreads), 0, st,
                           nodes, level, pv, tol);
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
