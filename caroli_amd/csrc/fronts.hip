// Front kernels of the multifrontal elimination on CDNA4 (gfx950).
//
// Every launch covers the fronts of one tree level (or one size class of them) times the points
// of the batch, so thousands of fronts fill the 256 CUs.
//  - small fronts (n <= 64 in complex f64, <= 90 in f32): front_border_kernel assembles the border
//    (F = z - H on the separator plus the children's border entries) into LDS, eliminates it there
//    with row pivoting restricted to the separator rows and the pivot test of
//    crates/caroli-dense/src/field.rs:61 / nested.rs:1167-1223, and forms S = F22 - L21 U12 with
//    F22 gathered from the children; front_lds_kernel and front_reg_kernel are the whole-front
//    variants the tuning can select.
//  - large fronts: gather_kernel assembles the border strips in HBM; f11_kernel + trsm_kernel,
//    border_kernel or the blocked kernels (by separator size and the tuning) factor F11 and form
//    L21, U12; schur_kernel forms S = F22 - L21 U12 on MFMA over 64 x 64 tiles, many workgroups per
//    front, with F22 gathered from the children.
//  - the root is a large front bordered by the lead slices (root_gather_kernel, then
//    assemble_kernel<root> for -H and z), whose Schur block is -G_LR; the trace kernels evaluate
//    Tr[Gamma_L G Gamma_R G^H] from it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <set>
#include <stdexcept>
#include <type_traits>

#include "cx.hpp"
#include "fronts.hpp"

namespace caroli::gpu {

namespace {

constexpr int kThreads = 256;


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

// Lane exchanges inside rows of 16 lanes by DPP, without a trip through LDS: quad_perm [1,0,3,2]
// and [2,3,0,1] swap neighbours and pairs, row_half_mirror reflects groups of 8 lanes and
// row_mirror groups of 16.
constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppHalfMirror = 0x141, kDppMirror = 0x140;
template <int kCtrl>
__device__ __forceinline__ int dpp(int x) {
    return __builtin_amdgcn_update_dpp(x, x, kCtrl, 0xF, 0xF, false);
}
template <int kCtrl>
__device__ __forceinline__ double dpp(double x) {
    return __hiloint2double(dpp<kCtrl>(__double2hiint(x)), dpp<kCtrl>(__double2loint(x)));
}

// Largest value and smallest key on ties over one wavefront (all lanes active); every lane gets
// the result. Four rounds by DPP inside rows of 16 lanes, then two across rows; the partners of
// each round see the same two pairs and keep the same one.
__device__ __forceinline__ void wave_argmax_key(double& v, int& key) {
    auto take = [&](double ov, int ok) {
        if (ov > v || (ov == v && ok < key)) {
            v = ov;
            key = ok;
        }
    };
    take(dpp<kDppXor1>(v), dpp<kDppXor1>(key));
    take(dpp<kDppXor2>(v), dpp<kDppXor2>(key));
    take(dpp<kDppHalfMirror>(v), dpp<kDppHalfMirror>(key));
    take(dpp<kDppMirror>(v), dpp<kDppMirror>(key));
    take(__shfl_xor(v, 16, 64), __shfl_xor(key, 16, 64));
    take(__shfl_xor(v, 32, 64), __shfl_xor(key, 32, 64));
}
// With a payload: idx (below 2^15) and pay (below 2^16) travel as one key, idx in the high half,
// so that ties still go to the smallest idx.
__device__ __forceinline__ void wave_argmax(double& v, int& idx, int& pay) {
    int key = (idx << 16) | pay;
    wave_argmax_key(v, key);
    idx = key >> 16;
    pay = key & 0xFFFF;
}

// Block-wide max of (value, -index): the largest value, the smallest index on ties, in every
// lane. sv / si hold one entry per wave; the caller separates consecutive calls by a barrier.
__device__ void block_argmax(double& v, int& idx, double* sv, int* si) {
    wave_argmax_key(v, idx);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        sv[wave] = v;
        si[wave] = idx;
    }
    __syncthreads();
    for (int w = 0; w < int(blockDim.x) / 64; ++w) {
        const double ov = sv[w];
        const int oi = si[w];
        if (ov > v || (ov == v && oi < idx)) {
            v = ov;
            idx = oi;
        }
    }
}

// The pivot test of field.rs:61 on the squared modulus big of the pivot: |p| > tol scale on the
// tested columns (the first ne1), else p nonzero and finite.
__device__ __forceinline__ bool pivot_ok(double big, bool tested, double tol, double scale) {
    const double ts = tol * scale;
    return tested ? big > ts * ts : (big > 0.0 && big < INFINITY);
}

// Assembly loops issue their global loads in batches of kBatch per lane, so that a lane has
// kBatch loads in flight instead of one (the loops are latency-bound otherwise). Indexes past the
// end are clamped for the loads; only the updates are guarded.
constexpr int kBatch = 8;
// Schur entries a lane of the small-front kernels forms at once (their children's loads in flight
// together).
constexpr int kSBatch = 2;

// dst(idx, src(idx)) for idx < count over the NT lanes, kBatch loads of a lane in flight together
// (indexes clamped for the loads, the stores guarded): the copies of fronts and triangles into
// LDS, which otherwise wait for each load before the next.
template <int NT, class T, class Src, class Dst>
__device__ __forceinline__ void copy_batched(int count, Src src, Dst dst) {
    for (int base = int(threadIdx.x); base < count; base += NT * kBatch) {
        T v[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) v[u] = src(min(base + u * NT, count - 1));
#pragma unroll
        for (int u = 0; u < kBatch; ++u)
            if (base + u * NT < count) dst(base + u * NT, v[u]);
    }
}

// Workgroups of a launch go round-robin to the 8 XCDs (each with its own L2): linear id l runs
// on XCD l % 8. Tile kernels renumber their (tile, point) pairs so that each XCD receives one
// contiguous range, and neighbouring tiles of one front and point share an L2.
constexpr int kXcds = 8;
struct TilePoint {
    int tile;
    int64_t point;
};
__device__ __forceinline__ TilePoint xcd_tile() {
    const uint64_t T = gridDim.x, total = T * gridDim.y;
    const uint64_t lin = blockIdx.x + uint64_t(blockIdx.y) * T;
    const uint64_t q = total / kXcds, r = total % kXcds, x = lin % kXcds, idx = lin / kXcds;
    const uint64_t logical = (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + idx;
    return {int(logical % T), int64_t(logical / T)};
}

// The H entries of a front: add(at, v) with at = a + b n and v = -H (conjugated as listed).
template <int NT, class Add>
__device__ __forceinline__ void add_entries(const Entry* ent, int beg, int end, const double2* vals, int n, Add add) {
    for (int base = beg + int(threadIdx.x); base < end; base += NT * kBatch) {
        Entry en[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) en[u] = ent[min(base + u * NT, end - 1)];
        double2 v[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) v[u] = vals[en[u].k];
#pragma unroll
        for (int u = 0; u < kBatch; ++u)
            if (base + u * NT < end) {
                double2 x = v[u];
                x.x = -x.x;
                x.y = en[u].conj ? x.y : -x.y;
                add(en[u].a + en[u].b * n, x);
            }
    }
}

// A child's Schur complement S (m x m, leading dimension ld) added at map[a] + map[b] n:
// add(at, value).
template <int NT, class TK, class Add>
__device__ __forceinline__ void extend_add(const TK* S, int ld, const int32_t* map, int m, int n, Add add) {
    const int m2 = m * m;
    for (int base = int(threadIdx.x); base < m2; base += NT * kBatch) {
        TK v[kBatch];
        int at[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int idx = min(base + u * NT, m2 - 1), a = idx % m, b = idx / m;
            v[u] = S[a + int64_t(b) * ld];
            at[u] = map[a] + map[b] * n;
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u)
            if (base + u * NT < m2) add(at[u], v[u]);
    }
}

// extend_add for a front in global memory: the front's old values are loaded for the whole batch
// before any is stored (the positions of one child are distinct).
template <int NT, class TF, class TK>
__device__ __forceinline__ void extend_add_global(TF* F, const TK* S, int ld, const int32_t* map, int m, int n) {
    const int m2 = m * m;
    for (int base = int(threadIdx.x); base < m2; base += NT * kBatch) {
        TK v[kBatch];
        int64_t at[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int idx = min(base + u * NT, m2 - 1), a = idx % m, b = idx / m;
            v[u] = S[a + int64_t(b) * ld];
            at[u] = map[a] + int64_t(map[b]) * n;
        }
        TF f[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) f[u] = F[at[u]];
#pragma unroll
        for (int u = 0; u < kBatch; ++u)
            if (base + u * NT < m2) F[at[u]] = cadd(f[u], convert<TF>(v[u]));
    }
}

// Assembly of one front (see the file comment). TF: type of the front, TK: type of the children.
// kFill = false: the front was written whole by root_gather_kernel (children, -Sigma, unit
// borders), and only -H and z are added here.
template <class TF, class TK, bool kRoot, bool kFill = true>
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
    if (kFill) {
        for (int64_t i = tid; i < n2; i += kThreads) F[i] = zero;
        __syncthreads();
    }
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
    if (kRoot && kFill) {
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
    if (kFill)
        for (int c = nd.kid_beg; c < nd.kid_end; ++c) {
            const KidDesc kd = kids[c];
            extend_add_global<kThreads>(F, kidbase + kd.src, kd.ld, maps + kd.map_beg, kd.size, n);
            __syncthreads();
        }
}

// The root front written whole, one workgroup per tile (Tile::pad: rows << 16 | cols) and point,
// every entry once: the children's Schur complements through their inverse maps, minus Sigma_L on
// the left slice block and Sigma_R on the right one, and the unit borders (rows r0 + t of columns
// k0 + t, rows k0 + t of columns b0 + t), which no child reaches (children map into the first
// nd + k positions). -H and z follow in assemble_kernel<kFill = false>. A root of 4,800 variables
// (cubic L = 40) spans 5,625 tiles instead of one workgroup per point.
template <class TK>
__global__ void __launch_bounds__(256)
    root_gather_kernel(const NodeDesc* __restrict__ nodes, const Tile* __restrict__ tiles,
                       const KidDesc* __restrict__ kids, const int32_t* __restrict__ maps, PointsView pv) {
    const TilePoint tp = xcd_tile();
    const Tile tl = tiles[tp.tile];
    const NodeDesc nd = nodes[tl.node];
    const int64_t p = tp.point;
    const TK* kidbase = reinterpret_cast<const TK*>(pv.arena) + p * pv.arena_stride;
    double2* F = reinterpret_cast<double2*>(pv.root_arena) + p * pv.root_stride + nd.off;
    const double2* sg = pv.sigma + p * pv.sigma_stride;
    const int n = nd.n, nl = nd.nl, nr = nd.nr, b0 = nd.nd, r0 = nd.nd + nd.right_at, k0 = nd.ne;
    const int rows = min(tl.pad >> 16, n - tl.i0), cols = min(tl.pad & 0xffff, n - tl.j0), cnt = rows * cols;
    for (int base = int(threadIdx.x); base < cnt; base += 256 * kBatch) {
        int ii[kBatch], jj[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int idx = min(base + u * 256, cnt - 1);
            ii[u] = tl.i0 + idx % rows;
            jj[u] = tl.j0 + idx / rows;
        }
        double2 acc[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) acc[u] = make_double2(0.0, 0.0);
        for (int c = nd.kid_beg; c < nd.kid_end; ++c) {  // per child: all map loads, then all values
            const KidDesc kd = kids[c];
            const int32_t* inv = maps + kd.inv_beg;
            const TK* S = kidbase + kd.src;
            int a[kBatch], b[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                a[u] = inv[ii[u]];
                b[u] = inv[jj[u]];
            }
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                double2 v = convert<double2>(S[max(a[u], 0) + int64_t(max(b[u], 0)) * kd.ld]);
                if (a[u] < 0 || b[u] < 0) v = make_double2(0.0, 0.0);
                acc[u] = cadd(acc[u], v);
            }
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int i = ii[u], j = jj[u];
            double2 f = acc[u];
            if (unsigned(i - b0) < unsigned(nl) && unsigned(j - b0) < unsigned(nl))
                f = csub(f, sg[(i - b0) + int64_t(j - b0) * nl]);
            if (unsigned(i - r0) < unsigned(nr) && unsigned(j - r0) < unsigned(nr))
                f = csub(f, sg[int64_t(nl) * nl + (i - r0) + int64_t(j - r0) * nr]);
            if (unsigned(j - k0) < unsigned(nr) && i == r0 + (j - k0)) f = make_double2(1.0, 0.0);
            if (unsigned(i - k0) < unsigned(nl) && j == b0 + (i - k0)) f = make_double2(1.0, 0.0);
            if (base + u * 256 < cnt) F[i + int64_t(j) * n] = f;
        }
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
            double v = cabs2(W[i + int64_t(k) * n]);
            if (v > big) {
                big = v;
                at = i;
            }
        }
        block_argmax(big, at, sv, si);
        const bool ok = pivot_ok(big, k < ne1, tol, scale);
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



// Readlane of a double, a float or a complex value (w uniform over the wave).
__device__ __forceinline__ double readlane(double x, int w) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), w),
                            __builtin_amdgcn_readlane(__double2loint(x), w));
}
__device__ __forceinline__ float readlane(float x, int w) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), w));
}
template <class T>
__device__ __forceinline__ T readlane(T x, int w) {
    T r;
    r.x = readlane(x.x, w);
    r.y = readlane(x.y, w);
    return r;
}

// The largest v over the wave and, among the lanes holding it, the smallest key, with that lane's
// payload and value; every lane receives them (w uniform). The values alone are reduced (DPP
// inside rows of 16 lanes, then across rows); a ballot finds the lane, a second reduction of the
// keys runs only on an exact tie.
template <class T>
__device__ __forceinline__ void wave_pick(double v, int key, int pay, T val, double& big, int& kout, int& pout,
                                          T& vout) {
    double m = v;
    m = fmax(m, dpp<kDppXor1>(m));
    m = fmax(m, dpp<kDppXor2>(m));
    m = fmax(m, dpp<kDppHalfMirror>(m));
    m = fmax(m, dpp<kDppMirror>(m));
    m = fmax(m, __shfl_xor(m, 16, 64));
    m = fmax(m, __shfl_xor(m, 32, 64));
    const bool top = v == m;
    unsigned long long mask = __ballot(top);
    if (__popcll(mask) > 1) {
        double dummy = 0.0;
        int k = top ? key : INT_MAX;
        wave_argmax_key(dummy, k);  // equal values: the smallest key
        mask = __ballot(top && key == k);
    }
    const int w = __ffsll(mask) - 1;
    big = m;
    kout = __builtin_amdgcn_readlane(key, w);
    pout = __builtin_amdgcn_readlane(pay, w);
    vout = readlane(val, w);
}

// Pivot search of the LDS eliminations at step k, by each wave on its own (no LDS exchange, no
// barrier): lane t takes the separator rows t, t + 64, ... (physical) whose logical position
// pos[t] lies in [k, lim); the largest squared modulus of colk[t] wins, the smallest logical
// position on exact ties. Every lane receives the winner's squared modulus (big; -1 when no row
// is left), logical position (at), physical row (ph) and the reciprocal of its entry (inv; each
// candidate forms its own before the reduction, off the chain of dependent steps). The values
// alone are reduced (DPP inside rows of 16 lanes, then across rows); a ballot finds the lane.
template <class T>
__device__ __forceinline__ void lds_pivot(const T* colk, const int32_t* pos, int ne, int k, int lim, double& big,
                                          int& at, int& ph, T& inv) {
    const int lane = threadIdx.x & 63;
    double v = -1.0;
    int lv = kLdsMaxN, pv = 0;
    T c = cx<T>(1.0, 0.0);
    for (int t = lane; t < ne; t += 64) {
        const int lp = pos[t];
        const T x = colk[t];
        const double a = cabs2(x);
        if (lp >= k && lp < lim && (a > v || (a == v && lp < lv))) {
            v = a;
            lv = lp;
            pv = t;
            c = x;
        }
    }
    wave_pick(v, lv, pv, crecip(c), big, at, ph, inv);
}

// Partial LU in LDS of the n rows of W (leading dimension ld; columns [0, ncol) updated) over its
// first ne columns, row pivoting restricted to the separator rows (the first ne) and the pivot
// test of field.rs:61 on columns [0, ne1) (later columns only for zero). G lanes share a row:
// lane t owns physical row t / G and the columns j = t (mod G); lg is the row's logical position
// on entry (its physical row) and on return. Interchanges only swap logical positions; LAPACK's pivots go to piv. Returns false (and
// sets the status) when a pivot fails; the whole workgroup returns alike.
// One barrier a step: every wave searches the pivot itself (lds_pivot) over the logical positions
// of the previous step (pos, double-buffered by step parity), and a row's multiplier goes to
// column k one step late, so that no wave's search of column k sees it change.
template <class T, int NT, int G>
__device__ __forceinline__ bool lds_lu(T* W, int ld, int n, int ncol, int ne, int ne1, double tol, int32_t* piv,
                                       int32_t* status, int& lg) {
    __shared__ double sv[NT / 64];
    __shared__ int32_t pos[2][kLdsMaxN];
    const int tid = threadIdx.x, g = tid % G, r = tid / G;
    double m = 0.0;
    for (int idx = tid; idx < ne1 * ne1; idx += NT) m = fmax(m, cabs(W[(idx % ne1) + (idx / ne1) * ld]));
    for (int t = tid; t < ne; t += NT) pos[0][t] = t;
    const double scale = block_max(m, sv);  // its barriers publish pos[0] too
    T lk = cx<T>(0.0, 0.0);                  // this row's multiplier of the previous step (lane g = 0)
    bool has = false;
    for (int k = 0; k < ne; ++k) {
        const int lim = k < ne1 ? ne1 : ne, par = k & 1;
        if (has) W[r + (k - 1) * ld] = lk;
        has = false;
        const T own = W[min(r, n - 1) + k * ld];
        double big;
        int at, ph;
        T inv;
        lds_pivot(W + k * ld, pos[par], ne, k, lim, big, at, ph, inv);
        if (!pivot_ok(big, k < ne1, tol, scale)) {
            if (tid == 0) *status = k < ne1 ? kFrontDelayed : kFrontSingular;
            return false;
        }
        if (tid == 0) piv[k] = at;
        if (r == ph)
            lg = k;
        else if (lg == k)
            lg = at;
        if (g == 0 && r < ne) pos[par ^ 1][r] = lg;
        if (r < n && lg > k) {
            const T l = cmul(own, inv);
            const int j0 = k + 1 + ((g - (k + 1)) % G + G) % G;
            for (int j = j0; j < ncol; j += G) W[r + j * ld] = csub(W[r + j * ld], cmul(l, W[ph + j * ld]));
            if (g == 0) {
                lk = l;
                has = true;
            }
        }
        __syncthreads();
    }
    if (has) W[r + (ne - 1) * ld] = lk;
    __syncthreads();
    if (tid == 0) *status = kFrontOk;
    return true;
}

// ---------------------------------------------------------------------------------------------
// Small fronts: assembly and elimination in LDS, one workgroup of NT lanes per (front, point).
//
// The front is assembled straight into LDS (no HBM copy of the assembled front; leading
// dimension n + 1 so that lanes on neighbouring columns hit different banks), eliminated there,
// and written back once: the whole factored front when the factors are kept for the solves
// (pv.keep), else only the Schur block the parent reads. G lanes share a row: lane t owns
// physical row t / G and the columns j = t (mod G) of it for the whole elimination (the G lanes
// of a row sit in one wavefront). A row interchange only swaps the logical positions two rows
// track, and the rows are written back in logical order, which leaves the factors of
// interchanging whole rows (LAPACK's convention, the one the solves use).
template <class T, int NT, int G, bool kAssemble>
__global__ void __launch_bounds__(NT)
    front_lds_kernel(const NodeDesc* __restrict__ nodes, const int32_t* __restrict__ level,
                     const Entry* __restrict__ ent, const int32_t* __restrict__ diag,
                     const KidDesc* __restrict__ kids, const int32_t* __restrict__ maps, PointsView pv, double tol) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    T* W = reinterpret_cast<T*>(lds_raw);
    const NodeDesc nd = nodes[level[blockIdx.x]];
    const int64_t p = blockIdx.y;
    T* F = reinterpret_cast<T*>(pv.arena) + p * pv.arena_stride + nd.off;
    int32_t* piv = pv.piv + p * pv.piv_stride + nd.piv_off;
    int32_t* status = pv.status + p * pv.status_stride + nd.slot;
    const int n = nd.n, ne = nd.ne, ne1 = nd.ne1, tid = threadIdx.x, ld = n + 1;
    const int nw = n * ld;
    if (kAssemble) {
        const T zero = cx<T>(0.0, 0.0);
        for (int i = tid; i < nw; i += NT) W[i] = zero;
        __syncthreads();
        add_entries<NT>(ent, nd.ent_beg, nd.ent_end, pv.vals[p], ld,
                        [&](int at, double2 v) { W[at] = cadd(W[at], convert<T>(v)); });
        __syncthreads();
        const T z = convert<T>(pv.z[p]);
        for (int d = nd.diag_beg + tid; d < nd.diag_end; d += NT) {
            const int q = diag[d];
            W[q + q * ld] = cadd(W[q + q * ld], z);
        }
        __syncthreads();
        const T* kidbase = reinterpret_cast<const T*>(pv.arena) + p * pv.arena_stride;
        for (int c = nd.kid_beg; c < nd.kid_end; ++c) {
            const KidDesc kd = kids[c];
            extend_add<NT>(kidbase + kd.src, kd.ld, maps + kd.map_beg, kd.size, ld,
                           [&](int at, T v) { W[at] = cadd(W[at], v); });
            __syncthreads();
        }
    } else {
        copy_batched<NT, T>(n * n, [&](int idx) { return F[idx]; },
                            [&](int idx, T v) { W[(idx % n) + (idx / n) * ld] = v; });
        __syncthreads();
    }
    const int g = tid % G, r = tid / G;  // this lane's column class and physical row
    int lg = r;                           // the row's logical position
    if (!lds_lu<T, NT, G>(W, ld, n, n, ne, ne1, tol, piv, status, lg)) return;
    // Row r goes to logical row lg: lanes of one wave cover G columns of 64 / G rows.
    if (r < n) {
        if (pv.keep) {
            for (int j = g; j < n; j += G) F[lg + j * n] = W[r + j * ld];
        } else if (lg >= ne) {
            for (int j = ne + ((g - ne) % G + G) % G; j < n; j += G) F[lg + j * n] = W[r + j * ld];
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Small fronts, border-only LDS. F22 (boundary rows and columns) gets nothing from H or z and, in
// the elimination, only the rank-ne update L21 U12, so LDS holds the border alone: Wl = all n rows
// of the separator columns (leading dimension n + 1) and Wr = the separator rows of the boundary
// columns (leading dimension ne + 1), about 2 ne n words instead of n^2 (a front of 64 with a
// separator of 14: 27 KB instead of 66 KB, so two and a half times the workgroups a CU holds).
//  1. assembly of the border: -H on the separator, z on its diagonal, the children's entries that
//     land in it (their F22 entries are not read here);
//  2. the partial LU over the separator columns with pivots among the separator rows, as lds_lu:
//     G lanes per physical row, logical positions swapped only, the pivot test of field.rs:61;
//     every row updates its separator columns only, and the separator rows' boundary columns
//     (U12) follow in one triangular solve, a column per lane;
//  3. S = F22 - L21 U12 entry by entry: F22 gathered from the children through their inverse maps
//     (or read from the given front), the product from LDS, S written once.
// The factors written back (when kept) are those of front_lds_kernel.
template <class T, int NT, int G, bool kAssemble>
__global__ void __launch_bounds__(NT)
    front_border_kernel(const NodeDesc* __restrict__ nodes, const int32_t* __restrict__ level,
                        const Entry* __restrict__ ent, const int32_t* __restrict__ diag,
                        const KidDesc* __restrict__ kids, const int32_t* __restrict__ maps, PointsView pv,
                        double tol) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ int32_t phys[kLdsMaxN];  // physical row of logical separator row k
    const NodeDesc nd = nodes[level[blockIdx.x]];
    const int64_t p = blockIdx.y;
    T* F = reinterpret_cast<T*>(pv.arena) + p * pv.arena_stride + nd.off;
    int32_t* piv = pv.piv + p * pv.piv_stride + nd.piv_off;
    int32_t* status = pv.status + p * pv.status_stride + nd.slot;
    const int n = nd.n, ne = nd.ne, ne1 = nd.ne1, nb = n - ne, tid = threadIdx.x;
    const int ldl = n + 1, ldr = ne + 1;
    T* Wl = reinterpret_cast<T*>(lds_raw);  // Wl[r + j ldl], r < n, j < ne
    T* Wr = Wl + ldl * ne;                   // Wr[r + (j - ne) ldr], r < ne, ne <= j < n
    const T zero = cx<T>(0.0, 0.0);
    // v added to the border entry (a, b) of the front (F22 entries are not held).
    auto add = [&](int a, int b, T v) {
        if (b < ne)
            Wl[a + b * ldl] = cadd(Wl[a + b * ldl], v);
        else if (a < ne)
            Wr[a + (b - ne) * ldr] = cadd(Wr[a + (b - ne) * ldr], v);
    };
    const int words = ldl * ne + ldr * nb;
    if (kAssemble) {
        for (int i = tid; i < words; i += NT) Wl[i] = zero;
        __syncthreads();
        add_entries<NT>(ent, nd.ent_beg, nd.ent_end, pv.vals[p], n,
                        [&](int q, double2 v) { add(q % n, q / n, convert<T>(v)); });
        __syncthreads();
        const T z = convert<T>(pv.z[p]);
        for (int d = nd.diag_beg + tid; d < nd.diag_end; d += NT) {
            const int q = diag[d];
            add(q, q, z);
        }
        __syncthreads();
        const T* kidbase = reinterpret_cast<const T*>(pv.arena) + p * pv.arena_stride;
        for (int c = nd.kid_beg; c < nd.kid_end; ++c) {  // the children's border entries
            const KidDesc kd = kids[c];
            const T* S = kidbase + kd.src;
            const int32_t* map = maps + kd.map_beg;
            const int m2 = kd.size * kd.size;
            for (int base = tid; base < m2; base += NT * kBatch) {
                int ia[kBatch], ib[kBatch];
                T v[kBatch];
#pragma unroll
                for (int u = 0; u < kBatch; ++u) {
                    const int idx = min(base + u * NT, m2 - 1), a = idx % kd.size, b = idx / kd.size;
                    ia[u] = map[a];
                    ib[u] = map[b];
                    v[u] = zero;
                    if (ia[u] < ne || ib[u] < ne) v[u] = S[a + int64_t(b) * kd.ld];
                }
#pragma unroll
                for (int u = 0; u < kBatch; ++u)
                    if (base + u * NT < m2) add(ia[u], ib[u], v[u]);
            }
            __syncthreads();
        }
    } else {
        copy_batched<NT, T>(n * ne, [&](int idx) { return F[idx]; },
                            [&](int idx, T v) { Wl[(idx % n) + (idx / n) * ldl] = v; });
        copy_batched<NT, T>(ne * nb, [&](int idx) { return F[(idx % ne) + int64_t(ne + idx / ne) * n]; },
                            [&](int idx, T v) { Wr[(idx % ne) + (idx / ne) * ldr] = v; });
        __syncthreads();
    }
    // 2. Partial LU of the border: lds_lu on Wl (every row, the separator columns only).
    const int g = tid % G, r = tid / G;
    int lg = r;
    if (!lds_lu<T, NT, G>(Wl, ldl, n, ne, ne, ne1, tol, piv, status, lg)) return;
    if (g == 0 && r < ne) phys[lg] = r;  // separator rows keep logical positions below ne
    __syncthreads();
    // U12 = L11^-1 P F12, a boundary column per lane: the separator rows in logical order k at
    // physical rows phys[k], L11[k][m] = Wl[phys[k] + m ldl]. Each entry takes the updates of
    // the pivots before it in pivot order, as the right-looking elimination would have applied
    // them, without a barrier per pivot.
    for (int jc = tid; jc < nb; jc += NT) {
        T* col = Wr + jc * ldr;
        for (int k = 1; k < ne; ++k) {
            const int pk = phys[k];
            T acc = col[pk];
            int m = 0;
            for (; m + 4 <= k; m += 4) {  // four terms' loads in flight together, summed in order
                int pm[4];
                T lv[4], uv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) pm[u] = phys[m + u];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    lv[u] = Wl[pk + (m + u) * ldl];
                    uv[u] = col[pm[u]];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) acc = csub(acc, cmul(lv[u], uv[u]));
            }
            for (; m < k; ++m) acc = csub(acc, cmul(Wl[pk + m * ldl], col[phys[m]]));
            col[pk] = acc;
        }
    }
    __syncthreads();
    // 3. S = F22 - L21 U12 (boundary rows are at their logical positions), written once.
    const T* kidbase = reinterpret_cast<const T*>(pv.arena) + p * pv.arena_stride;
    // kSBatch entries of a lane at a time, children two at a time: their inverse-map loads, then
    // their S loads, in flight together (indexes clamped, the sums guarded; each entry sums its
    // children in child order).
    const int nb2 = nb * nb;
    for (int base = tid; base < nb2; base += NT * kSBatch) {
        int ii[kSBatch], jj[kSBatch];
        T f[kSBatch];
#pragma unroll
        for (int u = 0; u < kSBatch; ++u) {
            const int idx = min(base + u * NT, nb2 - 1);
            ii[u] = ne + idx % nb;
            jj[u] = ne + idx / nb;
            f[u] = zero;
        }
        if (kAssemble) {
            for (int c = nd.kid_beg; c < nd.kid_end; c += 2) {
                const bool two = c + 1 < nd.kid_end;
                const KidDesc k0 = kids[c], k1 = kids[two ? c + 1 : c];
                const int32_t *inv0 = maps + k0.inv_beg, *inv1 = maps + k1.inv_beg;
                int a0[kSBatch], b0[kSBatch], a1[kSBatch], b1[kSBatch];
#pragma unroll
                for (int u = 0; u < kSBatch; ++u) {
                    a0[u] = inv0[ii[u]];
                    b0[u] = inv0[jj[u]];
                    a1[u] = inv1[ii[u]];
                    b1[u] = inv1[jj[u]];
                }
                T v0[kSBatch], v1[kSBatch];
#pragma unroll
                for (int u = 0; u < kSBatch; ++u) {
                    v0[u] = kidbase[k0.src + max(a0[u], 0) + int64_t(max(b0[u], 0)) * k0.ld];
                    v1[u] = kidbase[k1.src + max(a1[u], 0) + int64_t(max(b1[u], 0)) * k1.ld];
                }
#pragma unroll
                for (int u = 0; u < kSBatch; ++u) {
                    if (a0[u] >= 0 && b0[u] >= 0) f[u] = cadd(f[u], v0[u]);
                    if (two && a1[u] >= 0 && b1[u] >= 0) f[u] = cadd(f[u], v1[u]);
                }
            }
        } else {
#pragma unroll
            for (int u = 0; u < kSBatch; ++u) f[u] = F[ii[u] + int64_t(jj[u]) * n];
        }
#pragma unroll
        for (int u = 0; u < kSBatch; ++u)
            if (base + u * NT < nb2) {
                const int i = ii[u], j = jj[u];
                T acc = zero;
                for (int k = 0; k < ne; ++k) acc = cadd(acc, cmul(Wl[i + k * ldl], Wr[phys[k] + (j - ne) * ldr]));
                F[i + int64_t(j) * n] = csub(f[u], acc);
            }
    }
    // The factors (kept for the solves): the border in logical row order.
    if (pv.keep && r < n) {
        for (int j = g; j < ne; j += G) F[lg + int64_t(j) * n] = Wl[r + j * ldl];
        if (r < ne)
            for (int j = ne + g; j < n; j += G) F[lg + int64_t(j) * n] = Wr[r + (j - ne) * ldr];
    }
    if (tid == 0) *status = kFrontOk;
}

// ---------------------------------------------------------------------------------------------
// Small fronts (n <= 64), register-resident elimination. The front is assembled in LDS as in
// front_lds_kernel, then lane r of every wave takes row r, and wave w the columns j = w (mod NW),
// S of them, into registers. Step k of the partial LU:
//  - the wave holding column k finds the pivot among the rows of logical position [k, lim) by a
//    wave reduction (largest modulus, smallest logical position on ties), applies the pivot test
//    of field.rs:61, swaps the logical positions of the pivot row and of the row at k (the rows
//    stay in their lanes), turns column k of the rows below into the multipliers l = F[r, k] /
//    F[p, k], and publishes l and the pivot row through LDS (double-buffered by step parity:
//    one barrier a step);
//  - every wave then updates its columns j > k: F[r, j] -= l_r F[p, j], with F[p, j] read from
//    the pivot row's lane of the same wave by readlane (every wave holds every row).
// LDS carries 64 multipliers a step instead of the whole trailing block. The factors are
// written back in logical row order as front_lds_kernel leaves them (LAPACK's interchanges of
// whole rows, the layout the solves read).
template <class T>
__device__ __forceinline__ T lane_of(T v, int l);
template <>
__device__ __forceinline__ double2 lane_of(double2 v, int l) {
    auto one = [&](double x) {
        const int lo = __builtin_amdgcn_readlane(__double2loint(x), l);
        const int hi = __builtin_amdgcn_readlane(__double2hiint(x), l);
        return __hiloint2double(hi, lo);
    };
    return double2{one(v.x), one(v.y)};
}
template <>
__device__ __forceinline__ float2 lane_of(float2 v, int l) {
    return float2{__int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.x), l)),
                  __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.y), l))};
}

template <class T, int NW, int S, bool kAssemble>
__global__ void __launch_bounds__(64 * NW)
    front_reg_kernel(const NodeDesc* __restrict__ nodes, const int32_t* __restrict__ level,
                     const Entry* __restrict__ ent, const int32_t* __restrict__ diag,
                     const KidDesc* __restrict__ kids, const int32_t* __restrict__ maps, PointsView pv, double tol) {
    constexpr int NT = 64 * NW;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ T lbuf[2][64];
    __shared__ int32_t pinfo[2][2];
    __shared__ double smax[NW];
    T* W = reinterpret_cast<T*>(lds_raw);
    const NodeDesc nd = nodes[level[blockIdx.x]];
    const int64_t p = blockIdx.y;
    T* F = reinterpret_cast<T*>(pv.arena) + p * pv.arena_stride + nd.off;
    int32_t* piv = pv.piv + p * pv.piv_stride + nd.piv_off;
    int32_t* status = pv.status + p * pv.status_stride + nd.slot;
    const int n = nd.n, ne = nd.ne, ne1 = nd.ne1, tid = threadIdx.x, ld = n + 1;
    const T zero = cx<T>(0.0, 0.0);
    if (kAssemble) {
        for (int i = tid; i < n * ld; i += NT) W[i] = zero;
        __syncthreads();
        add_entries<NT>(ent, nd.ent_beg, nd.ent_end, pv.vals[p], ld,
                        [&](int at, double2 v) { W[at] = cadd(W[at], convert<T>(v)); });
        __syncthreads();
        const T z = convert<T>(pv.z[p]);
        for (int d = nd.diag_beg + tid; d < nd.diag_end; d += NT) {
            const int q = diag[d];
            W[q + q * ld] = cadd(W[q + q * ld], z);
        }
        __syncthreads();
        const T* kidbase = reinterpret_cast<const T*>(pv.arena) + p * pv.arena_stride;
        for (int c = nd.kid_beg; c < nd.kid_end; ++c) {
            const KidDesc kd = kids[c];
            extend_add<NT>(kidbase + kd.src, kd.ld, maps + kd.map_beg, kd.size, ld,
                           [&](int at, T v) { W[at] = cadd(W[at], v); });
            __syncthreads();
        }
    } else {
        copy_batched<NT, T>(n * n, [&](int idx) { return F[idx]; },
                            [&](int idx, T v) { W[(idx % n) + (idx / n) * ld] = v; });
        __syncthreads();
    }
    const int r = tid & 63, wave = tid >> 6;
    T w[S];
    double m = 0.0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int j = wave + s * NW;
        w[s] = (r < n && j < n) ? W[r + j * ld] : zero;
        if (r < ne1 && j < ne1) m = fmax(m, cabs(w[s]));
    }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if (r == 0) smax[wave] = m;
    __syncthreads();
    double scale = smax[0];
#pragma unroll
    for (int q = 1; q < NW; ++q) scale = fmax(scale, smax[q]);
    int lg = r;  // logical position of this lane's row
    // Column k sits in w[k / NW] of wave k % NW; it is selected (and written) by unrolled
    // compares so that w stays in registers.
    for (int k = 0; k < ne; ++k) {
        const int kw = k % NW, ks = k / NW, par = k & 1;
        if (wave == kw) {
            T colk = zero;
#pragma unroll
            for (int s = 0; s < S; ++s)
                if (s == ks) colk = w[s];
            const int lim = k < ne1 ? ne1 : ne;
            double big = -1.0;
            int at = kLdsMaxN, ph = 0;
            if (r < n && lg >= k && lg < lim) {
                big = cabs2(colk);
                at = lg;
                ph = r;
            }
            wave_argmax(big, at, ph);
            const bool ok = pivot_ok(big, k < ne1, tol, scale);
            const int phs = __builtin_amdgcn_readfirstlane(ph);
            const int lgn = r == phs ? k : (lg == k ? at : lg);
            T l = zero;
            if (ok) {
                const T inv = crecip(lane_of(colk, phs));
                if (r < n && lgn > k) {
                    l = cmul(colk, inv);
#pragma unroll
                    for (int s = 0; s < S; ++s)
                        if (s == ks) w[s] = l;
                }
            }
            lbuf[par][r] = l;
            if (r == 0) {
                pinfo[par][0] = ok ? phs : -1;
                pinfo[par][1] = at;
                if (ok)
                    piv[k] = at;
                else
                    *status = k < ne1 ? kFrontDelayed : kFrontSingular;
            }
        }
        __syncthreads();
        const int ph = __builtin_amdgcn_readfirstlane(pinfo[par][0]);
        if (ph < 0) return;  // the whole workgroup: the pivot test failed
        const int at = pinfo[par][1];
        lg = r == ph ? k : (lg == k ? at : lg);
        const T l = lbuf[par][r];
#pragma unroll
        for (int s = 0; s < S; ++s)
            if (wave + s * NW > k) w[s] = csub(w[s], cmul(l, lane_of(w[s], ph)));
    }
    if (r < n) {
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int j = wave + s * NW;
            if (j < n && (pv.keep || (lg >= ne && j >= ne))) F[lg + j * n] = w[s];
        }
    }
    if (tid == 0) *status = kFrontOk;
}

// ---------------------------------------------------------------------------------------------
// Large fronts, step 1: the border. Partial LU of F11 (row pivoting within the separator rows,
// the pivot test of field.rs:61), L21 = F21 U11^-1 and U12 = L11^-1 P F12, in place, panel by
// panel (kPanel columns), without touching F22: the Schur update S = F22 - L21 U12 is left to
// schur_kernel, which spreads it over many workgroups. The border is factored left-looking
// (Crout): a panel's columns and rows receive the updates of all earlier panels at once, as
// MFMA products with K = k0, when the panel comes up. Per panel [k0, k1):
//   1. panel columns, rows [k0, n):   F -= L[:, 0:k0] U[0:k0, panel];
//   2. the panel's separator rows [k0, ne) are factored in LDS (when they fit, else in place);
//   3. its interchanges (as one permutation of at most 2 kPanel rows) reach the other columns,
//      the L21 rows are solved against the panel's U11 block, the panel goes back to the front;
//   4. panel rows, columns [k1, n):   F -= L[panel, 0:k0] U[0:k0, k1:n];
//   5. the panel's U rows are solved with its unit-lower L11 block (one lane per column).
// Workgroup size of border_kernel: 4 wavefronts for the separators past kF11Lds, 8 for the root
// (one front per point: its panel loop is the longest chain of a sweep).
constexpr int kBorderThreads = 256, kRootThreads = 512;

// C -= A B for one 16 x 16 tile of the front F: C = F[i0:i0+16, j0:j0+16], A = F[i0:, 0:kend]
// (L columns), B = F[0:kend, j0:] (U rows), computed as C^T -= B^T A^T so that lanes read and
// write the front's columns contiguously; entries outside rows < rmax, columns < cmax are not
// stored. Loads are clamped inside the front and their values selected (a select of addresses
// would go through flat), 16 values of K per batch of loads.
template <class T>
__device__ __forceinline__ void mfma_gemm_tile(T* F, int n, int i0, int j0, int kend, int rmax, int cmax, int lane) {
    using M = Mfma<T>;
    typename M::acc cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
    const int li = lane % 16, lk = lane / 16;
    const int64_t ii = min(i0 + li, n - 1), jj0 = int64_t(min(j0 + li, n - 1)) * n;
    // Two register sets: the next 16 values of K load while the MFMAs run on the current ones.
    auto load = [&](int k, T (&a)[4], T (&b)[4]) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int kk = k + 4 * s + lk, kc = min(kk, kend - 1);
            a[s] = F[kc + jj0];
            b[s] = F[ii + int64_t(kc) * n];
            if (kk >= kend) {
                a[s] = cx<T>(0.0, 0.0);
                b[s] = cx<T>(0.0, 0.0);
            }
        }
    };
    auto mma = [&](const T (&a)[4], const T (&b)[4]) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            cr = M::mac(a[s].x, b[s].x, cr);
            cr = M::mac(-a[s].y, b[s].y, cr);
            ci = M::mac(a[s].x, b[s].y, ci);
            ci = M::mac(a[s].y, b[s].x, ci);
        }
    };
    T a0[4], b0[4], a1[4], b1[4];
    if (kend > 0) load(0, a0, b0);
    for (int k = 0; k < kend; k += 32) {
        if (k + 16 < kend) load(k + 16, a1, b1);
        mma(a0, b0);
        if (k + 16 >= kend) break;
        if (k + 32 < kend) load(k + 32, a0, b0);
        mma(a1, b1);
    }
    const int i = i0 + li;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int j = j0 + M::row(lane, q);
        if (i < rmax && j < cmax) {
            T& d = F[i + int64_t(j) * n];
            d.x -= cr[q];
            d.y -= ci[q];
        }
    }
}

// The interchanges k0 + c <-> sp[c] (c < w, in order) as one permutation: new row r holds the old
// row traced back through them.
__device__ __forceinline__ int traced_row(int r, const int32_t* sp, int k0, int w) {
    for (int c = w - 1; c >= 0; --c) {
        const int a = k0 + c, b = sp[c];
        if (r == a)
            r = b;
        else if (r == b)
            r = a;
    }
    return r;
}

// The border factorisation of one front; kLds: the panel's separator rows are held in LDS (P).
// mv: LDS table of the panel's row moves (count, then destination and source rows).
template <class T, bool kRoot, bool kLds, int NT>
__device__ __forceinline__ void border_factor(T* F, T* P, int32_t* spiv, int32_t* mv, int n, int ne, int ne1,
                                              int32_t* piv, int32_t* status, double scale, double tol) {
    __shared__ T dinv[kPanel];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    auto at = [&](int i, int j) -> T& { return F[i + int64_t(j) * n]; };
    for (int k0 = 0; k0 < ne; k0 += kPanel) {
        const int w = min(kPanel, ne - k0), R = ne - k0, k1 = k0 + w;
        // 1. Earlier panels' updates of this panel's columns.
        if (k0 > 0) {
            for (int t = wave; t * 16 < n - k0; t += NT / 64) mfma_gemm_tile(F, n, k0 + 16 * t, k0, k0, n, k1, lane);
            __syncthreads();
        }
        // Panel element (row i in [k0, ne), panel column c).
        auto pp = [&](int i, int c) -> T& { return kLds ? P[(i - k0) + c * R] : at(i, k0 + c); };
        if (kLds) {
            copy_batched<NT, T>(R * w, [&](int e) { return at(k0 + e % R, k0 + e / R); }, [&](int e, T v) { P[e] = v; });
            __syncthreads();
        }
        // 2. Panel factorisation on the separator rows.
        for (int c = 0; c < w; ++c) {
            const int k = k0 + c, lim = k < ne1 ? ne1 : ne;
            // Every wave finds the same pivot; the rows k and q are swapped only after every wave
            // has read column c (a barrier), so that no wave sees them half swapped.
            double v = -1.0;
            int qi = k;
            for (int i = k + lane; i < lim; i += 64) {
                const double a = cabs2(pp(i, c));
                if (a > v) {
                    v = a;
                    qi = i;
                }
            }
            double big;
            int q, unused;
            T unused_v;
            wave_pick(v, qi, 0, cx<T>(0.0, 0.0), big, q, unused, unused_v);
            const bool ok = pivot_ok(big, k < ne1, tol, scale);
            if (!ok) {
                if (tid == 0) *status = (k < ne1 && !kRoot) ? kFrontDelayed : kFrontSingular;
                return;
            }
            if (tid == 0) {
                piv[k] = q;
                spiv[c] = q;
            }
            __syncthreads();
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
        // The reciprocals of the panel's pivots, for the L21 rows of step 3.
        if (tid < w) dinv[tid] = crecip(pp(k0 + tid, tid));
        // The interchanges k0 + c <-> spiv[c], in order, as one permutation: row mv[1 + m] of the
        // front receives the old row mv[1 + 2 kPanel + m]. Wave 0 finds the rows that change, a
        // lane each: lane t < w the panel row k0 + t, lane w + c the row spiv[c] below the panel
        // (at its first occurrence), each traced back through the interchanges; a ballot numbers
        // the moves in lane order.
        if (wave == 0) {
            int dst = -1;
            if (lane < w) {
                dst = k0 + lane;
            } else if (lane < 2 * w) {
                const int c = lane - w, b = spiv[c];
                bool first = b >= k1;
                for (int c2 = 0; c2 < c; ++c2) first = first && spiv[c2] != b;
                if (first) dst = b;
            }
            const int src = dst >= 0 ? traced_row(dst, spiv, k0, w) : dst;
            const bool move = src != dst;
            const unsigned long long mask = __ballot(move);
            if (move) {
                const int m = __popcll(mask & ((1ull << lane) - 1));
                mv[1 + m] = dst;
                mv[1 + 2 * kPanel + m] = src;
            }
            if (lane == 0) mv[0] = __popcll(mask);
        }
        __syncthreads();
        // 3. The moves on the columns outside the panel (separator rows only), the panel back in
        //    place, and the L21 rows: l U11 = f with U11 the panel's upper triangle.
        //    All moves of one column are done by the lanes of one wave in one load and one store
        //    instruction (lane = column slot x move, the moves padded to a power of two), so every
        //    source entry of the column is read before any of them is written; kBatch columns of a
        //    lane are in flight together (different columns: no overlap).
        const int moves = mv[0];
        if (moves > 0) {
            int pm = 1;
            while (pm < moves) pm <<= 1;  // at most 2 kPanel = 32
            const int per = 64 / pm, mi = lane % pm, slot = lane / pm, cols = n - w;
            const bool on = mi < moves;
            const int dst = on ? mv[1 + mi] : 0, src = on ? mv[1 + 2 * kPanel + mi] : 0;
            auto col = [&](int jc) { return jc < k0 ? jc : jc + w; };
            const int stride = (NT / 64) * per;
            for (int base = wave * per + slot; base < cols; base += stride * kBatch) {
                T v[kBatch];
#pragma unroll
                for (int u = 0; u < kBatch; ++u) v[u] = at(src, col(min(base + u * stride, cols - 1)));
#pragma unroll
                for (int u = 0; u < kBatch; ++u)
                    if (on && base + u * stride < cols) at(dst, col(base + u * stride)) = v[u];
            }
        }
        __syncthreads();
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
                    l[c] = cmul(acc, dinv[c]);
                }
            }
#pragma unroll
            for (int c = 0; c < kPanel; ++c)
                if (c < w) at(i, k0 + c) = l[c];
        }
        __syncthreads();
        if (k1 >= n) break;
        // 4. Earlier panels' updates of this panel's rows right of it.
        if (k0 > 0) {
            for (int t = wave; t * 16 < n - k1; t += NT / 64) mfma_gemm_tile(F, n, k0, k1 + 16 * t, k0, k1, n, lane);
            __syncthreads();
        }
        // 5. U rows of the panel: unit-lower solve, one lane per column.
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
    }
    if (tid == 0) *status = kFrontOk;
}

template <class T, bool kRoot, int NT>
__global__ void __launch_bounds__(NT)
    border_kernel(const NodeDesc* __restrict__ nodes, const int32_t* __restrict__ level, const Entry* __restrict__ ent,
                  const int32_t* __restrict__ diag, PointsView pv, double tol, int lds_rows) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ double sv[NT / 64];
    __shared__ int32_t spiv[kPanel];
    __shared__ int32_t mv[1 + 4 * kPanel];
    const NodeDesc nd = nodes[level[blockIdx.x]];
    const int64_t p = blockIdx.y;
    T* F = (kRoot ? reinterpret_cast<T*>(pv.root_arena) + p * pv.root_stride
                  : reinterpret_cast<T*>(pv.arena) + p * pv.arena_stride) +
           nd.off;
    int32_t* piv = pv.piv + p * pv.piv_stride + nd.piv_off;
    int32_t* status = pv.status + p * pv.status_stride + nd.slot;
    const int n = nd.n, ne = nd.ne, ne1 = nd.ne1;
    if (!kRoot && nd.ent_end > nd.ent_beg) {  // -H on the separator (the children were gathered before)
        add_entries<NT>(ent, nd.ent_beg, nd.ent_end, pv.vals[p], n,
                                    [&](int at, double2 v) { F[at] = cadd(F[at], convert<T>(v)); });
        __syncthreads();
    }
    if (!kRoot && nd.diag_end > nd.diag_beg) {  // z on the separator's diagonal
        const T z = convert<T>(pv.z[p]);
        for (int d = nd.diag_beg + int(threadIdx.x); d < nd.diag_end; d += NT) {
            const int64_t q = diag[d];
            F[q + q * n] = cadd(F[q + q * n], z);
        }
        __syncthreads();
    }
    double m = 0.0;
    for (int idx = threadIdx.x; idx < ne1 * ne1; idx += NT)
        m = fmax(m, cabs(F[(idx % ne1) + int64_t(idx / ne1) * n]));
    const double scale = block_max(m, sv);
    if (ne <= lds_rows)
        border_factor<T, kRoot, true, NT>(F, reinterpret_cast<T*>(lds_raw), spiv, mv, n, ne, ne1, piv, status, scale,
                                          tol);
    else
        border_factor<T, kRoot, false, NT>(F, nullptr, spiv, mv, n, ne, ne1, piv, status, scale, tol);
}

// ---------------------------------------------------------------------------------------------
// Large fronts, step 0: the children's Schur complements gathered into the border of the front
// (rows or columns below ne; F22 is gathered by schur_kernel), one workgroup per tile of the two
// border strips (rows x cols in Tile::pad as (rows << 16) | cols, about kTile^2 entries, inside the
// border): F[i, j] = sum over children c of S_c[inv_c(i), inv_c(j)], written once (no zero fill,
// no read-modify-write).
template <class T>
__global__ void __launch_bounds__(256)
    gather_kernel(const NodeDesc* __restrict__ nodes, const Tile* __restrict__ tiles, const KidDesc* __restrict__ kids,
                  const int32_t* __restrict__ maps, PointsView pv) {
    constexpr int kGBatch = 4;  // entries per lane per batch of loads
    const TilePoint tp = xcd_tile();
    const Tile tl = tiles[tp.tile];
    const NodeDesc nd = nodes[tl.node];
    const int64_t p = tp.point;
    const T* kidbase = reinterpret_cast<const T*>(pv.arena) + p * pv.arena_stride;
    T* F = reinterpret_cast<T*>(pv.arena) + p * pv.arena_stride + nd.off;
    const int n = nd.n, ne = nd.ne;
    const int rows = min(tl.pad >> 16, n - tl.i0), cols = min(tl.pad & 0xffff, n - tl.j0), cnt = rows * cols;
    for (int base = int(threadIdx.x); base < cnt; base += 256 * kGBatch) {
        int ii[kGBatch], jj[kGBatch];
#pragma unroll
        for (int u = 0; u < kGBatch; ++u) {
            const int idx = min(base + u * 256, cnt - 1);
            ii[u] = tl.i0 + idx % rows;
            jj[u] = tl.j0 + idx / rows;
        }
        T acc[kGBatch];
#pragma unroll
        for (int u = 0; u < kGBatch; ++u) acc[u] = cx<T>(0.0, 0.0);
        // Children two at a time: the map loads of both, then the values of both, in flight
        // together; each entry sums its children in child order.
        for (int c = nd.kid_beg; c < nd.kid_end; c += 2) {
            const bool two = c + 1 < nd.kid_end;
            const KidDesc k0 = kids[c], k1 = kids[two ? c + 1 : c];
            const int32_t *inv0 = maps + k0.inv_beg, *inv1 = maps + k1.inv_beg;
            const T *S0 = kidbase + k0.src, *S1 = kidbase + k1.src;
            int a0[kGBatch], b0[kGBatch], a1[kGBatch], b1[kGBatch];
#pragma unroll
            for (int u = 0; u < kGBatch; ++u) {
                a0[u] = inv0[ii[u]];
                b0[u] = inv0[jj[u]];
                a1[u] = inv1[ii[u]];
                b1[u] = inv1[jj[u]];
            }
            T v0[kGBatch], v1[kGBatch];
#pragma unroll
            for (int u = 0; u < kGBatch; ++u) {
                v0[u] = S0[max(a0[u], 0) + int64_t(max(b0[u], 0)) * k0.ld];
                v1[u] = S1[max(a1[u], 0) + int64_t(max(b1[u], 0)) * k1.ld];
            }
#pragma unroll
            for (int u = 0; u < kGBatch; ++u) {
                acc[u] = cadd(acc[u], a0[u] < 0 || b0[u] < 0 ? cx<T>(0.0, 0.0) : v0[u]);
                if (two) acc[u] = cadd(acc[u], a1[u] < 0 || b1[u] < 0 ? cx<T>(0.0, 0.0) : v1[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < kGBatch; ++u)
            if (base + u * 256 < cnt && (ii[u] < ne || jj[u] < ne)) F[ii[u] + int64_t(jj[u]) * n] = acc[u];
    }
}

// ---------------------------------------------------------------------------------------------
// Large fronts with a small separator (ne <= kF11Lds): the border in two launches instead of
// border_kernel's panel loop.
//  - f11_kernel, one workgroup per front: -H and z added to the gathered border, then the LU of
//    F11 in LDS (lds_lu) written back with its pivots;
//  - trsm_kernel, one workgroup per kTrsmSpan columns of F12 (U12 = L11^-1 P F12) or rows of F21
//    (L21 = F21 U11^-1): the triangle in LDS, each column / row in the registers of 4 lanes.
template <class T, int NT, int G>
__global__ void __launch_bounds__(NT)
    f11_kernel(const NodeDesc* __restrict__ nodes, const int32_t* __restrict__ level, const Entry* __restrict__ ent,
               const int32_t* __restrict__ diag, PointsView pv, double tol) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    T* W = reinterpret_cast<T*>(lds_raw);
    const NodeDesc nd = nodes[level[blockIdx.x]];
    const int64_t p = blockIdx.y;
    T* F = reinterpret_cast<T*>(pv.arena) + p * pv.arena_stride + nd.off;
    int32_t* piv = pv.piv + p * pv.piv_stride + nd.piv_off;
    int32_t* status = pv.status + p * pv.status_stride + nd.slot;
    const int n = nd.n, ne = nd.ne, ne1 = nd.ne1, tid = threadIdx.x, ld = ne + 1;
    // F11 (the gathered children's part) into LDS, then -H and z added there; the entries in the
    // border strips F12 / F21 are added in place in HBM (trsm_kernel reads them next).
    copy_batched<NT, T>(ne * ne, [&](int idx) { return F[(idx % ne) + int64_t(idx / ne) * n]; },
                         [&](int idx, T v) { W[(idx % ne) + (idx / ne) * ld] = v; });
    __syncthreads();
    if (nd.ent_end > nd.ent_beg) {
        add_entries<NT>(ent, nd.ent_beg, nd.ent_end, pv.vals[p], n, [&](int at, double2 v) {
            const int a = at % n, b = at / n;
            if (a < ne && b < ne)
                W[a + b * ld] = cadd(W[a + b * ld], convert<T>(v));
            else
                F[at] = cadd(F[at], convert<T>(v));
        });
        __syncthreads();
    }
    if (nd.diag_end > nd.diag_beg) {  // the eliminated orbitals: all inside F11
        const T z = convert<T>(pv.z[p]);
        for (int d = nd.diag_beg + tid; d < nd.diag_end; d += NT) {
            const int q = diag[d];
            W[q + q * ld] = cadd(W[q + q * ld], z);
        }
        __syncthreads();
    }
    const int g = tid % G, r = tid / G;
    int lg = r;
    if (!lds_lu<T, NT, G>(W, ld, ne, ne, ne, ne1, tol, piv, status, lg)) return;
    if (r < ne)
        for (int j = g; j < ne; j += G) F[lg + int64_t(j) * n] = W[r + j * ld];
}



template <class T, int NS>
__global__ void __launch_bounds__(256)
    trsm_kernel(const NodeDesc* __restrict__ nodes, const Tile* __restrict__ tiles, PointsView pv) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ int32_t src[kLdsMaxN];
    const TilePoint tp = xcd_tile();
    const Tile tl = tiles[tp.tile];
    const NodeDesc nd = nodes[tl.node];
    const int64_t p = tp.point;
    if (pv.status[p * pv.status_stride + nd.slot] != kFrontOk) return;
    T* F = reinterpret_cast<T*>(pv.arena) + p * pv.arena_stride + nd.off;
    const int32_t* piv = pv.piv + p * pv.piv_stride + nd.piv_off;
    const int n = nd.n, ne = nd.ne, nb = n - ne, tid = threadIdx.x;
    // The triangle the solve reads, packed (ne (ne + 1) / 2 words, half the square): kind 0 the
    // strict lower L11 by columns, L[i][k] at lo(k) + i - k - 1 (i > k); kind 1 U11 by rows,
    // U[k][j] at up(k) + j - k (j >= k), with 1 / U[k][k] in place of the diagonal (the solve
    // only divides by it).
    T* A = reinterpret_cast<T*>(lds_raw);
    auto lo = [&](int k) { return k * ne - k * (k + 1) / 2; };
    auto up = [&](int k) { return k * ne - k * (k - 1) / 2; };
    const bool upper = tl.pad != 0;
    if (!upper && tid < 64) {  // the rows of P F12: row i is the front's row src[i] (ne <= 64)
        // Lane i holds src[i]; LAPACK's interchanges applied in order, two readlanes each.
        int v = tid;
        const int pk = tid < ne ? piv[tid] : 0;
        for (int k = 0; k < ne; ++k) {
            const int q = __builtin_amdgcn_readlane(pk, k);
            const int vk = __builtin_amdgcn_readlane(v, k), vq = __builtin_amdgcn_readlane(v, q);
            v = tid == k ? vq : (tid == q ? vk : v);
        }
        if (tid < ne) src[tid] = v;
    }
    auto wanted = [&](int i, int j) { return upper ? i <= j : i > j; };
    copy_batched<256, T>(ne * ne,
                         [&](int idx) {
                             const int i = idx % ne, j = idx / ne;
                             return wanted(i, j) ? F[i + int64_t(j) * n] : cx<T>(0.0, 0.0);
                         },
                         [&](int idx, T v) {
                             const int i = idx % ne, j = idx / ne;
                             if (wanted(i, j)) A[upper ? up(i) + j - i : lo(j) + i - j - 1] = i == j ? crecip(v) : v;
                         });
    __syncthreads();
    // Four lanes per column of U12 (kind 0) or row of L21 (kind 1), the workgroup's 64 of them in
    // one pass: lane g holds entries g + 4 s (s < NS, ne <= 4 NS) in registers, and the solve
    // broadcasts the pivot entry inside the four lanes (a quad permutation).
    const int grp = tid / 4, g = tid % 4;
    const int start = upper ? tl.i0 : tl.j0, end = min(start + kTrsmSpan, nb);
    const int col = min(start + grp, end - 1);
    T x[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int e = min(g + 4 * s, ne - 1);
        x[s] = upper ? F[(ne + col) + int64_t(e) * n] : F[src[e] + int64_t(ne + col) * n];
    }
    // The steps k = 4 kb + kk: kb unrolled (x[kb] stays a register), kk a loop, which keeps the
    // code a quarter of the fully unrolled size.
    if (!upper) {  // forward substitution with the unit-lower L11
#pragma unroll
        for (int kb = 0; kb < NS; ++kb)
#pragma unroll 1
            for (int kk = 0; kk < 4; ++kk) {
                const int k = 4 * kb + kk;
                if (k < ne - 1) {
                    T xk;
                    xk.x = __shfl(x[kb].x, kk, 4);
                    xk.y = __shfl(x[kb].y, kk, 4);
                    const int ck = lo(k) - k - 1;
#pragma unroll
                    for (int s = kb; s < NS; ++s) {
                        const int i = g + 4 * s;
                        if (i > k && i < ne) x[s] = csub(x[s], cmul(A[ck + i], xk));
                    }
                }
            }
    } else {  // x U11 = f, column by column
#pragma unroll
        for (int kb = 0; kb < NS; ++kb)
#pragma unroll 1
            for (int kk = 0; kk < 4; ++kk) {
                const int k = 4 * kb + kk;
                if (k < ne) {
                    const int rk = up(k) - k;
                    const T cand = cmul(x[kb], A[rk + k]);
                    T xk;
                    xk.x = __shfl(cand.x, kk, 4);
                    xk.y = __shfl(cand.y, kk, 4);
                    if (g == kk) x[kb] = xk;
#pragma unroll
                    for (int s = kb; s < NS; ++s) {
                        const int j = g + 4 * s;
                        if (j > k && j < ne) x[s] = csub(x[s], cmul(xk, A[rk + j]));
                    }
                }
            }
    }
    if (start + grp < end)
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int e = g + 4 * s;
            if (e < ne) {
                if (upper)
                    F[(ne + col) + int64_t(e) * n] = x[s];
                else
                    F[e + int64_t(ne + col) * n] = x[s];
            }
        }
}

// ---------------------------------------------------------------------------------------------
// Large fronts, step 2: S = F22 - L21 U12 (K = ne) over TS x TS tiles of the Schur block, one
// workgroup per (tile, point), on MFMA: 64 x 64 tiles with four waves or 32 x 32 tiles with one
// (the tuning's schur_tile). Each wave owns a 32 x 32 part (2 x 2 blocks of 16 x 16, complex = 4
// real MFMAs per block and K-step of 4); L21 and U12 are staged through LDS in chunks of kChunk
// columns / rows. Fronts whose border failed are skipped. (The compiler keeps the constant zero
// in 16-32 bytes of scratch here; the form without it, zero written as a literal, schedules the K
// loop worse: 10 % slower on the BHZ 200 x 64 batch.)
constexpr int kChunk = 16, kTileLd = kTile + 1;
constexpr int kChunk32 = 8, kSchurPad = 1;  // K chunk of the 32 x 32 Schur tiles (half the LDS of 16: more workgroups a CU); LDS row padding

// At least three waves a SIMD (the second bound): the compiler then keeps the accumulators in
// VGPRs instead of AGPRs, 162 registers in all instead of 226, with no spills (two waves a SIMD
// otherwise); 18.8 against 20.2 ms a BHZ 200 x 64 step.
template <class T, bool kRoot, bool kGather, int TS>
__global__ void __launch_bounds__(TS == 64 ? 256 : 64, 3)
    schur_kernel(const NodeDesc* __restrict__ nodes, const Tile* __restrict__ tiles, const KidDesc* __restrict__ kids,
                 const int32_t* __restrict__ maps, PointsView pv) {
    static_assert(TS == 64 || TS == 32, "Schur tiles are 64 x 64 or 32 x 32");
    using M = Mfma<T>;
    constexpr int NT = TS == 64 ? 256 : 64, kLd = TS + kSchurPad, kCh = TS == 64 ? kChunk : kChunk32;
    __shared__ T As[kCh * kLd];  // As[k * kLd + i] = L21[i0 + i, kc + k]
    __shared__ T Bs[kCh * kLd];  // Bs[k * kLd + j] = U12[kc + k, j0 + j]
    const TilePoint tp = xcd_tile();
    const Tile tl = tiles[tp.tile];
    const NodeDesc nd = nodes[tl.node];
    const int64_t p = tp.point;
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
    // The K loop in chunks of kCh: the next chunk's L21 / U12 entries are loaded into registers
    // (kStage per thread and operand) while the MFMAs run on the current one from LDS. Clamped
    // loads and a select of the value (a select of addresses would go through flat).
    constexpr int kStage = TS * kCh / NT;
    T sa[kStage], sb[kStage];
    auto fetch = [&](int kc) {
#pragma unroll
        for (int u = 0; u < kStage; ++u) {
            const int e = tid + u * NT;
            {
                const int r = e % TS, c = e / TS, i = tl.i0 + r, k = kc + c;
                T v = F[(ne + min(i, nb - 1)) + int64_t(min(k, ne - 1)) * n];
                if (!(i < nb && k < ne)) v = zero;
                sa[u] = v;
            }
            {
                const int c = e % kCh, r = e / kCh, j = tl.j0 + r, k = kc + c;
                T v = F[min(k, ne - 1) + int64_t(ne + min(j, nb - 1)) * n];
                if (!(j < nb && k < ne)) v = zero;
                sb[u] = v;
            }
        }
    };
    // The wave's 16 x 16 blocks that hold entries of S (a block past nb in rows or columns is
    // padding of a partial tile and is skipped; wave-uniform).
    bool live[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) live[x][y] = tl.j0 + wj + 16 * x < nb && tl.i0 + wi + 16 * y < nb;
    fetch(0);
    for (int kc = 0; kc < ne; kc += kCh) {
#pragma unroll
        for (int u = 0; u < kStage; ++u) {
            const int e = tid + u * NT;
            As[(e / TS) * kLd + e % TS] = sa[u];
            Bs[(e % kCh) * kLd + e / kCh] = sb[u];
        }
        __syncthreads();
        if (kc + kCh < ne) fetch(kc + kCh);
#pragma unroll
        for (int kk = 0; kk < kCh; kk += 4) {
            if (kc + kk >= ne) break;  // the rest of the chunk is padding (zeros)
            const int k = kk + lane / 16;
            T a[2], b[2];
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                a[s] = Bs[k * kLd + wj + 16 * s + lane % 16];  // U12^T operand
                b[s] = As[k * kLd + wi + 16 * s + lane % 16];  // L21^T operand
            }
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y)
                    if (live[x][y]) {
                        cr[x][y] = M::mac(a[x].x, b[y].x, cr[x][y]);
                        cr[x][y] = M::mac(-a[x].y, b[y].y, cr[x][y]);
                        ci[x][y] = M::mac(a[x].x, b[y].y, ci[x][y]);
                        ci[x][y] = M::mac(a[x].y, b[y].x, ci[x][y]);
                    }
        }
        __syncthreads();
    }
    // Block (x, y) of the wave: lane l, element jj holds (L21 U12)[i, j] with
    // i = wi + 16 y + l % 16 and j = wj + 16 x + row(l, jj). In two halves (x) to bound the
    // registers: F22 (read, or gathered from the children with all map loads before the values;
    // two children, the usual case, in one batch), then S = F22 - L21 U12 stored.
    const T* kidbase = reinterpret_cast<const T*>(pv.arena) + p * pv.arena_stride;
    int ir[2];  // this lane's rows of S (clamped)
#pragma unroll
    for (int y = 0; y < 2; ++y) ir[y] = min(tl.i0 + wi + 16 * y + lane % 16, nb - 1);
#pragma unroll
    for (int x = 0; x < 2; ++x) {
        int jc[4];
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) jc[jj] = min(tl.j0 + wj + 16 * x + M::row(lane, jj), nb - 1);
        T f22[2][4];
        if (kGather && nd.kid_end - nd.kid_beg == 2) {
            const KidDesc k0 = kids[nd.kid_beg], k1 = kids[nd.kid_beg + 1];
            const int32_t *inv0 = maps + k0.inv_beg + ne, *inv1 = maps + k1.inv_beg + ne;
            int a0[2], a1[2], b0[4], b1[4];
#pragma unroll
            for (int y = 0; y < 2; ++y) {
                a0[y] = inv0[ir[y]];
                a1[y] = inv1[ir[y]];
            }
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                b0[jj] = inv0[jc[jj]];
                b1[jj] = inv1[jc[jj]];
            }
            const T *S0 = kidbase + k0.src, *S1 = kidbase + k1.src;
            T v0[2][4], v1[2][4];
#pragma unroll
            for (int y = 0; y < 2; ++y)
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) {
                    v0[y][jj] = S0[max(a0[y], 0) + int64_t(max(b0[jj], 0)) * k0.ld];
                    v1[y][jj] = S1[max(a1[y], 0) + int64_t(max(b1[jj], 0)) * k1.ld];
                }
#pragma unroll
            for (int y = 0; y < 2; ++y)
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) {
                    T f = zero;
                    if (a0[y] >= 0 && b0[jj] >= 0) f = cadd(f, v0[y][jj]);
                    if (a1[y] >= 0 && b1[jj] >= 0) f = cadd(f, v1[y][jj]);
                    f22[y][jj] = f;
                }
        } else if (kGather) {
#pragma unroll
            for (int y = 0; y < 2; ++y)
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) f22[y][jj] = zero;
            for (int c = nd.kid_beg; c < nd.kid_end; ++c) {
                const KidDesc kd = kids[c];
                const int32_t* inv = maps + kd.inv_beg + ne;
                const T* S = kidbase + kd.src;
                int a[2], b[4];
#pragma unroll
                for (int y = 0; y < 2; ++y) a[y] = inv[ir[y]];
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) b[jj] = inv[jc[jj]];
#pragma unroll
                for (int y = 0; y < 2; ++y)
#pragma unroll
                    for (int jj = 0; jj < 4; ++jj) {
                        T v = S[max(a[y], 0) + int64_t(max(b[jj], 0)) * kd.ld];
                        if (a[y] < 0 || b[jj] < 0) v = zero;
                        f22[y][jj] = cadd(f22[y][jj], v);
                    }
            }
        } else {
#pragma unroll
            for (int y = 0; y < 2; ++y)
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) f22[y][jj] = F[(ne + ir[y]) + int64_t(ne + jc[jj]) * n];
        }
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            const int i = tl.i0 + wi + 16 * y + lane % 16;
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                const int j = tl.j0 + wj + 16 * x + M::row(lane, jj);
                if (i < nb && j < nb) {
                    T d = f22[y][jj];
                    d.x -= cr[x][y][jj];
                    d.y -= ci[x][y][jj];
                    F[(ne + i) + int64_t(ne + j) * n] = d;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Blocked fronts (large separators): the LU of F11 and the border strips L21, U12 in steps of kBlk
// columns, three launches per step over the fronts of a level that still have columns left, with
// many workgroups per front for the two wide parts:
//  - blocked_panel_kernel, one workgroup per front: the separator rows of the panel,
//    F[k0:ne, k0:k1] (k1 = k0 + kBlk), factored with row pivoting and the pivot test of
//    field.rs:61, in sub-panels of kSub columns (rank-1 updates inside a sub-panel, one rank-kSub
//    update of the rest of the panel after it; each pass also finds the next column's pivot
//    candidates). Interchanges swap whole rows of the panel only. At k0 = 0 it first adds -H and
//    z to the gathered border and keeps the pivot scale max|F11| after the front's pivots (two
//    words, read by the later steps);
//  - blocked_row_kernel, one workgroup per kRowCols columns outside the panel: the step's row
//    interchanges as one permutation of the separator rows (all loads staged through LDS before
//    any store), and right of the panel the U rows U[k0:k1, j] = L_kk^-1 F[k0:k1, j]; or per
//    kRowCols rows of the L21 strip, l = f U_kk^-1. Both solves run in the registers of 8 lanes
//    per column or row;
//  - blocked_update_kernel, one workgroup per 64 x 64 tile: F -= L[:, k0:k1] U[k0:k1, :] on MFMA
//    over the rows and columns beyond k1, except F22 (rows and columns >= ne): schur_kernel forms
//    S = F22 - L21 U12 from the children once, with K = ne, after the last step.
// The factors end as border_kernel leaves them (LAPACK's getrf layout with the interchanges
// applied to whole separator rows), so the solves and the root's trace read them alike.
constexpr int kPanelThreads = 512, kSub = 16;
constexpr int kUpdCols = 2;  // panel columns a lane updates at once in the rank-kSub update
// Separators above this many variables take the blocked path by default.
constexpr int kBlockedMinNe = 128;

__device__ __forceinline__ void put_scale(int32_t* at, double s) {
    at[0] = __double2loint(s);
    at[1] = __double2hiint(s);
}
__device__ __forceinline__ double get_scale(const int32_t* at) { return __hiloint2double(at[1], at[0]); }

template <class T, bool kRoot>
__global__ void __launch_bounds__(kPanelThreads)
    blocked_panel_kernel(const NodeDesc* __restrict__ nodes, const int32_t* __restrict__ list,
                         const Entry* __restrict__ ent, const int32_t* __restrict__ diag, PointsView pv, double tol,
                         int k0) {
    constexpr int NT = kPanelThreads;
    __shared__ double sv[NT / 64];
    __shared__ int si[NT / 64];
    __shared__ T su[kSub];         // the pivot row inside the sub-panel
    __shared__ T us[kSub * kBlk];  // U rows of a sub-panel right of it: us[t * kBlk + j]
    __shared__ T ls[kSub * kSub];  // the sub-panel's unit-lower triangle: ls[t + kSub * t2] = L[t, t2]
    const NodeDesc nd = nodes[list[blockIdx.x]];
    const int64_t p = blockIdx.y;
    T* F = (kRoot ? reinterpret_cast<T*>(pv.root_arena) + p * pv.root_stride
                  : reinterpret_cast<T*>(pv.arena) + p * pv.arena_stride) +
           nd.off;
    int32_t* piv = pv.piv + p * pv.piv_stride + nd.piv_off;
    int32_t* status = pv.status + p * pv.status_stride + nd.slot;
    const int n = nd.n, ne = nd.ne, ne1 = nd.ne1, tid = threadIdx.x;
    auto at = [&](int i, int j) -> T& { return F[i + int64_t(j) * n]; };
    double scale;
    if (k0 == 0) {
        if (!kRoot && nd.ent_end > nd.ent_beg) {  // -H on the separator (the children were gathered before)
            add_entries<NT>(ent, nd.ent_beg, nd.ent_end, pv.vals[p], n,
                            [&](int a, double2 v) { F[a] = cadd(F[a], convert<T>(v)); });
            __syncthreads();
        }
        if (!kRoot && nd.diag_end > nd.diag_beg) {  // z on the separator's diagonal
            const T z = convert<T>(pv.z[p]);
            for (int d = nd.diag_beg + tid; d < nd.diag_end; d += NT) {
                const int64_t q = diag[d];
                F[q + q * n] = cadd(F[q + q * n], z);
            }
            __syncthreads();
        }
        double m = 0.0;
        for (int idx = tid; idx < ne1 * ne1; idx += NT) m = fmax(m, cabs(at(idx % ne1, idx / ne1)));
        scale = block_max(m, sv);
        if (tid == 0) {
            put_scale(piv + ne, scale);
            *status = kFrontOk;
        }
    } else {
        if (*status != kFrontOk) return;
        scale = get_scale(piv + ne);
    }
    const int w = min(kBlk, ne - k0), k1 = k0 + w;
    // Each pass that updates the next column to eliminate also keeps each lane's pivot candidate
    // of it (largest |entry|, first row on ties), so that no separate pass searches it; the first
    // column of the panel is searched here.
    double big = -1.0;
    int q = k0;
    for (int i = k0 + tid; i < (k0 < ne1 ? ne1 : ne); i += NT) {
        const double v = cabs2(at(i, k0));
        if (v > big) {
            big = v;
            q = i;
        }
    }
    for (int s0 = 0; s0 < w; s0 += kSub) {
        const int ws = min(kSub, w - s0), c0 = k0 + s0;  // sub-panel columns [c0, c0 + ws)
        for (int c = 0; c < ws; ++c) {
            const int k = c0 + c;
            block_argmax(big, q, sv, si);
            const bool ok = pivot_ok(big, k < ne1, tol, scale);
            if (!ok) {
                if (tid == 0) *status = (k < ne1 && !kRoot) ? kFrontDelayed : kFrontSingular;
                return;
            }
            if (tid == 0) piv[k] = q;
            if (tid < w) {  // rows k and q interchanged across the panel; the new row k from column k on to su
                const T xk = at(k, k0 + tid), xq = at(q, k0 + tid);
                if (q != k) {
                    at(k, k0 + tid) = xq;
                    at(q, k0 + tid) = xk;
                }
                const int t = k0 + tid - k;
                if (t >= 0 && t < ws - c) su[t] = xq;
            }
            __syncthreads();
            const T inv = crecip(su[0]);
            const int kn = k + 1, limn = kn < ne1 ? ne1 : ne, last = c0 + ws - 1;
            const bool track = c + 1 < ws;
            big = -1.0;
            q = kn;
#pragma unroll 1
            for (int i = kn + tid; i < ne; i += NT) {  // separator rows only; L21 comes from the row step
                T v[kSub];
#pragma unroll
                for (int t = 0; t < kSub; ++t) v[t] = at(i, min(k + t, last));
                const T l = cmul(v[0], inv);
#pragma unroll
                for (int t = 1; t < kSub; ++t) v[t] = csub(v[t], cmul(l, su[t]));
                at(i, k) = l;
#pragma unroll
                for (int t = 1; t < kSub; ++t)
                    if (t < ws - c) at(i, k + t) = v[t];
                if (track && i < limn) {
                    const double a = cabs2(v[1]);
                    if (a > big) {
                        big = a;
                        q = i;
                    }
                }
            }
            __syncthreads();
        }
        const int r0 = c0 + ws, wr = k1 - r0;  // panel columns right of the sub-panel
        if (wr <= 0) break;
        if (tid < kSub * kSub) {  // L[c0:r0, c0:r0] to LDS (strictly lower part; zero elsewhere)
            const int t = tid % kSub, t2 = tid / kSub;
            ls[tid] = (t2 < t && t < ws) ? at(c0 + t, c0 + t2) : cx<T>(0.0, 0.0);
        }
        __syncthreads();
        if (tid < wr) {  // their sub-panel rows: unit-lower solve with L[c0:r0, c0:r0] in LDS, one lane per column
            const int j = r0 + tid;
#pragma unroll
            for (int t = 0; t < kSub; ++t) us[t * kBlk + tid] = t < ws ? at(c0 + t, j) : cx<T>(0.0, 0.0);
#pragma unroll 1
            for (int t = 0; t < ws - 1; ++t) {
                const T ut = us[t * kBlk + tid];
#pragma unroll 1
                for (int t3 = t + 1; t3 < ws; ++t3)
                    us[t3 * kBlk + tid] = csub(us[t3 * kBlk + tid], cmul(ls[t3 + kSub * t], ut));
            }
#pragma unroll 1
            for (int t = 1; t < ws; ++t) at(c0 + t, j) = us[t * kBlk + tid];
        }
        __syncthreads();
        const int limn = r0 < ne1 ? ne1 : ne;
        big = -1.0;
        q = r0;
#pragma unroll 1
        for (int i = r0 + tid; i < ne; i += NT) {  // rank-ws update of the separator rows below
            T l[kSub];
#pragma unroll
            for (int t = 0; t < kSub; ++t) l[t] = t < ws ? at(i, c0 + t) : cx<T>(0.0, 0.0);
            // kUpdCols columns of the row at a time, their loads in flight together.
#pragma unroll 1
            for (int j0 = 0; j0 < wr; j0 += kUpdCols) {
                T a[kUpdCols];
#pragma unroll
                for (int u = 0; u < kUpdCols; ++u) a[u] = at(i, r0 + min(j0 + u, wr - 1));
#pragma unroll
                for (int t = 0; t < kSub; ++t) {
#pragma unroll
                    for (int u = 0; u < kUpdCols; ++u)
                        a[u] = csub(a[u], cmul(l[t], us[t * kBlk + min(j0 + u, wr - 1)]));
                    __builtin_amdgcn_sched_barrier(0);  // bounds the LDS loads in flight (registers)
                }
#pragma unroll
                for (int u = 0; u < kUpdCols; ++u)
                    if (j0 + u < wr) at(i, r0 + j0 + u) = a[u];
                if (j0 == 0 && i < limn) {
                    const double v = cabs2(a[0]);
                    if (v > big) {
                        big = v;
                        q = i;
                    }
                }
            }
        }
        __syncthreads();
    }
}


template <class T, bool kRoot>
__global__ void __launch_bounds__(256)
    blocked_row_kernel(const NodeDesc* __restrict__ nodes, const Tile* __restrict__ tiles, PointsView pv, int k0) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ int32_t sp[kBlk], psrc[kBlk], opos[kBlk], osrc[kBlk];
    T* X = reinterpret_cast<T*>(lds_raw);  // new panel rows, w x kRowCols; then the triangle L_kk
    T* O = X + kBlk * kRowCols;            // new rows below the panel, w x kRowCols
    const TilePoint tp = xcd_tile();
    const Tile tl = tiles[tp.tile];
    const NodeDesc nd = nodes[tl.node];
    const int64_t p = tp.point;
    if (pv.status[p * pv.status_stride + nd.slot] != kFrontOk) return;
    T* F = (kRoot ? reinterpret_cast<T*>(pv.root_arena) + p * pv.root_stride
                  : reinterpret_cast<T*>(pv.arena) + p * pv.arena_stride) +
           nd.off;
    const int32_t* piv = pv.piv + p * pv.piv_stride + nd.piv_off;
    const int n = nd.n, ne = nd.ne, tid = threadIdx.x, w = min(kBlk, ne - k0), k1 = k0 + w;
    auto at = [&](int i, int j) -> T& { return F[i + int64_t(j) * n]; };
    constexpr int la = kBlk + 1;
    if (tl.pad == 2) {  // rows j0.. of the L21 strip: l U_kk = f, eight lanes per row (U_kk in LDS,
                        // 1 / U[k][k] in place of its diagonal)
        T* A = X;
        copy_batched<256, T>(w * w, [&](int e) { return at(k0 + e % w, k0 + e / w); },
                             [&](int e, T v) { A[(e % w) + (e / w) * la] = e % w == e / w ? crecip(v) : v; });
        __syncthreads();
        const int rows = min(kRowCols, n - tl.j0), grp = tid / 8, g = tid % 8, row = tl.j0 + min(grp, rows - 1);
        T x[8];
#pragma unroll
        for (int s = 0; s < 8; ++s) x[s] = at(row, k0 + min(g + 8 * s, w - 1));
#pragma unroll
        for (int kb = 0; kb < 8; ++kb)
#pragma unroll 1
            for (int kk = 0; kk < 8; ++kk) {  // a loop (x[kb] stays a register): an eighth of the code
                const int k = 8 * kb + kk;
                if (k < w) {
                    const T cand = cmul(x[kb], A[k + k * la]);
                    T xk;
                    xk.x = __shfl(cand.x, kk, 8);
                    xk.y = __shfl(cand.y, kk, 8);
                    if (g == kk) x[kb] = xk;
#pragma unroll
                    for (int s = 0; s < 8; ++s) {
                        const int j = g + 8 * s;
                        if (j > k && j < w) x[s] = csub(x[s], cmul(xk, A[k + j * la]));
                    }
                }
            }
        if (grp < rows)
#pragma unroll
            for (int s = 0; s < 8; ++s)
                if (g + 8 * s < w) at(row, k0 + g + 8 * s) = x[s];
        return;
    }
    const bool right = tl.pad != 0;
    const int j0 = tl.j0, cols = min(kRowCols, (right ? n : k0) - j0);
    if (tid < w) sp[tid] = piv[k0 + tid];
    __syncthreads();
    if (tid < w) {
        psrc[tid] = traced_row(k0 + tid, sp, k0, w);
    } else if (tid >= 64 && tid < 64 + w) {  // a row below the panel that an interchange reached
        const int t = tid - 64, pos = sp[t];
        bool first = pos >= k1;
        for (int c = 0; c < t; ++c) first = first && sp[c] != pos;
        opos[t] = first ? pos : -1;
        osrc[t] = first ? traced_row(pos, sp, k0, w) : pos;
    }
    __syncthreads();
    copy_batched<256, T>(w * cols, [&](int e) { return at(psrc[e % w], j0 + e / w); },
                         [&](int e, T v) { X[e % w + (e / w) * kBlk] = v; });
    copy_batched<256, T>(w * cols, [&](int e) { return opos[e % w] >= 0 ? at(osrc[e % w], j0 + e / w) : cx<T>(0.0, 0.0); },
                         [&](int e, T v) {
                             if (opos[e % w] >= 0) O[e % w + (e / w) * kBlk] = v;
                         });
    __syncthreads();
    for (int e = tid; e < w * cols; e += 256) {
        const int r = e % w, jc = e / w;
        if (opos[r] >= 0) at(opos[r], j0 + jc) = O[r + jc * kBlk];
        if (!right && psrc[r] != k0 + r) at(k0 + r, j0 + jc) = X[r + jc * kBlk];
    }
    if (!right) return;
    // U rows: eight lanes per column, lane g holding rows g + 8 s of it.
    const int grp = tid / 8, g = tid % 8, col = min(grp, cols - 1);
    T x[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) x[s] = X[min(g + 8 * s, w - 1) + col * kBlk];
    __syncthreads();
    T* A = X;  // L_kk, leading dimension la
    copy_batched<256, T>(w * w, [&](int e) { return at(k0 + e % w, k0 + e / w); },
                         [&](int e, T v) { A[(e % w) + (e / w) * la] = v; });
    __syncthreads();
#pragma unroll
    for (int kb = 0; kb < 8; ++kb)
#pragma unroll
        for (int kk = 0; kk < 8; ++kk) {
            const int k = 8 * kb + kk;
            if (k < w - 1) {
                T xk;
                xk.x = __shfl(x[kb].x, kk, 8);
                xk.y = __shfl(x[kb].y, kk, 8);
#pragma unroll
                for (int s = 0; s < 8; ++s) {
                    const int i = g + 8 * s;
                    if (i > k && i < w) x[s] = csub(x[s], cmul(A[i + k * la], xk));
                }
            }
        }
    if (grp < cols)
#pragma unroll
        for (int s = 0; s < 8; ++s)
            if (g + 8 * s < w) at(k0 + g + 8 * s, j0 + grp) = x[s];
}

template <class T, bool kRoot>
__global__ void __launch_bounds__(256)
    blocked_update_kernel(const NodeDesc* __restrict__ nodes, const Tile* __restrict__ tiles, PointsView pv, int k0) {
    using M = Mfma<T>;
    __shared__ T As[kChunk * kTileLd];  // As[k * kTileLd + i] = L[i0 + i, k0 + kc + k]
    __shared__ T Bs[kChunk * kTileLd];  // Bs[k * kTileLd + j] = U[k0 + kc + k, j0 + j]
    const TilePoint tp = xcd_tile();
    const Tile tl = tiles[tp.tile];
    const NodeDesc nd = nodes[tl.node];
    const int64_t p = tp.point;
    if (pv.status[p * pv.status_stride + nd.slot] != kFrontOk) return;
    T* F = (kRoot ? reinterpret_cast<T*>(pv.root_arena) + p * pv.root_stride
                  : reinterpret_cast<T*>(pv.arena) + p * pv.arena_stride) +
           nd.off;
    const int n = nd.n, ne = nd.ne, w = min(kBlk, ne - k0), tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
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
    // The K loop in chunks of kChunk, as in schur_kernel: the next chunk's L / U entries are loaded
    // into registers (kStage per thread and operand) while the MFMAs run on the current one from
    // LDS. Addresses are clamped into the front; entries outside it are replaced by zero.
    constexpr int kStage = kTile * kChunk / 256;
    T sa[kStage], sb[kStage];
    auto fetch = [&](int kc) {
#pragma unroll
        for (int u = 0; u < kStage; ++u) {
            const int e = tid + u * 256;
            {
                const int r = e % kTile, c = e / kTile, i = tl.i0 + r, k = kc + c;
                T v = F[min(i, n - 1) + int64_t(k0 + min(k, w - 1)) * n];
                if (!(i < n && k < w)) v = zero;
                sa[u] = v;
            }
            {
                const int c = e % kChunk, r = e / kChunk, j = tl.j0 + r, k = kc + c;
                T v = F[(k0 + min(k, w - 1)) + int64_t(min(j, n - 1)) * n];
                if (!(j < n && k < w)) v = zero;
                sb[u] = v;
            }
        }
    };
    // The wave's 16 x 16 blocks with entries to update: inside the front and not wholly in F22
    // (rows and columns >= ne, left to schur_kernel); the others are skipped (wave-uniform).
    bool live[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            const int ib = tl.i0 + wi + 16 * y, jb = tl.j0 + wj + 16 * x;
            live[x][y] = ib < n && jb < n && (ib < ne || jb < ne);
        }
    fetch(0);
    for (int kc = 0; kc < w; kc += kChunk) {
#pragma unroll
        for (int u = 0; u < kStage; ++u) {
            const int e = tid + u * 256;
            As[(e / kTile) * kTileLd + e % kTile] = sa[u];
            Bs[(e % kChunk) * kTileLd + e / kChunk] = sb[u];
        }
        __syncthreads();
        if (kc + kChunk < w) fetch(kc + kChunk);
#pragma unroll
        for (int kk = 0; kk < kChunk; kk += 4) {
            if (kc + kk >= w) break;  // the rest of the chunk is padding (zeros)
            const int k = kk + lane / 16;
            T a[2], b[2];
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                a[s] = Bs[k * kTileLd + wj + 16 * s + lane % 16];  // U^T operand
                b[s] = As[k * kTileLd + wi + 16 * s + lane % 16];  // L^T operand
            }
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y)
                    if (live[x][y]) {
                        cr[x][y] = M::mac(a[x].x, b[y].x, cr[x][y]);
                        cr[x][y] = M::mac(-a[x].y, b[y].y, cr[x][y]);
                        ci[x][y] = M::mac(a[x].x, b[y].y, ci[x][y]);
                        ci[x][y] = M::mac(a[x].y, b[y].x, ci[x][y]);
                    }
        }
        __syncthreads();
    }
    // Block (x, y) of the wave: lane l, element jj holds (L U)[i, j] with i = wi + 16 y + l % 16 and
    // j = wj + 16 x + row(l, jj) (tile-relative); F22 entries are left to schur_kernel.
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            const int i = tl.i0 + wi + 16 * y + lane % 16;
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                const int j = tl.j0 + wj + 16 * x + M::row(lane, jj);
                if (i < n && j < n && (i < ne || j < ne)) {
                    T d = F[i + int64_t(j) * n];
                    d.x -= cr[x][y][jj];
                    d.y -= ci[x][y][jj];
                    F[i + int64_t(j) * n] = d;
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

// T = |W_L^H G_LR W_R|_F^2 = Tr[Gamma_L G Gamma_R G^H] with Gamma = W W^H: columns of X = G_LR W_R
// in chunks that fit the LDS, then the squared moduli of W_L^H X. Work 8 nl nr cr + 8 cl nl cr
// flops instead of 8 nl nr (nl + nr).
constexpr int kTraceLds = 64 * 1024;
__global__ void __launch_bounds__(kThreads) trace_w_kernel(const NodeDesc* __restrict__ rootp, PointsView pv) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ double sv[kThreads / 64];
    double2* X = reinterpret_cast<double2*>(lds_raw);
    const NodeDesc nd = *rootp;
    const int64_t p = blockIdx.x;
    const double2* F = reinterpret_cast<const double2*>(pv.root_arena) + p * pv.root_stride + nd.off;
    const int n = nd.n, k0 = nd.ne, nl = nd.nl, nr = nd.nr, tid = threadIdx.x;
    const double2* wl = pv.w + p * pv.sigma_stride;
    const double2* wr = wl + int64_t(nl) * nl;
    const int cl = pv.wcols[2 * p], cr = pv.wcols[2 * p + 1];
    const int chunk = max(1, min(cr, kTraceLds / int(sizeof(double2)) / max(nl, 1)));
    // The border block holds -G_LR; the sign cancels in |.|^2.
    double acc = 0.0;
    for (int c0 = 0; c0 < cr; c0 += chunk) {
        const int cc = min(chunk, cr - c0);
        for (int idx = tid; idx < nl * cc; idx += kThreads) {
            const int a = idx % nl, c = c0 + idx / nl;
            double2 x = cx<double2>(0, 0);
            for (int b = 0; b < nr; ++b) x = cadd(x, cmul(F[(k0 + a) + int64_t(k0 + b) * n], wr[b + int64_t(c) * nr]));
            X[idx] = x;
        }
        __syncthreads();
        for (int idx = tid; idx < cl * cc; idx += kThreads) {
            const int d = idx % cl, c = idx / cl;
            double2 y = cx<double2>(0, 0);
            for (int a = 0; a < nl; ++a) {
                double2 w = wl[a + int64_t(d) * nl];
                w.y = -w.y;
                y = cadd(y, cmul(w, X[a + c * nl]));
            }
            acc += y.x * y.x + y.y * y.y;
        }
        __syncthreads();
    }
    const double t = block_sum(acc, sv);
    if (tid == 0) {
        const int32_t st = pv.status[p * pv.status_stride + nd.slot];
        pv.t_out[p] = st == kFrontOk ? t : NAN;
    }
}

// Trace as two batched complex GEMMs on MFMA when the leads carry many channels:
//   X = G_LR W_R (nl x cr)   then   T = |W_L^H X|_F^2,
// 64 x 64 tiles per workgroup as in schur_kernel (four 32 x 32 wave quarters, K staged through LDS
// in chunks of kChunk). W columns past a point's channel count are zero, so the tiles run to the
// largest count of the batch.
template <bool kConjA>
__device__ __forceinline__ void tile_zgemm(const double2* __restrict__ A, int lda, const double2* __restrict__ B,
                                           int ldb, int M, int N, int K, int i0, int j0, double2* As, double2* Bs,
                                           f64x4 (&cr)[2][2], f64x4 (&ci)[2][2]) {
    using M4 = Mfma<double2>;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wi = (wave & 1) * 32, wj = (wave >> 1) * 32;
    const double2 zero = cx<double2>(0.0, 0.0);
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            cr[x][y] = f64x4{0, 0, 0, 0};
            ci[x][y] = f64x4{0, 0, 0, 0};
        }
    for (int kc = 0; kc < K; kc += kChunk) {
        for (int e = tid; e < kTile * kChunk; e += 256) {
            // As[k][i] = op(A)(i0 + i, kc + k); clamped loads, values selected.
            int r, c;
            if (kConjA) {
                c = e % kChunk;
                r = e / kChunk;
            } else {
                r = e % kTile;
                c = e / kTile;
            }
            const int i = min(i0 + r, M - 1), k = min(kc + c, K - 1);
            double2 v = kConjA ? A[k + int64_t(i) * lda] : A[i + int64_t(k) * lda];
            if (kConjA) v.y = -v.y;
            if (!(i0 + r < M && kc + c < K)) v = zero;
            As[c * kTileLd + r] = v;
        }
        for (int e = tid; e < kTile * kChunk; e += 256) {
            const int c = e % kChunk, r = e / kChunk;
            const int j = min(j0 + r, N - 1), k = min(kc + c, K - 1);
            double2 v = B[k + int64_t(j) * ldb];
            if (!(j0 + r < N && kc + c < K)) v = zero;
            Bs[c * kTileLd + r] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kChunk; kk += 4) {
            const int k = kk + lane / 16;
            double2 a[2], b[2];
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                a[s] = Bs[k * kTileLd + wj + 16 * s + lane % 16];
                b[s] = As[k * kTileLd + wi + 16 * s + lane % 16];
            }
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y) {
                    cr[x][y] = M4::mac(a[x].x, b[y].x, cr[x][y]);
                    cr[x][y] = M4::mac(-a[x].y, b[y].y, cr[x][y]);
                    ci[x][y] = M4::mac(a[x].x, b[y].y, ci[x][y]);
                    ci[x][y] = M4::mac(a[x].y, b[y].x, ci[x][y]);
                }
        }
        __syncthreads();
    }
}

// X[p] = G_LR W_R; grid (tiles of nl x cr, points).
__global__ void __launch_bounds__(256)
    trace_x_kernel(const NodeDesc* __restrict__ rootp, PointsView pv, double2* __restrict__ X, int cr_max, int tiles_j) {
    __shared__ double2 As[kChunk * kTileLd];
    __shared__ double2 Bs[kChunk * kTileLd];
    const NodeDesc nd = *rootp;
    const int64_t p = blockIdx.y;
    const int nl = nd.nl, nr = nd.nr, n = nd.n, k0 = nd.ne;
    const int i0 = kTile * (blockIdx.x / tiles_j), j0 = kTile * (blockIdx.x % tiles_j);
    const double2* G = reinterpret_cast<const double2*>(pv.root_arena) + p * pv.root_stride + nd.off + k0 +
                       int64_t(k0) * n;  // the border block: -G_LR
    const double2* WR = pv.w + p * pv.sigma_stride + int64_t(nl) * nl;
    f64x4 cr[2][2], ci[2][2];
    tile_zgemm<false>(G, n, WR, nr, nl, cr_max, nr, i0, j0, As, Bs, cr, ci);
    double2* Xp = X + p * int64_t(nl) * cr_max;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, wi = (wave & 1) * 32, wj = (wave >> 1) * 32;
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            const int i = i0 + wi + 16 * y + lane % 16;
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                const int j = j0 + wj + 16 * x + Mfma<double2>::row(lane, jj);
                if (i < nl && j < cr_max) Xp[i + int64_t(j) * nl] = double2{cr[x][y][jj], ci[x][y][jj]};
            }
        }
}

// partial[p][tile] = sum of |W_L^H X|^2 over the tile; grid (tiles of cl x cr, points).
__global__ void __launch_bounds__(256)
    trace_y_kernel(const NodeDesc* __restrict__ rootp, PointsView pv, const double2* __restrict__ X, int cl_max,
                   int cr_max, int tiles_j, double* __restrict__ partial) {
    __shared__ double2 As[kChunk * kTileLd];
    __shared__ double2 Bs[kChunk * kTileLd];
    __shared__ double sv[4];
    const NodeDesc nd = *rootp;
    const int64_t p = blockIdx.y;
    const int nl = nd.nl;
    const int i0 = kTile * (blockIdx.x / tiles_j), j0 = kTile * (blockIdx.x % tiles_j);
    const double2* WL = pv.w + p * pv.sigma_stride;
    f64x4 cr[2][2], ci[2][2];
    tile_zgemm<true>(WL, nl, X + p * int64_t(nl) * cr_max, nl, cl_max, cr_max, nl, i0, j0, As, Bs, cr, ci);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, wi = (wave & 1) * 32, wj = (wave >> 1) * 32;
    double acc = 0.0;
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            const int i = i0 + wi + 16 * y + lane % 16;
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                const int j = j0 + wj + 16 * x + Mfma<double2>::row(lane, jj);
                if (i < cl_max && j < cr_max) acc += cr[x][y][jj] * cr[x][y][jj] + ci[x][y][jj] * ci[x][y][jj];
            }
        }
    const double t = block_sum(acc, sv);
    if (threadIdx.x == 0) partial[p * gridDim.x + blockIdx.x] = t;
}

__global__ void __launch_bounds__(64)
    trace_sum_kernel(const NodeDesc* __restrict__ rootp, PointsView pv, const double* __restrict__ partial, int tiles) {
    const NodeDesc nd = *rootp;
    const int64_t p = blockIdx.x;
    double acc = 0.0;
    for (int t = threadIdx.x; t < tiles; t += 64) acc += partial[p * tiles + t];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (threadIdx.x == 0) {
        const int32_t st = pv.status[p * pv.status_stride + nd.slot];
        pv.t_out[p] = st == kFrontOk ? acc : NAN;
    }
}

// Multi-terminal roots: T[a][b] = Tr[Gamma_a G Gamma_b G^H] = |W_a^H G W_b|_F^2 between every
// pair of labels (leads, or sectors of leads). X = G W over all labels' channels at once (the
// tiles of trace_x_kernel), then Y = W^H X with |Y|^2 stored entry by entry, then per pair of
// labels the sum over its block of Y, one thread per pair in a fixed order (no atomics).
__global__ void __launch_bounds__(256)
    label_x_kernel(const NodeDesc* __restrict__ rootp, PointsView pv, double2* __restrict__ X, int cols, int tiles_j) {
    __shared__ double2 As[kChunk * kTileLd];
    __shared__ double2 Bs[kChunk * kTileLd];
    const NodeDesc nd = *rootp;
    const int64_t p = blockIdx.y;
    const int k = nd.nl, n = nd.n, k0 = nd.ne;
    const int i0 = kTile * (blockIdx.x / tiles_j), j0 = kTile * (blockIdx.x % tiles_j);
    const double2* G = reinterpret_cast<const double2*>(pv.root_arena) + p * pv.root_stride + nd.off + k0 +
                       int64_t(k0) * n;  // the border block: -G on the kept orbitals
    const double2* W = pv.wlab + p * pv.wlab_stride;
    f64x4 cr[2][2], ci[2][2];
    tile_zgemm<false>(G, n, W, k, k, cols, k, i0, j0, As, Bs, cr, ci);
    double2* Xp = X + p * int64_t(k) * cols;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, wi = (wave & 1) * 32, wj = (wave >> 1) * 32;
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            const int i = i0 + wi + 16 * y + lane % 16;
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                const int j = j0 + wj + 16 * x + Mfma<double2>::row(lane, jj);
                if (i < k && j < cols) Xp[i + int64_t(j) * k] = double2{cr[x][y][jj], ci[x][y][jj]};
            }
        }
}

__global__ void __launch_bounds__(256)
    label_y_kernel(const NodeDesc* __restrict__ rootp, PointsView pv, const double2* __restrict__ X, int cols,
                   int tiles_j, double* __restrict__ ysq) {
    __shared__ double2 As[kChunk * kTileLd];
    __shared__ double2 Bs[kChunk * kTileLd];
    const NodeDesc nd = *rootp;
    const int64_t p = blockIdx.y;
    const int k = nd.nl;
    const int i0 = kTile * (blockIdx.x / tiles_j), j0 = kTile * (blockIdx.x % tiles_j);
    const double2* W = pv.wlab + p * pv.wlab_stride;
    f64x4 cr[2][2], ci[2][2];
    tile_zgemm<true>(W, k, X + p * int64_t(k) * cols, k, cols, cols, k, i0, j0, As, Bs, cr, ci);
    double* Yp = ysq + p * int64_t(cols) * cols;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, wi = (wave & 1) * 32, wj = (wave >> 1) * 32;
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            const int i = i0 + wi + 16 * y + lane % 16;
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                const int j = j0 + wj + 16 * x + Mfma<double2>::row(lane, jj);
                if (i < cols && j < cols)
                    Yp[i + int64_t(j) * cols] = cr[x][y][jj] * cr[x][y][jj] + ci[x][y][jj] * ci[x][y][jj];
            }
        }
}

__global__ void __launch_bounds__(64)
    label_sum_kernel(const NodeDesc* __restrict__ rootp, PointsView pv, const double* __restrict__ ysq, int cols) {
    const NodeDesc nd = *rootp;
    const int64_t p = blockIdx.x;
    const int L = pv.nlab;
    const int32_t* off = pv.lab_off + p * (L + 1);
    const double* Yp = ysq + p * int64_t(cols) * cols;
    const bool ok = pv.status[p * pv.status_stride + nd.slot] == kFrontOk;
    for (int pair = threadIdx.x; pair < L * L; pair += 64) {
        const int a = pair / L, b = pair % L;
        double acc = 0.0;
        for (int j = off[b]; j < off[b + 1]; ++j)
            for (int i = off[a]; i < off[a + 1]; ++i) acc += Yp[i + int64_t(j) * cols];
        pv.t_lab[p * int64_t(L) * L + pair] = ok ? acc : NAN;
    }
}

inline dim3 grid2(int32_t count, int32_t points) { return dim3(unsigned(count), unsigned(points)); }

// Dynamic LDS above the 64 KB default (gfx950 has 160 KB per workgroup), once per kernel.
void allow_lds(const void* kernel) {
    static std::mutex mu;
    static std::set<const void*> done;
    std::lock_guard<std::mutex> g(mu);
    if (done.insert(kernel).second)
        (void)hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, 128 * 1024);
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
                          const int32_t* diag, const KidDesc* kids, const int32_t* maps, const Tile* tiles,
                          int32_t tile_count, const PointsView& pv, int32_t points, int single, hipStream_t st) {
    if (count <= 0 || points <= 0) return;
    if (tile_count > 0) {  // the whole front by tiles, then -H and z
        if (single) {
            hipLaunchKernelGGL(root_gather_kernel<float2>, grid2(tile_count, points), dim3(256), 0, st, nodes, tiles, kids,
                               maps, pv);
            hipLaunchKernelGGL((assemble_kernel<double2, float2, true, false>), grid2(count, points), dim3(kThreads), 0,
                               st, nodes, level, ent, diag, kids, maps, pv);
        } else {
            hipLaunchKernelGGL(root_gather_kernel<double2>, grid2(tile_count, points), dim3(256), 0, st, nodes, tiles,
                               kids, maps, pv);
            hipLaunchKernelGGL((assemble_kernel<double2, double2, true, false>), grid2(count, points), dim3(kThreads), 0,
                               st, nodes, level, ent, diag, kids, maps, pv);
        }
        return;
    }
    if (single)
        hipLaunchKernelGGL((assemble_kernel<double2, float2, true>), grid2(count, points), dim3(kThreads), 0, st, nodes,
                           level, ent, diag, kids, maps, pv);
    else
        hipLaunchKernelGGL((assemble_kernel<double2, double2, true>), grid2(count, points), dim3(kThreads), 0, st, nodes,
                           level, ent, diag, kids, maps, pv);
}

namespace {
std::mutex tuning_mu;
bool tuning_set = false;
Tuning tuning_override;

/// The defaults and the CAROLI_* variables: CAROLI_SMALL_FRONTS=lds|reg, CAROLI_FRONT_KERNEL=
/// blocked (every large front and the root blocked) | border (no f11 / trsm split, none blocked) |
/// unblocked, CAROLI_BLOCKED_MIN_NE=<separator size above which fronts go blocked>,
/// CAROLI_SCHUR_TILE=64|32 (unset: by level).
Tuning env_tuning() {
    Tuning t;
    t.blocked_min_ne = kBlockedMinNe;
    if (const char* v = std::getenv("CAROLI_SMALL_FRONTS")) {
        if (std::strcmp(v, "lds") == 0) t.small_fronts = 1;
        if (std::strcmp(v, "reg") == 0) t.small_fronts = 2;
    }
    if (const char* v = std::getenv("CAROLI_FRONT_KERNEL")) {
        if (std::strcmp(v, "blocked") == 0) t.blocked_min_ne = 0;
        if (std::strcmp(v, "border") == 0) {
            t.split = 0;
            t.blocked_min_ne = 1 << 30;
        }
        if (std::strcmp(v, "unblocked") == 0) {
            t.unblocked = 1;
            t.blocked_min_ne = 1 << 30;
        }
    }
    if (const char* v = std::getenv("CAROLI_BLOCKED_MIN_NE"))
        if (*v) t.blocked_min_ne = std::max(0, std::atoi(v));
    if (const char* v = std::getenv("CAROLI_SCHUR_TILE")) {
        if (std::atoi(v) == 32) t.schur_tile = 32;
        if (std::atoi(v) == 64) t.schur_tile = 64;
    }
    return t;
}
}  // namespace

Tuning current_tuning() {
    std::lock_guard<std::mutex> g(tuning_mu);
    return tuning_set ? tuning_override : env_tuning();
}

void set_tuning_override(const Tuning* t) {
    std::lock_guard<std::mutex> g(tuning_mu);
    tuning_set = t != nullptr;
    if (t) tuning_override = *t;
}

bool unblocked_fronts() { return current_tuning().unblocked != 0; }

bool border_only() { return current_tuning().split == 0; }

int blocked_min_ne() { return current_tuning().blocked_min_ne; }

void launch_blocked_panel(const NodeDesc* nodes, const int32_t* list, int32_t count, const Entry* ent,
                          const int32_t* diag, const PointsView& pv, int32_t points, double tol, bool single, bool root,
                          int32_t k0, hipStream_t st) {
    if (count <= 0 || points <= 0) return;
    const dim3 g = grid2(count, points), b(kPanelThreads);
    auto go = [&](auto kernel) { hipLaunchKernelGGL(kernel, g, b, 0, st, nodes, list, ent, diag, pv, tol, k0); };
    if (root)
        go(blocked_panel_kernel<double2, true>);
    else if (single)
        go(blocked_panel_kernel<float2, false>);
    else
        go(blocked_panel_kernel<double2, false>);
}

void launch_blocked_row(const NodeDesc* nodes, const Tile* tiles, int32_t count, const PointsView& pv, int32_t points,
                        bool single, bool root, int32_t k0, hipStream_t st) {
    if (count <= 0 || points <= 0) return;
    // The moved rows (2 kBlk x kRowCols), then the triangle (kBlk x (kBlk + 1)).
    const size_t words = size_t(std::max(2 * kBlk * kRowCols, kBlk * (kBlk + 1)));
    const size_t lds = words * ((single && !root) ? sizeof(float2) : sizeof(double2));
    const dim3 g = grid2(count, points), b(256);
    auto go = [&](auto kernel) {
        allow_lds(reinterpret_cast<const void*>(kernel));
        hipLaunchKernelGGL(kernel, g, b, lds, st, nodes, tiles, pv, k0);
    };
    if (root)
        go(blocked_row_kernel<double2, true>);
    else if (single)
        go(blocked_row_kernel<float2, false>);
    else
        go(blocked_row_kernel<double2, false>);
}

void launch_blocked_update(const NodeDesc* nodes, const Tile* tiles, int32_t count, const PointsView& pv,
                           int32_t points, bool single, bool root, int32_t k0, hipStream_t st) {
    if (count <= 0 || points <= 0) return;
    const dim3 g = grid2(count, points), b(256);
    auto go = [&](auto kernel) { hipLaunchKernelGGL(kernel, g, b, 0, st, nodes, tiles, pv, k0); };
    if (root)
        go(blocked_update_kernel<double2, true>);
    else if (single)
        go(blocked_update_kernel<float2, false>);
    else
        go(blocked_update_kernel<double2, false>);
}

void launch_front_lds(const NodeDesc* nodes, const int32_t* level, int32_t count, int32_t nmax, int32_t border_words,
                      const Entry* ent, const int32_t* diag, const KidDesc* kids, const int32_t* maps,
                      const PointsView& pv, int32_t points, double tol, bool single, bool assemble, hipStream_t st) {
    if (count <= 0 || points <= 0) return;
    if (nmax > kLdsMaxN) throw std::invalid_argument("front too large for the LDS kernel");
    const size_t es = single ? sizeof(float2) : sizeof(double2);
    const int mode = small_front_mode();
    const size_t lds = mode == 0 ? size_t(border_words) * es : size_t(nmax) * size_t(nmax + 1) * es;
    // Lanes per row and workgroup size by the largest front: 16 rows x 4 (one wave), 32 x 2 (one
    // wave), 64 x 4, 128 x 2 (four waves). The default holds the border only in LDS
    // (front_border_kernel); CAROLI_SMALL_FRONTS=lds holds the whole front (front_lds_kernel), =reg
    // eliminates fronts up to 64 in registers (front_reg_kernel).
    auto go = [&](auto kernel, int nt) {
        allow_lds(reinterpret_cast<const void*>(kernel));
        hipLaunchKernelGGL(kernel, grid2(count, points), dim3(nt), lds, st, nodes, level, ent, diag, kids, maps, pv, tol);
    };
    auto pick = [&](auto t, auto a) {
        using T = decltype(t);
        constexpr bool A = decltype(a)::value;
        if (mode == 2 && nmax <= 64) {
            if (nmax <= 16)
                go(front_reg_kernel<T, 1, 16, A>, 64);
            else if (nmax <= 32)
                go(front_reg_kernel<T, 2, 16, A>, 128);
            else
                go(front_reg_kernel<T, 4, 16, A>, 256);
        } else if (mode == 0) {
            if (nmax <= 16)
                go(front_border_kernel<T, 64, 4, A>, 64);
            else if (nmax <= 32)
                go(front_border_kernel<T, 64, 2, A>, 64);
            else if (nmax <= 64)  // two waves, two lanes a row (four waves: 17.3 against 17.1 ms a C2 step)
                go(front_border_kernel<T, 128, 2, A>, 128);
            else
                go(front_border_kernel<T, 256, 2, A>, 256);
        } else if (nmax <= 16)
            go(front_lds_kernel<T, 64, 4, A>, 64);
        else if (nmax <= 32)
            go(front_lds_kernel<T, 64, 2, A>, 64);
        else if (nmax <= 64)
            go(front_lds_kernel<T, 256, 4, A>, 256);
        else
            go(front_lds_kernel<T, 256, 2, A>, 256);
    };
    if (single)
        assemble ? pick(float2{}, std::true_type{}) : pick(float2{}, std::false_type{});
    else
        assemble ? pick(double2{}, std::true_type{}) : pick(double2{}, std::false_type{});
}

int small_front_mode() { return current_tuning().small_fronts; }


int32_t border_lds_rows(int32_t max_ne, bool single) {
    const int32_t cap = kBorderLdsBytes / (kPanel * int32_t(single ? sizeof(float2) : sizeof(double2)));
    return max_ne <= cap ? max_ne : 0;
}

void launch_gather(const NodeDesc* nodes, const Tile* tiles, int32_t count, const KidDesc* kids, const int32_t* maps,
                   const PointsView& pv, int32_t points, bool single, hipStream_t st) {
    if (count <= 0 || points <= 0) return;
    if (single)
        hipLaunchKernelGGL(gather_kernel<float2>, grid2(count, points), dim3(256), 0, st, nodes, tiles, kids, maps, pv);
    else
        hipLaunchKernelGGL(gather_kernel<double2>, grid2(count, points), dim3(256), 0, st, nodes, tiles, kids, maps, pv);
}

void launch_border(const NodeDesc* nodes, const int32_t* level, int32_t count, int32_t max_ne, const Entry* ent,
                   const int32_t* diag, const PointsView& pv, int32_t points, double tol, bool single, bool root,
                   hipStream_t st) {
    if (count <= 0 || points <= 0) return;
    const int32_t rows = border_lds_rows(max_ne, single);
    const size_t lds = size_t(rows * kPanel) * (single ? sizeof(float2) : sizeof(double2));
    const dim3 g = grid2(count, points);
    auto go = [&](auto kernel, int nt) {
        allow_lds(reinterpret_cast<const void*>(kernel));
        hipLaunchKernelGGL(kernel, g, dim3(nt), lds, st, nodes, level, ent, diag, pv, tol, rows);
    };
    if (single)
        go(border_kernel<float2, false, kBorderThreads>, kBorderThreads);
    else if (root)
        go(border_kernel<double2, true, kRootThreads>, kRootThreads);
    else
        go(border_kernel<double2, false, kBorderThreads>, kBorderThreads);
}

void launch_f11(const NodeDesc* nodes, const int32_t* level, int32_t count, int32_t max_ne, const Entry* ent,
                const int32_t* diag, const PointsView& pv, int32_t points, double tol, bool single, hipStream_t st) {
    if (count <= 0 || points <= 0) return;
    if (max_ne > kLdsMaxN) throw std::invalid_argument("separator too large for f11_kernel");
    const size_t lds = size_t(max_ne) * size_t(max_ne + 1) * (single ? sizeof(float2) : sizeof(double2));
    auto go = [&](auto kernel, int nt) {
        allow_lds(reinterpret_cast<const void*>(kernel));
        hipLaunchKernelGGL(kernel, grid2(count, points), dim3(nt), lds, st, nodes, level, ent, diag, pv, tol);
    };
    // Complex f64 separators up to 64: two waves, two lanes a row; fewer waves shorten each
    // elimination step (6.7 -> 6.2 ms a BHZ 200 x 64 step against four waves and four lanes a row).
    if (single)
        max_ne <= 64 ? go(f11_kernel<float2, 256, 4>, 256) : go(f11_kernel<float2, 256, 2>, 256);
    else
        max_ne <= 64 ? go(f11_kernel<double2, 128, 2>, 128) : go(f11_kernel<double2, 256, 2>, 256);
}

void launch_trsm(const NodeDesc* nodes, const Tile* tiles, int32_t count, int32_t max_ne, const PointsView& pv,
                 int32_t points, bool single, hipStream_t st) {
    if (count <= 0 || points <= 0) return;
    if (max_ne > kF11Lds) throw std::invalid_argument("separator too large for trsm_kernel");
    const size_t lds = size_t(max_ne) * size_t(max_ne + 1) / 2 * (single ? sizeof(float2) : sizeof(double2));
    auto go = [&](auto kernel) {
        allow_lds(reinterpret_cast<const void*>(kernel));
        hipLaunchKernelGGL(kernel, grid2(count, points), dim3(256), lds, st, nodes, tiles, pv);
    };
    auto pick = [&](auto t) {  // register slots per lane: ne <= 4 NS
        using T = decltype(t);
        if (max_ne <= 16)
            go(trsm_kernel<T, 4>);
        else if (max_ne <= 32)
            go(trsm_kernel<T, 8>);
        else if (max_ne <= 48)
            go(trsm_kernel<T, 12>);
        else
            go(trsm_kernel<T, 16>);
    };
    single ? pick(float2{}) : pick(double2{});
}

void launch_schur(const NodeDesc* nodes, const Tile* tiles, int32_t count, const KidDesc* kids, const int32_t* maps,
                  const PointsView& pv, int32_t points, bool single, bool root, bool gather, int tile, hipStream_t st) {
    if (count <= 0 || points <= 0) return;
    if (tile != 64 && tile != 32) throw std::runtime_error("launch_schur: tiles are 64 or 32 wide");
    const dim3 g = grid2(count, points), b(tile == 64 ? 256 : 64);
    auto go = [&](auto kernel) { hipLaunchKernelGGL(kernel, g, b, 0, st, nodes, tiles, kids, maps, pv); };
    auto pick = [&](auto ts) {
        constexpr int TS = decltype(ts)::value;
        if (root)
            go(schur_kernel<double2, true, false, TS>);
        else if (single)
            gather ? go(schur_kernel<float2, false, true, TS>) : go(schur_kernel<float2, false, false, TS>);
        else
            gather ? go(schur_kernel<double2, false, true, TS>) : go(schur_kernel<double2, false, false, TS>);
    };
    tile == 64 ? pick(std::integral_constant<int, 64>{}) : pick(std::integral_constant<int, 32>{});
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

void launch_trace_labels(const NodeDesc* root, int32_t k, int32_t lab_cols, double2* xbuf, double* ysq,
                         const PointsView& pv, int32_t points, hipStream_t st) {
    if (points <= 0) return;
    if (lab_cols > 0) {
        const int ti = (k + kTile - 1) / kTile, tj = (lab_cols + kTile - 1) / kTile;
        hipLaunchKernelGGL(label_x_kernel, dim3(unsigned(ti * tj), unsigned(points)), dim3(256), 0, st, root, pv, xbuf,
                           lab_cols, tj);
        hipLaunchKernelGGL(label_y_kernel, dim3(unsigned(tj * tj), unsigned(points)), dim3(256), 0, st, root, pv, xbuf,
                           lab_cols, tj, ysq);
    }
    hipLaunchKernelGGL(label_sum_kernel, dim3(unsigned(points)), dim3(64), 0, st, root, pv, ysq, lab_cols);
}

void launch_trace(const NodeDesc* root, int32_t nl, int32_t cl_max, int32_t cr_max, double2* xbuf, double* partial,
                  const PointsView& pv, int32_t points, hipStream_t st) {
    if (points <= 0) return;
    if (std::max(cl_max, cr_max) > kTraceFewChannels && xbuf && partial) {
        const int ti = (nl + kTile - 1) / kTile, tj = (cr_max + kTile - 1) / kTile;
        hipLaunchKernelGGL(trace_x_kernel, dim3(unsigned(ti * tj), unsigned(points)), dim3(256), 0, st, root, pv, xbuf,
                           cr_max, tj);
        const int yi = (cl_max + kTile - 1) / kTile;
        hipLaunchKernelGGL(trace_y_kernel, dim3(unsigned(yi * tj), unsigned(points)), dim3(256), 0, st, root, pv, xbuf,
                           cl_max, cr_max, tj, partial);
        hipLaunchKernelGGL(trace_sum_kernel, dim3(unsigned(points)), dim3(64), 0, st, root, pv, partial, yi * tj);
        return;
    }
    if (nl <= kTraceRows) {
        allow_lds(reinterpret_cast<const void*>(trace_w_kernel));
        hipLaunchKernelGGL(trace_w_kernel, dim3(unsigned(points)), dim3(kThreads), kTraceLds, st, root, pv);
    } else {
        hipLaunchKernelGGL(trace_kernel, dim3(unsigned(points)), dim3(kThreads), 0, st, root, pv);
    }
}

}  // namespace caroli::gpu
