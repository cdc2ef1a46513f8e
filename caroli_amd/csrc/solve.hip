// Multifrontal solves with the stored front factors (complex f64 right-hand sides).
//
// A front F = [[F11, F12], [F21, F22]] was factored in place as P F11 = L11 U11 (row pivots within
// the separator rows), L21 = F21 U11^-1, U12 = L11^-1 P F12. For A x = b:
//   forward   y = L11^-1 P b_e,        b_b -= L21 y          (children before parents)
//   backward  x_e = U11^-1 (y - U12 x_b)                     (parents before children)
// and for A^T x = b (the adjoint and the refinement's dual solve):
//   forward   g = U11^-T b_e,           b_b -= U12^T g
//   backward  x_e = P^T L11^-T (g - L21^T x_b).
// The root is a dense LU of its leading nd + k block. A forward step hands its boundary part to
// its parent as an update vector, which the parent adds in the order of its children (no
// atomics): every solve is bit-reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <set>
#include <type_traits>
#include <stdexcept>

#include "cx.hpp"
#include "fronts.hpp"

namespace caroli::gpu {

namespace {

constexpr int kThreads = 256;

// The point of grid index q (SolveView::pts).
__device__ __forceinline__ int64_t point_of(const SolveView& sv, int64_t q) { return sv.pts ? sv.pts[q] : q; }

template <class TF>
__device__ __forceinline__ double2 widen(TF a) {
    return double2{double(a.x), double(a.y)};
}

// Adds the children's update vectors (the boundary columns c0 .. c0 + cw of each, from the
// forward solve) into y at their positions in this front (the children's maps), one child after
// the other: a child's positions are distinct, so within a child the lanes never collide, and the
// order of the sums does not depend on the schedule. Ends with a barrier.
template <int NT>
__device__ __forceinline__ void gather_updates(const NodeDesc& nd, const SolveView& sv, int64_t p, int c0, int cw,
                                               double2* y, int ldy) {
    const double2* U = sv.upd + p * sv.upd_stride;
    for (int k = nd.kid_beg; k < nd.kid_end; ++k) {
        const KidDesc kd = sv.kids[k];
        const double2* Uc = U + int64_t(sv.c) * kd.upd + int64_t(c0) * kd.size;
        const int32_t* mp = sv.maps + kd.map_beg;
        for (int idx = threadIdx.x; idx < kd.size * cw; idx += NT) {
            const int a = idx % kd.size, col = idx / kd.size;
            double2& t = y[mp[a] + col * ldy];
            t = cadd(t, Uc[a + int64_t(col) * kd.size]);
        }
        __syncthreads();
    }
}

// This front's update vector: rows ne .. n - 1 of y (columns c0 .. c0 + cw) to its slot.
template <int NT>
__device__ __forceinline__ void store_update(const NodeDesc& nd, const SolveView& sv, int64_t p, int c0, int cw,
                                             const double2* y, int ldy) {
    const int nb = nd.n - nd.ne;
    double2* U = sv.upd + p * sv.upd_stride + int64_t(sv.c) * nd.upd_off + int64_t(c0) * nb;
    for (int idx = threadIdx.x; idx < nb * cw; idx += NT) {
        const int b = idx % nb, col = idx / nb;
        U[b + int64_t(col) * nb] = y[nd.ne + b + col * ldy];
    }
}

// Transposed forward step (A^T x = b with right-hand sides off the kept orbitals), one workgroup
// per front and point with all columns in LDS (n * c <= kSolveLds).
template <class TF>
__global__ void __launch_bounds__(kThreads)
    solve_forward_t_kernel(const NodeDesc* __restrict__ nodes, const int32_t* __restrict__ level, PointsView pv,
                           SolveView sv) {
    extern __shared__ double2 y[];  // n x c, column-major
    const NodeDesc nd = nodes[level[blockIdx.x]];
    const int64_t p = point_of(sv, blockIdx.y);
    const TF* F = reinterpret_cast<const TF*>(pv.arena) + p * pv.arena_stride + nd.off;
    const int32_t* vars = sv.vars + nd.var_off;
    double2* X = sv.x + p * sv.x_stride;
    const int n = nd.n, ne = nd.ne, nb = n - ne, c = sv.c, tid = threadIdx.x;
    auto f = [&](int i, int j) { return widen(F[i + int64_t(j) * n]); };
    // y over all n rows: b_e on the separator, zero on the boundary, plus the children's updates.
    for (int idx = tid; idx < n * c; idx += kThreads) {
        const int r = idx % n, col = idx / n;
        y[idx] = r < ne ? X[vars[r] + int64_t(col) * sv.ldx] : double2{0.0, 0.0};
    }
    __syncthreads();
    gather_updates<kThreads>(nd, sv, p, 0, c, y, n);
    for (int k = 0; k < ne; ++k) {  // U11^T g = b_e (lower, non-unit)
        for (int col = tid; col < c; col += kThreads) y[k + col * n] = cmul(y[k + col * n], crecip(f(k, k)));
        __syncthreads();
        for (int idx = tid; idx < (ne - k - 1) * c; idx += kThreads) {
            int i = k + 1 + idx % (ne - k - 1), col = idx / (ne - k - 1);
            y[i + col * n] = csub(y[i + col * n], cmul(f(k, i), y[k + col * n]));
        }
        __syncthreads();
    }
    for (int idx = tid; idx < nb * c; idx += kThreads) {  // u_b = (updates) - U12^T g
        int b = idx % nb, col = idx / nb;
        double2 acc{0.0, 0.0};
        for (int k = 0; k < ne; ++k) acc = cadd(acc, cmul(f(k, ne + b), y[k + col * n]));
        y[ne + b + col * n] = csub(y[ne + b + col * n], acc);
    }
    __syncthreads();
    store_update<kThreads>(nd, sv, p, 0, c, y, n);
    for (int idx = tid; idx < ne * c; idx += kThreads) X[vars[idx % ne] + int64_t(idx / ne) * sv.ldx] = y[idx % ne + (idx / ne) * n];
}

// ---------------------------------------------------------------------------------------------
// Blocked solves of A x = b (the refinement and the adjoint's G columns). A workgroup holds CC
// right-hand-side columns of its front in LDS (grid.z covers the columns) and reads each factor
// entry once, coalesced: the triangles go by diagonal blocks of kDiag rows, each solved by wave 0
// with one row per lane and the pivot values broadcast by readlane, and the rows past a diagonal
// block are updated by all lanes as a matrix-vector product (rows over lanes, the K range split
// across groups of lanes when the rows are few).
constexpr int kDiag = 64;
constexpr int kPre = 8;  // factor columns loaded ahead in a diagonal block
// Factor entries a lane loads at once in the matrix-vector products (the loads of a batch are in
// flight together; the products are latency-bound otherwise).
template <class TF>
struct GemvBatch;
template <>
struct GemvBatch<float2> {
    static constexpr int value = 16;
};
template <>
struct GemvBatch<double2> {
    static constexpr int value = 16;
};
// Waves per SIMD the blocked solve kernels are compiled for (at most 128 registers a lane, which
// holds 4 columns without spilling): the solves are latency-bound and need the occupancy more
// than wider column chunks.
constexpr int kSolveWaves = 4;
constexpr int kSolveCols = 4;
constexpr int kSolveColsWide = 8;  // columns per workgroup past kSolveCols right-hand sides

// Workgroup -> (front, point, column chunk) with the chunk fastest and each XCD (workgroups go
// round-robin over the 8 XCDs) taking one contiguous range, so that the chunks of a front and
// point, which read the same factors, share an L2. grid = (count * chunks, points).
struct SolveItem {
    int front, chunk;
    int64_t point;
};
__device__ __forceinline__ SolveItem solve_item(int chunks) {
    const uint64_t T = gridDim.x, total = T * gridDim.y;
    const uint64_t lin = blockIdx.x + uint64_t(blockIdx.y) * T;
    const uint64_t q = total / 8, r = total % 8, x = lin % 8, idx = lin / 8;
    const uint64_t logical = (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + idx;
    const uint64_t fc = logical % T;
    return {int(fc / chunks), int(fc % chunks), int64_t(logical / T)};
}

__device__ __forceinline__ double lane_value(double v, int k) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), k);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), k);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double2 lane_value(double2 v, int k) { return double2{lane_value(v.x, k), lane_value(v.y, k)}; }

// y[r] -= sum_{k in [k0, k1)} F[r + k ld] y[k] for r in [r0, r1) (CC columns of y, leading
// dimension ldy). red: NT x CC words of LDS for the partial sums. Ends with a barrier.
template <class TF, int NT, int CC>
__device__ __forceinline__ void gemv_sub(const TF* __restrict__ F, int ld, int r0, int r1, int k0, int k1, double2* y,
                                         int ldy, double2* red) {
    const int R = r1 - r0, K = k1 - k0, tid = threadIdx.x;
    if (R <= 0 || K <= 0) return;  // uniform
    // Lanes per row group: the rows rounded up to a power of two below 64 (a wave then covers
    // several groups, with no lane reading a duplicate row), to whole waves above.
    int RP = 64;
    if (R < 64)
        for (RP = 1; RP < R; RP <<= 1) {
        }
    else
        RP = min(NT, (R + 63) / 64 * 64);
    const int G = NT / RP, rr = tid % RP, g = tid / RP;
    const int kc = (K + G - 1) / G, ka = k0 + min(g, G - 1) * kc, kz = g < G ? min(k1, ka + kc) : ka;
    for (int base = 0; base < R; base += RP) {
        const int r = r0 + base + rr;
        const int64_t rc = min(r, r1 - 1);
        double2 acc[CC];
#pragma unroll
        for (int col = 0; col < CC; ++col) acc[col] = double2{0.0, 0.0};
        // Batches of factor entries in their own type, the last one clamped and guarded.
        constexpr int BT = GemvBatch<TF>::value;
        for (int k = ka; k < kz; k += BT) {
            TF l[BT];
#pragma unroll
            for (int u = 0; u < BT; ++u) l[u] = F[rc + int64_t(min(k + u, kz - 1)) * ld];
#pragma unroll
            for (int u = 0; u < BT; ++u)
                if (k + u < kz) {
                    const double2 lv = widen(l[u]);
#pragma unroll
                    for (int col = 0; col < CC; ++col) acc[col] = cadd(acc[col], cmul(lv, y[k + u + col * ldy]));
                }
        }
        if (G == 1) {
            if (g == 0 && r < r1)
#pragma unroll
                for (int col = 0; col < CC; ++col) y[r + col * ldy] = csub(y[r + col * ldy], acc[col]);
        } else {
            if (g < G)
#pragma unroll
                for (int col = 0; col < CC; ++col) red[tid + col * NT] = acc[col];
            __syncthreads();
            if (g == 0 && r < r1)
#pragma unroll
                for (int col = 0; col < CC; ++col) {
                    double2 s = red[tid + col * NT];
                    for (int q = 1; q < G; ++q) s = cadd(s, red[tid + q * RP + col * NT]);
                    y[r + col * ldy] = csub(y[r + col * ldy], s);
                }
        }
        __syncthreads();
    }
}

// Wave 0: unit-lower diagonal block [kb, kb + w) of F (rows kb + lane), forward substitution.
template <class TF, int CC>
__device__ __forceinline__ void diag_lower(const TF* __restrict__ F, int ld, int kb, int w, double2* y, int ldy) {
    const int lane = threadIdx.x;
    const bool mine = lane < w;
    const int64_t ic = kb + min(lane, w - 1);
    double2 v[CC];
#pragma unroll
    for (int col = 0; col < CC; ++col) v[col] = mine ? y[kb + lane + col * ldy] : double2{0.0, 0.0};
    for (int k0 = 0; k0 < w; k0 += kPre) {
        double2 l[kPre];
#pragma unroll
        for (int u = 0; u < kPre; ++u) l[u] = widen(F[ic + int64_t(kb + min(k0 + u, w - 1)) * ld]);
#pragma unroll
        for (int u = 0; u < kPre; ++u) {
            const int k = k0 + u;
            if (k < w) {
#pragma unroll
                for (int col = 0; col < CC; ++col) {
                    const double2 x = lane_value(v[col], k);
                    if (lane > k) v[col] = csub(v[col], cmul(l[u], x));
                }
            }
        }
    }
    if (mine)
#pragma unroll
        for (int col = 0; col < CC; ++col) y[kb + lane + col * ldy] = v[col];
}

// Wave 0: upper diagonal block [kb, kb + w) of F, back substitution (non-unit diagonal).
template <class TF, int CC>
__device__ __forceinline__ void diag_upper(const TF* __restrict__ F, int ld, int kb, int w, double2* y, int ldy) {
    const int lane = threadIdx.x;
    const bool mine = lane < w;
    const int64_t ic = kb + min(lane, w - 1);
    double2 v[CC];
#pragma unroll
    for (int col = 0; col < CC; ++col) v[col] = mine ? y[kb + lane + col * ldy] : double2{0.0, 0.0};
    for (int top = w - 1; top >= 0; top -= kPre) {
        double2 u[kPre];
#pragma unroll
        for (int s = 0; s < kPre; ++s) u[s] = widen(F[ic + int64_t(kb + max(top - s, 0)) * ld]);
#pragma unroll
        for (int s = 0; s < kPre; ++s) {
            const int k = top - s;
            if (k >= 0) {
                const double2 inv = crecip(lane_value(u[s], k));  // 1 / U[kb + k, kb + k]
#pragma unroll
                for (int col = 0; col < CC; ++col) {
                    const double2 x = cmul(lane_value(v[col], k), inv);
                    if (lane == k)
                        v[col] = x;
                    else if (lane < k)
                        v[col] = csub(v[col], cmul(u[s], x));
                }
            }
        }
    }
    if (mine)
#pragma unroll
        for (int col = 0; col < CC; ++col) y[kb + lane + col * ldy] = v[col];
}

// Forward substitution with the unit-lower columns [0, ne) of F over the rows [0, nrows).
template <class TF, int NT, int CC>
__device__ __forceinline__ void lower_blocked(const TF* F, int ld, int ne, int nrows, double2* y, int ldy, double2* red) {
    for (int kb = 0; kb < ne; kb += kDiag) {
        const int w = min(kDiag, ne - kb);
        if (threadIdx.x < 64) diag_lower<TF, CC>(F, ld, kb, w, y, ldy);
        __syncthreads();
        gemv_sub<TF, NT, CC>(F, ld, kb + w, nrows, kb, kb + w, y, ldy, red);
    }
}

// Back substitution with the upper triangle of F[0:ne, 0:ne].
template <class TF, int NT, int CC>
__device__ __forceinline__ void upper_blocked(const TF* F, int ld, int ne, double2* y, int ldy, double2* red) {
    for (int kb = (ne - 1) / kDiag * kDiag; kb >= 0; kb -= kDiag) {
        const int w = min(kDiag, ne - kb);
        if (threadIdx.x < 64) diag_upper<TF, CC>(F, ld, kb, w, y, ldy);
        __syncthreads();
        gemv_sub<TF, NT, CC>(F, ld, 0, kb, kb, kb + w, y, ldy, red);
    }
}

// P y: LAPACK's interchanges in order, one lane per column.
__device__ __forceinline__ void apply_pivots(const int32_t* __restrict__ piv, int ne, double2* y, int ldy, int cw) {
    if (int(threadIdx.x) < cw) {
        double2* yc = y + threadIdx.x * ldy;
        for (int k = 0; k < ne; ++k) {
            const int q = piv[k];
            if (q > k && q < ne) {  // pivots of a front that delayed are stale; its point is not used
                const double2 t = yc[k];
                yc[k] = yc[q];
                yc[q] = t;
            }
        }
    }
    __syncthreads();
}

// Forward step of a front: with the children's update vectors added (in the order of the
// children), y = L11^-1 P b_e on the separator, and the front's own update vector u_b - L21 y
// (u_b: what the children passed on to its boundary) stored for its parent.
template <class TF, int NT, int CC>
__global__ void __launch_bounds__(NT, kSolveWaves)
    fwd_blocked_kernel(const NodeDesc* __restrict__ nodes, const int32_t* __restrict__ level, PointsView pv, SolveView sv) {
    extern __shared__ double2 ys[];  // n x CC, then NT x CC partial sums
    const SolveItem it = solve_item((sv.c + CC - 1) / CC);
    const NodeDesc nd = nodes[level[it.front]];
    const int64_t p = point_of(sv, it.point);
    const int c0 = it.chunk * CC, cw = min(CC, sv.c - c0);
    if (cw <= 0) return;
    const TF* F = reinterpret_cast<const TF*>(pv.arena) + p * pv.arena_stride + nd.off;
    const int32_t* piv = pv.piv + p * pv.piv_stride + nd.piv_off;
    const int32_t* vars = sv.vars + nd.var_off;
    double2* X = sv.x + p * sv.x_stride + int64_t(c0) * sv.ldx;
    const int n = nd.n, ne = nd.ne, tid = threadIdx.x;
    double2* red = ys + n * CC;
    for (int idx = tid; idx < n * CC; idx += NT) {
        const int r = idx % n, col = idx / n;
        ys[idx] = (r < ne && col < cw) ? X[vars[r] + int64_t(col) * sv.ldx] : double2{0.0, 0.0};
    }
    __syncthreads();
    gather_updates<NT>(nd, sv, p, c0, cw, ys, n);
    apply_pivots(piv, ne, ys, n, cw);
    lower_blocked<TF, NT, CC>(F, n, ne, n, ys, n, red);
    for (int idx = tid; idx < ne * cw; idx += NT) {
        const int r = idx % ne, col = idx / ne;
        X[vars[r] + int64_t(col) * sv.ldx] = ys[r + col * n];
    }
    store_update<NT>(nd, sv, p, c0, cw, ys, n);
}

// Backward step of a front: x_e = U11^-1 (y - U12 x_b), the boundary values final.
template <class TF, int NT, int CC>
__global__ void __launch_bounds__(NT, kSolveWaves)
    bwd_blocked_kernel(const NodeDesc* __restrict__ nodes, const int32_t* __restrict__ level, PointsView pv, SolveView sv) {
    extern __shared__ double2 ys[];
    const SolveItem it = solve_item((sv.c + CC - 1) / CC);
    const NodeDesc nd = nodes[level[it.front]];
    const int64_t p = point_of(sv, it.point);
    const int c0 = it.chunk * CC, cw = min(CC, sv.c - c0);
    if (cw <= 0) return;
    const TF* F = reinterpret_cast<const TF*>(pv.arena) + p * pv.arena_stride + nd.off;
    const int32_t* vars = sv.vars + nd.var_off;
    double2* X = sv.x + p * sv.x_stride + int64_t(c0) * sv.ldx;
    const int n = nd.n, ne = nd.ne, tid = threadIdx.x;
    double2* red = ys + n * CC;
    for (int idx = tid; idx < n * CC; idx += NT) {
        const int r = idx % n, col = idx / n;
        ys[idx] = col < cw ? X[vars[r] + int64_t(col) * sv.ldx] : double2{0.0, 0.0};
    }
    __syncthreads();
    gemv_sub<TF, NT, CC>(F, n, 0, ne, ne, n, ys, n, red);
    upper_blocked<TF, NT, CC>(F, n, ne, ys, n, red);
    for (int idx = tid; idx < ne * cw; idx += NT) {
        const int r = idx % ne, col = idx / ne;
        X[vars[r] + int64_t(col) * sv.ldx] = ys[r + col * n];
    }
}

// The root: P, L, U of its leading ne x ne block (complex f64).
template <int NT, int CC>
__global__ void __launch_bounds__(NT, kSolveWaves)
    root_blocked_kernel(const NodeDesc* __restrict__ nodes, const int32_t* __restrict__ level, PointsView pv, SolveView sv) {
    extern __shared__ double2 ys[];
    const SolveItem it = solve_item((sv.c + CC - 1) / CC);
    const NodeDesc nd = nodes[level[0]];
    const int64_t p = point_of(sv, it.point);
    const int c0 = it.chunk * CC, cw = min(CC, sv.c - c0);
    if (cw <= 0) return;
    const double2* F = reinterpret_cast<const double2*>(pv.root_arena) + p * pv.root_stride + nd.off;
    const int32_t* piv = pv.piv + p * pv.piv_stride + nd.piv_off;
    const int32_t* vars = sv.vars + nd.var_off;
    double2* X = sv.x + p * sv.x_stride + int64_t(c0) * sv.ldx;
    const int n = nd.n, ne = nd.ne, tid = threadIdx.x;
    double2* red = ys + ne * CC;
    for (int idx = tid; idx < ne * CC; idx += NT) {
        const int r = idx % ne, col = idx / ne;
        ys[idx] = col < cw ? X[vars[r] + int64_t(col) * sv.ldx] : double2{0.0, 0.0};
    }
    __syncthreads();
    if (sv.gather) gather_updates<NT>(nd, sv, p, c0, cw, ys, ne);
    apply_pivots(piv, ne, ys, ne, cw);
    lower_blocked<double2, NT, CC>(F, n, ne, ne, ys, ne, red);
    upper_blocked<double2, NT, CC>(F, n, ne, ys, ne, red);
    for (int idx = tid; idx < ne * cw; idx += NT) {
        const int r = idx % ne, col = idx / ne;
        X[vars[r] + int64_t(col) * sv.ldx] = ys[r + col * ne];
    }
}

// ---------------------------------------------------------------------------------------------
// Blocked transposed solves (A^T x = b) for right-hand sides on the kept orbitals: the root's
// (P^-1 L U)^T and, parents before children, each front's backward step
// x_e = P^T L11^-T (x_e - L21^T x_b). The transposed products read the factors down their
// columns: a wave takes one output and its lanes run down the column (coalesced loads), then a
// wave reduction in a fixed lane order, so the result does not depend on the schedule.

// y[o] -= sum_{r in [r0, r1)} F[r + o ld] y[r] for o in [o0, o1) (CC columns of y, leading
// dimension ldy; the two row ranges are disjoint). Ends with a barrier.
template <class TF, int NT, int CC>
__device__ __forceinline__ void gemv_t_sub(const TF* __restrict__ F, int ld, int r0, int r1, int o0, int o1,
                                           double2* y, int ldy) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (r1 > r0) {
        for (int o = o0 + wave; o < o1; o += NT / 64) {
            const TF* col = F + int64_t(o) * ld;
            double2 acc[CC];
#pragma unroll
            for (int c = 0; c < CC; ++c) acc[c] = double2{0.0, 0.0};
            int r = r0 + lane;
            for (; r + 3 * 64 < r1; r += 4 * 64) {
                double2 f[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) f[u] = widen(col[r + u * 64]);
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int c = 0; c < CC; ++c) acc[c] = cadd(acc[c], cmul(f[u], y[r + u * 64 + c * ldy]));
            }
            for (; r < r1; r += 64) {
                const double2 f = widen(col[r]);
#pragma unroll
                for (int c = 0; c < CC; ++c) acc[c] = cadd(acc[c], cmul(f, y[r + c * ldy]));
            }
#pragma unroll
            for (int c = 0; c < CC; ++c)
                for (int off = 32; off > 0; off >>= 1) {
                    acc[c].x += __shfl_xor(acc[c].x, off, 64);
                    acc[c].y += __shfl_xor(acc[c].y, off, 64);
                }
            if (lane == 0)
#pragma unroll
                for (int c = 0; c < CC; ++c) y[o + c * ldy] = csub(y[o + c * ldy], acc[c]);
        }
    }
    __syncthreads();
}

// Wave 0: L^T z = y on the diagonal block [kb, kb + w) (L unit lower, so L^T unit upper), back
// substitution: once z_i is final (lane i), the lanes j < i subtract L[kb + i, kb + j] z_i; lane j
// reads its own column of L.
template <class TF, int CC>
__device__ __forceinline__ void diag_lower_t(const TF* __restrict__ F, int ld, int kb, int w, double2* y, int ldy) {
    const int lane = threadIdx.x;
    const bool mine = lane < w;
    const int64_t jc = int64_t(kb + min(lane, w - 1)) * ld;
    double2 v[CC];
#pragma unroll
    for (int c = 0; c < CC; ++c) v[c] = mine ? y[kb + lane + c * ldy] : double2{0.0, 0.0};
    for (int top = w - 1; top >= 0; top -= kPre) {
        double2 l[kPre];
#pragma unroll
        for (int s = 0; s < kPre; ++s) l[s] = widen(F[kb + max(top - s, 0) + jc]);
#pragma unroll
        for (int s = 0; s < kPre; ++s) {
            const int i = top - s;
            if (i >= 0) {
#pragma unroll
                for (int c = 0; c < CC; ++c) {
                    const double2 z = lane_value(v[c], i);
                    if (lane < i) v[c] = csub(v[c], cmul(l[s], z));
                }
            }
        }
    }
    if (mine)
#pragma unroll
        for (int c = 0; c < CC; ++c) y[kb + lane + c * ldy] = v[c];
}

// Wave 0: U^T a = y on the diagonal block [kb, kb + w) (U^T lower, non-unit), forward
// substitution: a_k = y_k / U[kb + k, kb + k] (lane k), then the lanes j > k subtract
// U[kb + k, kb + j] a_k; lane j reads its own column of U.
template <class TF, int CC>
__device__ __forceinline__ void diag_upper_t(const TF* __restrict__ F, int ld, int kb, int w, double2* y, int ldy) {
    const int lane = threadIdx.x;
    const bool mine = lane < w;
    const int64_t jc = int64_t(kb + min(lane, w - 1)) * ld;
    double2 v[CC];
#pragma unroll
    for (int c = 0; c < CC; ++c) v[c] = mine ? y[kb + lane + c * ldy] : double2{0.0, 0.0};
    for (int k0 = 0; k0 < w; k0 += kPre) {
        double2 u[kPre];
#pragma unroll
        for (int s = 0; s < kPre; ++s) u[s] = widen(F[kb + min(k0 + s, w - 1) + jc]);
#pragma unroll
        for (int s = 0; s < kPre; ++s) {
            const int k = k0 + s;
            if (k < w) {
                const double2 inv = crecip(lane_value(u[s], k));  // 1 / U[kb + k, kb + k]
#pragma unroll
                for (int c = 0; c < CC; ++c) {
                    const double2 a = cmul(lane_value(v[c], k), inv);
                    if (lane == k)
                        v[c] = a;
                    else if (lane > k)
                        v[c] = csub(v[c], cmul(u[s], a));
                }
            }
        }
    }
    if (mine)
#pragma unroll
        for (int c = 0; c < CC; ++c) y[kb + lane + c * ldy] = v[c];
}

// L^T z = y with the unit-lower columns [0, ne) of F (rows [0, ne)), bottom block first.
template <class TF, int NT, int CC>
__device__ __forceinline__ void lower_t_blocked(const TF* F, int ld, int ne, double2* y, int ldy) {
    for (int kb = (ne - 1) / kDiag * kDiag; kb >= 0; kb -= kDiag) {
        const int w = min(kDiag, ne - kb);
        gemv_t_sub<TF, NT, CC>(F, ld, kb + w, ne, kb, kb + w, y, ldy);
        if (threadIdx.x < 64) diag_lower_t<TF, CC>(F, ld, kb, w, y, ldy);
        __syncthreads();
    }
}

// U^T a = y with the upper triangle of F[0:ne, 0:ne], top block first.
template <class TF, int NT, int CC>
__device__ __forceinline__ void upper_t_blocked(const TF* F, int ld, int ne, double2* y, int ldy) {
    for (int kb = 0; kb < ne; kb += kDiag) {
        const int w = min(kDiag, ne - kb);
        gemv_t_sub<TF, NT, CC>(F, ld, 0, kb, kb, kb + w, y, ldy);
        if (threadIdx.x < 64) diag_upper_t<TF, CC>(F, ld, kb, w, y, ldy);
        __syncthreads();
    }
}

// P^T y: LAPACK's interchanges in reverse order, one lane per column.
__device__ __forceinline__ void apply_pivots_t(const int32_t* __restrict__ piv, int ne, double2* y, int ldy, int cw) {
    if (int(threadIdx.x) < cw) {
        double2* yc = y + threadIdx.x * ldy;
        for (int k = ne - 1; k >= 0; --k) {
            const int q = piv[k];
            if (q > k && q < ne) {
                const double2 t = yc[k];
                yc[k] = yc[q];
                yc[q] = t;
            }
        }
    }
    __syncthreads();
}

// Transposed backward step of a front: x_e = P^T L11^-T (x_e - L21^T x_b), the boundary values
// final (set by the ancestors).
template <class TF, int NT, int CC>
__global__ void __launch_bounds__(NT, kSolveWaves)
    bwd_t_blocked_kernel(const NodeDesc* __restrict__ nodes, const int32_t* __restrict__ level, PointsView pv,
                         SolveView sv) {
    extern __shared__ double2 ys[];  // n x CC
    const SolveItem it = solve_item((sv.c + CC - 1) / CC);
    const NodeDesc nd = nodes[level[it.front]];
    const int64_t p = point_of(sv, it.point);
    const int c0 = it.chunk * CC, cw = min(CC, sv.c - c0);
    if (cw <= 0) return;
    const TF* F = reinterpret_cast<const TF*>(pv.arena) + p * pv.arena_stride + nd.off;
    const int32_t* piv = pv.piv + p * pv.piv_stride + nd.piv_off;
    const int32_t* vars = sv.vars + nd.var_off;
    double2* X = sv.x + p * sv.x_stride + int64_t(c0) * sv.ldx;
    const int n = nd.n, ne = nd.ne, tid = threadIdx.x;
    for (int idx = tid; idx < n * CC; idx += NT) {
        const int r = idx % n, col = idx / n;
        ys[idx] = col < cw ? X[vars[r] + int64_t(col) * sv.ldx] : double2{0.0, 0.0};
    }
    __syncthreads();
    gemv_t_sub<TF, NT, CC>(F, n, ne, n, 0, ne, ys, n);
    lower_t_blocked<TF, NT, CC>(F, n, ne, ys, n);
    apply_pivots_t(piv, ne, ys, n, cw);
    for (int idx = tid; idx < ne * cw; idx += NT) {
        const int r = idx % ne, col = idx / ne;
        X[vars[r] + int64_t(col) * sv.ldx] = ys[r + col * n];
    }
}

// Transposed solve with the root's LU (its leading ne x ne block, complex f64): U^T a = b,
// L^T z = a, x = P^T z; the children's update vectors first when the transposed forward solve ran.
template <int NT, int CC>
__global__ void __launch_bounds__(NT, kSolveWaves)
    root_t_blocked_kernel(const NodeDesc* __restrict__ nodes, const int32_t* __restrict__ level, PointsView pv,
                          SolveView sv) {
    extern __shared__ double2 ys[];  // ne x CC
    const SolveItem it = solve_item((sv.c + CC - 1) / CC);
    const NodeDesc nd = nodes[level[0]];
    const int64_t p = point_of(sv, it.point);
    const int c0 = it.chunk * CC, cw = min(CC, sv.c - c0);
    if (cw <= 0) return;
    const double2* F = reinterpret_cast<const double2*>(pv.root_arena) + p * pv.root_stride + nd.off;
    const int32_t* piv = pv.piv + p * pv.piv_stride + nd.piv_off;
    const int32_t* vars = sv.vars + nd.var_off;
    double2* X = sv.x + p * sv.x_stride + int64_t(c0) * sv.ldx;
    const int n = nd.n, ne = nd.ne, tid = threadIdx.x;
    for (int idx = tid; idx < ne * CC; idx += NT) {
        const int r = idx % ne, col = idx / ne;
        ys[idx] = col < cw ? X[vars[r] + int64_t(col) * sv.ldx] : double2{0.0, 0.0};
    }
    __syncthreads();
    if (sv.gather) gather_updates<NT>(nd, sv, p, c0, cw, ys, ne);
    upper_t_blocked<double2, NT, CC>(F, n, ne, ys, ne);
    lower_t_blocked<double2, NT, CC>(F, n, ne, ys, ne);
    apply_pivots_t(piv, ne, ys, ne, cw);
    for (int idx = tid; idx < ne * cw; idx += NT) {
        const int r = idx % ne, col = idx / ne;
        X[vars[r] + int64_t(col) * sv.ldx] = ys[r + col * ne];
    }
}

// B[orbs[a], col] = W[a, col] (or conj) for a < m, per point; B must be zero elsewhere.
__global__ void __launch_bounds__(kThreads)
    scatter_kernel(SolveView sv, const int32_t* __restrict__ orbs, int32_t m, const double2* __restrict__ w,
                   int64_t w_stride, int32_t conj) {
    const int64_t p = point_of(sv, blockIdx.y);
    double2* X = sv.x + p * sv.x_stride;
    const double2* W = w + p * w_stride;
    for (int idx = blockIdx.x * kThreads + threadIdx.x; idx < m * sv.c; idx += gridDim.x * kThreads) {
        int a = idx % m, col = idx / m;
        double2 v = W[a + int64_t(col) * m];
        if (conj) v.y = -v.y;
        X[orbs[a] + int64_t(col) * sv.ldx] = v;
    }
}

constexpr int kResBatch = 8;
constexpr int kResCols = 4;  // columns of a row per thread in the residual

// D = B - A X on the rows of a block in FP64, A = z - H - Sigma_L - Sigma_R (the residual of the
// refinement), or D = B - A^T X with transpose (the residual of the dual solve). Sigma_L acts on
// the left-slice rows (positions left_pos), Sigma_R on the right. H is Hermitian, so the rows of
// H^T are the conjugated rows of H.
__global__ void __launch_bounds__(kThreads)
    residual_kernel(PointsView pv, SolveView xv, const double2* __restrict__ b, double2* __restrict__ d,
                    const int32_t* __restrict__ rows, int32_t nrows, const int64_t* __restrict__ row_ptr,
                    const uint32_t* __restrict__ cols, const int32_t* __restrict__ left_pos,
                    const int32_t* __restrict__ right_pos, const int32_t* __restrict__ left_orbs, int32_t nl,
                    const int32_t* __restrict__ right_orbs, int32_t nr, int32_t transpose) {
    const int64_t p = point_of(xv, blockIdx.y);
    const double2* X = xv.x + p * xv.x_stride;
    const double2* B = b + p * xv.x_stride;
    double2* D = d + p * xv.x_stride;
    const double2* vals = pv.vals[p];
    const double2 z = pv.z[p];
    const double2* sl = pv.sigma + p * pv.sigma_stride;
    const double2* sr = sl + int64_t(nl) * nl;
    // Sigma (m x m, column-major) entry (row, a), or (a, row) transposed.
    const int64_t sa = transpose ? 1 : 0;
    // One thread per row and kResCols columns: the row's entries are loaded once for them.
    const int chunks = (xv.c + kResCols - 1) / kResCols;
    for (int idx = blockIdx.x * kThreads + threadIdx.x; idx < nrows * chunks; idx += gridDim.x * kThreads) {
        const int r = rows[idx % nrows], c0 = (idx / nrows) * kResCols, cw = min(kResCols, xv.c - c0);
        const double2* x = X + int64_t(c0) * xv.ldx;
        const int64_t ldx = xv.ldx;
        double2 acc[kResCols];
#pragma unroll
        for (int c = 0; c < kResCols; ++c) acc[c] = c < cw ? cmul(z, x[r + c * ldx]) : double2{0.0, 0.0};
        // Loads in batches of kResBatch (indexes clamped, only the sums guarded), so that a lane
        // has a batch of loads in flight instead of one.
        const int64_t e0 = row_ptr[r], e1 = row_ptr[r + 1];
        for (int64_t e = e0; e < e1; e += kResBatch) {
            uint32_t cc[kResBatch];
            double2 v[kResBatch];
#pragma unroll
            for (int u = 0; u < kResBatch; ++u) {
                const int64_t q = min(e + u, e1 - 1);
                cc[u] = cols[q];
                v[u] = vals[q];
                if (transpose) v[u] = cconj(v[u]);
            }
#pragma unroll
            for (int c = 0; c < kResCols; ++c)
                if (c < cw) {
                    double2 xv[kResBatch];
#pragma unroll
                    for (int u = 0; u < kResBatch; ++u) xv[u] = x[cc[u] + c * ldx];
#pragma unroll
                    for (int u = 0; u < kResBatch; ++u)
                        if (e + u < e1) acc[c] = csub(acc[c], cmul(v[u], xv[u]));
                }
        }
        // Sigma (m x m, column-major) row `pos` (column `pos` transposed) against x on the slice.
        auto slice = [&](const double2* sg, int pos, const int32_t* orbs, int m) {
            for (int a0 = 0; a0 < m; a0 += kResBatch) {
                int32_t o[kResBatch];
                double2 sv[kResBatch];
#pragma unroll
                for (int u = 0; u < kResBatch; ++u) {
                    const int a = min(a0 + u, m - 1);
                    o[u] = orbs[a];
                    sv[u] = sg[sa ? a + int64_t(pos) * m : pos + int64_t(a) * m];
                }
#pragma unroll
                for (int c = 0; c < kResCols; ++c)
                    if (c < cw) {
                        double2 xs[kResBatch];
#pragma unroll
                        for (int u = 0; u < kResBatch; ++u) xs[u] = x[o[u] + c * ldx];
#pragma unroll
                        for (int u = 0; u < kResBatch; ++u)
                            if (a0 + u < m) acc[c] = csub(acc[c], cmul(sv[u], xs[u]));
                    }
            }
        };
        const int lp = left_pos[r], rp = right_pos[r];
        if (lp >= 0) slice(sl, lp, left_orbs, nl);
        if (rp >= 0) slice(sr, rp, right_orbs, nr);
#pragma unroll
        for (int c = 0; c < kResCols; ++c)
            if (c < cw) {
                const int64_t at = r + (c0 + c) * ldx;
                D[at] = csub(B[at], acc[c]);
            }
    }
}

// The goal-oriented estimate of the refinement, per point of xv.pts, every sum in a fixed order
// (the result does not depend on the schedule):
//  - with wl: t'[d, c] = sum_a conj(W_L[a, d]) X[left_orbs[a], c] (+ sum_rows phi[r, d] R[r, c]
//    with phi): the transmission amplitudes W_L^H x_L, corrected by the dual solution phi
//    (A^T phi = E_L conj(W_L)) against the residual R = b - A x;
//  - with nrm: sum_rows |nrm[r, c]|^2 per column c of nrm.
// goal_part_kernel: grid (chunks, points), each workgroup over kGoalRows block rows (the left-slice
// term in chunk 0), items (the cl cr pairs, then the nc columns) summed by G lanes over
// interleaved rows and the G partial sums in lane order, to part[(q chunks + chunk) items + item].
// goal_sum_kernel: per point, the chunks' partial sums in chunk order to out (t' at
// out[0 .. 2 cl cr), the norms at out[nrm_at ..], out_stride doubles per point).
constexpr int kGoalRows = 2048;
struct GoalArgs {
    SolveView xv;
    const double2* res;
    const double2* phi;
    int64_t phi_stride;
    const double2* wl;
    int64_t wl_stride;
    int32_t nl, cl;
    const int32_t* left_orbs;
    const double2* nrm;
    int64_t nrm_stride;
    int32_t nc, nrm_at;
    const int32_t* rows;
    int32_t nrows;
    double2* part;
    int32_t chunks;
    double* out;
    int64_t out_stride;
};

__global__ void __launch_bounds__(kThreads) goal_part_kernel(GoalArgs ga) {
    __shared__ double2 red[kThreads];
    const SolveView& xv = ga.xv;
    const int64_t q = blockIdx.y, p = point_of(xv, q);
    const int tid = threadIdx.x, cr = xv.c, ld = xv.ldx, chunk = blockIdx.x;
    const int i0 = chunk * kGoalRows, i1 = min(ga.nrows, i0 + kGoalRows);
    const int pairs = ga.wl ? ga.cl * cr : 0, norms = ga.nrm ? ga.nc : 0, items = pairs + norms;
    const double2* X = xv.x + p * xv.x_stride;
    const double2* R = ga.res ? ga.res + p * xv.x_stride : nullptr;
    const double2* F = ga.phi ? ga.phi + p * ga.phi_stride : nullptr;
    const double2* W = ga.wl ? ga.wl + p * ga.wl_stride : nullptr;
    const double2* V = ga.nrm ? ga.nrm + p * ga.nrm_stride : nullptr;
    double2* out = ga.part + (q * ga.chunks + chunk) * int64_t(items);
    const int G = max(1, kThreads / max(items, 1)), slots = kThreads / G;
    constexpr int B = 4;
    for (int t0 = 0; t0 < items; t0 += slots) {
        const int t = t0 + tid / G, g = tid % G;
        double2 acc{0.0, 0.0};
        if (tid / G < slots && t < items) {
            if (t < pairs) {
                const int dd = t % ga.cl, c = t / ga.cl;
                if (chunk == 0)
                    for (int a = g; a < ga.nl; a += G)
                        acc = cadd(acc, cmul(cconj(W[a + int64_t(dd) * ga.nl]), X[ga.left_orbs[a] + int64_t(c) * ld]));
                if (F)
                    for (int i = i0 + g; i < i1; i += G * B) {
                        int rr[B];
#pragma unroll
                        for (int u = 0; u < B; ++u) rr[u] = ga.rows[min(i + u * G, i1 - 1)];
                        double2 f[B], v[B];
#pragma unroll
                        for (int u = 0; u < B; ++u) {
                            f[u] = F[rr[u] + int64_t(dd) * ld];
                            v[u] = R[rr[u] + int64_t(c) * ld];
                        }
#pragma unroll
                        for (int u = 0; u < B; ++u)
                            if (i + u * G < i1) acc = cadd(acc, cmul(f[u], v[u]));
                    }
            } else {
                const int c = t - pairs;
                for (int i = i0 + g; i < i1; i += G * B) {
                    int rr[B];
#pragma unroll
                    for (int u = 0; u < B; ++u) rr[u] = ga.rows[min(i + u * G, i1 - 1)];
                    double2 v[B];
#pragma unroll
                    for (int u = 0; u < B; ++u) v[u] = V[rr[u] + int64_t(c) * ld];
#pragma unroll
                    for (int u = 0; u < B; ++u)
                        if (i + u * G < i1) acc.x += v[u].x * v[u].x + v[u].y * v[u].y;
                }
            }
        }
        red[tid] = acc;
        __syncthreads();
        if (g == 0 && tid / G < slots && t < items) {
            double2 s = red[tid];
            for (int k = 1; k < G; ++k) s = cadd(s, red[tid + k]);
            out[t] = s;
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kThreads) goal_sum_kernel(GoalArgs ga) {
    const int64_t q = blockIdx.x, p = point_of(ga.xv, q);
    const int pairs = ga.wl ? ga.cl * ga.xv.c : 0, norms = ga.nrm ? ga.nc : 0, items = pairs + norms;
    double* O = ga.out + p * ga.out_stride;
    for (int t = threadIdx.x; t < items; t += kThreads) {
        double2 s{0.0, 0.0};
        for (int k = 0; k < ga.chunks; ++k) s = cadd(s, ga.part[(q * ga.chunks + k) * int64_t(items) + t]);
        if (t < pairs) {
            O[2 * t] = s.x;
            O[2 * t + 1] = s.y;
        } else {
            O[ga.nrm_at + (t - pairs)] = s.x;
        }
    }
}

// X += D on the block rows; maxes[2p] = max |D|, maxes[2p + 1] = max |X| (the relative
// correction is their ratio), as the bit patterns of non-negative doubles, which order like the
// integers. A NaN correction counts as infinite. maxes is zero on entry.
__global__ void __launch_bounds__(kThreads)
    update_kernel(SolveView xv, const double2* __restrict__ d, const int32_t* __restrict__ rows, int32_t nrows,
                  unsigned long long* __restrict__ maxes) {
    __shared__ double sd[kThreads / 64], sx[kThreads / 64];
    const int64_t p = point_of(xv, blockIdx.y);
    double2* X = xv.x + p * xv.x_stride;
    const double2* D = d + p * xv.x_stride;
    double md = 0.0, mx = 0.0;
    for (int idx = blockIdx.x * kThreads + threadIdx.x; idx < nrows * xv.c; idx += gridDim.x * kThreads) {
        const int64_t at = rows[idx % nrows] + int64_t(idx / nrows) * xv.ldx;
        const double2 dv = D[at], v = cadd(X[at], dv);
        X[at] = v;
        const double a = cabs(dv);
        md = fmax(md, a == a ? a : INFINITY);
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
        atomicMax(&maxes[2 * p], static_cast<unsigned long long>(__double_as_longlong(md)));
        atomicMax(&maxes[2 * p + 1], static_cast<unsigned long long>(__double_as_longlong(mx)));
    }
}

// out[a, col] = X[orbs[a], col] per point (m x c, column-major).
__global__ void __launch_bounds__(kThreads)
    gather_kernel(SolveView xv, const int32_t* __restrict__ orbs, int32_t m, double2* __restrict__ out) {
    const int64_t p = point_of(xv, blockIdx.y);
    const double2* X = xv.x + p * xv.x_stride;
    double2* O = out + p * int64_t(m) * xv.c;
    for (int idx = blockIdx.x * kThreads + threadIdx.x; idx < m * xv.c; idx += gridDim.x * kThreads)
        O[idx] = X[orbs[idx % m] + int64_t(idx / m) * xv.ldx];
}

// The adjoint's small product t = W_L^H psi_L per point (cl x cr, column-major, stride cl * cr):
// one thread per entry, the dot product over the nl left-slice orbitals in a fixed order.
__global__ void __launch_bounds__(kThreads)
    adjoint_t_kernel(SolveView psi, const int32_t* __restrict__ left_orbs, int32_t nl, const double2* __restrict__ wl,
                     int32_t cl, double2* __restrict__ t) {
    const int64_t p = blockIdx.y;
    const int e = blockIdx.x * kThreads + threadIdx.x;
    if (e >= cl * psi.c) return;
    const int d = e % cl, c = e / cl;
    const double2* W = wl + p * int64_t(nl) * cl + int64_t(d) * nl;
    const double2* X = psi.x + p * psi.x_stride + int64_t(c) * psi.ldx;
    double2 acc{0.0, 0.0};
    for (int a = 0; a < nl; ++a) {
        const double2 w = W[a], x = X[left_orbs[a]];
        acc.x += w.x * x.x + w.y * x.y;  // conj(w) x
        acc.y += w.x * x.y - w.y * x.x;
    }
    t[p * int64_t(cl) * psi.c + e] = acc;
}

// dT/d(eps_a) = 2 Re sum_{c,d} psi[a, c] conj(t[d, c]) phi[a, d] on the rows of a block, one
// thread per (row, point): no accumulation across threads, so the result does not depend on the
// schedule. grad has n_orb entries per point.
__global__ void __launch_bounds__(kThreads)
    adjoint_grad_kernel(SolveView psi, SolveView phi, const double2* __restrict__ t, const int32_t* __restrict__ rows,
                        int32_t nrows, double* __restrict__ grad, int64_t n_orb) {
    const int64_t p = blockIdx.y;
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= nrows) return;
    const int r = rows[i], cr = psi.c, cl = phi.c;
    const double2* X = psi.x + p * psi.x_stride;
    const double2* Y = phi.x + p * phi.x_stride;
    const double2* T = t + p * int64_t(cl) * cr;
    double acc = 0.0;
    for (int c = 0; c < cr; ++c) {
        double2 u{0.0, 0.0};  // u = sum_d conj(t[d, c]) phi[r, d]
        for (int d = 0; d < cl; ++d) {
            const double2 tv = T[d + int64_t(c) * cl], y = Y[r + int64_t(d) * phi.ldx];
            u.x += tv.x * y.x + tv.y * y.y;
            u.y += tv.x * y.y - tv.y * y.x;
        }
        const double2 x = X[r + int64_t(c) * psi.ldx];
        acc += x.x * u.x - x.y * u.y;
    }
    grad[p * n_orb + r] = 2.0 * acc;
}

}  // namespace

void launch_adjoint(const SolveView& psi, const SolveView& phi, const int32_t* left_orbs, int32_t nl,
                    const double2* wl, double2* t, const int32_t* rows, int32_t nrows, double* grad, int64_t n_orb,
                    int32_t points, hipStream_t st) {
    if (points <= 0 || psi.c <= 0 || phi.c <= 0) return;
    const int entries = phi.c * psi.c;
    hipLaunchKernelGGL(adjoint_t_kernel, dim3(unsigned((entries + kThreads - 1) / kThreads), unsigned(points)),
                       dim3(kThreads), 0, st, psi, left_orbs, nl, wl, phi.c, t);
    hipLaunchKernelGGL(adjoint_grad_kernel, dim3(unsigned((nrows + kThreads - 1) / kThreads), unsigned(points)),
                       dim3(kThreads), 0, st, psi, phi, t, rows, nrows, grad, n_orb);
}

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
                     int32_t nr, int32_t points, hipStream_t st, bool transpose) {
    if (points <= 0 || xv.c <= 0) return;
    hipLaunchKernelGGL(residual_kernel, rows_grid(int64_t(nrows) * ((xv.c + kResCols - 1) / kResCols), points),
                       dim3(kThreads), 0, st, pv, xv, b, d,
                       rows, nrows, row_ptr, cols, left_pos, right_pos, left_orbs, nl, right_orbs, nr,
                       transpose ? 1 : 0);
}

int32_t goal_chunks(int32_t nrows) { return std::max<int32_t>(1, (nrows + kGoalRows - 1) / kGoalRows); }

size_t goal_part_words(int32_t nrows, int32_t points, int32_t pairs, int32_t norms) {
    return size_t(goal_chunks(nrows)) * size_t(points) * size_t(pairs + norms);
}

void launch_goal(const SolveView& xv, const double2* res, const double2* phi, int64_t phi_stride, const double2* wl,
                 int64_t wl_stride, int32_t nl, int32_t cl, const int32_t* left_orbs, const double2* nrm,
                 int64_t nrm_stride, int32_t nc, int32_t nrm_at, const int32_t* rows, int32_t nrows, double2* part,
                 double* out, int64_t out_stride, int32_t points, hipStream_t st) {
    if (points <= 0) return;
    GoalArgs ga{xv,   res,    phi,   phi_stride, wl,   wl_stride,           nl,  cl,  left_orbs, nrm,
                nrm_stride, nc, nrm_at, rows, nrows, part, goal_chunks(nrows), out, out_stride};
    if ((wl ? cl * xv.c : 0) + (nrm ? nc : 0) == 0) return;
    hipLaunchKernelGGL(goal_part_kernel, dim3(unsigned(ga.chunks), unsigned(points)), dim3(kThreads), 0, st, ga);
    hipLaunchKernelGGL(goal_sum_kernel, dim3(unsigned(points)), dim3(kThreads), 0, st, ga);
}

void launch_update(const SolveView& xv, const double2* d, const int32_t* rows, int32_t nrows, double* maxes,
                   int32_t points, int32_t all_points, hipStream_t st) {
    if (points <= 0 || xv.c <= 0) return;
    if (hipMemsetAsync(maxes, 0, size_t(all_points) * 2 * sizeof(double), st) != hipSuccess)
        throw std::runtime_error("hipMemsetAsync of the refinement maxima failed");
    hipLaunchKernelGGL(update_kernel, rows_grid(int64_t(nrows) * xv.c, points), dim3(kThreads), 0, st, xv, d, rows,
                       nrows, reinterpret_cast<unsigned long long*>(maxes));
}

void launch_gather(const SolveView& xv, const int32_t* orbs, int32_t m, double2* out, int32_t points, hipStream_t st) {
    if (points <= 0 || m <= 0 || xv.c <= 0) return;
    hipLaunchKernelGGL(gather_kernel, rows_grid(int64_t(m) * xv.c, points), dim3(kThreads), 0, st, xv, orbs, m, out);
}

using SolveKernel = void (*)(const NodeDesc*, const int32_t*, PointsView, SolveView);

// Shape of a blocked solve launch over fronts of at most `rows` rows and c columns: threads,
// columns per workgroup (c itself up to two, kSolveCols up to four, else kSolveColsWide; fewer when
// the LDS would pass 64 KB; one column fits up to kSolveMaxLds) and LDS bytes (the columns, then
// the partial sums).
struct SolveShape {
    int nt, cc;
    size_t lds;
};
SolveShape solve_shape(int rows, int c, int nt, bool single) {
    auto bytes = [&](int cc) { return size_t(rows + nt) * cc * sizeof(double2); };
    // kSolveColsWide for complex f32 factors only (complex f64 ones would spill at that width).
    int cc = c <= 2 ? std::max(c, 1) : (c <= kSolveCols || !single) ? kSolveCols : kSolveColsWide;
    if (cc == kSolveColsWide && bytes(cc) > 64 * 1024) cc = kSolveCols;
    if (bytes(cc) > 64 * 1024) cc = 1;
    return {nt, cc, bytes(cc)};
}
SolveShape solve_shape(int rows, int c, bool single) { return solve_shape(rows, c, rows <= 128 ? 64 : 256, single); }

template <template <class, int, int> class K, class TF, int NT>
SolveKernel by_cc(int cc) {
    if constexpr (std::is_same<TF, float2>::value)
        if (cc == kSolveColsWide) return K<TF, NT, kSolveColsWide>::fn;
    return cc == kSolveCols ? K<TF, NT, kSolveCols>::fn : cc == 2 ? K<TF, NT, 2>::fn : K<TF, NT, 1>::fn;
}
template <class TF, int NT, int CC>
struct Fwd {
    static constexpr SolveKernel fn = fwd_blocked_kernel<TF, NT, CC>;
};
template <class TF, int NT, int CC>
struct Bwd {
    static constexpr SolveKernel fn = bwd_blocked_kernel<TF, NT, CC>;
};
template <class, int NT, int CC>
struct Root {
    static constexpr SolveKernel fn = root_blocked_kernel<NT, CC>;
};
template <class TF, int NT, int CC>
struct BwdT {
    static constexpr SolveKernel fn = bwd_t_blocked_kernel<TF, NT, CC>;
};
template <class, int NT, int CC>
struct RootT {
    static constexpr SolveKernel fn = root_t_blocked_kernel<NT, CC>;
};
template <template <class, int, int> class K, class TF>
SolveKernel pick(const SolveShape& s) {
    return s.nt == 64 ? by_cc<K, TF, 64>(s.cc) : by_cc<K, TF, 256>(s.cc);
}

// Dynamic LDS past 64 KB has to be allowed per kernel, once.
void allow_solve_lds(SolveKernel k, size_t lds) {
    if (lds <= 64 * 1024) return;
    static std::mutex mu;
    static std::set<const void*> done;
    std::lock_guard<std::mutex> g(mu);
    if (done.insert(reinterpret_cast<const void*>(k)).second)
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  kSolveMaxLds);
}

bool solve_fits(int rows) { return solve_shape(rows, 1, 256, false).lds <= size_t(kSolveMaxLds); }

void launch_solve_forward(const NodeDesc* nodes, const int32_t* level, int32_t count, int32_t max_n,
                          const PointsView& pv, const SolveView& sv, int32_t points, bool single, bool transpose,
                          hipStream_t st) {
    if (count <= 0 || points <= 0 || sv.c <= 0) return;
    if (transpose) {
        // all n rows of a front's columns in LDS (the separator, and the boundary's update vector)
        if (int64_t(max_n) * sv.c > kSolveLds) throw std::runtime_error("transposed forward solve beyond its LDS");
        SolveKernel k = single ? solve_forward_t_kernel<float2> : solve_forward_t_kernel<double2>;
        hipLaunchKernelGGL(k, dim3(unsigned(count), unsigned(points)), dim3(kThreads),
                           size_t(kSolveLds) * sizeof(double2), st, nodes, level, pv, sv);
        return;
    }
    const SolveShape s = solve_shape(max_n, sv.c, single);
    SolveKernel k = single ? pick<Fwd, float2>(s) : pick<Fwd, double2>(s);
    allow_solve_lds(k, s.lds);
    hipLaunchKernelGGL(k, dim3(unsigned(count * ((sv.c + s.cc - 1) / s.cc)), unsigned(points)), dim3(s.nt), s.lds, st,
                       nodes, level, pv, sv);
}

void launch_solve_backward(const NodeDesc* nodes, const int32_t* level, int32_t count, int32_t max_n,
                           const PointsView& pv, const SolveView& sv, int32_t points, bool single, bool transpose,
                           hipStream_t st) {
    if (count <= 0 || points <= 0 || sv.c <= 0) return;
    const SolveShape s = solve_shape(max_n, sv.c, single);
    SolveKernel k = transpose ? (single ? pick<BwdT, float2>(s) : pick<BwdT, double2>(s))
                              : (single ? pick<Bwd, float2>(s) : pick<Bwd, double2>(s));
    allow_solve_lds(k, s.lds);
    hipLaunchKernelGGL(k, dim3(unsigned(count * ((sv.c + s.cc - 1) / s.cc)), unsigned(points)), dim3(s.nt), s.lds, st,
                       nodes, level, pv, sv);
}

void launch_solve_root(const NodeDesc* nodes, const int32_t* level, int32_t root_ne, const PointsView& pv,
                       const SolveView& sv, int32_t points, bool transpose, hipStream_t st) {
    if (points <= 0 || sv.c <= 0) return;
    const SolveShape s = solve_shape(root_ne, sv.c, 256, false);
    const dim3 grid(unsigned((sv.c + s.cc - 1) / s.cc), unsigned(points));
    const SolveKernel k = transpose ? by_cc<RootT, double2, 256>(s.cc) : by_cc<Root, double2, 256>(s.cc);
    allow_solve_lds(k, s.lds);
    hipLaunchKernelGGL(k, grid, dim3(256), s.lds, st, nodes, level, pv, sv);
}

}  // namespace caroli::gpu
