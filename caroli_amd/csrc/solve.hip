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
#include <mutex>
#include <set>
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

// Transposed solves (A^T x = b, the adjoint's second set of columns), one workgroup per front and
// point with all columns in LDS (ne * c <= kSolveLds).
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

template <class TF>
__global__ void __launch_bounds__(kThreads)
    solve_backward_t_kernel(const NodeDesc* __restrict__ nodes, const int32_t* __restrict__ level, PointsView pv,
                            SolveView sv) {
    extern __shared__ double2 y[];  // ne x c
    const NodeDesc nd = nodes[level[blockIdx.x]];
    const int64_t p = point_of(sv, blockIdx.y);
    const TF* F = reinterpret_cast<const TF*>(pv.arena) + p * pv.arena_stride + nd.off;
    const int32_t* piv = pv.piv + p * pv.piv_stride + nd.piv_off;
    const int32_t* vars = sv.vars + nd.var_off;
    double2* X = sv.x + p * sv.x_stride;
    const int n = nd.n, ne = nd.ne, nb = n - ne, c = sv.c, tid = threadIdx.x;
    auto f = [&](int i, int j) { return widen(F[i + int64_t(j) * n]); };
    // y = x_e - L21^T x_b
    for (int idx = tid; idx < ne * c; idx += kThreads) {
        int k = idx % ne, col = idx / ne;
        double2 acc = X[vars[k] + int64_t(col) * sv.ldx];
        for (int b = 0; b < nb; ++b) {
            double2 xb = X[vars[ne + b] + int64_t(col) * sv.ldx];
            acc = csub(acc, cmul(f(ne + b, k), xb));
        }
        y[idx] = acc;
    }
    __syncthreads();
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
    for (int idx = tid; idx < ne * c; idx += kThreads) X[vars[idx % ne] + int64_t(idx / ne) * sv.ldx] = y[idx];
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
constexpr int kGemvBatch = 4;  // factor entries a lane loads at once in the matrix-vector products
// Waves per SIMD the blocked solve kernels are compiled for (at most 128 registers a lane, which
// holds 4 columns without spilling): the solves are latency-bound and need the occupancy more
// than wider column chunks.
constexpr int kSolveWaves = 4;
constexpr int kSolveCols = 4;

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
    const int RP = min(NT, (R + 63) / 64 * 64), G = NT / RP, rr = tid % RP, g = tid / RP;
    const int kc = (K + G - 1) / G, ka = k0 + min(g, G - 1) * kc, kz = g < G ? min(k1, ka + kc) : ka;
    for (int base = 0; base < R; base += RP) {
        const int r = r0 + base + rr;
        const int64_t rc = min(r, r1 - 1);
        double2 acc[CC];
#pragma unroll
        for (int col = 0; col < CC; ++col) acc[col] = double2{0.0, 0.0};
        int k = ka;
        for (; k + kGemvBatch <= kz; k += kGemvBatch) {
            double2 l[kGemvBatch];
#pragma unroll
            for (int u = 0; u < kGemvBatch; ++u) l[u] = widen(F[rc + int64_t(k + u) * ld]);
#pragma unroll
            for (int u = 0; u < kGemvBatch; ++u)
#pragma unroll
                for (int col = 0; col < CC; ++col) acc[col] = cadd(acc[col], cmul(l[u], y[k + u + col * ldy]));
        }
        for (; k < kz; ++k) {
            const double2 l = widen(F[rc + int64_t(k) * ld]);
#pragma unroll
            for (int col = 0; col < CC; ++col) acc[col] = cadd(acc[col], cmul(l, y[k + col * ldy]));
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

// Transposed dense solve with the root's LU (its leading ne x ne block, complex f64).
__global__ void __launch_bounds__(kThreads)
    solve_root_t_kernel(const NodeDesc* __restrict__ nodes, const int32_t* __restrict__ level, PointsView pv, SolveView sv) {
    extern __shared__ double2 y[];
    const NodeDesc nd = nodes[level[0]];
    const int64_t p = point_of(sv, blockIdx.x);
    const double2* F = reinterpret_cast<const double2*>(pv.root_arena) + p * pv.root_stride + nd.off;
    const int32_t* piv = pv.piv + p * pv.piv_stride + nd.piv_off;
    const int32_t* vars = sv.vars + nd.var_off;
    double2* X = sv.x + p * sv.x_stride;
    const int n = nd.n, ne = nd.ne, c = sv.c, tid = threadIdx.x;
    auto f = [&](int i, int j) { return F[i + int64_t(j) * n]; };
    for (int idx = tid; idx < ne * c; idx += kThreads) y[idx] = X[vars[idx % ne] + int64_t(idx / ne) * sv.ldx];
    __syncthreads();
    if (sv.gather) gather_updates<kThreads>(nd, sv, p, 0, c, y, ne);
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
    {
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

// D = B - (z - H - Sigma_L - Sigma_R) X on the rows of a block, in FP64 (the residual of the
// refinement). Sigma_L acts on the left-slice rows (positions left_pos), Sigma_R on the right.
__global__ void __launch_bounds__(kThreads)
    residual_kernel(PointsView pv, SolveView xv, const double2* __restrict__ b, double2* __restrict__ d,
                    const int32_t* __restrict__ rows, int32_t nrows, const int64_t* __restrict__ row_ptr,
                    const uint32_t* __restrict__ cols, const int32_t* __restrict__ left_pos,
                    const int32_t* __restrict__ right_pos, const int32_t* __restrict__ left_orbs, int32_t nl,
                    const int32_t* __restrict__ right_orbs, int32_t nr) {
    const int64_t p = point_of(xv, blockIdx.y);
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
                     int32_t nr, int32_t points, hipStream_t st) {
    if (points <= 0 || xv.c <= 0) return;
    hipLaunchKernelGGL(residual_kernel, rows_grid(int64_t(nrows) * xv.c, points), dim3(kThreads), 0, st, pv, xv, b, d,
                       rows, nrows, row_ptr, cols, left_pos, right_pos, left_orbs, nl, right_orbs, nr);
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
// columns per workgroup (kSolveCols, or one when c is one or the LDS would pass 64 KB; one column
// fits up to kSolveMaxLds) and LDS bytes (the columns, then the partial sums).
struct SolveShape {
    int nt, cc;
    size_t lds;
};
SolveShape solve_shape(int rows, int c, int nt) {
    auto bytes = [&](int cc) { return size_t(rows + nt) * cc * sizeof(double2); };
    int cc = c <= 1 ? 1 : kSolveCols;
    if (bytes(cc) > 64 * 1024) cc = 1;
    return {nt, cc, bytes(cc)};
}
SolveShape solve_shape(int rows, int c) { return solve_shape(rows, c, rows <= 128 ? 64 : 256); }

template <template <class, int, int> class K, class TF, int NT>
SolveKernel by_cc(int cc) {
    return cc == kSolveCols ? K<TF, NT, kSolveCols>::fn : K<TF, NT, 1>::fn;
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

bool solve_fits(int rows) { return solve_shape(rows, 1, 256).lds <= size_t(kSolveMaxLds); }

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
    const SolveShape s = solve_shape(max_n, sv.c);
    SolveKernel k = single ? pick<Fwd, float2>(s) : pick<Fwd, double2>(s);
    allow_solve_lds(k, s.lds);
    hipLaunchKernelGGL(k, dim3(unsigned(count * ((sv.c + s.cc - 1) / s.cc)), unsigned(points)), dim3(s.nt), s.lds, st,
                       nodes, level, pv, sv);
}

void launch_solve_backward(const NodeDesc* nodes, const int32_t* level, int32_t count, int32_t max_n,
                           const PointsView& pv, const SolveView& sv, int32_t points, bool single, bool transpose,
                           hipStream_t st) {
    if (count <= 0 || points <= 0 || sv.c <= 0) return;
    if (transpose) {
        SolveKernel k = single ? solve_backward_t_kernel<float2> : solve_backward_t_kernel<double2>;
        hipLaunchKernelGGL(k, dim3(unsigned(count), unsigned(points)), dim3(kThreads),
                           size_t(kSolveLds) * sizeof(double2), st, nodes, level, pv, sv);
        return;
    }
    const SolveShape s = solve_shape(max_n, sv.c);
    SolveKernel k = single ? pick<Bwd, float2>(s) : pick<Bwd, double2>(s);
    allow_solve_lds(k, s.lds);
    hipLaunchKernelGGL(k, dim3(unsigned(count * ((sv.c + s.cc - 1) / s.cc)), unsigned(points)), dim3(s.nt), s.lds, st,
                       nodes, level, pv, sv);
}

void launch_solve_root(const NodeDesc* nodes, const int32_t* level, int32_t root_ne, const PointsView& pv,
                       const SolveView& sv, int32_t points, bool transpose, hipStream_t st) {
    if (points <= 0 || sv.c <= 0) return;
    if (transpose) {
        hipLaunchKernelGGL(solve_root_t_kernel, dim3(unsigned(points)), dim3(kThreads),
                           size_t(kSolveLds) * sizeof(double2), st, nodes, level, pv, sv);
        return;
    }
    const SolveShape s = solve_shape(root_ne, sv.c, 256);
    const dim3 grid(unsigned((sv.c + s.cc - 1) / s.cc), unsigned(points));
    const SolveKernel k = by_cc<Root, double2, 256>(s.cc);
    allow_solve_lds(k, s.lds);
    hipLaunchKernelGGL(k, grid, dim3(256), s.lds, st, nodes, level, pv, sv);
}

}  // namespace caroli::gpu
