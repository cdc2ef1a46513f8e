// Device-side descriptors of fronts and the launchers of the front kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace caroli::gpu {

/// One front of an elimination tree, as the kernels see it. The front F (n x n, column-major,
/// leading dimension n) sits at `off` complex words into a point's arena. Columns [0, ne) are
/// eliminated with row pivoting; the pivots of columns [0, ne1) are tested against the largest
/// entry of F[0:ne1, 0:ne1] (a failure delays the front), those of [ne1, ne) only for zero.
struct NodeDesc {
    int64_t off;
    int32_t n, ne, ne1;
    int32_t ent_beg, ent_end;    // H entries of the separator
    int32_t diag_beg, diag_end;  // positions receiving z
    int32_t kid_beg, kid_end;    // children
    int32_t root;                // 1: subtract the self-energies and add the borders
    int32_t nl, nr, right_at, m; // root: lead slices (from position nd), border size m
    int32_t nd;                  // root: variables delayed up to it
    int32_t slot;                // status / pivot slot of the front within a point
    int64_t piv_off;             // first pivot index of the front within a point
    int64_t var_off;             // first of its n orbital indices in the plan's variable list
    int64_t upd_off;             // its boundary's first entry in a point's update vectors (sum of the
                                 // earlier fronts' n - ne)
};

struct KidDesc;

/// Right-hand sides of the multifrontal solves: per point, c columns of n_orb complex f64
/// (column-major, leading dimension ldx), overwritten by the solution.
struct SolveView {
    double2* x;
    int64_t x_stride;
    int32_t ldx, c;
    const int32_t* vars;  // orbital index of every front variable (NodeDesc::var_off)
    const int32_t* pts;   // points of the launch (grid index q -> point pts[q]); null: all, in order
    // The forward solve's update vectors: per point (stride upd_stride) and front, its boundary's
    // c columns (n - ne rows each) at c * NodeDesc::upd_off, written by the front and gathered by
    // its parent through the children's maps (kids, maps), in the order of the children.
    double2* upd;
    int64_t upd_stride;
    const KidDesc* kids;
    const int32_t* maps;
    int32_t gather;  // 1: the forward solve ran, the root gathers its children's updates
};

/// An H entry: F[a, b] -= conj ? conj(vals[k]) : vals[k].
struct Entry {
    int32_t a, b, k, conj;
};

/// A child's Schur complement (size x size, leading dimension ld) at `src` in the kid arena,
/// added at the front positions maps[map_beg + i]; maps[inv_beg + q] is the child's index of
/// front position q (-1 when the child does not touch it), n entries.
struct KidDesc {
    int64_t src;
    int32_t ld, size, map_beg, inv_beg;
    int64_t upd;  // the child's NodeDesc::upd_off (its update vector in the forward solve)
};

/// Per-point inputs and outputs of one launch.
struct PointsView {
    const double2* const* vals;  // device CSR values of each point's system
    const double2* z;            // energy of each point
    const double2* sigma;        // Sigma_L (nl^2) then Sigma_R (nr^2) of each point
    const double2* gamma;        // Gamma_L then Gamma_R of each point
    const double2* w;            // W_L (nl x cl, at 0) and W_R (nr x cr, at nl^2): Gamma = W W^H
    const int32_t* wcols;        // cl, cr of each point
    int64_t sigma_stride;
    void* arena;                 // fronts of each point
    int64_t arena_stride;        // in complex words of the front type
    void* root_arena;            // root fronts (complex f64)
    int64_t root_stride;
    int32_t* piv;
    int64_t piv_stride;
    int32_t* status;  // per point: one slot per front
    int32_t status_stride;
    double* t_out;    // per point
    int32_t keep;     // 1: small fronts write their factors back (for the solves), 0: only S
    // Multi-terminal roots (nlab > 0): the root's border block holds -G on all kept orbitals;
    // wlab holds the channels of every label (kept x lab_cols per point, wlab_stride apart, zero
    // past a point's columns), label l in columns lab_off[l] .. lab_off[l + 1] (nlab + 1 per
    // point), and t_lab receives T between labels (nlab x nlab per point, row: into, column: from).
    const double2* wlab;
    const int32_t* lab_off;
    int64_t wlab_stride;
    int32_t lab_cols, nlab;
    double* t_lab;
};

/// A kTile x kTile tile of the Schur block of front `node` (rows i0.., columns j0.. of S).
struct Tile {
    int32_t node, i0, j0, pad;
};

/// Status codes written by the kernels.
enum : int32_t { kFrontOk = 0, kFrontDelayed = 5, kFrontSingular = 1 };

/// Largest front eliminated in LDS, per precision (complex f64: 64 KB of LDS, f32: 63 KB), and
/// the largest the LDS kernel accepts at all (its row permutation).
constexpr int kLdsFrontC128 = 64;
constexpr int kLdsFrontC64 = 90;
constexpr int kLdsMaxN = 128;
/// Panel width of the border factorisation, LDS budget of its panel, Schur tile size.
constexpr int kPanel = 16;
constexpr int kBorderLdsBytes = 64 * 1024;
/// Separators up to kF11Lds variables are factored in LDS (f11_kernel) and their border strips
/// solved kTrsmSpan columns / rows per workgroup (trsm_kernel).
constexpr int kF11Lds = 64;
constexpr int kTrsmSpan = 64;
constexpr int kTile = 64;
/// Blocked fronts: separators above blocked_min_ne() variables are factored kBlk columns per step
/// (panel, row interchanges and U rows, trailing update; many workgroups per front), kRowCols
/// columns per workgroup of the row step. Each front's pivots are followed by kPivExtra words
/// (the pivot scale of the blocked steps).
constexpr int kBlk = 64;
constexpr int kRowCols = 32;
constexpr int kPivExtra = 2;

// Launchers (fronts.hip). `level` lists node indices into `nodes`; grid = count x points.
void launch_assemble_c128(const NodeDesc* nodes, const int32_t* level, int32_t count, const Entry* ent,
                          const int32_t* diag, const KidDesc* kids, const int32_t* maps, const PointsView& pv,
                          int32_t points, hipStream_t st);
void launch_assemble_c64(const NodeDesc* nodes, const int32_t* level, int32_t count, const Entry* ent,
                         const int32_t* diag, const KidDesc* kids, const int32_t* maps, const PointsView& pv,
                         int32_t points, hipStream_t st);
/// Root assembly (complex f64 root, children in the front precision: c64 when single != 0). With
/// tile_count > 0 the root is written by tiles (root_gather_kernel: children, Sigma, unit borders)
/// and then gets -H and z; else one workgroup per root and point does all of it.
void launch_assemble_root(const NodeDesc* nodes, const int32_t* level, int32_t count, const Entry* ent,
                          const int32_t* diag, const KidDesc* kids, const int32_t* maps, const Tile* tiles,
                          int32_t tile_count, const PointsView& pv, int32_t points, int single, hipStream_t st);
/// Small fronts (n <= nmax <= kLdsMaxN): assembly (or a copy of the given front when !assemble)
/// and elimination in LDS. border_words: the largest (n + 1) ne + (ne + 1) (n - ne) of the fronts
/// (LDS words of the border-only kernel, the default).
void launch_front_lds(const NodeDesc* nodes, const int32_t* level, int32_t count, int32_t nmax, int32_t border_words,
                      const Entry* ent, const int32_t* diag, const KidDesc* kids, const int32_t* maps,
                      const PointsView& pv, int32_t points, double tol, bool single, bool assemble, hipStream_t st);
/// LDS words of a small front in the border-only kernel.
inline int32_t border_lds_words(int32_t n, int32_t ne) { return (n + 1) * ne + (ne + 1) * (n - ne); }
/// Large fronts: the children's contributions to the border (rows or columns below ne) gathered
/// over `tiles`; the border LU (F11, L21, U12; non-root fronts first add -H and z); the Schur
/// update S = F22 - L21 U12 over `tiles`, F22 gathered from the children (gather) or read from the
/// front.
void launch_gather(const NodeDesc* nodes, const Tile* tiles, int32_t count, const KidDesc* kids, const int32_t* maps,
                   const PointsView& pv, int32_t points, bool single, hipStream_t st);
void launch_border(const NodeDesc* nodes, const int32_t* level, int32_t count, int32_t max_ne, const Entry* ent,
                   const int32_t* diag, const PointsView& pv, int32_t points, double tol, bool single, bool root,
                   hipStream_t st);
void launch_schur(const NodeDesc* nodes, const Tile* tiles, int32_t count, const KidDesc* kids, const int32_t* maps,
                  const PointsView& pv, int32_t points, bool single, bool root, bool gather, int tile, hipStream_t st);
/// Large fronts with ne <= kF11Lds: the LU of F11 in LDS (after -H and z), then the triangular
/// solves over `tiles` (pad 0: kTrsmSpan columns of U12 from j0; pad 1: kTrsmSpan rows of L21
/// from i0).
void launch_f11(const NodeDesc* nodes, const int32_t* level, int32_t count, int32_t max_ne, const Entry* ent,
                const int32_t* diag, const PointsView& pv, int32_t points, double tol, bool single, hipStream_t st);
void launch_trsm(const NodeDesc* nodes, const Tile* tiles, int32_t count, int32_t max_ne, const PointsView& pv,
                 int32_t points, bool single, hipStream_t st);
/// Kernel choices per front class, process-wide. The CAROLI_* variables give the defaults;
/// cg_set_tuning and cg_batch_autotune override them (set_tuning_override(nullptr) goes back).
/// With the tuning's schur_tile 0, a level whose Schur blocks have at most this many rows takes
/// 32 x 32 tiles, others 64 x 64.
constexpr int kSchurTile32MaxNb = 256;
struct Tuning {
    int32_t small_fronts = 0;      // 0 front_border_kernel, 1 front_lds_kernel, 2 front_reg_kernel (n <= 64)
    int32_t split = 1;             // separators up to kF11Lds by f11_kernel + trsm_kernel, else border_kernel
    int32_t blocked_min_ne = 128;  // separators above this many variables by the blocked kernels
    int32_t unblocked = 0;         // every large front by factor_global_kernel
    int32_t schur_tile = 0;        // Schur tiles of 64 x 64 (four waves), 32 x 32 (one wave) or by level (0)
    /// The fields that decide a plan's kernel lists (small_fronts is read at each launch).
    bool same_plan(const Tuning& o) const {
        return split == o.split && blocked_min_ne == o.blocked_min_ne && unblocked == o.unblocked &&
               schur_tile == o.schur_tile;
    }
};
Tuning current_tuning();
void set_tuning_override(const Tuning* t);
/// Small-front kernel: 0 border-only LDS (front_border_kernel, the default), 1 the whole front in
/// LDS (CAROLI_SMALL_FRONTS=lds), 2 in registers for fronts up to 64 (CAROLI_SMALL_FRONTS=reg).
int small_front_mode();
/// CAROLI_FRONT_KERNEL=border: every large front through border_kernel (no f11/trsm split).
bool border_only();
/// Smallest separator (exclusive) of the blocked path: CAROLI_FRONT_KERNEL=blocked takes every
/// large front and the root there, =border / =unblocked none.
int blocked_min_ne();
void launch_blocked_panel(const NodeDesc* nodes, const int32_t* list, int32_t count, const Entry* ent,
                          const int32_t* diag, const PointsView& pv, int32_t points, double tol, bool single, bool root,
                          int32_t k0, hipStream_t st);
void launch_blocked_row(const NodeDesc* nodes, const Tile* tiles, int32_t count, const PointsView& pv, int32_t points,
                        bool single, bool root, int32_t k0, hipStream_t st);
void launch_blocked_update(const NodeDesc* nodes, const Tile* tiles, int32_t count, const PointsView& pv,
                           int32_t points, bool single, bool root, int32_t k0, hipStream_t st);
/// The whole elimination by rank-1 updates in global memory (CAROLI_FRONT_KERNEL=unblocked).
void launch_factor_unblocked(const NodeDesc* nodes, const int32_t* level, int32_t count, const PointsView& pv,
                             int32_t points, double tol, bool single, bool root, hipStream_t st);
bool unblocked_fronts();
/// Caroli trace of the root border block: t_out[p] = Tr[Gamma_L G Gamma_R G^H], evaluated as
/// |W_L^H G_LR W_R|_F^2 with cl_max / cr_max the largest channel counts of the launch: two batched
/// MFMA GEMMs (scratch xbuf of points x nl x cr_max, partial of points x tiles) past
/// kTraceFewChannels channels, else one workgroup per point (nl <= kTraceRows) or from Gamma.
constexpr int kTraceRows = 4096;
constexpr int kTraceFewChannels = 16;
void launch_trace(const NodeDesc* root, int32_t nl, int32_t cl_max, int32_t cr_max, double2* xbuf, double* partial,
                  const PointsView& pv, int32_t points, hipStream_t st);
/// T between labelled groups of lead channels of a multi-terminal root (PointsView::wlab): X = G W
/// (scratch xbuf of points x k x lab_cols), |W^H X|^2 entry by entry (ysq, points x lab_cols^2),
/// then the sums over each pair of labels into t_lab. k: the kept orbitals.
void launch_trace_labels(const NodeDesc* root, int32_t k, int32_t lab_cols, double2* xbuf, double* ysq,
                         const PointsView& pv, int32_t points, hipStream_t st);
/// Sizes of the scratch of launch_trace's GEMM path.
inline size_t trace_x_words(int32_t nl, int32_t cr_max) { return size_t(nl) * size_t(cr_max); }
inline size_t trace_partials(int32_t cl_max, int32_t cr_max) {
    return size_t((cl_max + kTile - 1) / kTile) * size_t((cr_max + kTile - 1) / kTile);
}

/// Largest ne * c of a front handled by the transposed solve kernels (LDS words of complex f64).
constexpr int kSolveLds = 4096;
/// LDS a blocked solve workgroup may take; solve_fits(rows): fronts (or a root block) of `rows`
/// rows fit in it (one right-hand-side column at a time at least).
constexpr int kSolveMaxLds = 128 * 1024;
bool solve_fits(int rows);

/// Multifrontal solves with the stored factors, one tree level per launch (grid = fronts x points
/// x column chunks). forward: children first; A x = b (transpose = 0, blocked kernels sized by
/// max_n, the level's largest front) or A^T x = b (transpose = 1). root: the dense root (its ne =
/// root_ne). backward: parents first. `single`: the fronts hold complex f32 factors.
void launch_solve_forward(const NodeDesc* nodes, const int32_t* level, int32_t count, int32_t max_n,
                          const PointsView& pv, const SolveView& sv, int32_t points, bool single, bool transpose,
                          hipStream_t st);
void launch_solve_root(const NodeDesc* nodes, const int32_t* level, int32_t root_ne, const PointsView& pv,
                       const SolveView& sv, int32_t points, bool transpose, hipStream_t st);
void launch_solve_backward(const NodeDesc* nodes, const int32_t* level, int32_t count, int32_t max_n,
                           const PointsView& pv, const SolveView& sv, int32_t points, bool single, bool transpose,
                           hipStream_t st);

/// Refinement helpers (all per point, on c columns of n_orb rows).
void launch_scatter(const SolveView& sv, const int32_t* orbs, int32_t m, const double2* w, int64_t w_stride, bool conj,
                    int32_t points, hipStream_t st);
void launch_residual(const PointsView& pv, const SolveView& xv, const double2* b, double2* d, const int32_t* rows,
                     int32_t nrows, const int64_t* row_ptr, const uint32_t* cols, const int32_t* left_pos,
                     const int32_t* right_pos, const int32_t* left_orbs, int32_t nl, const int32_t* right_orbs,
                     int32_t nr, int32_t points, hipStream_t st, bool transpose = false);
/// The refinement's goal-oriented estimate per point of xv.pts (fixed-order sums): with wl,
/// t' = W_L^H x_L (+ phi^T res with phi) (cl x cr complex at out[0 .. 2 cl cr), W_L nl x cl per
/// point at stride wl_stride, phi with the leading dimension of xv); with nrm, the squared 2-norms
/// of its nc columns over the block rows at out[nrm_at ..]. out_stride doubles per point. part:
/// scratch of goal_part_words(nrows, points, pairs, norms) complex words.
void launch_goal(const SolveView& xv, const double2* res, const double2* phi, int64_t phi_stride, const double2* wl,
                 int64_t wl_stride, int32_t nl, int32_t cl, const int32_t* left_orbs, const double2* nrm,
                 int64_t nrm_stride, int32_t nc, int32_t nrm_at, const int32_t* rows, int32_t nrows, double2* part,
                 double* out, int64_t out_stride, int32_t points, hipStream_t st);
size_t goal_part_words(int32_t nrows, int32_t points, int32_t pairs, int32_t norms);
/// X += D; maxes[2p], maxes[2p + 1] receive max |D| and max |X| of point p (2 doubles per point,
/// all_points of them cleared first).
void launch_update(const SolveView& xv, const double2* d, const int32_t* rows, int32_t nrows, double* maxes,
                   int32_t points, int32_t all_points, hipStream_t st);
void launch_gather(const SolveView& xv, const int32_t* orbs, int32_t m, double2* out, int32_t points, hipStream_t st);
/// The adjoint's contraction per point: t = W_L^H psi_L (cl x cr, cl = phi.c, cr = psi.c; W_L is
/// nl x cl per point) and grad[p n_orb + a] = 2 Re sum_{c,d} psi[a, c] conj(t[d, c]) phi[a, d] on
/// the `rows` of the block (psi = G E_R W_R, phi = G^T E_L conj(W_L)).
void launch_adjoint(const SolveView& psi, const SolveView& phi, const int32_t* left_orbs, int32_t nl,
                    const double2* wl, double2* t, const int32_t* rows, int32_t nrows, double* grad, int64_t n_orb,
                    int32_t points, hipStream_t st);

}  // namespace caroli::gpu
