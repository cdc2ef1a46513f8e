// Host side of the GPU engine: device plans of the elimination trees, the batched level sweep,
// the re-planned sweep of points whose separators delayed, and the device bookkeeping.
//
// The numeric phase of crates/caroli-green/src/nested.rs (Tree::root/fronts/front :246-374,
// root_with_leads :806-845, lead_green :783-802, caroli_trace lib.rs:108-113) runs on the device;
// the host only plans (integers) and moves inputs and results.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <cmath>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>

#include "engine.hpp"
#include "fronts.hpp"

namespace caroli::gpu {

namespace {

struct DeviceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define HIP_CHECK(x)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess) throw DeviceError(std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Pivot tests of crates/caroli-dense/src/field.rs:61 (c64) and :111 (c32). The refined mixed
// mode uses complex f32 fronts only as the preconditioner of an FP64 refinement: it accepts
// pivots down to kTolMixed (still 16 f32 ulps above rounding), so that fewer points fall back to
// complex f64; the refinement decides whether the factors are good enough.
constexpr double kTolDouble = 1e-11;
constexpr double kTolSingle = 1e-5;
constexpr double kTolMixed = 1e-6;

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) {
        o.p = nullptr;
        o.bytes = 0;
    }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    void ensure(size_t b) {
        if (b <= bytes) return;
        if (p) HIP_CHECK(hipFree(p));
        p = nullptr;
        bytes = 0;
        HIP_CHECK(hipMalloc(&p, std::max<size_t>(b, 256)));
        bytes = std::max<size_t>(b, 256);
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

template <class T>
void upload(DevBuf& d, const std::vector<T>& v, hipStream_t st) {
    d.ensure(v.size() * sizeof(T));
    if (v.empty()) return;
    // The host vector may die right after: wait for the copy.
    HIP_CHECK(hipMemcpyAsync(d.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
    HIP_CHECK(hipStreamSynchronize(st));
}

/// The fronts of one launch group (a tree level in one precision, or the root) split by kernel:
/// small fronts in LDS size classes, large fronts with the tiles of their Schur blocks.
struct LevelSet {
    // LDS classes by front size: n <= 16, 32, 48, then up to the precision's limit (the LDS a
    // workgroup takes, and so the workgroups a CU holds, follow the class's largest front).
    static constexpr int kClasses = 4;
    std::vector<int32_t> lds[kClasses];
    int32_t lds_n[kClasses] = {0, 0, 0, 0};       // largest n of each class
    int32_t lds_border[kClasses] = {0, 0, 0, 0};  // largest border_lds_words of each class
    std::vector<int32_t> large;
    int32_t large_ne = 0;  // largest ne of the large fronts
    std::vector<Tile> tiles, gtiles;  // Schur tiles, border tiles (gather)
    // Large fronts split by the border method: small separators (f11 + trsm), others (border).
    std::vector<int32_t> f11, big;
    int32_t f11_ne = 0, big_ne = 0;
    std::vector<Tile> ttiles;  // trsm tiles of the f11 fronts
    // Blocked fronts (largest separator first): per step s of kBlk columns, the first blk_count[s]
    // of them take part; their row and update tiles are rtiles / utiles [off[s], off[s + 1]).
    std::vector<int32_t> blk, blk_count, rt_off, ut_off;
    std::vector<Tile> rtiles, utiles;
    std::vector<double> panel_flops, panel_words, row_flops, row_words, upd_flops, upd_words;
    // Classified for the unblocked fallback (every large front assembled whole and eliminated by
    // rank-1 updates): its lists are the only ones the launches may follow.
    bool unblocked = false;
    int32_t schur_tile = 64;  // the width of the Schur tiles listed in `tiles`
    // Algorithmic work (complex words, flops) for the kernel timers.
    double lds_flops[kClasses] = {}, lds_words[kClasses] = {}, lds_s_words[kClasses] = {}, lds_f_words[kClasses] = {};
    double asm_words = 0, gather_words = 0, border_flops = 0, border_words = 0, schur_flops = 0, schur_words = 0;
    double f11_flops = 0, f11_words = 0, trsm_flops = 0, trsm_words = 0;
    /// Lists the blocked steps of `blk` (sorted here by separator size, largest first).
    void plan_blocked(const std::vector<NodeDesc>& nodes);
};

void LevelSet::plan_blocked(const std::vector<NodeDesc>& nodes) {
    std::stable_sort(blk.begin(), blk.end(), [&](int32_t a, int32_t b) { return nodes[a].ne > nodes[b].ne; });
    const int steps = blk.empty() ? 0 : (nodes[blk[0]].ne + kBlk - 1) / kBlk;
    rt_off.assign(1, 0);
    ut_off.assign(1, 0);
    for (int s = 0; s < steps; ++s) {
        const int k0 = s * kBlk;
        int32_t count = 0;
        double pf = 0, pw = 0, rf = 0, rw = 0, uf = 0, uw = 0;
        for (int32_t q : blk) {
            const NodeDesc& d = nodes[q];
            if (d.ne <= k0) break;
            ++count;
            const int w = std::min(kBlk, d.ne - k0), k1 = k0 + w, n = d.n, ne = d.ne;
            const double R = ne - k0;
            pf += 8.0 * (R * w * w / 2.0);
            pw += 2.0 * R * w;
            for (int j0 = 0; j0 < k0; j0 += kRowCols) rtiles.push_back({q, k0, j0, 0});
            for (int j0 = k1; j0 < n; j0 += kRowCols) rtiles.push_back({q, k0, j0, 1});
            for (int i0 = ne; i0 < n; i0 += kRowCols) rtiles.push_back({q, k0, i0, 2});
            rf += 8.0 * w * w / 2.0 * (n - k1 + n - ne);
            rw += 2.0 * w * (n - k1 + n - ne) + 4.0 * w * k0;
            for (int j0 = k1; j0 < n; j0 += kTile)
                for (int i0 = k1; i0 < n; i0 += kTile)
                    if (i0 < ne || j0 < ne) utiles.push_back({q, i0, j0, 0});
            const double area = double(n - k1) * (ne - k1) + double(ne - k1) * (n - ne);
            uf += 8.0 * w * area;
            uw += 2.0 * area + w * (2.0 * n - k1 - k1);
        }
        blk_count.push_back(count);
        rt_off.push_back(int32_t(rtiles.size()));
        ut_off.push_back(int32_t(utiles.size()));
        panel_flops.push_back(pf);
        panel_words.push_back(pw);
        row_flops.push_back(rf);
        row_words.push_back(rw);
        upd_flops.push_back(uf);
        upd_words.push_back(uw);
    }
}

/// Device pointers of a LevelSet's lists.
struct DevLevel {
    const int32_t* lds[LevelSet::kClasses] = {};
    const int32_t* large = nullptr;
    const Tile* tiles = nullptr;
    const Tile* gtiles = nullptr;
    const int32_t* f11 = nullptr;
    const int32_t* big = nullptr;
    const Tile* ttiles = nullptr;
    const int32_t* blk = nullptr;
    const Tile* rtiles = nullptr;
    const Tile* utiles = nullptr;
};

/// Splits the fronts `list` of `nodes` by kernel (root: the bordered root front, always large).
/// Gather tiles covering the rectangle [r0, r1) x [c0, c1) of a front exactly: kTile along the
/// short side, and along the long side as many multiples of kTile as keep a tile near kTile^2
/// entries (a border strip of 20 columns: tiles of 192 x 20). The extents go in Tile::pad as
/// (rows << 16) | cols.
void add_gather_tiles(std::vector<Tile>& out, int32_t node, int r0, int r1, int c0, int c1) {
    const int R = r1 - r0, C = c1 - c0;
    if (R <= 0 || C <= 0) return;
    int th, tw;
    if (R >= C) {
        tw = std::min(C, kTile);
        th = std::min(R, std::max(kTile, kTile * kTile / tw / kTile * kTile));
    } else {
        th = std::min(R, kTile);
        tw = std::min(C, std::max(kTile, kTile * kTile / th / kTile * kTile));
    }
    for (int j0 = c0; j0 < c1; j0 += tw)
        for (int i0 = r0; i0 < r1; i0 += th) {
            const int h = std::min(th, r1 - i0), w = std::min(tw, c1 - j0);
            out.push_back({node, i0, j0, (h << 16) | w});
        }
}

LevelSet classify(const std::vector<NodeDesc>& nodes, const std::vector<KidDesc>& kids, const std::vector<int32_t>& list,
                  bool single, bool root) {
    LevelSet L;
    const int limit = single ? kLdsFrontC64 : kLdsFrontC128;
    const bool unblocked = unblocked_fronts(), split = !border_only();
    L.unblocked = unblocked;
    // Schur tile width: the tuning's, or by level (0): 32 when every Schur block of the level's
    // large fronts has at most kSchurTile32MaxNb rows, else 64.
    int ts = current_tuning().schur_tile;
    if (ts == 0) {
        int nb_max = 0;
        for (int32_t q : list)
            if (root || nodes[q].n > limit) nb_max = std::max(nb_max, nodes[q].n - nodes[q].ne);
        ts = nb_max <= kSchurTile32MaxNb ? 32 : 64;
    }
    L.schur_tile = ts;
    const int blocked_above = blocked_min_ne();
    for (int32_t q : list) {
        const NodeDesc& d = nodes[q];
        const double ne = d.ne, nb = d.n - d.ne, n2 = double(d.n) * d.n;
        double kid_words = 0;
        for (int c = d.kid_beg; c < d.kid_end; ++c) kid_words += double(kids[c].size) * kids[c].size;
        const double fl = 8.0 / 3.0 * ne * ne * ne + 8.0 * ne * ne * nb + 8.0 * ne * nb * nb;
        if (!root && d.n <= limit) {
            const int c = d.n <= 16 ? 0 : d.n <= 32 ? 1 : d.n <= 48 ? 2 : 3;
            L.lds[c].push_back(q);
            L.lds_n[c] = std::max(L.lds_n[c], d.n);
            L.lds_border[c] = std::max(L.lds_border[c], border_lds_words(d.n, d.ne));
            L.lds_flops[c] += fl;
            L.lds_words[c] += kid_words + 2.0 * (d.ent_end - d.ent_beg);
            L.lds_s_words[c] += nb * nb;
            L.lds_f_words[c] += n2;
            continue;
        }
        L.large.push_back(q);
        L.large_ne = std::max(L.large_ne, d.ne);
        L.asm_words += n2 + kid_words + 2.0 * (d.ent_end - d.ent_beg);
        if (unblocked) {
            L.border_flops += fl;
            L.border_words += 2.0 * n2;
            continue;
        }
        // The children's words split between the border (gathered first) and F22 (by the Schur
        // tiles) in proportion to the areas.
        const double border_share = (n2 - nb * nb) / std::max(n2, 1.0);
        L.gather_words += (n2 - nb * nb) + border_share * kid_words;
        if (d.ne > blocked_above) {
            L.blk.push_back(q);
        } else if (!root && split && d.ne <= kF11Lds) {
            L.f11.push_back(q);
            L.f11_ne = std::max(L.f11_ne, d.ne);
            L.f11_flops += 8.0 / 3.0 * ne * ne * ne;
            L.f11_words += 2.0 * ne * ne;
            L.trsm_flops += 8.0 * ne * ne * nb;
            L.trsm_words += 4.0 * ne * nb + 2.0 * ne * ne * std::ceil(nb / kTrsmSpan);
            for (int o = 0; o < d.n - d.ne; o += kTrsmSpan) {
                L.ttiles.push_back({q, 0, o, 0});
                L.ttiles.push_back({q, o, 0, 1});
            }
        } else {
            L.big.push_back(q);
            L.big_ne = std::max(L.big_ne, d.ne);
            L.border_flops += 8.0 / 3.0 * ne * ne * ne + 8.0 * ne * ne * nb;
            L.border_words += 2.0 * (n2 - nb * nb);
        }
        L.schur_flops += 8.0 * ne * nb * nb;
        L.schur_words += nb * nb + 2.0 * ne * nb + (root ? nb * nb : (1.0 - border_share) * kid_words);
        for (int i0 = 0; i0 < d.n - d.ne; i0 += ts)
            for (int j0 = 0; j0 < d.n - d.ne; j0 += ts) L.tiles.push_back({q, i0, j0, 0});
        if (!root) {
            add_gather_tiles(L.gtiles, q, 0, d.n, 0, d.ne);     // separator columns, all rows
            add_gather_tiles(L.gtiles, q, 0, d.ne, d.ne, d.n);  // separator rows, boundary columns
        } else {
            add_gather_tiles(L.gtiles, q, 0, d.n, 0, d.n);  // the whole root (root_gather_kernel)
        }
    }
    L.plan_blocked(nodes);
    return L;
}

/// Appends a LevelSet's lists to the device list / tile arrays; `at` records the offsets.
struct SetOffsets {
    int32_t lds[LevelSet::kClasses], large, tiles, gtiles, f11, big, ttiles, blk, rtiles, utiles;
};
SetOffsets append_set(const LevelSet& L, std::vector<int32_t>& lv, std::vector<Tile>& tl) {
    SetOffsets o{};
    for (int c = 0; c < LevelSet::kClasses; ++c) {
        o.lds[c] = int32_t(lv.size());
        lv.insert(lv.end(), L.lds[c].begin(), L.lds[c].end());
    }
    o.large = int32_t(lv.size());
    lv.insert(lv.end(), L.large.begin(), L.large.end());
    o.tiles = int32_t(tl.size());
    tl.insert(tl.end(), L.tiles.begin(), L.tiles.end());
    o.gtiles = int32_t(tl.size());
    tl.insert(tl.end(), L.gtiles.begin(), L.gtiles.end());
    o.f11 = int32_t(lv.size());
    lv.insert(lv.end(), L.f11.begin(), L.f11.end());
    o.big = int32_t(lv.size());
    lv.insert(lv.end(), L.big.begin(), L.big.end());
    o.ttiles = int32_t(tl.size());
    tl.insert(tl.end(), L.ttiles.begin(), L.ttiles.end());
    o.blk = int32_t(lv.size());
    lv.insert(lv.end(), L.blk.begin(), L.blk.end());
    o.rtiles = int32_t(tl.size());
    tl.insert(tl.end(), L.rtiles.begin(), L.rtiles.end());
    o.utiles = int32_t(tl.size());
    tl.insert(tl.end(), L.utiles.begin(), L.utiles.end());
    return o;
}
DevLevel resolve(const SetOffsets& o, const int32_t* lv, const Tile* tl) {
    DevLevel d;
    for (int c = 0; c < LevelSet::kClasses; ++c) d.lds[c] = lv + o.lds[c];
    d.large = lv + o.large;
    d.tiles = tl + o.tiles;
    d.gtiles = tl + o.gtiles;
    d.f11 = lv + o.f11;
    d.big = lv + o.big;
    d.ttiles = tl + o.ttiles;
    d.blk = lv + o.blk;
    d.rtiles = tl + o.rtiles;
    d.utiles = tl + o.utiles;
    return d;
}
DevLevel resolve(const SetOffsets& o, const DevBuf& lv, const DevBuf& tl) {
    return resolve(o, lv.as<int32_t>(), tl.as<Tile>());
}

/// Descriptors of one block's fronts (tree nodes, then the root), in host memory.
struct HostPlan {
    std::vector<NodeDesc> nodes;
    std::vector<Entry> ent;
    std::vector<int32_t> diag;
    std::vector<KidDesc> kids;
    std::vector<int32_t> maps;
    /// Node indices per level, and the same split by kernel per front precision (0: complex
    /// f64, 1: complex f32); the root's own split.
    std::vector<std::vector<int32_t>> level_all;
    std::vector<LevelSet> sets[2];
    LevelSet root_set;
    int64_t arena_stride = 0, root_stride = 0, piv_stride = 0;
    int32_t slots = 0;
    int32_t root = 0;  // index of the root descriptor
    /// Orbital of every front variable (NodeDesc::var_off), for the solves.
    std::vector<int32_t> vars;
    /// The block's orbitals, and each orbital's position in the left / right slice (-1 if none).
    std::vector<int32_t> rows, left_pos, right_pos, left_orbs, right_orbs;
    int64_t upd_total = 0;  // sum of the fronts' n - ne (a point's update vectors, per column)
    int32_t max_ne = 0;  // largest ne of a front or the root
    int32_t max_n = 0;   // largest n of a front (the root excluded)
    /// Factor entries one triangular sweep of the solves reads (per front n ne - ne (ne - 1) / 2:
    /// L11 and L21 forward, U11 and U12 backward; the root excluded).
    double tri_words = 0.0;
};

struct DevPlan {
    DevBuf nodes, ent, diag, kids, maps, levels, tiles, vars, rows, left_pos, right_pos, row_ptr, cols, left_orbs,
        right_orbs;
    // Ranges into `levels` of each level's node list (for the solves); the kernel split of every
    // level per front precision, and of the root.
    std::vector<std::pair<int32_t, int32_t>> all;
    std::vector<int32_t> all_max_n;  // largest front of each level
    std::vector<DevLevel> sets[2];
    DevLevel root_set;
    int32_t root_at = 0;  // position of the root index in `levels`
};

}  // namespace

struct Structure {
    std::vector<HostPlan> host;
    std::vector<DevPlan> dev;
    Tuning tuning;  // the kernel choices its plans were classified under
};

namespace {

/// Root front of a block with `delayed` variables before the kept ones (nested.rs:262-280,
/// :806-838), bordered by unit columns on the right slice and unit rows on the left slice:
/// eliminating the first nd + k variables leaves -G_LR in the trailing block.
void root_desc(const Block& blk, const SparseMatrix& h, const std::vector<size_t>& delayed, NodeDesc& d,
               std::vector<Entry>& ent, std::vector<int32_t>& diag) {
    const int nd = int(delayed.size()), k = int(blk.kept.size());
    const int nl = int(blk.left_orbitals.size()), nr = int(blk.right_orbitals.size());
    std::vector<size_t> vars = delayed;
    vars.insert(vars.end(), blk.kept.begin(), blk.kept.end());
    Positions kept(blk.kept);
    d.ent_beg = int32_t(ent.size());
    for (int a = 0; a < k; ++a) {
        size_t i = blk.kept[a];
        for (int64_t e = h.row_ptr[i]; e < h.row_ptr[i + 1]; ++e) {
            int q = kept.get(h.cols[e]);
            if (q >= 0) ent.push_back({nd + a, nd + q, int32_t(e), 0});
        }
    }
    d.ent_end = int32_t(ent.size());
    d.diag_beg = int32_t(diag.size());
    for (int a = 0; a < k; ++a) diag.push_back(nd + a);
    d.diag_end = int32_t(diag.size());
    d.m = std::max(nl, nr);
    d.n = nd + k + d.m;
    d.ne = nd + k;
    d.ne1 = nd;
    d.nd = nd;
    d.root = 1;
    d.nl = nl;
    d.nr = nr;
    d.right_at = blk.right;
}

/// Appends the inverse maps of kids [beg, end) of a front of n variables to `maps`: the child's
/// index of every front position, -1 where the child contributes nothing.
void add_inverse_maps(std::vector<KidDesc>& kids, int32_t beg, int32_t end, std::vector<int32_t>& maps, int n) {
    for (int32_t c = beg; c < end; ++c) {
        KidDesc& kd = kids[c];
        kd.inv_beg = int32_t(maps.size());
        maps.resize(maps.size() + size_t(n), -1);
        for (int a = 0; a < kd.size; ++a) maps[size_t(kd.inv_beg) + size_t(maps[size_t(kd.map_beg) + a])] = a;
    }
}

HostPlan build_plan(const Block& blk, const SparseMatrix& h) {
    const Tree& t = blk.tree;
    HostPlan hp;
    const int nn = int(t.nodes.size());
    int64_t off = 0, piv = 0;
    for (int k = 0; k < nn; ++k) {
        const Plan& pl = t.plans[k];
        NodeDesc d{};
        d.off = off;
        d.n = int32_t(pl.vars.size());
        d.ne = pl.ne;
        d.ne1 = pl.ne;
        d.upd_off = hp.upd_total;
        hp.upd_total += d.n - d.ne;
        d.ent_beg = int32_t(hp.ent.size());
        for (auto& e : pl.entries) hp.ent.push_back({e.a, e.b, int32_t(e.k), e.conj ? 1 : 0});
        d.ent_end = int32_t(hp.ent.size());
        d.diag_beg = int32_t(hp.diag.size());
        hp.diag.insert(hp.diag.end(), pl.diagonal.begin(), pl.diagonal.end());
        d.diag_end = int32_t(hp.diag.size());
        d.kid_beg = int32_t(hp.kids.size());
        for (size_t q = 0; q < t.nodes[k].children.size(); ++q) {
            const NodeDesc& c = hp.nodes[t.nodes[k].children[q]];
            KidDesc kd{};
            kd.src = c.off + c.ne + int64_t(c.ne) * c.n;
            kd.ld = c.n;
            kd.size = c.n - c.ne;
            kd.upd = c.upd_off;
            kd.map_beg = int32_t(hp.maps.size());
            hp.maps.insert(hp.maps.end(), pl.child_pos[q].begin(), pl.child_pos[q].end());
            hp.kids.push_back(kd);
        }
        d.kid_end = int32_t(hp.kids.size());
        add_inverse_maps(hp.kids, d.kid_beg, d.kid_end, hp.maps, d.n);
        d.slot = k;
        d.piv_off = piv;
        d.var_off = int64_t(hp.vars.size());
        for (size_t o : pl.vars) hp.vars.push_back(int32_t(o));
        hp.max_ne = std::max(hp.max_ne, d.ne);
        hp.max_n = std::max(hp.max_n, d.n);
        hp.tri_words += double(d.n) * d.ne - 0.5 * double(d.ne) * (d.ne - 1);
        off += int64_t(d.n) * d.n;
        piv += d.ne + kPivExtra;
        hp.nodes.push_back(d);
    }
    hp.arena_stride = std::max<int64_t>(off, 1);
    NodeDesc r{};
    root_desc(blk, h, {}, r, hp.ent, hp.diag);
    r.off = 0;
    r.kid_beg = int32_t(hp.kids.size());
    Positions kept(blk.kept);
    for (int top : t.top) {
        const NodeDesc& c = hp.nodes[top];
        KidDesc kd{};
        kd.src = c.off + c.ne + int64_t(c.ne) * c.n;
        kd.ld = c.n;
        kd.size = c.n - c.ne;
        kd.upd = c.upd_off;
        kd.map_beg = int32_t(hp.maps.size());
        for (size_t o : t.nodes[top].boundary) {
            int q = kept.get(o);
            if (q < 0) throw std::logic_error("top boundary outside the kept orbitals");
            hp.maps.push_back(q);
        }
        hp.kids.push_back(kd);
    }
    r.kid_end = int32_t(hp.kids.size());
    add_inverse_maps(hp.kids, r.kid_beg, r.kid_end, hp.maps, r.n);
    r.slot = nn;
    r.piv_off = piv;
    r.var_off = int64_t(hp.vars.size());
    for (size_t o : blk.kept) hp.vars.push_back(int32_t(o));
    hp.max_ne = std::max(hp.max_ne, r.ne);
    piv += r.ne + kPivExtra;
    hp.root = nn;
    // Orbitals of the block and their slice positions (residuals of the refinement).
    for (auto& node : t.nodes) hp.rows.insert(hp.rows.end(), node.eliminate.begin(), node.eliminate.end());
    hp.rows.insert(hp.rows.end(), blk.kept.begin(), blk.kept.end());
    std::sort(hp.rows.begin(), hp.rows.end());
    hp.rows.erase(std::unique(hp.rows.begin(), hp.rows.end()), hp.rows.end());
    hp.left_pos.assign(h.n, -1);
    hp.right_pos.assign(h.n, -1);
    for (size_t a = 0; a < blk.left_orbitals.size(); ++a) hp.left_pos[blk.left_orbitals[a]] = int32_t(a);
    for (size_t a = 0; a < blk.right_orbitals.size(); ++a) hp.right_pos[blk.right_orbitals[a]] = int32_t(a);
    for (size_t o : blk.left_orbitals) hp.left_orbs.push_back(int32_t(o));
    for (size_t o : blk.right_orbitals) hp.right_orbs.push_back(int32_t(o));
    hp.nodes.push_back(r);
    hp.root_stride = int64_t(r.n) * r.n;
    hp.piv_stride = std::max<int64_t>(piv, 1);
    hp.slots = nn + 1;
    hp.level_all.assign(t.levels, {});
    for (int k = 0; k < nn; ++k) hp.level_all[t.height[k]].push_back(k);
    for (int prec = 0; prec < 2; ++prec)
        for (auto& l : hp.level_all) hp.sets[prec].push_back(classify(hp.nodes, hp.kids, l, prec == 1, false));
    hp.root_set = classify(hp.nodes, hp.kids, {hp.root}, false, true);
    return hp;
}

void upload_plan(const HostPlan& hp, DevPlan& dp, const SparseMatrix& h, hipStream_t st) {
    upload(dp.vars, hp.vars, st);
    upload(dp.rows, hp.rows, st);
    upload(dp.left_pos, hp.left_pos, st);
    upload(dp.right_pos, hp.right_pos, st);
    upload(dp.row_ptr, h.row_ptr, st);
    upload(dp.cols, h.cols, st);
    upload(dp.left_orbs, hp.left_orbs, st);
    upload(dp.right_orbs, hp.right_orbs, st);
    upload(dp.nodes, hp.nodes, st);
    upload(dp.ent, hp.ent, st);
    upload(dp.diag, hp.diag, st);
    upload(dp.kids, hp.kids, st);
    upload(dp.maps, hp.maps, st);
    std::vector<int32_t> lv;
    std::vector<Tile> tl;
    for (auto& l : hp.level_all) {
        dp.all.push_back({int32_t(lv.size()), int32_t(l.size())});
        int32_t mn = 0;
        for (int32_t q : l) mn = std::max(mn, hp.nodes[q].n);
        dp.all_max_n.push_back(mn);
        lv.insert(lv.end(), l.begin(), l.end());
    }
    std::vector<SetOffsets> offs[2];
    for (int prec = 0; prec < 2; ++prec)
        for (auto& L : hp.sets[prec]) offs[prec].push_back(append_set(L, lv, tl));
    SetOffsets root_off = append_set(hp.root_set, lv, tl);
    dp.root_at = int32_t(lv.size());
    lv.push_back(hp.root);
    upload(dp.levels, lv, st);
    upload(dp.tiles, tl, st);
    for (int prec = 0; prec < 2; ++prec)
        for (auto& o : offs[prec]) dp.sets[prec].push_back(resolve(o, dp.levels, dp.tiles));
    dp.root_set = resolve(root_off, dp.levels, dp.tiles);
}

/// Device copies of the per-point inputs of one launch unit.
struct Inputs {
    DevBuf vals_ptrs, z, sigma, gamma, w, wcols;
    int64_t stride = 0;
    int32_t cl_max = 0, cr_max = 0;  // largest channel counts of the points
    // Multi-terminal points: the channels of every label (kept x lab_cols each) and the labels'
    // column ranges (nlab + 1 each).
    DevBuf wlab, lab_off;
    int64_t wlab_stride = 0;
    int32_t lab_cols = 0, nlab = 0;
};

/// Kernel families timed by cg_kernel_timing (border_kernel stands for factor_global_kernel when
/// CAROLI_FRONT_KERNEL=unblocked; the small-front family is reported under the name of the kernel
/// CAROLI_SMALL_FRONTS selects).
enum Family {
    kFrontLds,
    kGather,
    kAssemble,
    kF11,
    kTrsm,
    kBorder,
    kSchur,
    kRootAssemble,
    kRootBorder,
    kRootSchur,
    kTrace,
    kSolve,
    kBlkPanel,
    kBlkRow,
    kBlkUpdate,
    kSolveT,
    kRefineVec,
    kSolveKept,
    kFamilies
};
const char* const kFamilyNames[kFamilies] = {"front_lds_kernel",    "gather_kernel",         "assemble_kernel",
                                             "f11_kernel",          "trsm_kernel",           "border_kernel",
                                             "schur_kernel",        "assemble_kernel<root>", "border_kernel<root>",
                                             "schur_kernel<root>",  "trace_kernel",          "solve_kernels",
                                             "blocked_panel_kernel", "blocked_row_kernel",   "blocked_update_kernel",
                                             "solve_t_kernels",      "refine_vector_kernels", "solve_kept_kernels"};

/// A stream with the device buffers of one sweep in flight: the arenas, pivots, statuses, results
/// and the trace scratch. Several lanes let the sweeps of several launch units (spin blocks, halves
/// of a block's points, systems of other sizes) run at once: their top tree levels hold few
/// fronts each and leave CUs idle alone.
/// Pinned host memory (hipHostMalloc), for the small copies a refinement iteration makes while
/// other lanes' work is in flight.
template <class T>
struct Pinned {
    T* p = nullptr;
    size_t count = 0;
    Pinned() = default;
    Pinned(const Pinned&) = delete;
    Pinned& operator=(const Pinned&) = delete;
    ~Pinned() {
        if (p) (void)hipHostFree(p);
    }
    void ensure(size_t n) {
        if (n <= count) return;
        if (p) HIP_CHECK(hipHostFree(p));
        p = nullptr;
        count = 0;
        HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(n, 1) * sizeof(T), hipHostMallocDefault));
        count = std::max<size_t>(n, 1);
    }
};

/// The refinement's vectors of one launch unit, kept in its lane between runs: the solution x, the
/// right-hand side b, the residual and correction (d, r, used in turn), the dual
/// solution phi and its residual s, the per-point maxima and estimates (and their partial sums),
/// the list of iterating points, the update vectors; and the pinned host ends.
struct RefineBufs {
    DevBuf x, b, d, r, phi, s, maxes, goal, part, act, upd;
    Pinned<int32_t> h_act;
    Pinned<double> h_maxes, h_goal, h_goal0;
};

struct Lane {
    hipStream_t stream = nullptr;
    DevBuf arena, root_arena, piv, status, tout, trace_x, trace_p, tlab;
    RefineBufs rf;
};
constexpr int kLanes = 4;
constexpr size_t kMinUnitPoints = 32;

/// Device vectors of the batched adjoint, kept between calls: W_R, W_L, psi, phi, t, gradients.
struct AdjointBufs {
    DevBuf wr, wl, psi, phi, t, grad;
};

struct Engine {
    std::mutex mu;
    AdjointBufs adjoint;
    double tol_single = kTolSingle;  // pivot test of complex f32 fronts in the current run
    bool ok = false;
    std::map<uint64_t, std::weak_ptr<Structure>> registry;
    Lane lanes[kLanes];
    Lane* lane = &lanes[0];  // the lane work is issued to (lane 0 outside the concurrent sweeps)
    Inputs scratch_in;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // Per-kernel timing.
    bool timing = false;
    struct Fam {
        int64_t launches = 0;
        double seconds = 0, flops = 0, bytes = 0;
    } fams[kFamilies];
    struct Pending {
        int fam;
        hipEvent_t a, b;
        double flops, bytes;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;

    Engine() {
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return;
        for (auto& l : lanes)
            if (hipStreamCreateWithFlags(&l.stream, hipStreamNonBlocking) != hipSuccess) return;
        if (hipEventCreate(&ev0) != hipSuccess || hipEventCreate(&ev1) != hipSuccess) return;
        ok = true;
    }
    static Engine& get() {
        static Engine e;
        return e;
    }
    hipEvent_t take() {
        if (pool.empty()) {
            hipEvent_t e;
            HIP_CHECK(hipEventCreate(&e));
            return e;
        }
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    /// Runs `launch`, bracketed by events when timing is on.
    template <class F>
    void timed(int fam, double flops, double bytes, F&& launch) {
        if (!timing) {
            launch();
            return;
        }
        hipEvent_t a = take(), b = take();
        HIP_CHECK(hipEventRecord(a, lane->stream));
        launch();
        HIP_CHECK(hipEventRecord(b, lane->stream));
        pending.push_back({fam, a, b, flops, bytes});
    }
    /// Folds the finished timed launches into the counters (after a stream synchronisation).
    void collect() {
        for (auto& p : pending) {
            float ms = 0.f;
            HIP_CHECK(hipEventElapsedTime(&ms, p.a, p.b));
            auto& f = fams[p.fam];
            f.launches += 1;
            f.seconds += ms * 1e-3;
            f.flops += p.flops;
            f.bytes += p.bytes;
            pool.push_back(p.a);
            pool.push_back(p.b);
        }
        pending.clear();
    }
};

/// The plans of a system's structure, shared by the systems of one structure key, and re-planned
/// when the kernel tuning in force classifies fronts differently from the one they were built under.
std::shared_ptr<Structure> structure_of(Engine& eng, const cg_system* s) {
    std::lock_guard<std::mutex> g(s->mu);
    const Tuning tu = current_tuning();
    if (s->structure && s->structure->tuning.same_plan(tu)) return s->structure;
    auto it = eng.registry.find(s->key);
    if (it != eng.registry.end())
        if (auto sp = it->second.lock())
            if (sp->tuning.same_plan(tu)) {
                s->structure = sp;
                return sp;
            }
    auto sp = std::make_shared<Structure>();
    sp->tuning = tu;
    for (auto& blk : s->sym.blocks) sp->host.push_back(build_plan(blk, s->spec.h));
    sp->dev.resize(sp->host.size());
    for (size_t b = 0; b < sp->host.size(); ++b) upload_plan(sp->host[b], sp->dev[b], s->spec.h, eng.lane->stream);
    HIP_CHECK(hipStreamSynchronize(eng.lane->stream));
    eng.registry[s->key] = sp;
    s->structure = sp;
    return sp;
}

const double2* vals_of(Engine& eng, const cg_system* s) {
    std::lock_guard<std::mutex> g(s->mu);
    if (!s->d_vals) {
        size_t bytes = std::max<size_t>(s->spec.h.vals.size(), 1) * sizeof(double2);
        HIP_CHECK(hipMalloc(&s->d_vals, bytes));
        if (!s->spec.h.vals.empty())
            HIP_CHECK(hipMemcpy(s->d_vals, s->spec.h.vals.data(), s->spec.h.vals.size() * sizeof(double2), hipMemcpyHostToDevice));
    }
    return static_cast<const double2*>(s->d_vals);
}

void bind_inputs(const Inputs& in, PointsView& pv) {
    pv.vals = in.vals_ptrs.as<const double2* const>();
    pv.z = in.z.as<const double2>();
    pv.sigma = in.sigma.as<const double2>();
    pv.gamma = in.gamma.as<const double2>();
    pv.w = in.w.as<const double2>();
    pv.wcols = in.wcols.as<const int32_t>();
    pv.sigma_stride = in.stride;
    pv.wlab = in.wlab.as<const double2>();
    pv.lab_off = in.lab_off.as<const int32_t>();
    pv.wlab_stride = in.wlab_stride;
    pv.lab_cols = in.lab_cols;
    pv.nlab = in.nlab;
}

/// Per-point inputs of one block: energies, self-energies and broadenings, Hamiltonian values.
void stage_points(Engine& eng, const std::vector<const cg_system*>& sys, const std::vector<const cg_sigma*>& sig, int b,
                  int nl, int nr, Inputs& in, PointsView& pv) {
    const size_t P = sys.size();
    const int64_t stride = int64_t(nl) * nl + int64_t(nr) * nr;
    in.cl_max = in.cr_max = 0;
    std::vector<const double2*> vp(P);
    std::vector<double2> z(P), sg(P * stride), gm(P * stride), wv(P * stride, double2{0.0, 0.0});
    std::vector<int32_t> wc(2 * P);
    for (size_t p = 0; p < P; ++p) {
        vp[p] = vals_of(eng, sys[p]);
        z[p] = double2{sig[p]->energy, 0.0};
        auto& blk = sig[p]->blocks[b];
        auto put = [&](std::vector<double2>& dst, const CMat& a, const CMat& c) {
            double2* o = dst.data() + p * stride;
            for (size_t e = 0; e < a.a.size(); ++e) o[e] = double2{a.a[e].real(), a.a[e].imag()};
            o += a.a.size();
            for (size_t e = 0; e < c.a.size(); ++e) o[e] = double2{c.a[e].real(), c.a[e].imag()};
        };
        put(sg, blk.sl, blk.sr);
        put(gm, blk.gl, blk.gr);
        // W_L (nl x cl) at 0 and W_R (nr x cr) at nl^2, column-major.
        double2* o = wv.data() + p * stride;
        for (size_t e = 0; e < blk.wl.a.size(); ++e) o[e] = double2{blk.wl.a[e].real(), blk.wl.a[e].imag()};
        o += int64_t(nl) * nl;
        for (size_t e = 0; e < blk.wr.a.size(); ++e) o[e] = double2{blk.wr.a[e].real(), blk.wr.a[e].imag()};
        wc[2 * p] = blk.wl.cols;
        wc[2 * p + 1] = blk.wr.cols;
        in.cl_max = std::max(in.cl_max, blk.wl.cols);
        in.cr_max = std::max(in.cr_max, blk.wr.cols);
    }
    upload(in.vals_ptrs, vp, eng.lane->stream);
    upload(in.z, z, eng.lane->stream);
    upload(in.sigma, sg, eng.lane->stream);
    upload(in.gamma, gm, eng.lane->stream);
    upload(in.w, wv, eng.lane->stream);
    upload(in.wcols, wc, eng.lane->stream);
    in.stride = stride;
    in.nlab = 0;
    in.lab_cols = 0;
    in.wlab_stride = 0;
    if (P > 0 && sig[0]->multi) {
        // W of every label on the kept orbitals (nl of them), padded to the widest point.
        const int32_t L = int32_t(sig[0]->labels.size());
        int32_t cols = 0;
        for (size_t p = 0; p < P; ++p) cols = std::max(cols, sig[p]->mblocks[b].wab.cols);
        std::vector<double2> wl(P * size_t(nl) * size_t(cols), double2{0.0, 0.0});
        std::vector<int32_t> off(P * size_t(L + 1));
        for (size_t p = 0; p < P; ++p) {
            const auto& mb = sig[p]->mblocks[b];
            double2* o = wl.data() + p * size_t(nl) * size_t(cols);
            for (size_t e = 0; e < mb.wab.a.size(); ++e) o[e] = double2{mb.wab.a[e].real(), mb.wab.a[e].imag()};
            std::copy(mb.lab_off.begin(), mb.lab_off.end(), off.begin() + p * size_t(L + 1));
        }
        upload(in.wlab, wl, eng.lane->stream);
        upload(in.lab_off, off, eng.lane->stream);
        in.nlab = L;
        in.lab_cols = cols;
        in.wlab_stride = int64_t(nl) * cols;
    }
    bind_inputs(in, pv);
}

/// The blocked fronts of a LevelSet, step by step (their Schur blocks come with the other large
/// fronts' tiles afterwards).
void run_blocked(Engine& eng, const LevelSet& L, const DevLevel& D, const NodeDesc* nodes, const Entry* ent,
                 const int32_t* diag, const PointsView& pv, int32_t P, double tol, bool fsingle, bool root) {
    const double es = root ? 16.0 : fsingle ? 8.0 : 16.0;
    for (size_t s = 0; s < L.blk_count.size(); ++s) {
        const int32_t k0 = int32_t(s) * kBlk;
        eng.timed(kBlkPanel, P * L.panel_flops[s], P * es * L.panel_words[s], [&] {
            launch_blocked_panel(nodes, D.blk, L.blk_count[s], ent, diag, pv, P, tol, fsingle, root, k0,
                                 eng.lane->stream);
        });
        const int32_t nr = L.rt_off[s + 1] - L.rt_off[s], nu = L.ut_off[s + 1] - L.ut_off[s];
        if (nr > 0)
            eng.timed(kBlkRow, P * L.row_flops[s], P * es * L.row_words[s], [&] {
                launch_blocked_row(nodes, D.rtiles + L.rt_off[s], nr, pv, P, fsingle, root, k0, eng.lane->stream);
            });
        if (nu > 0)
            eng.timed(kBlkUpdate, P * L.upd_flops[s], P * es * L.upd_words[s], [&] {
                launch_blocked_update(nodes, D.utiles + L.ut_off[s], nu, pv, P, fsingle, root, k0, eng.lane->stream);
            });
    }
}

/// Launches the fronts of one LevelSet for the P points of `pv`: the LDS classes, then the large
/// fronts (assembly, border, Schur tiles). `root`: the bordered root front (complex f64 whatever
/// the precision of the fronts below it).
void run_level(Engine& eng, const LevelSet& L, const DevLevel& D, const NodeDesc* nodes, const Entry* ent,
               const int32_t* diag, const KidDesc* kids, const int32_t* maps, const PointsView& pv, int32_t P,
               bool single, bool root) {
    const double es = single ? 8.0 : 16.0, fes = root ? 16.0 : es;
    const double tol = single && !root ? eng.tol_single : kTolDouble;
    for (int c = 0; c < LevelSet::kClasses; ++c) {
        const int32_t count = int32_t(L.lds[c].size());
        if (count == 0) continue;
        const double words = L.lds_words[c] + (pv.keep ? L.lds_f_words[c] : L.lds_s_words[c]);
        eng.timed(kFrontLds, P * L.lds_flops[c], P * es * words, [&] {
            launch_front_lds(nodes, D.lds[c], count, L.lds_n[c], L.lds_border[c], ent, diag, kids, maps, pv, P, tol,
                             single, true,
                             eng.lane->stream);
        });
    }
    const int32_t count = int32_t(L.large.size());
    if (count == 0) return;
    // The plan's own classification, not the tuning in force now: a plan classified for the
    // unblocked path has no tile lists, and one classified for the tiled path no whole-front list.
    const bool fsingle = single && !root, unblocked = L.unblocked;
    if (root || unblocked) {  // the whole front assembled in HBM
        eng.timed(root ? kRootAssemble : kAssemble, 0.0, P * fes * L.asm_words, [&] {
            if (root)
                launch_assemble_root(nodes, D.large, count, ent, diag, kids, maps, D.gtiles, int32_t(L.gtiles.size()), pv,
                                     P, single ? 1 : 0, eng.lane->stream);
            else if (single)
                launch_assemble_c64(nodes, D.large, count, ent, diag, kids, maps, pv, P, eng.lane->stream);
            else
                launch_assemble_c128(nodes, D.large, count, ent, diag, kids, maps, pv, P, eng.lane->stream);
        });
    } else {  // only the border; F22 is gathered by the Schur tiles
        eng.timed(kGather, 0.0, P * fes * L.gather_words, [&] {
            launch_gather(nodes, D.gtiles, int32_t(L.gtiles.size()), kids, maps, pv, P, fsingle, eng.lane->stream);
        });
    }
    if (unblocked) {
        eng.timed(root ? kRootBorder : kBorder, P * L.border_flops, P * fes * L.border_words, [&] {
            launch_factor_unblocked(nodes, D.large, count, pv, P, tol, fsingle, root, eng.lane->stream);
        });
        return;
    }
    run_blocked(eng, L, D, nodes, ent, diag, pv, P, tol, fsingle, root);
    if (!L.f11.empty()) {
        eng.timed(kF11, P * L.f11_flops, P * fes * L.f11_words, [&] {
            launch_f11(nodes, D.f11, int32_t(L.f11.size()), L.f11_ne, ent, diag, pv, P, tol, fsingle, eng.lane->stream);
        });
        eng.timed(kTrsm, P * L.trsm_flops, P * fes * L.trsm_words, [&] {
            launch_trsm(nodes, D.ttiles, int32_t(L.ttiles.size()), L.f11_ne, pv, P, fsingle, eng.lane->stream);
        });
    }
    if (!L.big.empty())
        eng.timed(root ? kRootBorder : kBorder, P * L.border_flops, P * fes * L.border_words, [&] {
            launch_border(nodes, D.big, int32_t(L.big.size()), L.big_ne, ent, diag, pv, P, tol, fsingle, root,
                          eng.lane->stream);
        });
    eng.timed(root ? kRootSchur : kSchur, P * L.schur_flops, P * fes * L.schur_words, [&] {
        launch_schur(nodes, D.tiles, int32_t(L.tiles.size()), kids, maps, pv, P, fsingle, root, !root, L.schur_tile,
                     eng.lane->stream);
    });
}

/// The trace of the root's border block for the P points of `pv` (scratch of the GEMM path from
/// the engine).
void launch_trace_for(Engine& eng, const NodeDesc* root, int32_t nl, const Inputs& in, const PointsView& pv,
                      int32_t P) {
    if (in.nlab > 0) {  // multi-terminal: T between every pair of labels
        if (!pv.t_lab || !pv.lab_off) throw std::logic_error("a multi-terminal trace without its buffers");
        const size_t cols = size_t(std::max(in.lab_cols, 1));
        eng.lane->trace_x.ensure(size_t(P) * size_t(nl) * cols * sizeof(double2));
        eng.lane->trace_p.ensure(size_t(P) * cols * cols * sizeof(double));
        launch_trace_labels(root, nl, in.lab_cols, eng.lane->trace_x.as<double2>(), eng.lane->trace_p.as<double>(), pv,
                            P, eng.lane->stream);
        return;
    }
    double2* xbuf = nullptr;
    double* partial = nullptr;
    if (std::max(in.cl_max, in.cr_max) > kTraceFewChannels) {
        eng.lane->trace_x.ensure(size_t(P) * trace_x_words(nl, in.cr_max) * sizeof(double2));
        eng.lane->trace_p.ensure(size_t(P) * trace_partials(in.cl_max, in.cr_max) * sizeof(double));
        xbuf = eng.lane->trace_x.as<double2>();
        partial = eng.lane->trace_p.as<double>();
    }
    launch_trace(root, nl, in.cl_max, in.cr_max, xbuf, partial, pv, P, eng.lane->stream);
}

/// One sweep over the static plan of block b for the points of `pv`.
void sweep(Engine& eng, const HostPlan& hp, const DevPlan& dp, const PointsView& pv, int32_t P, bool single,
           const Inputs& in) {
    const NodeDesc* nodes = dp.nodes.as<NodeDesc>();
    const Entry* ent = dp.ent.as<Entry>();
    const int32_t* diag = dp.diag.as<int32_t>();
    const KidDesc* kids = dp.kids.as<KidDesc>();
    const int32_t* maps = dp.maps.as<int32_t>();
    const int prec = single ? 1 : 0;
    for (size_t l = 0; l < hp.level_all.size(); ++l)
        run_level(eng, hp.sets[prec][l], dp.sets[prec][l], nodes, ent, diag, kids, maps, pv, P, single, false);
    run_level(eng, hp.root_set, dp.root_set, nodes, ent, diag, kids, maps, pv, P, single, true);
    const NodeDesc& r = hp.nodes[hp.root];
    const double nl = r.nl, nr = r.nr;
    eng.timed(kTrace, 0.0, P * 16.0 * (nl * nl + nr * nr + nl * nr), [&] {
        launch_trace_for(eng, nodes + hp.root, r.nl, in, pv, P);
    });
    HIP_CHECK(hipGetLastError());
}

/// What a front handed to its parent in the re-planned sweep.
struct Handed {
    std::vector<size_t> vars;
    int64_t src;
    int32_t ld, size;
};

/// Re-planned sweep of one point of one block whose separators delayed: level by level, the
/// fronts above delayed children are assembled from the variables actually handed over
/// (nested.rs:331-344), exactly as the reference's unplanned path. Returns the block's T or NaN
/// and the status; for a multi-terminal point, T between its labels in lab_out (nlab^2).
int32_t replanned_point(Engine& eng, const cg_system* s, const cg_sigma* g, int b, bool single, double& t_out,
                        double* lab_out = nullptr) {
    const Block& blk = s->sym.blocks[b];
    const Tree& t = blk.tree;
    const int nn = int(t.nodes.size());
    const size_t es = single ? sizeof(float2) : sizeof(double2);
    std::vector<const cg_system*> one_s{s};
    std::vector<const cg_sigma*> one_g{g};
    PointsView pv{};
    stage_points(eng, one_s, one_g, b, int(blk.left_orbitals.size()), int(blk.right_orbitals.size()), eng.scratch_in, pv);
    DevBuf arena, piv, status, tout, root, dn, dent, ddiag, dkids, dmaps, tlab;
    int64_t used = 0, piv_used = 0;
    const int32_t nlab = eng.scratch_in.nlab;
    if (nlab > 0) {
        if (!lab_out) throw std::logic_error("a multi-terminal point without its label output");
        tlab.ensure(sizeof(double) * size_t(nlab) * size_t(nlab));
        pv.t_lab = tlab.as<double>();
    }
    auto grow = [&](DevBuf& buf, int64_t need_bytes, int64_t keep_bytes) {
        if (size_t(need_bytes) <= buf.bytes) return;
        DevBuf nb;
        nb.ensure(size_t(std::max<int64_t>(need_bytes, 2 * int64_t(buf.bytes))));
        if (keep_bytes > 0) HIP_CHECK(hipMemcpyAsync(nb.p, buf.p, size_t(keep_bytes), hipMemcpyDeviceToDevice, eng.lane->stream));
        std::swap(buf.p, nb.p);
        std::swap(buf.bytes, nb.bytes);
    };
    status.ensure(sizeof(int32_t) * (nn + 1));
    tout.ensure(sizeof(double));
    std::vector<Handed> handed(nn);
    std::vector<int32_t> st_host(nn + 1, 0);
    pv.status = status.as<int32_t>();
    pv.status_stride = nn + 1;
    pv.t_out = tout.as<double>();
    pv.arena_stride = 0;
    pv.piv_stride = 0;
    pv.root_stride = 0;
    for (int l = 0; l < t.levels + 1; ++l) {
        const bool root_level = l == t.levels;
        const int64_t prev_used = used, prev_piv = piv_used;
        std::vector<NodeDesc> nodes;
        std::vector<Entry> ent;
        std::vector<int32_t> diag, maps, list;
        std::vector<KidDesc> kids;
        std::vector<std::vector<size_t>> fvars;
        std::vector<int> which;
        auto add_kids = [&](NodeDesc& d, const std::vector<int>& children, const Positions& pos) {
            d.kid_beg = int32_t(kids.size());
            for (int c : children) {
                KidDesc kd{};
                kd.src = handed[c].src;
                kd.ld = handed[c].ld;
                kd.size = handed[c].size;
                kd.map_beg = int32_t(maps.size());
                for (size_t o : handed[c].vars) {
                    int q = pos.get(o);
                    if (q < 0) throw std::logic_error("handed variable missing from the parent front");
                    maps.push_back(q);
                }
                kids.push_back(kd);
            }
            d.kid_end = int32_t(kids.size());
            // n entries each: for the root that is past the delayed and kept variables of `pos`
            // (its border rows and columns, which no child reaches).
            add_inverse_maps(kids, d.kid_beg, d.kid_end, maps, d.n);
        };
        if (!root_level) {
            for (int k = 0; k < nn; ++k) {
                if (t.height[k] != l) continue;
                const Node& node = t.nodes[k];
                std::vector<const std::vector<size_t>*> hv;
                for (int c : node.children) hv.push_back(&handed[c].vars);
                auto [vars, ne] = front_vars(node, hv);
                Positions pos(vars);
                NodeDesc d{};
                d.n = int32_t(vars.size());
                d.ne = ne;
                d.ne1 = ne;
                d.off = used;
                d.slot = k;
                d.piv_off = piv_used;
                used += int64_t(d.n) * d.n;
                piv_used += ne + kPivExtra;
                d.ent_beg = int32_t(ent.size());
                for (auto& e : separator_entries(s->spec.h, node, pos)) ent.push_back({e.a, e.b, int32_t(e.k), e.conj ? 1 : 0});
                d.ent_end = int32_t(ent.size());
                d.diag_beg = int32_t(diag.size());
                for (size_t o : node.eliminate) diag.push_back(pos.get(o));
                d.diag_end = int32_t(diag.size());
                add_kids(d, node.children, pos);
                list.push_back(int32_t(nodes.size()));
                nodes.push_back(d);
                fvars.push_back(vars);
                which.push_back(k);
            }
        } else {
            Positions in_kept(blk.kept);
            std::vector<size_t> delayed;
            for (int top : t.top)
                for (size_t o : handed[top].vars)
                    if (in_kept.get(o) < 0) delayed.push_back(o);
            std::sort(delayed.begin(), delayed.end());
            delayed.erase(std::unique(delayed.begin(), delayed.end()), delayed.end());
            NodeDesc d{};
            root_desc(blk, s->spec.h, delayed, d, ent, diag);
            std::vector<size_t> vars = delayed;
            vars.insert(vars.end(), blk.kept.begin(), blk.kept.end());
            Positions pos(vars);
            add_kids(d, t.top, pos);
            d.off = 0;
            d.slot = nn;
            d.piv_off = piv_used;
            piv_used += d.ne + kPivExtra;
            list.push_back(0);
            nodes.push_back(d);
        }
        grow(arena, used * int64_t(es), prev_used * int64_t(es));
        grow(piv, piv_used * 4, prev_piv * 4);
        pv.arena = arena.p;
        pv.piv = piv.as<int32_t>();
        upload(dn, nodes, eng.lane->stream);
        upload(dent, ent, eng.lane->stream);
        upload(ddiag, diag, eng.lane->stream);
        upload(dkids, kids, eng.lane->stream);
        upload(dmaps, maps, eng.lane->stream);
        const NodeDesc* dd = dn.as<NodeDesc>();
        if (!root_level) {
            // The same kernels as the static sweep, on this level's re-planned fronts.
            LevelSet L = classify(nodes, kids, list, single, false);
            std::vector<int32_t> lv;
            std::vector<Tile> tl;
            SetOffsets o = append_set(L, lv, tl);
            DevBuf dl, dt;
            upload(dl, lv, eng.lane->stream);
            upload(dt, tl, eng.lane->stream);
            run_level(eng, L, resolve(o, dl, dt), dd, dent.as<Entry>(), ddiag.as<int32_t>(), dkids.as<KidDesc>(),
                      dmaps.as<int32_t>(), pv, 1, single, false);
            HIP_CHECK(hipGetLastError());
            HIP_CHECK(hipMemcpyAsync(st_host.data(), status.p, sizeof(int32_t) * (nn + 1), hipMemcpyDeviceToHost, eng.lane->stream));
            HIP_CHECK(hipStreamSynchronize(eng.lane->stream));
            // A delayed front moves up as assembled (nested.rs:371-372): the fronts whose pivot test
            // failed part-way are assembled again over their partial elimination.
            std::vector<int32_t> failed;
            for (size_t q = 0; q < nodes.size(); ++q)
                if (st_host[which[q]] != kFrontOk) failed.push_back(int32_t(q));
            if (!failed.empty()) {
                DevBuf df;
                upload(df, failed, eng.lane->stream);
                if (single)
                    launch_assemble_c64(dd, df.as<int32_t>(), int32_t(failed.size()), dent.as<Entry>(), ddiag.as<int32_t>(),
                                        dkids.as<KidDesc>(), dmaps.as<int32_t>(), pv, 1, eng.lane->stream);
                else
                    launch_assemble_c128(dd, df.as<int32_t>(), int32_t(failed.size()), dent.as<Entry>(), ddiag.as<int32_t>(),
                                         dkids.as<KidDesc>(), dmaps.as<int32_t>(), pv, 1, eng.lane->stream);
                HIP_CHECK(hipGetLastError());
                HIP_CHECK(hipStreamSynchronize(eng.lane->stream));
            }
            for (size_t q = 0; q < nodes.size(); ++q) {
                int k = which[q];
                const NodeDesc& d = nodes[q];
                if (st_host[k] == kFrontOk) {
                    handed[k] = {t.nodes[k].boundary, d.off + d.ne + int64_t(d.ne) * d.n, d.n, d.n - d.ne};
                } else {  // delayed (or nothing eliminated): the whole front moves up
                    handed[k] = {fvars[q], d.off, d.n, d.n};
                }
            }
        } else {
            const NodeDesc& d = nodes[0];
            root.ensure(size_t(int64_t(d.n) * d.n) * sizeof(double2));
            pv.root_arena = root.p;
            LevelSet L = classify(nodes, kids, list, single, true);
            std::vector<int32_t> lv;
            std::vector<Tile> tl;
            SetOffsets o = append_set(L, lv, tl);
            DevBuf dl, dt;
            upload(dl, lv, eng.lane->stream);
            upload(dt, tl, eng.lane->stream);
            run_level(eng, L, resolve(o, dl, dt), dd, dent.as<Entry>(), ddiag.as<int32_t>(), dkids.as<KidDesc>(),
                      dmaps.as<int32_t>(), pv, 1, single, true);
            launch_trace_for(eng, dd, nodes[0].nl, eng.scratch_in, pv, 1);
            HIP_CHECK(hipGetLastError());
            int32_t rs = 0;
            HIP_CHECK(hipMemcpyAsync(&rs, static_cast<int32_t*>(status.p) + nn, sizeof(int32_t), hipMemcpyDeviceToHost, eng.lane->stream));
            if (nlab > 0) {
                HIP_CHECK(hipMemcpyAsync(lab_out, tlab.p, sizeof(double) * size_t(nlab) * size_t(nlab),
                                         hipMemcpyDeviceToHost, eng.lane->stream));
                t_out = 0.0;
            } else {
                HIP_CHECK(hipMemcpyAsync(&t_out, tout.p, sizeof(double), hipMemcpyDeviceToHost, eng.lane->stream));
            }
            HIP_CHECK(hipStreamSynchronize(eng.lane->stream));
            if (nlab > 0)
                for (int32_t q = 0; q < nlab * nlab; ++q) t_out += lab_out[q];
            if (rs != kFrontOk) {
                t_out = NAN;
                return CG_ERR_SINGULAR;
            }
            if (!std::isfinite(t_out)) return CG_ERR_SINGULAR;
        }
    }
    return CG_OK;
}

/// Multifrontal solve of the columns of `sv` with the factors the last sweep left in the arenas:
/// forward elimination (skipped when the right-hand sides live on the kept orbitals only, where
/// it does nothing; each front's boundary update goes to its slot of sv.upd and its parent gathers
/// it), the dense root, the backward substitution.
void solve(Engine& eng, const HostPlan& hp, const DevPlan& dp, const PointsView& pv, const SolveView& sv_in, int32_t P,
           bool single, bool transpose, bool kept_only) {
    const NodeDesc* nodes = dp.nodes.as<NodeDesc>();
    const int32_t* lv = dp.levels.as<int32_t>();
    SolveView sv = sv_in;
    sv.kids = dp.kids.as<KidDesc>();
    sv.maps = dp.maps.as<int32_t>();
    sv.gather = kept_only ? 0 : 1;
    if (!kept_only && !sv.upd) throw DeviceError("a forward solve needs its update vectors");
    // Algorithmic work: every factor entry of the sweeps read once (the root's LU in complex f64),
    // the right-hand sides read and written once per sweep, 8 flops per complex multiply-add.
    const double es = single ? sizeof(float2) : sizeof(double2), rne = hp.nodes[hp.root].ne;
    const double sweeps = kept_only ? 1.0 : 2.0, words = sweeps * hp.tri_words;
    const double bytes = double(P) * (words * es + rne * rne * sizeof(double2) +
                                      2.0 * sweeps * double(sv_in.ldx) * sv_in.c * sizeof(double2));
    const double flops = double(P) * 8.0 * (words + rne * rne) * sv_in.c;
    eng.timed(transpose ? kSolveT : kept_only ? kSolveKept : kSolve, flops, bytes, [&] {
        if (!kept_only)
            for (size_t l = 0; l < dp.all.size(); ++l)
                launch_solve_forward(nodes, lv + dp.all[l].first, dp.all[l].second, dp.all_max_n[l], pv, sv, P, single,
                                     transpose, eng.lane->stream);
        launch_solve_root(nodes, lv + dp.root_at, hp.nodes[hp.root].ne, pv, sv, P, transpose, eng.lane->stream);
        for (size_t l = dp.all.size(); l-- > 0;)
            launch_solve_backward(nodes, lv + dp.all[l].first, dp.all[l].second, dp.all_max_n[l], pv, sv, P, single,
                                  transpose, eng.lane->stream);
    });
    HIP_CHECK(hipGetLastError());
}

/// Refinement of x = G E_R W_R with the complex f32 factors M and FP64 residuals: x0 = M^-1 b,
/// then x_{k+1} = x_k + d_k with d_k = M^-1 (b - A x_k). T = |t|_F^2 with the transmission
/// amplitudes t = W_L^H x_L, formed on the device per iteration. A point is done when its
/// relative correction r_k = max|d_k| / max|x_{k+1}| is below kRefineTol, or below kRefineEstFrom
/// with the error estimated at the observed contraction, r_k^2 / r_{k-1}, below kRefineEst.
///
/// Optionally (refine_goal()) the dual solution phi0 = M^-T g, g = E_L conj(W_L) (A^T phi = g),
/// is solved once and every iterate is evaluated through the dual-corrected amplitudes
///     t'_k = W_L^H x_k,L + phi0^T (b - A x_k),
/// whose error is t - t'_k = s0^T e_k exactly, with s0 = g - A^T phi0 the dual residual (FP64) and
/// e_k = x - x_k. By Cauchy-Schwarz |t - t'_k|[d, c] <= |s0_d| |e_k,c|, and |e_{k+1}| is at most
/// rho / (1 - rho) times the correction |d_k| just applied (rho the contraction per iteration);
/// t'_{k+1} - t'_k = s0^T d_k, so the error of t'_{k+1} is also about rho / (1 - rho) times that
/// change. A point is then also done when either, carried to T as sum 2 |t'| delta + delta^2, is
/// below kGoalTol (the bound) or kGoalEst (the estimate), with rho taken as kGoalSafety times the
/// larger observed contraction of the corrections and of the amplitudes' changes (r_0 after the
/// first iteration).
constexpr double kRefineTol = 1e-10;
constexpr double kRefineEst = 1e-12;
constexpr double kRefineEstFrom = 1e-7;
constexpr double kGoalTol = 1e-11;
constexpr double kGoalEst = 1e-12;
constexpr double kGoalSafety = 10.0;
constexpr int kRefineMaxIter = 12;
/// From the third iteration on, a point whose correction shrinks by less than 1 / kRefineRate per
/// iteration is left to the complex f64 recomputation (f32 factors of a well-conditioned point
/// contract by 1e-3 or better; a slow one would hold the whole batch in the loop).
constexpr double kRefineRate = 0.25;
/// Once some point of the batch goes to the complex f64 recomputation anyway (a batch of a few
/// points costs one latency-bound sweep, whatever their number), a point still iterating after
/// kRefineJoinFrom iterations joins it when, at its observed contraction, it would need more than
/// one further iteration to reach kRefineTol: the slowest points would otherwise hold every unit
/// of the group in the loop for several more solves.
constexpr int kRefineJoinFrom = 3;

struct Refined {
    std::vector<double> t;
    std::vector<char> ok;
    int64_t iters = 0;
    double worst = 0.0;
    int64_t goal_done = 0;  // points accepted by the goal-oriented bound
};

/// CAROLI_REFINE_TRACE=1 prints, per refinement iteration, how many points are still iterating
/// and the range of their corrections (stderr).
bool refine_trace() {
    const char* v = std::getenv("CAROLI_REFINE_TRACE");
    return v && std::strcmp(v, "1") == 0;
}

/// CAROLI_REFINE_GOAL=1 adds the dual solution and its acceptance tests; by default the primal
/// corrections alone decide (the dual solve costs about what it saves: BHZ 200x64, 264 points,
/// 139 against 136 ms a step; graphene 512x64 with about 20 channels, 666 against 527 ms).
bool refine_goal() {
    const char* v = std::getenv("CAROLI_REFINE_GOAL");
    return v && std::strcmp(v, "1") == 0;
}

/// The refinement of one launch unit (block b of P points) from the factors its sweep left in
/// its lane (complex f32 fronts). Its work is issued on the lane's stream, so that the refinements
/// of the units of a group (one lane each) run at once: start(), then issue() / collect() per
/// iteration (with the lane's stream synchronised in between) until issue() returns false, then
/// finish().
class Refinement {
  public:
    /// recompute: shared by the batch, set once some point goes to the complex f64 recomputation.
    Refinement(Engine& eng, Lane& lane, const HostPlan& hp, const DevPlan& dp, const PointsView& pv,
               std::vector<const cg_sigma*> sig, int b, int64_t n_orb, const std::vector<char>& skip, bool& recompute)
        : eng_(eng), lane_(lane), hp_(hp), dp_(dp), pv_(pv), sig_(std::move(sig)), b_(b), n_orb_(n_orb),
          recompute_(recompute) {
        P_ = int32_t(sig_.size());
        out_.t.assign(P_, 0.0);
        out_.ok.assign(P_, 1);
        for (auto* g : sig_) {
            c_ = std::max(c_, g->blocks[b].wr.cols);
            cl_ = std::max(cl_, g->blocks[b].wl.cols);
        }
        if (cl_ == 0) c_ = 0;  // no channel on one side: T = 0
        state_.assign(P_, 0);
        for (int32_t p = 0; p < P_; ++p)
            if (skip[p]) {
                state_[p] = 2;
                recompute_ = true;
            }
        ratio_.assign(P_, INFINITY);
        last_.assign(P_, INFINITY);
        if (c_ > 0 && (!solve_fits(hp.max_n) || !solve_fits(hp.nodes[hp.root].ne))) {
            out_.ok.assign(P_, 0);  // left to the complex f64 recomputation
            c_ = 0;
            recompute_ = true;
        }
        goal_ = c_ > 0 && refine_goal();
        nl_ = int(hp.left_orbs.size());
        nr_ = int(hp.right_orbs.size());
        nrows_ = int(hp.rows.size());
        gstride_ = 2 * int64_t(cl_) * c_ + c_ + cl_;
        wr_ = pv.w ? pv.w + int64_t(nl_) * nl_ : nullptr;
    }

    void start() {
        if (c_ == 0) return;
        eng_.lane = &lane_;
        RefineBufs& r = lane_.rf;
        hipStream_t st = lane_.stream;
        bytes_ = size_t(P_) * size_t(n_orb_) * c_ * sizeof(double2);
        const size_t dual_bytes = size_t(P_) * size_t(n_orb_) * cl_ * sizeof(double2);
        r.x.ensure(bytes_);
        r.b.ensure(bytes_);
        r.d.ensure(bytes_);
        r.r.ensure(bytes_);
        if (goal_) {
            r.phi.ensure(dual_bytes);
            r.s.ensure(dual_bytes);
        }
        r.maxes.ensure(size_t(P_) * 2 * sizeof(double));
        r.goal.ensure(size_t(P_) * gstride_ * sizeof(double));
        r.part.ensure(goal_part_words(nrows_, P_, cl_ * c_, std::max(c_, cl_)) * sizeof(double2));
        r.act.ensure(size_t(P_) * sizeof(int32_t));
        r.upd.ensure(size_t(P_) * size_t(hp_.upd_total) * c_ * sizeof(double2));
        r.h_act.ensure(size_t(P_));
        r.h_maxes.ensure(size_t(P_) * 2);
        r.h_goal.ensure(size_t(P_) * gstride_);
        HIP_CHECK(hipMemsetAsync(r.b.p, 0, bytes_, st));
        HIP_CHECK(hipMemsetAsync(r.goal.p, 0, size_t(P_) * gstride_ * sizeof(double), st));
        xv_ = SolveView{r.x.as<double2>(), n_orb_ * c_, int32_t(n_orb_), c_, dp_.vars.as<int32_t>()};
        xv_.upd = r.upd.as<double2>();
        xv_.upd_stride = hp_.upd_total * c_;
        SolveView bv = xv_;
        bv.x = r.b.as<double2>();
        // W_L (nl x cl) and W_R (nr x cr) of every point are resident with the inputs (zero past a
        // point's own columns).
        launch_scatter(bv, dp_.right_orbs.as<int32_t>(), nr_, wr_, pv_.sigma_stride, false, P_, st);
        HIP_CHECK(hipMemcpyAsync(r.x.p, r.b.p, bytes_, hipMemcpyDeviceToDevice, st));
        solve(eng_, hp_, dp_, pv_, xv_, P_, true, false, true);  // x0 = M^-1 b (root and backward only)
        res_ = r.d.as<double2>();
        spare_ = r.r.as<double2>();
        if (goal_) {
            // phi0 = M^-T g, g = E_L conj(W_L) (root and backward only); then s0 = g - A^T phi0
            // in FP64.
            HIP_CHECK(hipMemsetAsync(r.phi.p, 0, dual_bytes, st));
            SolveView fv{r.phi.as<double2>(), n_orb_ * cl_, int32_t(n_orb_), cl_, dp_.vars.as<int32_t>()};
            launch_scatter(fv, dp_.left_orbs.as<int32_t>(), nl_, pv_.w, pv_.sigma_stride, true, P_, st);
            HIP_CHECK(hipMemcpyAsync(r.s.p, r.phi.p, dual_bytes, hipMemcpyDeviceToDevice, st));
            solve(eng_, hp_, dp_, pv_, fv, P_, true, true, true);
            launch_residual(pv_, fv, r.s.as<double2>(), r.s.as<double2>(), dp_.rows.as<int32_t>(), nrows_,
                            dp_.row_ptr.as<int64_t>(), dp_.cols.as<uint32_t>(), dp_.left_pos.as<int32_t>(),
                            dp_.right_pos.as<int32_t>(), dp_.left_orbs.as<int32_t>(), nl_,
                            dp_.right_orbs.as<int32_t>(), nr_, P_, st, true);
            // |s0_d|^2 per column d (kept in the estimate buffer for every later iteration).
            SolveView sv = fv;
            launch_goal(sv, nullptr, nullptr, 0, nullptr, 0, nl_, 0, dp_.left_orbs.as<int32_t>(), r.s.as<double2>(),
                        int64_t(n_orb_) * cl_, cl_, int32_t(2 * int64_t(cl_) * c_ + c_), dp_.rows.as<int32_t>(),
                        nrows_, r.part.as<double2>(), r.goal.as<double>(), gstride_, P_, st);
        }
        residual(P_, xv_);  // r0 = b - A x0, the first iteration's right-hand side
        if (goal_) {  // t'_0, the first term of the sequence of dual-corrected amplitudes
            launch_goal(xv_, res_, r.phi.as<double2>(), int64_t(n_orb_) * cl_, pv_.w, pv_.sigma_stride, nl_, cl_,
                        dp_.left_orbs.as<int32_t>(), nullptr, 0, 0, 0, dp_.rows.as<int32_t>(), nrows_,
                        r.part.as<double2>(), r.goal.as<double>(), gstride_, P_, st);
            r.h_goal0.ensure(size_t(P_) * gstride_);
            HIP_CHECK(hipMemcpyAsync(r.h_goal0.p, r.goal.p, size_t(P_) * gstride_ * sizeof(double),
                                     hipMemcpyDeviceToHost, st));
        }
        HIP_CHECK(hipGetLastError());
    }

    /// Issues one iteration on the points still iterating; false when none is.
    bool issue() {
        if (c_ == 0 || it_ >= kRefineMaxIter) return false;
        RefineBufs& r = lane_.rf;
        A_ = 0;
        for (int32_t p = 0; p < P_; ++p)
            if (state_[p] == 0) r.h_act.p[A_++] = p;
        if (A_ == 0) return false;
        eng_.lane = &lane_;
        hipStream_t st = lane_.stream;
        HIP_CHECK(hipMemcpyAsync(r.act.p, r.h_act.p, size_t(A_) * sizeof(int32_t), hipMemcpyHostToDevice, st));
        SolveView xa = xv_, da = xv_;
        xa.pts = r.act.as<int32_t>();
        da.pts = r.act.as<int32_t>();
        da.x = res_;
        solve(eng_, hp_, dp_, pv_, da, A_, true, false, false);  // d_k = M^-1 r_k, in place
        eng_.timed(kRefineVec, 0.0, 0.0, [&] {
            launch_update(xa, res_, dp_.rows.as<int32_t>(), nrows_, r.maxes.as<double>(), A_, P_, st);
            std::swap(res_, spare_);
            residual(A_, xa);  // r_{k+1} = b - A x_{k+1}
            // t'_{k+1} and |d_k|^2 per column.
            launch_goal(xa, res_, goal_ ? r.phi.as<double2>() : nullptr, int64_t(n_orb_) * cl_, pv_.w,
                        pv_.sigma_stride, nl_, cl_, dp_.left_orbs.as<int32_t>(), spare_, int64_t(n_orb_) * c_, c_,
                        int32_t(2 * int64_t(cl_) * c_), dp_.rows.as<int32_t>(), nrows_, r.part.as<double2>(),
                        r.goal.as<double>(), gstride_, A_, st);
        });
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpyAsync(r.h_maxes.p, r.maxes.p, size_t(P_) * 2 * sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipMemcpyAsync(r.h_goal.p, r.goal.p, size_t(P_) * gstride_ * sizeof(double), hipMemcpyDeviceToHost, st));
        return true;
    }

    /// After the lane's stream is synchronised: the convergence test of the issued iteration.
    void collect() {
        const double* maxes = lane_.rf.h_maxes.p;
        const double* gl = lane_.rf.h_goal.p;
        for (int32_t p = 0; p < P_; ++p) {
            if (state_[p] != 0) continue;
            const double md = maxes[2 * p], mx = maxes[2 * p + 1];
            ratio_[p] = mx > 0.0 ? md / mx : (md > 0.0 ? INFINITY : 0.0);
        }
        ++out_.iters;
        if (refine_trace()) {
            int n_active = 0, n_done = 0, n_gone = 0;
            double worst = 0.0, best = INFINITY;
            for (int32_t p = 0; p < P_; ++p) {
                n_active += state_[p] == 0;
                n_done += state_[p] == 1;
                n_gone += state_[p] == 2;
                if (state_[p] == 0) {
                    worst = std::max(worst, ratio_[p]);
                    best = std::min(best, ratio_[p]);
                }
            }
            std::fprintf(stderr, "[refine] block %d iteration %d: %d iterating (correction %.1e .. %.1e), %d done, %d given up\n",
                         b_, it_, n_active, best, worst, n_done, n_gone);
        }
        const int64_t npair = int64_t(cl_) * c_;
        if (goal_ && it_ == 0) {  // t'_0 from start()
            prev_t_.assign(lane_.rf.h_goal0.p, lane_.rf.h_goal0.p + size_t(P_) * gstride_);
            prev_dt_.assign(P_, INFINITY);
        }
        std::vector<double> trace_bounds, trace_est;
        for (int32_t p = 0; p < P_; ++p) {
            if (state_[p] != 0) continue;
            const double* g = gl + size_t(p) * gstride_;
            double tsum = 0.0;
            for (int64_t e = 0; e < npair; ++e) tsum += g[2 * e] * g[2 * e] + g[2 * e + 1] * g[2 * e + 1];
            const double rr = ratio_[p];
            bool goal_ok = false;
            if (goal_ && rr < INFINITY && std::isfinite(tsum)) {
                double* pt = prev_t_.data() + size_t(p) * gstride_;
                double dtn = 0.0;  // |t'_{k+1} - t'_k|_F = |s0^T d_k|_F
                for (int64_t e = 0; e < 2 * npair; ++e) dtn += (g[e] - pt[e]) * (g[e] - pt[e]);
                dtn = std::sqrt(dtn);
                // Contraction per iteration: of the corrections, and of the amplitudes' differences.
                const double rho_x = last_[p] < INFINITY ? (last_[p] > 0.0 ? rr / last_[p] : 0.0) : rr;
                const double rho_t = prev_dt_[p] < INFINITY && prev_dt_[p] > 0.0 ? dtn / prev_dt_[p] : 0.0;
                const double rho = kGoalSafety * std::max(rho_x, rho_t);
                if (rho < 0.5) {
                    // |e_{k+1}| <= rho |e_k| <= rho (|d_k| + |e_{k+1}|), so |e_{k+1}| <= rho / (1 - rho) |d_k|;
                    // likewise for the error s0^T e_{k+1} of t'_{k+1} against t'_{k+1} - t'_k = s0^T d_k.
                    const double grow = rho / (1.0 - rho);
                    const double* dn = g + 2 * npair;
                    const double* sn = dn + c_;
                    double bound = 0.0, est = 0.0;
                    for (int c = 0; c < c_; ++c)
                        for (int d = 0; d < cl_; ++d) {
                            const int64_t e = d + int64_t(c) * cl_;
                            const double tv = std::hypot(g[2 * e], g[2 * e + 1]);
                            const double del = std::sqrt(sn[d]) * grow * std::sqrt(dn[c]);
                            bound += 2.0 * tv * del + del * del;
                            const double dif = grow * std::hypot(g[2 * e] - pt[2 * e], g[2 * e + 1] - pt[2 * e + 1]);
                            est += 2.0 * tv * dif + dif * dif;
                        }
                    goal_ok = bound <= kGoalTol || est <= kGoalEst;
                    if (refine_trace()) {
                        trace_bounds.push_back(bound);
                        trace_est.push_back(est);
                    }
                }
                std::copy(g, g + 2 * npair, pt);
                prev_dt_[p] = dtn;
            }
            if (goal_ok || rr <= kRefineTol ||
                (last_[p] < INFINITY && rr <= kRefineEstFrom && rr * rr <= kRefineEst * last_[p])) {
                state_[p] = 1;
                out_.t[p] = tsum;
                out_.goal_done += goal_ok && rr > kRefineTol;
                out_.worst = std::max(out_.worst, rr);
            } else if (!(rr < INFINITY) || (it_ >= 2 && rr > kRefineRate * last_[p]) ||
                       (recompute_ && it_ >= kRefineJoinFrom && last_[p] < INFINITY &&
                        rr * (rr / last_[p]) > kRefineTol)) {
                state_[p] = 2;
                recompute_ = true;
            } else {
                last_[p] = rr;
            }
        }
        if (refine_trace() && !trace_bounds.empty()) {
            std::sort(trace_bounds.begin(), trace_bounds.end());
            std::sort(trace_est.begin(), trace_est.end());
            const size_t m = trace_bounds.size();
            std::fprintf(stderr,
                         "[refine] block %d iteration %d, %zu points: |T' - T| bound %.1e .. %.1e (median %.1e), "
                         "estimate %.1e .. %.1e (median %.1e)\n",
                         b_, it_, m, trace_bounds[0], trace_bounds[m - 1], trace_bounds[m / 2], trace_est[0],
                         trace_est[m - 1], trace_est[m / 2]);
        }
        ++it_;
    }

    /// T per point (points not converged: ok = 0).
    Refined finish() {
        if (c_ == 0) return std::move(out_);
        for (int32_t p = 0; p < P_; ++p) out_.ok[p] = state_[p] == 1;
        return std::move(out_);
    }

  private:
    /// r = b - A x on the `points` of xv into res_.
    void residual(int32_t points, const SolveView& xv) {
        RefineBufs& r = lane_.rf;
        launch_residual(pv_, xv, r.b.as<double2>(), res_, dp_.rows.as<int32_t>(), nrows_, dp_.row_ptr.as<int64_t>(),
                        dp_.cols.as<uint32_t>(), dp_.left_pos.as<int32_t>(), dp_.right_pos.as<int32_t>(),
                        dp_.left_orbs.as<int32_t>(), nl_, dp_.right_orbs.as<int32_t>(), nr_, points, lane_.stream);
    }

    Engine& eng_;
    Lane& lane_;
    const HostPlan& hp_;
    const DevPlan& dp_;
    PointsView pv_;
    std::vector<const cg_sigma*> sig_;
    int b_;
    int64_t n_orb_;
    int32_t P_ = 0, A_ = 0;
    int c_ = 0, cl_ = 0, nl_ = 0, nr_ = 0, nrows_ = 0, it_ = 0;
    bool goal_ = false;
    bool& recompute_;
    size_t bytes_ = 0;
    int64_t gstride_ = 0;  // doubles per point of the estimate buffer: t' (2 cl c), |d_c|^2 (c), |s0_d|^2 (cl)
    SolveView xv_{};
    const double2* wr_ = nullptr;  // W_R of point 0 in the resident inputs (stride pv_.sigma_stride)
    double2* res_ = nullptr;    // the residual of the current iterate (the next solve's right-hand side)
    double2* spare_ = nullptr;  // the other residual buffer (holds the last correction)
    Refined out_;
    // Per point: 0 iterating, 1 converged (later corrections only add rounding noise), 2 given up
    // (unusable factors, a correction that is not finite or shrinks too slowly): left to the
    // complex f64 recomputation.
    std::vector<char> state_;
    std::vector<double> ratio_, last_;
    // The last dual-corrected amplitudes per point (gstride_ doubles each, t' first) and the
    // norm of their last change.
    std::vector<double> prev_t_, prev_dt_;
};

}  // namespace

bool device_available() { return Engine::get().ok; }

void release(const cg_system* s) {
    if (s->d_vals) {
        (void)hipFree(s->d_vals);
        s->d_vals = nullptr;
    }
}

}  // namespace caroli::gpu

/// A resident batch: launch units (structure, block, chunk of points) with their inputs in HBM.
/// A sweep captured as a HIP graph, with the device buffers it was captured on (the lane's
/// arenas, pivots, statuses, results and trace scratch), its point count and the small-front
/// kernel in force: a replay is valid while they are the same.
struct SweepGraph {
    hipGraphExec_t exec = nullptr;
    std::array<const void*, 10> key{};
    int eager_runs = 0;
    SweepGraph() = default;
    SweepGraph(SweepGraph&& o) noexcept : exec(o.exec), key(o.key), eager_runs(o.eager_runs) { o.exec = nullptr; }
    SweepGraph& operator=(SweepGraph&& o) noexcept {
        reset();
        exec = o.exec;
        key = o.key;
        eager_runs = o.eager_runs;
        o.exec = nullptr;
        return *this;
    }
    SweepGraph(const SweepGraph&) = delete;
    ~SweepGraph() { reset(); }
    void reset() {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
    }
};

struct cg_batch {
    struct Unit {
        std::shared_ptr<caroli::gpu::Structure> structure;
        int block;
        std::vector<int64_t> points;
        caroli::gpu::Inputs in;
        SweepGraph graph;
    };
    std::vector<const cg_system*> systems;
    std::vector<const cg_sigma*> sigmas;
    int32_t precision = CG_DOUBLE;
    std::vector<Unit> units;
    double flops = 0, bytes = 0;
    int lanes = 1;       // units in flight at once (caroli::gpu::kLanes when their buffers fit)
    bool eager = false;  // sweeps without HIP graphs (the autotuner's timed runs)
    int32_t nlab = 0;    // multi-terminal points: T is nlab x nlab per point
};

namespace caroli::gpu {

/// Device bytes a launch unit of P points of plan hp needs in a lane (front arena in the front
/// precision, root arena, pivots, statuses, results).
struct LaneNeed {
    size_t arena, root, piv, status, tout;
};
LaneNeed lane_need(const HostPlan& hp, size_t P, bool single) {
    const size_t es = single ? sizeof(float2) : sizeof(double2);
    return {P * size_t(hp.arena_stride) * es, P * size_t(hp.root_stride) * sizeof(double2), P * size_t(hp.piv_stride) * 4,
            P * size_t(hp.slots) * 4, P * sizeof(double)};
}
void ensure_lane(Lane& ln, const LaneNeed& nd) {
    ln.arena.ensure(nd.arena);
    ln.root_arena.ensure(nd.root);
    ln.piv.ensure(nd.piv);
    ln.status.ensure(nd.status);
    ln.tout.ensure(nd.tout);
}

int32_t create_batch(const cg_system* const* systems, const cg_sigma* const* sigmas, int64_t count, int32_t precision,
                     cg_batch** out) {
    Engine& eng = Engine::get();
    if (!eng.ok) return CG_ERR_DEVICE;
    std::lock_guard<std::mutex> lock(eng.mu);
    auto bt = std::make_unique<cg_batch>();
    bt->systems.assign(systems, systems + count);
    bt->sigmas.assign(sigmas, sigmas + count);
    bt->precision = precision;
    bt->nlab = count > 0 && sigmas[0]->multi ? int32_t(sigmas[0]->labels.size()) : 0;
    const bool single = precision != CG_DOUBLE;
    const size_t es = single ? sizeof(float2) : sizeof(double2);
    std::map<uint64_t, std::vector<int64_t>> groups;
    for (int64_t p = 0; p < count; ++p) groups[systems[p]->key].push_back(p);
    struct Need {
        size_t arena, root, piv, status, tout, refine;
    };
    std::vector<Need> need;  // device bytes of each unit
    // The points of a (structure, block) are split so that at least kLanes units can be in flight
    // (the top tree levels hold few fronts per point and leave CUs idle with one or two units),
    // down to kMinUnitPoints points a unit; the refinements of the mixed mode run on the units'
    // lanes as well.
    size_t blocks_total = 0;
    for (auto& [key, idx] : groups) blocks_total += structure_of(eng, systems[idx[0]])->host.size();
    const size_t splits = (kLanes + blocks_total - 1) / std::max<size_t>(blocks_total, 1);
    for (auto& [key, idx] : groups) {
        auto sp = structure_of(eng, systems[idx[0]]);
        for (auto p : idx) structure_of(eng, systems[p]);
        Work w = algorithmic_work(systems[idx[0]]->sym, systems[idx[0]]->spec.h.vals.size());
        bt->flops += w.flops * double(idx.size());
        bt->bytes += w.bytes * double(idx.size()) * (single ? 0.5 : 1.0);
        for (size_t b = 0; b < sp->host.size(); ++b) {
            const HostPlan& hp = sp->host[b];
            const Block& blk = systems[idx[0]]->sym.blocks[b];
            // Points per launch: bounded by the scratch budget and the grid's y dimension. Mixed
            // precision also holds, per point and channel, the refinement's solution, right-hand
            // side, correction and update vectors (complex f64) beside the factors.
            size_t per_point = size_t(hp.arena_stride) * es + size_t(hp.root_stride) * sizeof(double2) +
                               size_t(hp.piv_stride) * 4 + size_t(hp.slots) * 4 + 64;
            size_t refine_point = 0;
            if (precision == CG_MIXED) {
                int32_t c = 0, cl = 0;
                for (auto p : idx) {
                    c = std::max(c, sigmas[p]->blocks[b].wr.cols);
                    cl = std::max(cl, sigmas[p]->blocks[b].wl.cols);
                }
                const size_t n_orb = systems[idx[0]]->spec.h.n;
                // x, b and two residual / correction buffers, the update vectors; the dual solution
                // and its residual; the estimates.
                refine_point = (4 * n_orb + size_t(hp.upd_total)) * size_t(c) * sizeof(double2) +
                               2 * n_orb * size_t(cl) * sizeof(double2) + (2 * size_t(cl) * c + c + cl) * sizeof(double) +
                               goal_part_words(int32_t(n_orb), 1, cl * c, std::max(c, cl)) * sizeof(double2);
                per_point += refine_point;
            }
            const size_t budget = size_t(96) << 30;
            // As many parts as keep kMinUnitPoints each (33 points stay one unit, not 32 + 1).
            const size_t parts = std::max<size_t>(1, std::min(splits, idx.size() / kMinUnitPoints));
            const size_t even = (idx.size() + parts - 1) / parts;
            size_t chunk = std::max<size_t>(1, std::min<size_t>({budget / per_point, 65535, idx.size(), even}));
            for (size_t c0 = 0; c0 < idx.size(); c0 += chunk) {
                size_t P = std::min(chunk, idx.size() - c0);
                cg_batch::Unit u;
                u.structure = sp;
                u.block = int(b);
                std::vector<const cg_system*> ss;
                std::vector<const cg_sigma*> gg;
                for (size_t q = 0; q < P; ++q) {
                    u.points.push_back(idx[c0 + q]);
                    ss.push_back(systems[idx[c0 + q]]);
                    gg.push_back(sigmas[idx[c0 + q]]);
                }
                PointsView pv{};
                stage_points(eng, ss, gg, int(b), int(blk.left_orbitals.size()), int(blk.right_orbitals.size()), u.in, pv);
                const LaneNeed ln = lane_need(hp, P, single);
                need.push_back({ln.arena, ln.root, ln.piv, ln.status, ln.tout, refine_point * P});
                bt->units.push_back(std::move(u));
            }
        }
    }
    // Lanes: as many as there are units, up to kLanes, while the largest units fit the concurrent
    // budget together.
    std::vector<size_t> sizes;
    for (auto& nd : need) sizes.push_back(nd.arena + nd.root + nd.piv + nd.status + nd.tout + nd.refine);
    std::sort(sizes.rbegin(), sizes.rend());
    const size_t concurrent_budget = size_t(160) << 30;
    int lanes = int(std::min<size_t>(kLanes, sizes.size()));
    auto fits = [&](int L) {
        size_t t = 0;
        for (int q = 0; q < L; ++q) t += sizes[q];
        return t <= concurrent_budget;
    };
    while (lanes > 1 && !fits(lanes)) --lanes;
    bt->lanes = std::max(lanes, 1);
    for (size_t ui = 0; ui < need.size(); ++ui)
        ensure_lane(eng.lanes[ui % size_t(bt->lanes)],
                    {need[ui].arena, need[ui].root, need[ui].piv, need[ui].status, need[ui].tout});
    *out = bt.release();
    return CG_OK;
}

/// CAROLI_GRAPHS=0 runs every sweep eagerly (no HIP graphs).
bool graphs_enabled() {
    const char* v = std::getenv("CAROLI_GRAPHS");
    return !(v && std::strcmp(v, "0") == 0);
}

/// The sweep of a unit on the current lane: eager on the unit's first run (it sizes the lane's
/// scratch and sets kernel attributes), captured into a HIP graph on the next and replayed after
/// that, as long as the lane's buffers and the point count are the ones it was captured with.
/// The sweep issues kernel launches only (no memsets or copies), so the graph holds kernel nodes.
void sweep_graph(Engine& eng, cg_batch::Unit& u, const HostPlan& hp, const DevPlan& dp, const PointsView& pv,
                 int32_t P, bool single) {
    Lane& ln = *eng.lane;
    // The small-front kernel is chosen at launch (the tuning's small_fronts): a graph captured
    // under another choice is captured anew.
    const std::array<const void*, 10> key{ln.arena.p,   ln.root_arena.p, ln.piv.p,
                                          ln.status.p,  ln.tout.p,       ln.trace_x.p,
                                          ln.trace_p.p, ln.tlab.p,       reinterpret_cast<const void*>(intptr_t(P)),
                                          reinterpret_cast<const void*>(intptr_t(small_front_mode()))};
    SweepGraph& g = u.graph;
    if (g.exec && g.key == key) {
        HIP_CHECK(hipGraphLaunch(g.exec, ln.stream));
        return;
    }
    g.reset();
    if (g.eager_runs++ == 0) {
        sweep(eng, hp, dp, pv, P, single, u.in);
        return;
    }
    hipGraph_t graph = nullptr;
    HIP_CHECK(hipStreamBeginCapture(ln.stream, hipStreamCaptureModeThreadLocal));
    sweep(eng, hp, dp, pv, P, single, u.in);
    HIP_CHECK(hipStreamEndCapture(ln.stream, &graph));
    const hipError_t rc = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIP_CHECK(rc);
    g.key = key;
    HIP_CHECK(hipGraphLaunch(g.exec, ln.stream));
}

/// The batch's points; out_lab: for multi-terminal points, T between labels (nlab^2 per point,
/// summed over blocks), with out_t the sum of each point's entries.
int32_t run_batch_impl(cg_batch* bt, double* out_t, double* out_lab, int32_t* out_status, cg_batch_stats* stats) {
    Engine& eng = Engine::get();
    if (!eng.ok) return CG_ERR_DEVICE;
    const bool mixed = bt->precision == CG_MIXED;
    const bool single = bt->precision == CG_SINGLE || mixed;
    const int64_t count = int64_t(bt->systems.size());
    cg_batch_stats local{};
    local.front_flops = bt->flops;
    local.front_bytes = bt->bytes;
    std::vector<int32_t> status(count, CG_OK);
    std::vector<char> redo(count, 0);  // mixed: recompute in complex f64
    bool recompute = false;            // mixed: some point goes to the recomputation
    for (int64_t p = 0; p < count; ++p) out_t[p] = 0.0;
    const int32_t nlab = bt->nlab;
    const size_t lab2 = size_t(nlab) * size_t(nlab);
    if (nlab > 0) {
        if (!out_lab || mixed) return CG_ERR_INVALID;
        std::fill(out_lab, out_lab + size_t(count) * lab2, 0.0);
    }
    {
        std::lock_guard<std::mutex> lock(eng.mu);
        eng.tol_single = mixed ? kTolMixed : kTolSingle;
        // A unit planned under a tuning that classifies fronts differently from the one in force
        // (cg_set_tuning after cg_batch_create) is re-planned here, and its captured graph dropped.
        for (auto& u : bt->units) {
            auto sp = structure_of(eng, bt->systems[u.points[0]]);
            if (sp != u.structure) {
                u.structure = sp;
                u.graph = SweepGraph{};
            }
        }
        // Units go out in groups of bt->lanes, one lane (stream and buffers) each, so that their
        // sweeps overlap; results are read back once the group is done.
        // Per-kernel timing runs the units one at a time, so that each timed launch has the GPU to
        // itself and its duration is its own.
        const size_t L = eng.timing ? 1 : size_t(bt->lanes);
        for (size_t u0 = 0; u0 < bt->units.size(); u0 += L) {
            const size_t u1 = std::min(u0 + L, bt->units.size());
            std::vector<PointsView> pvs(u1 - u0);
            std::vector<std::vector<double>> ts(u1 - u0);
            std::vector<std::vector<int32_t>> sts(u1 - u0);
            std::vector<std::vector<double>> labs(u1 - u0);
            auto w0 = std::chrono::steady_clock::now();
            for (size_t ui = u0; ui < u1; ++ui) {
                auto& u = bt->units[ui];
                const HostPlan& hp = u.structure->host[u.block];
                const DevPlan& dp = u.structure->dev[u.block];
                eng.lane = &eng.lanes[ui - u0];
                // The lane's buffers cover the units create_batch gave it; run one at a time (kernel
                // timing, mixed precision), every unit goes to lane 0, which grows to it here (the
                // previous group has finished: its streams were synchronised).
                ensure_lane(*eng.lane, lane_need(hp, u.points.size(), single));
                PointsView& pv = pvs[ui - u0];
                bind_inputs(u.in, pv);
                pv.arena = eng.lane->arena.p;
                pv.arena_stride = hp.arena_stride;
                pv.root_arena = eng.lane->root_arena.p;
                pv.root_stride = hp.root_stride;
                pv.piv = eng.lane->piv.as<int32_t>();
                pv.piv_stride = hp.piv_stride;
                pv.status = eng.lane->status.as<int32_t>();
                pv.status_stride = hp.slots;
                pv.t_out = eng.lane->tout.as<double>();
                if (nlab > 0) {
                    eng.lane->tlab.ensure(u.points.size() * lab2 * sizeof(double));
                    pv.t_lab = eng.lane->tlab.as<double>();
                }
                pv.keep = mixed ? 1 : 0;  // the refinement solves with the factors
                if (!eng.timing && !bt->eager && graphs_enabled())
                    sweep_graph(eng, u, hp, dp, pv, int32_t(u.points.size()), single);
                else
                    sweep(eng, hp, dp, pv, int32_t(u.points.size()), single, u.in);
            }
            for (size_t ui = u0; ui < u1; ++ui) {
                auto& u = bt->units[ui];
                const HostPlan& hp = u.structure->host[u.block];
                Lane& ln = eng.lanes[ui - u0];
                const size_t P = u.points.size();
                ts[ui - u0].resize(P);
                sts[ui - u0].resize(P * size_t(hp.slots));
                HIP_CHECK(hipMemcpyAsync(ts[ui - u0].data(), ln.tout.p, P * sizeof(double), hipMemcpyDeviceToHost,
                                         ln.stream));
                HIP_CHECK(hipMemcpyAsync(sts[ui - u0].data(), ln.status.p, sts[ui - u0].size() * 4,
                                         hipMemcpyDeviceToHost, ln.stream));
                if (nlab > 0) {
                    labs[ui - u0].resize(P * lab2);
                    HIP_CHECK(hipMemcpyAsync(labs[ui - u0].data(), ln.tlab.p, P * lab2 * sizeof(double),
                                             hipMemcpyDeviceToHost, ln.stream));
                }
            }
            for (size_t ui = u0; ui < u1; ++ui) HIP_CHECK(hipStreamSynchronize(eng.lanes[ui - u0].stream));
            eng.lane = &eng.lanes[0];
            eng.collect();
            local.seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - w0).count();
            if (mixed) {
                // The refinements of the group's units at once, one lane each: every iteration is
                // issued on all lanes, then the lanes are synchronised and each unit tests its
                // points; a unit whose points are all done or given up stops issuing.
                auto t0 = std::chrono::steady_clock::now();
                std::vector<std::unique_ptr<Refinement>> refs;
                for (int64_t p = 0; p < count; ++p) recompute |= redo[p] != 0;
                for (size_t ui = u0; ui < u1; ++ui) {
                    auto& u = bt->units[ui];
                    const HostPlan& hp = u.structure->host[u.block];
                    const size_t P = u.points.size();
                    const std::vector<int32_t>& st = sts[ui - u0];
                    std::vector<const cg_sigma*> sig;
                    for (auto p : u.points) sig.push_back(bt->sigmas[p]);
                    std::vector<char> skip(P, 0);
                    for (size_t q = 0; q < P; ++q)
                        for (int32_t k = 0; k <= hp.root; ++k) skip[q] |= st[q * hp.slots + k] != kFrontOk;
                    refs.push_back(std::make_unique<Refinement>(eng, eng.lanes[ui - u0], hp, u.structure->dev[u.block],
                                                                pvs[ui - u0], std::move(sig), u.block,
                                                                int64_t(bt->systems[u.points[0]]->spec.h.n), skip,
                                                                recompute));
                    refs.back()->start();
                }
                auto sync_group = [&] {
                    for (size_t ui = u0; ui < u1; ++ui) HIP_CHECK(hipStreamSynchronize(eng.lanes[ui - u0].stream));
                };
                auto since = [&](std::chrono::steady_clock::time_point a) {
                    return std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count() * 1e3;
                };
                if (refine_trace()) {
                    sync_group();
                    std::fprintf(stderr, "[refine] sweep %.2f ms, start %.2f ms\n",
                                 std::chrono::duration<double>(t0 - w0).count() * 1e3, since(t0));
                }
                for (;;) {
                    const auto ti = std::chrono::steady_clock::now();
                    std::vector<char> issued(refs.size(), 0);
                    bool any = false;
                    for (size_t k = 0; k < refs.size(); ++k) any |= (issued[k] = refs[k]->issue()) != 0;
                    if (!any) break;
                    sync_group();
                    for (size_t k = 0; k < refs.size(); ++k)
                        if (issued[k]) refs[k]->collect();
                    if (refine_trace()) std::fprintf(stderr, "[refine] iteration %.2f ms\n", since(ti));
                }
                eng.lane = &eng.lanes[0];
                eng.collect();
                local.seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                for (size_t ui = u0; ui < u1; ++ui) {
                    auto& u = bt->units[ui];
                    const HostPlan& hp = u.structure->host[u.block];
                    const size_t P = u.points.size();
                    const std::vector<int32_t>& st = sts[ui - u0];
                    Refined rf = refs[ui - u0]->finish();
                    local.refinement_iters += rf.iters;
                    local.goal_points += rf.goal_done;
                    for (size_t q = 0; q < P; ++q) {
                        int64_t p = u.points[q];
                        bool delayed = false;
                        for (int32_t k = 0; k < hp.root; ++k) delayed |= st[q * hp.slots + k] != kFrontOk;
                        if (delayed || st[q * hp.slots + hp.root] != kFrontOk || !rf.ok[q] || !std::isfinite(rf.t[q])) {
                            redo[p] = 1;
                            local.delayed_points += delayed;
                        } else {
                            local.max_refine_resid = std::max(local.max_refine_resid, rf.worst);
                            out_t[p] += rf.t[q];
                        }
                    }
                }
                continue;
            }
            for (size_t ui = u0; ui < u1; ++ui) {
                auto& u = bt->units[ui];
                const HostPlan& hp = u.structure->host[u.block];
                const size_t P = u.points.size();
                std::vector<double>& t = ts[ui - u0];
                const std::vector<int32_t>& st = sts[ui - u0];
                if (nlab > 0)  // the point's T is the sum of its label matrix (NaN when any entry is)
                    for (size_t q = 0; q < P; ++q) {
                        t[q] = 0.0;
                        for (size_t e = 0; e < lab2; ++e) t[q] += labs[ui - u0][q * lab2 + e];
                    }
                std::vector<double> lab_tmp(lab2);
                for (size_t q = 0; q < P; ++q) {
                    int64_t p = u.points[q];
                    bool delayed = false;
                    int32_t root_status = st[q * hp.slots + hp.root];
                    for (int32_t k = 0; k < hp.root; ++k) delayed |= st[q * hp.slots + k] != kFrontOk;
                    if (delayed) {
                        double tb = NAN;
                        auto t0 = std::chrono::steady_clock::now();
                        int32_t rc = replanned_point(eng, bt->systems[p], bt->sigmas[p], u.block, single, tb,
                                                     nlab > 0 ? lab_tmp.data() : nullptr);
                        local.seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                        ++local.delayed_points;
                        if (rc != CG_OK) status[p] = rc;
                        out_t[p] += tb;
                        for (size_t e = 0; e < lab2; ++e) out_lab[size_t(p) * lab2 + e] += lab_tmp[e];
                    } else if (root_status != kFrontOk || !std::isfinite(t[q])) {
                        status[p] = CG_ERR_SINGULAR;
                    } else {
                        out_t[p] += t[q];
                        for (size_t e = 0; e < lab2; ++e) out_lab[size_t(p) * lab2 + e] += labs[ui - u0][q * lab2 + e];
                    }
                }
            }
        }
    }
    // Mixed precision: points whose c32 fronts delayed or whose refinement did not converge are
    // recomputed in complex f64 on the device.
    std::vector<int64_t> again;
    for (int64_t p = 0; p < count; ++p)
        if (redo[p]) again.push_back(p);
    local.recomputed_points = int64_t(again.size());
    if (!again.empty()) {
        std::vector<const cg_system*> ss;
        std::vector<const cg_sigma*> gg;
        for (auto p : again) {
            ss.push_back(bt->systems[p]);
            gg.push_back(bt->sigmas[p]);
        }
        std::vector<double> t(again.size());
        std::vector<int32_t> st(again.size());
        cg_batch_stats s2{};
        const auto tr = std::chrono::steady_clock::now();
        int32_t rc = batch_transmission(ss.data(), gg.data(), int64_t(again.size()), CG_DOUBLE, t.data(), st.data(), &s2);
        if (refine_trace())
            std::fprintf(stderr, "[refine] complex f64 recomputation of %zu points %.2f ms\n", again.size(),
                         std::chrono::duration<double>(std::chrono::steady_clock::now() - tr).count() * 1e3);
        if (rc != CG_OK) return rc;
        local.seconds += s2.seconds;
        for (size_t q = 0; q < again.size(); ++q) {
            out_t[again[q]] = t[q];
            status[again[q]] = st[q];
        }
    }
    {
        std::lock_guard<std::mutex> lock(eng.mu);
        HIP_CHECK(hipStreamSynchronize(eng.lane->stream));
        eng.collect();  // timers of the refinement and of re-planned points
    }
    for (int64_t p = 0; p < count; ++p) {
        if (status[p] != CG_OK) {
            out_t[p] = NAN;
            if (nlab > 0) std::fill(out_lab + size_t(p) * lab2, out_lab + size_t(p + 1) * lab2, NAN);
        }
        if (out_status) out_status[p] = status[p];
    }
    if (stats) *stats = local;
    return CG_OK;
}

int32_t run_batch(cg_batch* bt, double* out_t, int32_t* out_status, cg_batch_stats* stats) {
    if (bt->nlab > 0) return CG_ERR_INVALID;  // multi-terminal batches give a matrix per point
    return run_batch_impl(bt, out_t, nullptr, out_status, stats);
}

int32_t run_batch_labels(cg_batch* bt, double* out_lab, int32_t* out_status, cg_batch_stats* stats) {
    if (bt->nlab <= 0) return CG_ERR_INVALID;
    std::vector<double> t(bt->systems.size());
    return run_batch_impl(bt, t.data(), out_lab, out_status, stats);
}

int32_t batch_transmission(const cg_system* const* systems, const cg_sigma* const* sigmas, int64_t count,
                           int32_t precision, double* out_t, int32_t* out_status, cg_batch_stats* stats) {
    cg_batch* bt = nullptr;
    int32_t rc = create_batch(systems, sigmas, count, precision, &bt);
    if (rc != CG_OK) return rc;
    std::unique_ptr<cg_batch> hold(bt);
    return run_batch(bt, out_t, out_status, stats);
}

void destroy_batch(cg_batch* b) { delete b; }

namespace {
cg_tuning to_c(const Tuning& t) { return {t.small_fronts, t.split, t.blocked_min_ne, t.unblocked, t.schur_tile}; }
Tuning from_c(const cg_tuning& c) {
    Tuning t;
    t.small_fronts = c.small_fronts;
    t.split = c.split;
    t.blocked_min_ne = c.blocked_min_ne;
    t.unblocked = c.unblocked;
    t.schur_tile = c.schur_tile;
    return t;
}
}  // namespace

int32_t get_tuning(cg_tuning* out) {
    *out = to_c(current_tuning());
    return CG_OK;
}

int32_t set_tuning(const cg_tuning* t) {
    if (!t) {
        set_tuning_override(nullptr);
        return CG_OK;
    }
    if (t->small_fronts < 0 || t->small_fronts > 2 || t->split < 0 || t->split > 1 || t->blocked_min_ne < 0 ||
        t->unblocked < 0 || t->unblocked > 1 || (t->schur_tile != 64 && t->schur_tile != 32 && t->schur_tile != 0))
        return CG_ERR_INVALID;
    const Tuning tu = from_c(*t);
    set_tuning_override(&tu);
    return CG_OK;
}

/// Coordinate descent over the kernel choices on the batch's own points: every candidate puts the
/// batch's units on plans classified under it (their captured graphs dropped), runs one eager
/// sweep to settle and keeps the faster of the next two; a candidate replaces the best so far when
/// 2 % faster and when every point ends with the status it had under the tuning in force at the
/// start (a candidate that fails points does not win on speed). If a device error interrupts the
/// search, the tuning in force before the call is restored before the error is returned.
int32_t autotune_batch(cg_batch* bt, cg_tuning* out, double* out_seconds) {
    Engine& eng = Engine::get();
    if (!eng.ok) return CG_ERR_DEVICE;
    const int64_t count = int64_t(bt->systems.size());
    std::vector<double> t(size_t(std::max<int64_t>(count, 1)));
    std::vector<int32_t> st(t.size()), st0;
    auto apply = [&](const Tuning& tu) {
        std::lock_guard<std::mutex> lock(eng.mu);
        set_tuning_override(&tu);
        for (auto& u : bt->units) {
            u.structure = structure_of(eng, bt->systems[u.points[0]]);
            u.graph = SweepGraph{};
        }
        for (int64_t p = 0; p < count; ++p) structure_of(eng, bt->systems[p]);
    };
    int32_t rc = CG_OK;
    std::vector<double> lab(t.size() * size_t(bt->nlab) * size_t(bt->nlab));
    auto run = [&] {
        return bt->nlab > 0 ? run_batch_labels(bt, lab.data(), st.data(), nullptr) : run_batch(bt, t.data(), st.data(), nullptr);
    };
    auto measure = [&](const Tuning& tu) {
        apply(tu);
        bt->eager = true;
        rc = run();
        double sec = INFINITY;
        for (int rep = 0; rep < 2 && rc == CG_OK; ++rep) {  // the faster of two timed sweeps
            auto t0 = std::chrono::steady_clock::now();
            rc = run();
            sec = std::min(sec, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
        }
        bt->eager = false;
        return sec;
    };
    const Tuning start = current_tuning();
    Tuning best = start;
    double best_s = INFINITY;
    try {
        best_s = measure(best);
        st0 = st;
        auto consider = [&](const Tuning& c) {
            if (rc != CG_OK || (c.same_plan(best) && c.small_fronts == best.small_fronts)) return;
            const double sec = measure(c);
            if (rc == CG_OK && st == st0 && sec < 0.98 * best_s) {
                best = c;
                best_s = sec;
            }
        };
        for (int sf : {0, 1, 2}) {
            Tuning c = best;
            c.small_fronts = sf;
            consider(c);
        }
        if (!best.unblocked) {
            for (int sp : {1, 0}) {
                Tuning c = best;
                c.split = sp;
                consider(c);
            }
            for (int bm : {64, 128, 256}) {
                Tuning c = best;
                c.blocked_min_ne = bm;
                consider(c);
            }
            for (int ts : {0, 64, 32}) {
                Tuning c = best;
                c.schur_tile = ts;
                consider(c);
            }
        }
        apply(best);  // the winner stays in force, the batch on its plans (graphs captured anew)
    } catch (...) {
        bt->eager = false;
        try {
            apply(start);
        } catch (...) {
        }
        throw;
    }
    if (rc != CG_OK) return rc;
    if (out) *out = to_c(best);
    if (out_seconds) *out_seconds = best_s;
    return CG_OK;
}

namespace {

/// One front update of cg_front_schur: the front's descriptor, the kernel path its size takes
/// under the tuning in force, and the host copies of its lists. `lv` holds the front's index (0)
/// first; `tl` the Schur tiles first, then the triangular-solve or blocked tiles.
struct FrontJob {
    enum Path { kSmall, kUnblocked, kBlocked, kSplit, kBorder } path = kSmall;
    NodeDesc d{};
    std::vector<int32_t> lv{0};
    std::vector<Tile> tl;
    int32_t schur_tiles = 0, trsm_tiles = 0, schur_tile = 64;
    LevelSet blocked;
    SetOffsets offsets{};
};

FrontJob plan_front(int64_t n, int64_t ne, int64_t ne1, bool single) {
    FrontJob j;
    j.d.off = 0;
    j.d.n = int32_t(n);
    j.d.ne = int32_t(ne);
    j.d.ne1 = int32_t(ne1);
    j.d.slot = 0;
    j.d.piv_off = 0;
    if (n <= (single ? kLdsFrontC64 : kLdsFrontC128)) {
        j.path = FrontJob::kSmall;
        return j;
    }
    if (unblocked_fronts()) {
        j.path = FrontJob::kUnblocked;
        return j;
    }
    j.schur_tile = current_tuning().schur_tile;
    if (j.schur_tile == 0) j.schur_tile = n - ne <= kSchurTile32MaxNb ? 32 : 64;
    for (int i0 = 0; i0 < n - ne; i0 += j.schur_tile)
        for (int j0 = 0; j0 < n - ne; j0 += j.schur_tile) j.tl.push_back({0, i0, j0, 0});
    j.schur_tiles = int32_t(j.tl.size());
    if (ne > blocked_min_ne()) {
        j.path = FrontJob::kBlocked;
        j.blocked.blk = {0};
        j.blocked.plan_blocked({j.d});
        j.offsets = append_set(j.blocked, j.lv, j.tl);
    } else if (ne <= kF11Lds && !border_only()) {
        j.path = FrontJob::kSplit;
        for (int o = 0; o < n - ne; o += kTrsmSpan) {
            j.tl.push_back({0, 0, o, 0});
            j.tl.push_back({0, o, 0, 1});
        }
        j.trsm_tiles = int32_t(j.tl.size()) - j.schur_tiles;
    } else {
        j.path = FrontJob::kBorder;
    }
    return j;
}

/// The kernels of a planned front update on the front in pv.arena, with the device copies of its
/// descriptor (`nodes`), lists (`lv`) and tiles (`tl`).
void launch_front_job(Engine& eng, const FrontJob& j, const NodeDesc* nodes, const int32_t* lv, const Tile* tl,
                      const PointsView& pv, bool single) {
    const double tol = single ? kTolSingle : kTolDouble;
    const int32_t n = j.d.n, ne = j.d.ne;
    hipStream_t st = eng.lane->stream;
    switch (j.path) {
        case FrontJob::kSmall:
            launch_front_lds(nodes, lv, 1, n, border_lds_words(n, ne), nullptr, nullptr, nullptr, nullptr, pv, 1, tol,
                             single, false, st);
            return;
        case FrontJob::kUnblocked:
            launch_factor_unblocked(nodes, lv, 1, pv, 1, tol, single, false, st);
            return;
        case FrontJob::kBlocked:
            run_blocked(eng, j.blocked, resolve(j.offsets, lv, tl), nodes, nullptr, nullptr, pv, 1, tol, single, false);
            break;
        case FrontJob::kSplit:
            launch_f11(nodes, lv, 1, ne, nullptr, nullptr, pv, 1, tol, single, st);
            launch_trsm(nodes, tl + j.schur_tiles, j.trsm_tiles, ne, pv, 1, single, st);
            break;
        case FrontJob::kBorder:
            launch_border(nodes, lv, 1, ne, nullptr, nullptr, pv, 1, tol, single, false, st);
            break;
    }
    launch_schur(nodes, tl, j.schur_tiles, nullptr, nullptr, pv, 1, single, false, false, j.schur_tile, st);
}

}  // namespace

int32_t front_schur(const double* f, int64_t n, int64_t ne, int64_t ne1, int32_t precision, double* out_s,
                    int32_t* out_status) {
    Engine& eng = Engine::get();
    if (!eng.ok) return CG_ERR_DEVICE;
    std::lock_guard<std::mutex> lock(eng.mu);
    const bool single = precision == CG_SINGLE;
    const FrontJob job = plan_front(n, ne, ne1, single);
    DevBuf fb, nodes, lv, tl, pivb, stb;
    const size_t nn = size_t(n) * size_t(n);
    if (single) {
        std::vector<float2> h(nn);
        for (size_t e = 0; e < nn; ++e) h[e] = float2{float(f[2 * e]), float(f[2 * e + 1])};
        upload(fb, h, eng.lane->stream);
    } else {
        std::vector<double2> h(nn);
        for (size_t e = 0; e < nn; ++e) h[e] = double2{f[2 * e], f[2 * e + 1]};
        upload(fb, h, eng.lane->stream);
    }
    upload(nodes, std::vector<NodeDesc>{job.d}, eng.lane->stream);
    upload(lv, job.lv, eng.lane->stream);
    upload(tl, job.tl, eng.lane->stream);
    pivb.ensure(size_t(ne + kPivExtra) * 4);
    stb.ensure(4);
    PointsView pv{};
    pv.arena = fb.p;
    pv.piv = pivb.as<int32_t>();
    pv.status = stb.as<int32_t>();
    pv.status_stride = 1;
    pv.keep = 1;
    launch_front_job(eng, job, nodes.as<NodeDesc>(), lv.as<int32_t>(), tl.as<Tile>(), pv, single);
    HIP_CHECK(hipGetLastError());
    int32_t st = 0;
    HIP_CHECK(hipMemcpyAsync(&st, stb.p, 4, hipMemcpyDeviceToHost, eng.lane->stream));
    const int64_t nb = n - ne;
    if (single) {
        std::vector<float2> h(nn);
        HIP_CHECK(hipMemcpyAsync(h.data(), fb.p, nn * sizeof(float2), hipMemcpyDeviceToHost, eng.lane->stream));
        HIP_CHECK(hipStreamSynchronize(eng.lane->stream));
        for (int64_t j = 0; j < nb; ++j)
            for (int64_t i = 0; i < nb; ++i) {
                float2 v = h[size_t(ne + i) + size_t(ne + j) * size_t(n)];
                out_s[2 * (i + j * nb)] = v.x;
                out_s[2 * (i + j * nb) + 1] = v.y;
            }
    } else {
        std::vector<double2> h(nn);
        HIP_CHECK(hipMemcpyAsync(h.data(), fb.p, nn * sizeof(double2), hipMemcpyDeviceToHost, eng.lane->stream));
        HIP_CHECK(hipStreamSynchronize(eng.lane->stream));
        for (int64_t j = 0; j < nb; ++j)
            for (int64_t i = 0; i < nb; ++i) {
                double2 v = h[size_t(ne + i) + size_t(ne + j) * size_t(n)];
                out_s[2 * (i + j * nb)] = v.x;
                out_s[2 * (i + j * nb) + 1] = v.y;
            }
    }
    *out_status = st;
    return CG_OK;
}

// ---- the allocation-free front update -----------------------------------------------------------
//
// cg_front_schur_noalloc works in one file-scope device array, `scratch`, that the GPU holds from
// the library's load to the end of the process (one copy per GPU, at its own address):
//   [0, kNoAllocFront)                 the front in its working precision (n <= kNoAllocMaxN)
//   [kNoAllocFront, 2 kNoAllocFront)   complex f64 staging (complex f32 fronts: input and S)
//   [2 kNoAllocFront, + kNoAllocDesc)  descriptor, pivots, status, front list, tiles
// Every call on a GPU uses the whole array: calls that overlap on one GPU race.

constexpr size_t kNoAllocFront = size_t(kNoAllocMaxN) * size_t(kNoAllocMaxN) * sizeof(double2);
constexpr size_t kNoAllocDesc = size_t(1) << 20;
constexpr size_t kNoAllocBytes = 2 * kNoAllocFront + kNoAllocDesc;

static_assert(kNoAllocBytes == CG_FRONT_NOALLOC_SCRATCH_BYTES, "the header states the work area's size");

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

/// Device address of `scratch` on the current GPU, looked up once per GPU (the array's name in
/// host code is a stand-in, not a device address).
unsigned char* scratch_on_device() {
    static std::mutex mu;
    static std::map<int, unsigned char*> at;
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> g(mu);
    auto it = at.find(dev);
    if (it != at.end()) return it->second;
    void* ptr = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw DeviceError("hipGetSymbolAddress returned no address for the scratch array");
    at[dev] = static_cast<unsigned char*>(ptr);
    return at[dev];
}

/// Clears the front's status word and, for a complex f32 front, converts the complex f64 input.
__global__ void noalloc_in_kernel(const double2* __restrict__ src, float2* __restrict__ dst, int64_t count,
                                  int32_t* __restrict__ status) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i == 0) *status = kFrontOk;
    for (int64_t e = i; e < count; e += int64_t(gridDim.x) * blockDim.x) dst[e] = float2{float(src[e].x), float(src[e].y)};
}

/// The trailing nb x nb block of a complex f32 front of order n, as complex f64, column-major.
__global__ void noalloc_out_kernel(const float2* __restrict__ front, int64_t n, int64_t ne, double2* __restrict__ dst) {
    const int64_t nb = n - ne, count = nb * nb;
    for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < count; e += int64_t(gridDim.x) * blockDim.x) {
        const int64_t i = e % nb, j = e / nb;
        const float2 v = front[(ne + i) + (ne + j) * n];
        dst[e] = double2{double(v.x), double(v.y)};
    }
}

int32_t grid_for(int64_t count) { return int32_t(std::max<int64_t>(1, std::min<int64_t>((count + 255) / 256, 2048))); }

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

int32_t front_schur_noalloc(const double* f, int64_t n, int64_t ne, int64_t ne1, int32_t precision, double* out_s,
                            int32_t* out_status) {
    Engine& eng = Engine::get();
    if (!eng.ok) return CG_ERR_DEVICE;
    if (n > kNoAllocMaxN) return CG_ERR_CAPACITY;
    std::lock_guard<std::mutex> lock(eng.mu);
    unsigned char* base = scratch_on_device();
    const bool single = precision == CG_SINGLE;
    const FrontJob job = plan_front(n, ne, ne1, single);
    // Descriptor region: node, pivots, status, front list, tiles.
    const size_t node_at = 0, piv_at = align256(sizeof(NodeDesc)), status_at = piv_at + align256(size_t(ne + kPivExtra) * 4),
                 lv_at = status_at + 256, tl_at = lv_at + align256(job.lv.size() * sizeof(int32_t)),
                 desc_end = tl_at + job.tl.size() * sizeof(Tile);
    if (desc_end > kNoAllocDesc) return CG_ERR_CAPACITY;
    unsigned char* front = base;
    double2* stage = reinterpret_cast<double2*>(base + kNoAllocFront);
    unsigned char* desc = base + 2 * kNoAllocFront;
    NodeDesc* d_node = reinterpret_cast<NodeDesc*>(desc + node_at);
    int32_t* d_piv = reinterpret_cast<int32_t*>(desc + piv_at);
    int32_t* d_status = reinterpret_cast<int32_t*>(desc + status_at);
    int32_t* d_lv = reinterpret_cast<int32_t*>(desc + lv_at);
    Tile* d_tl = reinterpret_cast<Tile*>(desc + tl_at);
    hipStream_t st = eng.lane->stream;
    // Timing events are not created here (the call allocates nothing): the timers pause.
    const bool timing = eng.timing;
    eng.timing = false;
    struct Restore {
        Engine& e;
        bool t;
        ~Restore() { e.timing = t; }
    } restore{eng, timing};
    const int64_t nn = n * n, nb = n - ne;
    if (single) {
        HIP_CHECK(hipMemcpyAsync(stage, f, size_t(nn) * sizeof(double2), hipMemcpyHostToDevice, st));
        noalloc_in_kernel<<<grid_for(nn), 256, 0, st>>>(stage, reinterpret_cast<float2*>(front), nn, d_status);
    } else {
        HIP_CHECK(hipMemcpyAsync(front, f, size_t(nn) * sizeof(double2), hipMemcpyHostToDevice, st));
        noalloc_in_kernel<<<1, 256, 0, st>>>(nullptr, nullptr, 0, d_status);
    }
    HIP_CHECK(hipMemcpyAsync(d_node, &job.d, sizeof(NodeDesc), hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemcpyAsync(d_lv, job.lv.data(), job.lv.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
    if (!job.tl.empty())
        HIP_CHECK(hipMemcpyAsync(d_tl, job.tl.data(), job.tl.size() * sizeof(Tile), hipMemcpyHostToDevice, st));
    PointsView pv{};
    pv.arena = front;
    pv.piv = d_piv;
    pv.status = d_status;
    pv.status_stride = 1;
    pv.keep = 1;
    launch_front_job(eng, job, d_node, d_lv, d_tl, pv, single);
    if (single) {
        noalloc_out_kernel<<<grid_for(nb * nb), 256, 0, st>>>(reinterpret_cast<const float2*>(front), n, ne, stage);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpyAsync(out_s, stage, size_t(nb * nb) * sizeof(double2), hipMemcpyDeviceToHost, st));
    } else {
        HIP_CHECK(hipGetLastError());
        if (nb > 0)
            HIP_CHECK(hipMemcpy2DAsync(out_s, size_t(nb) * sizeof(double2),
                                       front + size_t(ne + ne * n) * sizeof(double2), size_t(n) * sizeof(double2),
                                       size_t(nb) * sizeof(double2), size_t(nb), hipMemcpyDeviceToHost, st));
    }
    int32_t status = 0;
    HIP_CHECK(hipMemcpyAsync(&status, d_status, 4, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    *out_status = status;
    return CG_OK;
}

int32_t kernel_timing(int32_t enable) {
    Engine& eng = Engine::get();
    if (!eng.ok) return CG_ERR_DEVICE;
    std::lock_guard<std::mutex> lock(eng.mu);
    eng.timing = enable != 0;
    if (enable)
        for (auto& f : eng.fams) f = Engine::Fam{};
    return CG_OK;
}

int32_t kernel_stats(cg_kernel_stat* out, int32_t max_records) {
    Engine& eng = Engine::get();
    std::lock_guard<std::mutex> lock(eng.mu);
    int32_t n = 0;
    for (int f = 0; f < kFamilies && n < max_records; ++f) {
        cg_kernel_stat& r = out[n++];
        // The small-front family carries the name of the kernel that ran it.
        const char* small[3] = {"front_border_kernel", "front_lds_kernel", "front_reg_kernel"};
        std::snprintf(r.name, sizeof(r.name), "%s", f == kFrontLds ? small[small_front_mode()] : kFamilyNames[f]);
        r.launches = eng.fams[f].launches;
        r.seconds = eng.fams[f].seconds;
        r.flops = eng.fams[f].flops;
        r.bytes = eng.fams[f].bytes;
    }
    return n;
}

// Adjoint of T = Tr[Gamma_L G Gamma_R G^H], G = (z - H - Sigma)^-1: for an on-site energy eps_a,
// dG/deps_a = G e_a e_a^T G, so dT/deps_a = 2 Re (G Gamma_R G^H Gamma_L G)_aa. With Gamma = W W^H,
// psi = G E_R W_R (the columns on the right slice) and phi = G^T E_L conj(W_L) (rows on the left
// slice, by a solve with A^T), t = W_L^H psi_L:  dT/deps_a = 2 Re sum_cd psi_ac conj(t_dc) phi_ad.
// Both solves reuse the complex f64 factors of one forward sweep: the root solve, then the
// backward substitution through every front (the right-hand sides live on the lead slices).
namespace {

/// Block b of one point (s, g) factored on lane 0 with its factors kept, for the solves of the
/// adjoint and the scattering states. CG_ERR_INVALID when a separator was delayed (the solves
/// follow the static plan), CG_ERR_SINGULAR for a singular root.
int32_t factor_point(Engine& eng, const cg_system* s, const cg_sigma* g, size_t b, const HostPlan& hp,
                     const DevPlan& dp, PointsView& pv) {
    const Block& blk = s->sym.blocks[b];
    stage_points(eng, {s}, {g}, int(b), int(blk.left_orbitals.size()), int(blk.right_orbitals.size()), eng.scratch_in, pv);
    eng.lane->arena.ensure(size_t(hp.arena_stride) * sizeof(double2));
    eng.lane->root_arena.ensure(size_t(hp.root_stride) * sizeof(double2));
    eng.lane->piv.ensure(size_t(hp.piv_stride) * 4);
    eng.lane->status.ensure(size_t(hp.slots) * 4);
    eng.lane->tout.ensure(sizeof(double));
    pv.arena = eng.lane->arena.p;
    pv.arena_stride = hp.arena_stride;
    pv.root_arena = eng.lane->root_arena.p;
    pv.root_stride = hp.root_stride;
    pv.piv = eng.lane->piv.as<int32_t>();
    pv.piv_stride = hp.piv_stride;
    pv.status = eng.lane->status.as<int32_t>();
    pv.status_stride = hp.slots;
    pv.t_out = eng.lane->tout.as<double>();
    if (eng.scratch_in.nlab > 0) {  // a multi-terminal point: T between its labels
        eng.lane->tlab.ensure(sizeof(double) * size_t(eng.scratch_in.nlab) * size_t(eng.scratch_in.nlab));
        pv.t_lab = eng.lane->tlab.as<double>();
    }
    pv.keep = 1;
    sweep(eng, hp, dp, pv, 1, false, eng.scratch_in);
    std::vector<int32_t> st(hp.slots);
    HIP_CHECK(hipMemcpyAsync(st.data(), eng.lane->status.p, st.size() * 4, hipMemcpyDeviceToHost, eng.lane->stream));
    HIP_CHECK(hipStreamSynchronize(eng.lane->stream));
    eng.collect();
    for (int32_t k = 0; k < hp.root; ++k)
        if (st[k] != kFrontOk) return CG_ERR_INVALID;
    if (st[hp.root] != kFrontOk) return CG_ERR_SINGULAR;
    return CG_OK;
}

/// x = G E_p W (or G^T E_p W with transpose) with the stored factors of factor_point: W (m x c) on
/// the slice orbitals `orbs` of the block, x (n_orb x c, complex, column-major) to the host.
void solve_columns(Engine& eng, const HostPlan& hp, const DevPlan& dp, const PointsView& pv, int64_t n_orb,
                   const CMat& w, int m, const DevBuf& orbs, bool transpose, std::vector<double2>& host) {
    const int c = w.cols;
    std::vector<double2> wh(size_t(m) * c);
    for (int col = 0; col < c; ++col)
        for (int a = 0; a < m; ++a) wh[a + size_t(col) * m] = double2{w(a, col).real(), w(a, col).imag()};
    DevBuf dw, xb;
    upload(dw, wh, eng.lane->stream);
    const size_t bytes = size_t(n_orb) * c * sizeof(double2);
    xb.ensure(bytes);
    HIP_CHECK(hipMemsetAsync(xb.p, 0, bytes, eng.lane->stream));
    SolveView xv{xb.as<double2>(), n_orb * c, int32_t(n_orb), c, dp.vars.as<int32_t>()};
    launch_scatter(xv, orbs.as<int32_t>(), m, dw.as<double2>(), int64_t(m) * c, transpose, 1, eng.lane->stream);
    solve(eng, hp, dp, pv, xv, 1, false, transpose, true);
    host.resize(size_t(n_orb) * c);
    HIP_CHECK(hipMemcpyAsync(host.data(), xb.p, bytes, hipMemcpyDeviceToHost, eng.lane->stream));
    HIP_CHECK(hipStreamSynchronize(eng.lane->stream));
    eng.collect();
}

}  // namespace

int32_t backward_batch(const cg_system* const* systems, const cg_sigma* const* sigmas, int64_t count, double* out_t,
                       double* out_grad, int32_t* out_status) {
    Engine& eng = Engine::get();
    if (!eng.ok) return CG_ERR_DEVICE;
    if (count <= 0) return CG_OK;
    for (int64_t p = 1; p < count; ++p)
        if (systems[p]->key != systems[0]->key) return CG_ERR_INVALID;  // one structure per call
    std::lock_guard<std::mutex> lock(eng.mu);
    auto sp = structure_of(eng, systems[0]);
    for (int64_t p = 0; p < count; ++p) structure_of(eng, systems[p]);
    const int64_t n_orb = int64_t(systems[0]->spec.h.n);
    std::fill(out_grad, out_grad + count * n_orb, 0.0);
    std::fill(out_t, out_t + count, 0.0);
    std::vector<int32_t> status(count, CG_OK);
    eng.lane = &eng.lanes[0];
    Lane& ln = *eng.lane;
    hipStream_t st = ln.stream;
    for (size_t b = 0; b < sp->host.size(); ++b) {
        const HostPlan& hp = sp->host[b];
        const DevPlan& dp = sp->dev[b];
        const Block& blk = systems[0]->sym.blocks[b];
        const int nl = int(hp.left_orbs.size()), nr = int(hp.right_orbs.size()), nrows = int(hp.rows.size());
        int cl = 0, cr = 0;
        for (int64_t p = 0; p < count; ++p) {
            cl = std::max(cl, sigmas[p]->blocks[b].wl.cols);
            cr = std::max(cr, sigmas[p]->blocks[b].wr.cols);
        }
        if (cl == 0 || cr == 0) continue;  // no channel: T and the gradient of the block are 0
        if (!solve_fits(hp.max_n) || !solve_fits(hp.nodes[hp.root].ne))
            return CG_ERR_INVALID;  // a front's rows (one column at a time) must fit the solves' LDS
        // Points per chunk: the stored fronts, the root and the two sets of columns within 64 GB.
        const size_t per_point = size_t(hp.arena_stride) * sizeof(double2) + size_t(hp.root_stride) * sizeof(double2) +
                                 size_t(hp.piv_stride) * 4 + size_t(hp.slots) * 4 +
                                 size_t(n_orb) * (size_t(cl) + size_t(cr) + 1) * sizeof(double2);
        const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>({count, int64_t((size_t(64) << 30) / per_point), 65535}));
        for (int64_t c0 = 0; c0 < count; c0 += chunk) {
            const int32_t P = int32_t(std::min(chunk, count - c0));
            std::vector<const cg_system*> ss(systems + c0, systems + c0 + P);
            std::vector<const cg_sigma*> gg(sigmas + c0, sigmas + c0 + P);
            // 1. The fronts of every point, factors kept (complex f64).
            PointsView pv{};
            stage_points(eng, ss, gg, int(b), int(blk.left_orbitals.size()), int(blk.right_orbitals.size()),
                         eng.scratch_in, pv);
            ensure_lane(ln, lane_need(hp, size_t(P), false));
            pv.arena = ln.arena.p;
            pv.arena_stride = hp.arena_stride;
            pv.root_arena = ln.root_arena.p;
            pv.root_stride = hp.root_stride;
            pv.piv = ln.piv.as<int32_t>();
            pv.piv_stride = hp.piv_stride;
            pv.status = ln.status.as<int32_t>();
            pv.status_stride = hp.slots;
            pv.t_out = ln.tout.as<double>();
            pv.keep = 1;
            sweep(eng, hp, dp, pv, P, false, eng.scratch_in);
            std::vector<int32_t> fs(size_t(P) * hp.slots);
            HIP_CHECK(hipMemcpyAsync(fs.data(), ln.status.p, fs.size() * 4, hipMemcpyDeviceToHost, st));
            // 2. psi = G E_R W_R and phi = G^T E_L conj(W_L) for all points (W zero-padded to the
            //    largest channel counts), the root and the backward substitution only.
            std::vector<double2> wr(size_t(P) * nr * cr, double2{0.0, 0.0}), wl(size_t(P) * nl * cl, double2{0.0, 0.0});
            for (int32_t q = 0; q < P; ++q) {
                const auto& sb = gg[q]->blocks[b];
                for (int col = 0; col < sb.wr.cols; ++col)
                    for (int a = 0; a < nr; ++a)
                        wr[size_t(q) * nr * cr + a + size_t(col) * nr] = double2{sb.wr(a, col).real(), sb.wr(a, col).imag()};
                for (int col = 0; col < sb.wl.cols; ++col)
                    for (int a = 0; a < nl; ++a)
                        wl[size_t(q) * nl * cl + a + size_t(col) * nl] = double2{sb.wl(a, col).real(), sb.wl(a, col).imag()};
            }
            AdjointBufs& ab = eng.adjoint;
            upload(ab.wr, wr, st);
            upload(ab.wl, wl, st);
            const size_t psi_bytes = size_t(P) * size_t(n_orb) * cr * sizeof(double2);
            const size_t phi_bytes = size_t(P) * size_t(n_orb) * cl * sizeof(double2);
            ab.psi.ensure(psi_bytes);
            ab.phi.ensure(phi_bytes);
            ab.t.ensure(size_t(P) * cl * cr * sizeof(double2));
            ab.grad.ensure(size_t(P) * size_t(n_orb) * sizeof(double));
            HIP_CHECK(hipMemsetAsync(ab.psi.p, 0, psi_bytes, st));
            HIP_CHECK(hipMemsetAsync(ab.phi.p, 0, phi_bytes, st));
            HIP_CHECK(hipMemsetAsync(ab.grad.p, 0, size_t(P) * size_t(n_orb) * sizeof(double), st));
            SolveView psi{ab.psi.as<double2>(), n_orb * cr, int32_t(n_orb), cr, dp.vars.as<int32_t>()};
            SolveView phi{ab.phi.as<double2>(), n_orb * cl, int32_t(n_orb), cl, dp.vars.as<int32_t>()};
            launch_scatter(psi, dp.right_orbs.as<int32_t>(), nr, ab.wr.as<double2>(), int64_t(nr) * cr, false, P, st);
            launch_scatter(phi, dp.left_orbs.as<int32_t>(), nl, ab.wl.as<double2>(), int64_t(nl) * cl, true, P, st);
            solve(eng, hp, dp, pv, psi, P, false, false, true);
            solve(eng, hp, dp, pv, phi, P, false, true, true);
            // 3. The contraction on the GPU: t = W_L^H psi_L and the gradient rows of the block.
            launch_adjoint(psi, phi, dp.left_orbs.as<int32_t>(), nl, ab.wl.as<double2>(), ab.t.as<double2>(),
                           dp.rows.as<int32_t>(), nrows, ab.grad.as<double>(), n_orb, P, st);
            HIP_CHECK(hipGetLastError());
            std::vector<double2> t(size_t(P) * cl * cr);
            std::vector<double> grad(size_t(P) * size_t(n_orb));
            HIP_CHECK(hipMemcpyAsync(t.data(), ab.t.p, t.size() * sizeof(double2), hipMemcpyDeviceToHost, st));
            HIP_CHECK(hipMemcpyAsync(grad.data(), ab.grad.p, grad.size() * sizeof(double), hipMemcpyDeviceToHost, st));
            HIP_CHECK(hipStreamSynchronize(st));
            eng.collect();
            for (int32_t q = 0; q < P; ++q) {
                const int64_t p = c0 + q;
                bool delayed = false;
                for (int32_t k = 0; k < hp.root; ++k) delayed |= fs[size_t(q) * hp.slots + k] != kFrontOk;
                if (delayed) {
                    status[p] = CG_ERR_INVALID;  // delayed separators: the solves follow the static plan
                    continue;
                }
                if (fs[size_t(q) * hp.slots + hp.root] != kFrontOk) {
                    status[p] = CG_ERR_SINGULAR;
                    continue;
                }
                double tb = 0.0;
                for (size_t e = 0; e < size_t(cl) * cr; ++e) {
                    const double2 v = t[size_t(q) * cl * cr + e];
                    tb += v.x * v.x + v.y * v.y;
                }
                out_t[p] += tb;
                for (int32_t row : hp.rows) out_grad[p * n_orb + row] += grad[size_t(q) * n_orb + row];
            }
        }
    }
    for (int64_t p = 0; p < count; ++p) {
        if (status[p] != CG_OK) {
            out_t[p] = NAN;
            std::fill(out_grad + p * n_orb, out_grad + (p + 1) * n_orb, NAN);
        }
        if (out_status) out_status[p] = status[p];
    }
    return CG_OK;
}

int32_t backward(const cg_system* s, const cg_sigma* g, double* out_t, double* out_grad) {
    int32_t st = CG_OK;
    const int32_t rc = backward_batch(&s, &g, 1, out_t, out_grad, &st);
    return rc != CG_OK ? rc : st;
}

int32_t multi_block_solve(const cg_system* s, const cg_sigma* g, int b, CMat* green, std::vector<cplx>* states) {
    Engine& eng = Engine::get();
    if (!eng.ok) return CG_ERR_DEVICE;
    std::lock_guard<std::mutex> lock(eng.mu);
    auto sp = structure_of(eng, s);
    const HostPlan& hp = sp->host[size_t(b)];
    const DevPlan& dp = sp->dev[size_t(b)];
    PointsView pv{};
    const int32_t rc = factor_point(eng, s, g, size_t(b), hp, dp, pv);
    if (rc != CG_OK) return rc;
    if (green) {
        // The root's border block holds -G on the kept orbitals (k0 = the root's ne).
        const NodeDesc& r = hp.nodes[hp.root];
        const int k = r.nl, n = r.n, k0 = r.ne;
        std::vector<double2> blk(size_t(k) * size_t(k));
        const double2* src = static_cast<const double2*>(eng.lane->root_arena.p) + k0 + int64_t(k0) * n;
        HIP_CHECK(hipMemcpy2DAsync(blk.data(), size_t(k) * sizeof(double2), src, size_t(n) * sizeof(double2),
                                   size_t(k) * sizeof(double2), size_t(k), hipMemcpyDeviceToHost, eng.lane->stream));
        HIP_CHECK(hipStreamSynchronize(eng.lane->stream));
        *green = CMat(k, k);
        for (size_t e = 0; e < blk.size(); ++e) green->a[e] = -cplx(blk[e].x, blk[e].y);
    }
    if (states) {
        const CMat& w = g->mblocks[size_t(b)].wab;
        if (!solve_fits(hp.max_n) || !solve_fits(hp.nodes[hp.root].ne)) return CG_ERR_INVALID;
        states->clear();
        if (w.cols > 0) {
            std::vector<double2> x;
            solve_columns(eng, hp, dp, pv, int64_t(s->spec.h.n), w, int(hp.left_orbs.size()), dp.left_orbs, false, x);
            states->resize(x.size());
            for (size_t e = 0; e < x.size(); ++e) (*states)[e] = cplx(x[e].x, x[e].y);
        }
    }
    return CG_OK;
}

int32_t scattering_states(const cg_system* s, const cg_sigma* g, int64_t* cols, double* out_l, double* out_r) {
    int64_t cl_total = 0, cr_total = 0;
    for (const auto& sb : g->blocks) {
        cl_total += sb.wl.cols;
        cr_total += sb.wr.cols;
    }
    cols[0] = cl_total;
    cols[1] = cr_total;
    if (!out_l && !out_r) return CG_OK;
    Engine& eng = Engine::get();
    if (!eng.ok) return CG_ERR_DEVICE;
    std::lock_guard<std::mutex> lock(eng.mu);
    auto sp = structure_of(eng, s);
    const int64_t n_orb = int64_t(s->spec.h.n);
    if (out_l) std::fill(out_l, out_l + 2 * n_orb * cl_total, 0.0);
    if (out_r) std::fill(out_r, out_r + 2 * n_orb * cr_total, 0.0);
    int64_t c0l = 0, c0r = 0;  // first column of the block's states
    for (size_t b = 0; b < sp->host.size(); ++b) {
        const HostPlan& hp = sp->host[b];
        const DevPlan& dp = sp->dev[b];
        const auto& sb = g->blocks[b];
        PointsView pv{};
        const int32_t rc = factor_point(eng, s, g, b, hp, dp, pv);
        if (rc != CG_OK) return rc;
        if (!solve_fits(hp.max_n) || !solve_fits(hp.nodes[hp.root].ne)) return CG_ERR_INVALID;
        std::vector<double2> x;
        if (out_l && sb.wl.cols > 0) {
            solve_columns(eng, hp, dp, pv, n_orb, sb.wl, int(hp.left_orbs.size()), dp.left_orbs, false, x);
            std::memcpy(out_l + 2 * n_orb * c0l, x.data(), x.size() * sizeof(double2));
        }
        if (out_r && sb.wr.cols > 0) {
            solve_columns(eng, hp, dp, pv, n_orb, sb.wr, int(hp.right_orbs.size()), dp.right_orbs, false, x);
            std::memcpy(out_r + 2 * n_orb * c0r, x.data(), x.size() * sizeof(double2));
        }
        c0l += sb.wl.cols;
        c0r += sb.wr.cols;
    }
    return CG_OK;
}

}  // namespace caroli::gpu
