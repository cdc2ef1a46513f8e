// Symbolic phase of nested dissection (see symbolic.hpp for the reference lines it follows).
#include "symbolic.hpp"

#include <algorithm>
#include <climits>
#include <stdexcept>

namespace caroli {

namespace {

// nested.rs:1051-1093: bisection along the longest extent (ties: lowest axis), separator layers
// as thick as the hopping range, leaves of at most kLeafOrbitals orbitals. Post-order.
int split(const SystemSpec& s, std::vector<size_t> sites, SiteTree& out) {
    if (sites.empty()) return -1;
    int orbitals = 0;
    for (size_t i : sites) orbitals += s.site_orbitals[i];
    int best = 0;
    int64_t blo = 0, bhi = 0;
    for (int d = 0; d < s.dim; ++d) {
        int64_t lo = LLONG_MAX, hi = LLONG_MIN;
        for (size_t i : sites) {
            lo = std::min(lo, s.cells[i][d]);
            hi = std::max(hi, s.cells[i][d]);
        }
        if (d == 0 || hi - lo > bhi - blo) {
            best = d;
            blo = lo;
            bhi = hi;
        }
    }
    int d = best;
    int64_t range = s.range[d];
    if (orbitals <= kLeafOrbitals || bhi - blo < range) {
        out.nodes.push_back({std::move(sites), {}});
        return int(out.nodes.size()) - 1;
    }
    int64_t mid = blo + (bhi - blo + 1 - range) / 2;
    std::vector<size_t> a, sep, b;
    for (size_t i : sites) {
        int64_t c = s.cells[i][d];
        if (c < mid)
            a.push_back(i);
        else if (c < mid + range)
            sep.push_back(i);
        else
            b.push_back(i);
    }
    std::vector<int> children;
    int ca = split(s, std::move(a), out);
    if (ca >= 0) children.push_back(ca);
    int cb = split(s, std::move(b), out);
    if (cb >= 0) children.push_back(cb);
    out.nodes.push_back({std::move(sep), std::move(children)});
    return int(out.nodes.size()) - 1;
}

bool contains(const std::vector<size_t>& sorted, size_t x) { return std::binary_search(sorted.begin(), sorted.end(), x); }

// nested.rs:1097-1122
void symbolic(const SparseMatrix& h, std::vector<Node>& nodes) {
    const size_t never = SIZE_MAX;
    std::vector<size_t> rank(h.n, never), last_rank(nodes.size());
    size_t next = 0;
    for (size_t k = 0; k < nodes.size(); ++k) {
        for (size_t o : nodes[k].eliminate) rank[o] = next++;
        last_rank[k] = next;
    }
    for (size_t k = 0; k < nodes.size(); ++k) {
        std::vector<size_t> bnd;
        for (size_t o : nodes[k].eliminate)
            for (int64_t e = h.row_ptr[o]; e < h.row_ptr[o + 1]; ++e) bnd.push_back(h.cols[e]);
        for (int c : nodes[k].children) bnd.insert(bnd.end(), nodes[c].boundary.begin(), nodes[c].boundary.end());
        bnd.erase(std::remove_if(bnd.begin(), bnd.end(), [&](size_t o) { return !(rank[o] == never || rank[o] >= last_rank[k]); }),
                  bnd.end());
        std::sort(bnd.begin(), bnd.end());
        bnd.erase(std::unique(bnd.begin(), bnd.end()), bnd.end());
        nodes[k].boundary = std::move(bnd);
    }
}

}  // namespace

SiteTree site_tree(const SystemSpec& s, const std::vector<size_t>& interior) {
    SiteTree t;
    int top = split(s, interior, t);
    if (top >= 0) t.top.push_back(top);
    return t;
}

Positions::Positions(const std::vector<size_t>& vars) {
    v.reserve(vars.size());
    for (size_t k = 0; k < vars.size(); ++k) v.emplace_back(vars[k], int(k));
    std::sort(v.begin(), v.end());
}

int Positions::get(size_t x) const {
    auto it = std::lower_bound(v.begin(), v.end(), std::make_pair(x, INT_MIN));
    return (it != v.end() && it->first == x) ? it->second : -1;
}

// nested.rs:147-159
std::pair<std::vector<size_t>, int> front_vars(const Node& node, const std::vector<const std::vector<size_t>*>& handed) {
    std::vector<size_t> elim = node.eliminate;
    for (auto* vars : handed)
        for (size_t o : *vars)
            if (!contains(node.boundary, o)) elim.push_back(o);
    std::sort(elim.begin(), elim.end());
    elim.erase(std::unique(elim.begin(), elim.end()), elim.end());
    int ne = int(elim.size());
    elim.insert(elim.end(), node.boundary.begin(), node.boundary.end());
    return {elim, ne};
}

// nested.rs:164-179
std::vector<PlanEntry> separator_entries(const SparseMatrix& h, const Node& node, const Positions& pos) {
    std::vector<PlanEntry> out;
    for (size_t i : node.eliminate) {
        int pi = pos.get(i);
        for (int64_t e = h.row_ptr[i]; e < h.row_ptr[i + 1]; ++e) {
            size_t j = h.cols[e];
            if (j == i) {
                out.push_back({pi, pi, e, false});
                continue;
            }
            int pj = pos.get(j);
            if (pj < 0) continue;
            bool later = contains(node.eliminate, j) ? j > i : contains(node.boundary, j);
            if (later) {
                out.push_back({pi, pj, e, false});
                out.push_back({pj, pi, e, true});
            }
        }
    }
    return out;
}

// nested.rs:74-96 and Plan::new (nested.rs:127-141)
Tree build_tree(const SiteTree& st, const SystemSpec& s, const std::vector<char>& mine) {
    Tree t;
    t.top = st.top;
    for (auto& [sites, children] : st.nodes) {
        Node n;
        for (size_t i : sites)
            for (int o = 0; o < s.site_orbitals[i]; ++o) {
                size_t orb = s.site_offsets[i] + size_t(o);
                if (mine[orb]) n.eliminate.push_back(orb);
            }
        std::sort(n.eliminate.begin(), n.eliminate.end());
        n.children = children;
        t.nodes.push_back(std::move(n));
    }
    symbolic(s.h, t.nodes);
    t.height.assign(t.nodes.size(), 0);
    for (size_t k = 0; k < t.nodes.size(); ++k) {
        for (int c : t.nodes[k].children) t.height[k] = std::max(t.height[k], t.height[c] + 1);
        t.levels = std::max(t.levels, t.height[k] + 1);
        const Node& node = t.nodes[k];
        std::vector<const std::vector<size_t>*> handed;
        for (int c : node.children) handed.push_back(&t.nodes[c].boundary);
        Plan p;
        std::tie(p.vars, p.ne) = front_vars(node, handed);
        Positions pos(p.vars);
        p.entries = separator_entries(s.h, node, pos);
        for (size_t o : node.eliminate) p.diagonal.push_back(pos.get(o));
        for (auto* vars : handed) {
            std::vector<int> where;
            for (size_t o : *vars) {
                int q = pos.get(o);
                if (q < 0) throw std::logic_error("child variable missing from the parent front");
                where.push_back(q);
            }
            p.child_pos.push_back(std::move(where));
        }
        t.plans.push_back(std::move(p));
    }
    return t;
}

namespace {
// nested.rs:378-420: union-find whose root is the smallest member.
struct Components {
    std::vector<size_t> p;
    explicit Components(size_t n) : p(n) {
        for (size_t i = 0; i < n; ++i) p[i] = i;
    }
    size_t find(size_t i) {
        while (p[i] != i) {
            p[i] = p[p[i]];
            i = p[i];
        }
        return i;
    }
    void join(size_t a, size_t b) {
        size_t ra = find(a), rb = find(b);
        if (ra != rb) p[std::max(ra, rb)] = std::min(ra, rb);
    }
    void join_matrix(const CMat& m, const std::vector<size_t>& rows, const std::vector<size_t>& cols) {
        for (int b = 0; b < m.cols; ++b)
            for (int a = 0; a < m.rows; ++a)
                if (m(a, b) != cplx(0.0)) join(rows[a], cols[b]);
    }
};

CMat restrict_to(const CMat& m, const std::vector<size_t>& k) {
    CMat r(int(k.size()), int(k.size()));
    for (size_t b = 0; b < k.size(); ++b)
        for (size_t a = 0; a < k.size(); ++a) r(int(a), int(b)) = m(int(k[a]), int(k[b]));
    return r;
}
}  // namespace

// nested.rs:536-609 (the gauge is trivial without a field)
Symbolic symbolic_phase(const SystemSpec& s) {
    const size_t n = s.h.n;
    size_t nsites = s.cells.size();
    if (nsites == 0) throw std::invalid_argument("empty region");
    int64_t first = LLONG_MAX, last = LLONG_MIN;
    for (auto& c : s.cells) {
        first = std::min(first, c[s.axis]);
        last = std::max(last, c[s.axis]);
    }
    std::vector<size_t> left_orbitals, right_orbitals, interior;
    for (size_t i = 0; i < nsites; ++i) {
        int64_t c = s.cells[i][s.axis];
        for (int o = 0; o < s.site_orbitals[i]; ++o) {
            if (c == first) left_orbitals.push_back(s.site_offsets[i] + o);
            if (c == last) right_orbitals.push_back(s.site_offsets[i] + o);
        }
        if (c != first && c != last) interior.push_back(i);
    }
    if (int(left_orbitals.size()) != s.left_h00.rows || int(right_orbitals.size()) != s.right_h00.rows)
        throw std::invalid_argument("lead cells do not match the end slices");
    SiteTree st = site_tree(s, interior);

    Components comp(n);
    for (size_t r = 0; r < n; ++r)
        for (int64_t e = s.h.row_ptr[r]; e < s.h.row_ptr[r + 1]; ++e)
            if (s.h.vals[e] != cplx(0.0)) comp.join(r, s.h.cols[e]);
    comp.join_matrix(s.left_h00, left_orbitals, left_orbitals);
    comp.join_matrix(s.left_h01, left_orbitals, left_orbitals);
    comp.join_matrix(s.right_h00, right_orbitals, right_orbitals);
    comp.join_matrix(s.right_h01, right_orbitals, right_orbitals);
    std::vector<size_t> roots(n);
    for (size_t i = 0; i < n; ++i) roots[i] = comp.find(i);
    std::vector<size_t> labels = roots;
    std::sort(labels.begin(), labels.end());
    labels.erase(std::unique(labels.begin(), labels.end()), labels.end());

    Symbolic sym;
    for (size_t label : labels) {
        std::vector<char> mine(n);
        for (size_t i = 0; i < n; ++i) mine[i] = roots[i] == label;
        auto pick = [&](const std::vector<size_t>& orbs) {
            std::vector<size_t> k;
            for (size_t q = 0; q < orbs.size(); ++q)
                if (mine[orbs[q]]) k.push_back(q);
            return k;
        };
        auto kl = pick(left_orbitals), kr = pick(right_orbitals);
        if (kl.empty() || kr.empty()) continue;
        Block b;
        b.left_h00 = restrict_to(s.left_h00, kl);
        b.left_h01 = restrict_to(s.left_h01, kl);
        b.right_h00 = restrict_to(s.right_h00, kr);
        b.right_h01 = restrict_to(s.right_h01, kr);
        for (size_t q : kl) b.left_orbitals.push_back(left_orbitals[q]);
        for (size_t q : kr) b.right_orbitals.push_back(right_orbitals[q]);
        b.same_lead = b.left_h00 == b.right_h00 && b.left_h01 == b.right_h01;
        b.kept = b.left_orbitals;
        bool same = b.left_orbitals == b.right_orbitals;
        b.right = same ? 0 : int(b.left_orbitals.size());
        if (!same) b.kept.insert(b.kept.end(), b.right_orbitals.begin(), b.right_orbitals.end());
        b.tree = build_tree(st, s, mine);
        sym.blocks.push_back(std::move(b));
    }
    return sym;
}

// nested_multi.rs:60-161. One difference from the reference, deliberate: an orbital of a site
// that a lead couples to, but which no lead reaches itself, is kept at the root with its z - H
// entries (the reference leaves it out of both the tree and the root); results agree whenever
// every orbital of such sites is coupled, as in the BHZ, square, graphene and cubic leads.
Symbolic multi_symbolic(const SystemSpec& s, const std::vector<AttachedLead>& leads) {
    const size_t n = s.h.n, nsites = s.cells.size();
    if (nsites == 0) throw std::invalid_argument("empty region");
    if (leads.empty()) throw std::invalid_argument("no leads");
    std::vector<char> coupled(n, 0);
    for (auto& l : leads) {
        if (l.h00.rows == 0 || l.h00.rows != l.h00.cols || l.h01.rows != l.h00.rows || l.h01.cols != l.h00.rows ||
            l.coupling.rows != int(l.orbitals.size()) || l.coupling.cols != l.h00.rows || l.orbitals.empty())
            throw std::invalid_argument("malformed lead");
        for (size_t o : l.orbitals) {
            if (o >= n) throw std::invalid_argument("lead orbital outside the device");
            coupled[o] = 1;
        }
    }
    std::vector<size_t> interior;
    std::vector<char> at_root(n, 0);
    for (size_t i = 0; i < nsites; ++i) {
        bool any = false;
        for (int o = 0; o < s.site_orbitals[i]; ++o) any |= coupled[s.site_offsets[i] + o] != 0;
        if (!any) {
            interior.push_back(i);
        } else {
            for (int o = 0; o < s.site_orbitals[i]; ++o) at_root[s.site_offsets[i] + o] = 1;
        }
    }
    SiteTree st = site_tree(s, interior);

    // Components over the device orbitals, then the orbitals of each lead cell (nested_multi.rs:79-97).
    std::vector<size_t> offsets;
    size_t total = n;
    for (auto& l : leads) {
        offsets.push_back(total);
        total += size_t(l.h00.rows);
    }
    Components comp(total);
    for (size_t r = 0; r < n; ++r)
        for (int64_t e = s.h.row_ptr[r]; e < s.h.row_ptr[r + 1]; ++e)
            if (s.h.vals[e] != cplx(0.0)) comp.join(r, s.h.cols[e]);
    for (size_t p = 0; p < leads.size(); ++p) {
        std::vector<size_t> cell(size_t(leads[p].h00.rows));
        for (size_t k = 0; k < cell.size(); ++k) cell[k] = offsets[p] + k;
        comp.join_matrix(leads[p].h00, cell, cell);
        comp.join_matrix(leads[p].h01, cell, cell);
        comp.join_matrix(leads[p].coupling, leads[p].orbitals, cell);
    }
    std::vector<size_t> roots(total);
    for (size_t i = 0; i < total; ++i) roots[i] = comp.find(i);
    std::vector<size_t> labels = roots;
    std::sort(labels.begin(), labels.end());
    labels.erase(std::unique(labels.begin(), labels.end()), labels.end());

    Symbolic sym;
    for (size_t label : labels) {
        auto mine = [&](size_t o) { return roots[o] == label; };
        Block b;
        b.multi = true;
        bool reached = false;
        for (size_t p = 0; p < leads.size(); ++p) {
            const AttachedLead& l = leads[p];
            std::vector<size_t> kc, kd;
            for (int k = 0; k < l.h00.rows; ++k)
                if (mine(offsets[p] + size_t(k))) kc.push_back(size_t(k));
            for (size_t k = 0; k < l.orbitals.size(); ++k)
                if (mine(l.orbitals[k])) kd.push_back(k);
            AttachedLead r;
            if (!kc.empty() && !kd.empty()) {
                reached = true;
                r.h00 = restrict_to(l.h00, kc);
                r.h01 = restrict_to(l.h01, kc);
                for (size_t k : kd) r.orbitals.push_back(l.orbitals[k]);
                r.coupling = CMat(int(kd.size()), int(kc.size()));
                for (size_t y = 0; y < kc.size(); ++y)
                    for (size_t x = 0; x < kd.size(); ++x) r.coupling(int(x), int(y)) = l.coupling(int(kd[x]), int(kc[y]));
            } else {
                kc.clear();
            }
            b.leads.push_back(std::move(r));
            b.lead_cells.push_back(std::move(kc));
        }
        if (!reached) continue;
        std::vector<char> in_block(n);
        for (size_t o = 0; o < n; ++o) {
            in_block[o] = mine(o);
            if (in_block[o] && at_root[o]) b.kept.push_back(o);
        }
        b.left_orbitals = b.kept;
        b.right_orbitals = b.kept;
        b.right = 0;
        b.same_lead = true;
        b.tree = build_tree(st, s, in_block);
        sym.blocks.push_back(std::move(b));
    }
    if (sym.blocks.empty()) throw std::invalid_argument("no lead reaches the device");
    return sym;
}

}  // namespace caroli
