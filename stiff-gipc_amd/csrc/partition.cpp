// CEMAS partitions on the host: the slot maps of a vertex partition and a
// partitioner into parts of at most one bank of vertices.
//
// Slot maps (app/common/sim_bootstrap.cpp:90-113, set_mas_partition): part
// p owns the bank slots [B p, B p + B); its vertices take them in increasing
// vertex order.  partId_map_real[slot] is the vertex in a slot (-1 for an
// empty slot), real_map_partId[vertex] its slot.  The reference's loop
// restarts the in-part counter whenever the part id changes from one vertex
// to the next and then scans the slots for vertices 0, 1, 2, ... in turn, so
// it is only correct for part ids in non-decreasing order (the "_sorted"
// meshes MeshProcess writes); for those the rule here gives the same arrays,
// and it stays a valid map for any order.
//
// Partitioner (replaces METIS_PartGraphKway in
// MeshProcess/metis_partition/src/node_edge_model.cpp:5-28 as called by
// metis_sort.cpp:71-110; METIS is not part of this library): recursive
// multilevel bisection into k = ceil(n / (B - 1)) parts, the fill the
// reference's retry loop settles at (part count grown until the largest part
// fits a bank).  Each bisection coarsens the subset's graph by heavy-edge
// matching, grows side 0 greedily from four seeds on the coarsest graph,
// keeps the best Fiduccia-Mattheyses-refined cut and refines it again on
// every finer level; side 0 may hold any weight that lets both sides' parts
// fit a bank.  Parts are numbered in the order of the recursion's leaves, so
// consecutive part ids are close in the graph: the coarse MAS levels group
// consecutive parts (level-1 domains), and a partition renumbered at random
// roughly doubles the PCG iterations on the reference's bunny mesh.
// Deterministic (fixed matching orders); every part holds 1..B vertices.
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <queue>
#include <utility>
#include <vector>

#include "../../include/sgipc.h"

namespace {

// Weighted graph (CSR, symmetric) of the multilevel bisection.
struct WGraph {
    int n = 0;
    std::vector<int> xadj, adj, ew, vw;
    int total() const
    {
        int t = 0;
        for (int w : vw) t += w;
        return t;
    }
};

// Deterministic permutation (LCG shuffle) used as the matching visit order.
std::vector<int> shuffled(int n, uint32_t seed)
{
    std::vector<int> p(n);
    for (int i = 0; i < n; ++i) p[i] = i;
    uint32_t x = seed * 2654435761u + 1u;
    for (int i = n - 1; i > 0; --i) {
        x = x * 1664525u + 1013904223u;
        std::swap(p[i], p[(int)(x % (uint32_t)(i + 1))]);
    }
    return p;
}

// Heavy-edge matching: each unmatched vertex joins its unmatched neighbour
// of largest edge weight (combined weight at most cap).  Returns false when
// the graph shrinks by less than 10 %.
bool coarsen(const WGraph& g, int cap, uint32_t seed, WGraph& c, std::vector<int>& cmap)
{
    const int n = g.n;
    std::vector<int> match(n, -1);
    const std::vector<int> order = shuffled(n, seed);
    for (int u : order) {
        if (match[u] >= 0) continue;
        int best = -1, bw = -1;
        for (int e = g.xadj[u]; e < g.xadj[u + 1]; ++e) {
            const int v = g.adj[e];
            if (match[v] >= 0 || v == u || g.vw[u] + g.vw[v] > cap) continue;
            if (g.ew[e] > bw) {
                bw = g.ew[e];
                best = v;
            }
        }
        match[u] = best >= 0 ? best : u;
        if (best >= 0) match[best] = u;
    }
    cmap.assign(n, -1);
    int nc = 0;
    for (int u = 0; u < n; ++u)
        if (cmap[u] < 0) {
            cmap[u] = nc;
            cmap[match[u]] = nc;
            ++nc;
        }
    if (nc > 0.9 * n) return false;
    c.n = nc;
    c.vw.assign(nc, 0);
    for (int u = 0; u < n; ++u) c.vw[cmap[u]] += g.vw[u];
    // coarse adjacency: merge the two members' lists (marker array per coarse vertex)
    std::vector<int> members(2 * nc, -1);
    for (int u = 0; u < n; ++u) {
        const int cu = cmap[u];
        members[2 * cu + (members[2 * cu] < 0 ? 0 : 1)] = u;
    }
    std::vector<int> mark(nc, -1), pos(nc, 0);
    c.xadj.assign(nc + 1, 0);
    c.adj.clear();
    c.ew.clear();
    for (int cu = 0; cu < nc; ++cu) {
        const int start = (int)c.adj.size();
        for (int h = 0; h < 2; ++h) {
            const int u = members[2 * cu + h];
            if (u < 0) continue;
            for (int e = g.xadj[u]; e < g.xadj[u + 1]; ++e) {
                const int cv = cmap[g.adj[e]];
                if (cv == cu) continue;
                if (mark[cv] != cu) {
                    mark[cv] = cu;
                    pos[cv] = (int)c.adj.size();
                    c.adj.push_back(cv);
                    c.ew.push_back(g.ew[e]);
                } else {
                    c.ew[pos[cv]] += g.ew[e];
                }
            }
        }
        c.xadj[cu + 1] = (int)c.adj.size();
        (void)start;
    }
    return true;
}

int cut_of(const WGraph& g, const std::vector<char>& side)
{
    long c = 0;
    for (int u = 0; u < g.n; ++u)
        for (int e = g.xadj[u]; e < g.xadj[u + 1]; ++e)
            if (side[u] != side[g.adj[e]]) c += g.ew[e];
    return (int)(c / 2);
}

// distance of a side weight from [lo, hi]
int excess(int w, int lo, int hi) { return w < lo ? lo - w : (w > hi ? w - hi : 0); }

// Fiduccia-Mattheyses refinement of a bisection: moves of the highest gain
// that keep side 0's weight within [lo, hi] (or bring it closer), each
// vertex once per pass, kept up to the best state seen: within the bounds
// first, then the smallest cut, then closest to target.
void refine(const WGraph& g, std::vector<char>& side, int lo, int hi, int target, int passes)
{
    const int n = g.n;
    std::vector<int> gain(n), ver(n, 0);
    std::vector<char> locked(n);
    int w0 = 0;
    for (int u = 0; u < n; ++u)
        if (side[u] == 0) w0 += g.vw[u];
    for (int pass = 0; pass < passes; ++pass) {
        using Item = std::pair<int, std::pair<int, int>>;  // gain, (-vertex, version)
        std::priority_queue<Item> q[2];
        for (int u = 0; u < n; ++u) {
            int ed = 0, id = 0;
            for (int e = g.xadj[u]; e < g.xadj[u + 1]; ++e) (side[g.adj[e]] == side[u] ? id : ed) += g.ew[e];
            gain[u] = ed - id;
            locked[u] = 0;
            if (ed > 0) q[(int)side[u]].push({gain[u], {-u, ver[u]}});
        }
        int cut = cut_of(g, side), best_cut = cut, best_len = 0, w_best = w0;
        int best_dev = std::abs(w0 - target) + (1 << 20) * excess(w0, lo, hi);
        std::vector<int> moves;
        const int limit = std::max(50, n / 20);
        for (;;) {
            // candidate from each side that keeps the balance
            int pick = -1, pg = 0;
            for (int sd = 0; sd < 2; ++sd) {
                while (!q[sd].empty()) {
                    const Item& t = q[sd].top();
                    const int u = -t.second.first;
                    if (locked[u] || side[u] != sd || t.second.second != ver[u]) {
                        q[sd].pop();
                        continue;
                    }
                    break;
                }
                if (q[sd].empty()) continue;
                const int u = -q[sd].top().second.first;
                const int nw = sd == 0 ? w0 - g.vw[u] : w0 + g.vw[u];
                // within the bounds, or closer to them than before
                if ((nw < lo || nw > hi) && excess(nw, lo, hi) >= excess(w0, lo, hi)) continue;
                if (pick < 0 || gain[u] > pg || (gain[u] == pg && std::abs(nw - target) < std::abs(w0 - target))) {
                    pick = u;
                    pg = gain[u];
                }
            }
            if (pick < 0) break;
            const int u = pick, sd = side[u];
            q[sd].pop();
            locked[u] = 1;
            side[u] = (char)(1 - sd);
            w0 += sd == 0 ? -g.vw[u] : g.vw[u];
            cut -= gain[u];
            moves.push_back(u);
            for (int e = g.xadj[u]; e < g.xadj[u + 1]; ++e) {
                const int v = g.adj[e];
                if (locked[v]) continue;
                gain[v] += side[v] == side[u] ? -2 * g.ew[e] : 2 * g.ew[e];
                ++ver[v];
                q[(int)side[v]].push({gain[v], {-v, ver[v]}});
            }
            const int dev = std::abs(w0 - target) + (1 << 20) * excess(w0, lo, hi);
            if (dev < best_dev - (1 << 19) || (excess(w0, lo, hi) == 0 && best_dev < (1 << 20) &&
                                               (cut < best_cut || (cut == best_cut && dev < best_dev)))) {
                best_cut = cut;
                best_dev = dev;
                best_len = (int)moves.size();
                w_best = w0;
            } else if ((int)moves.size() - best_len > limit) {
                break;
            }
        }
        for (int i = (int)moves.size() - 1; i >= best_len; --i) side[moves[i]] ^= 1;
        w0 = w_best;
        if (best_len == 0) break;
    }
}

// Greedy graph growing: side 0 grows from seed by the boundary vertex of
// largest gain until it weighs at least target (restarting at the first
// vertex left when a component is exhausted).
std::vector<char> grow(const WGraph& g, int seed, int target)
{
    const int n = g.n;
    std::vector<char> side(n, 1);
    std::vector<int> gain(n, 0), ver(n, 0);
    using Item = std::pair<int, std::pair<int, int>>;
    std::priority_queue<Item> q;
    int w0 = 0, next = 0;
    auto add = [&](int u) {
        side[u] = 0;
        w0 += g.vw[u];
        for (int e = g.xadj[u]; e < g.xadj[u + 1]; ++e) {
            const int v = g.adj[e];
            if (side[v] == 0) continue;
            gain[v] += 2 * g.ew[e];
            ++ver[v];
            q.push({gain[v], {-v, ver[v]}});
        }
    };
    for (int u = 0; u < n; ++u) {
        int d = 0;
        for (int e = g.xadj[u]; e < g.xadj[u + 1]; ++e) d += g.ew[e];
        gain[u] = -d;
    }
    add(seed);
    while (w0 < target) {
        int u = -1;
        while (!q.empty()) {
            const Item t = q.top();
            q.pop();
            const int v = -t.second.first;
            if (side[v] == 1 && t.second.second == ver[v]) {
                u = v;
                break;
            }
        }
        if (u < 0) {
            while (next < n && side[next] == 0) ++next;
            if (next == n) break;
            u = next;
        }
        add(u);
    }
    return side;
}

int far_vertex(const WGraph& g, int src)
{
    std::vector<int> dist(g.n, -1), q{src};
    dist[src] = 0;
    int last = src;
    for (size_t h = 0; h < q.size(); ++h) {
        const int u = q[h];
        last = u;
        for (int e = g.xadj[u]; e < g.xadj[u + 1]; ++e)
            if (dist[g.adj[e]] < 0) {
                dist[g.adj[e]] = dist[u] + 1;
                q.push_back(g.adj[e]);
            }
    }
    return last;
}

// Multilevel bisection: coarsen by heavy-edge matching, grow from a few
// seeds on the coarsest graph and keep the best refined cut, then project
// and refine level by level.
std::vector<char> bisect(const WGraph& g0, int lo, int hi, int target, uint32_t seed)
{
    std::vector<WGraph> gs{g0};
    std::vector<std::vector<int>> maps;
    const int cap = std::max(1, (hi - lo) / 2 + 1);
    while (gs.back().n > 96) {
        WGraph c;
        std::vector<int> cm;
        if (!coarsen(gs.back(), std::max(cap, 1), seed + (uint32_t)gs.size(), c, cm)) break;
        gs.push_back(std::move(c));
        maps.push_back(std::move(cm));
    }
    const WGraph& gc = gs.back();
    std::vector<char> best;
    int best_cut = -1;
    const int f = far_vertex(gc, 0);
    const int seeds[4] = {f, far_vertex(gc, f), 0, gc.n / 2};
    for (int sd : seeds) {
        std::vector<char> side = grow(gc, sd, target);
        refine(gc, side, lo, hi, target, 4);
        int w0 = 0;
        for (int u = 0; u < gc.n; ++u)
            if (side[u] == 0) w0 += gc.vw[u];
        const int c = cut_of(gc, side) + (w0 < lo || w0 > hi ? 1 << 28 : 0);
        if (best_cut < 0 || c < best_cut) {
            best_cut = c;
            best = side;
        }
    }
    for (int l = (int)maps.size() - 1; l >= 0; --l) {
        std::vector<char> fine(gs[l].n);
        for (int u = 0; u < gs[l].n; ++u) fine[u] = best[maps[l][u]];
        best.swap(fine);
        refine(gs[l], best, lo, hi, target, 2);
    }
    // the bounds are a guarantee (the parts below must fit a bank): unit
    // weights here, so vertices of the heavy side move, largest gain first
    int w0 = 0;
    for (int u = 0; u < g0.n; ++u) w0 += best[u] == 0 ? 1 : 0;
    if (w0 < lo || w0 > hi) {
        const int from = w0 > hi ? 0 : 1;
        std::vector<std::pair<int, int>> cand;
        for (int u = 0; u < g0.n; ++u) {
            if (best[u] != from) continue;
            int gsum = 0;
            for (int e = g0.xadj[u]; e < g0.xadj[u + 1]; ++e) gsum += best[g0.adj[e]] != from ? g0.ew[e] : -g0.ew[e];
            cand.push_back({-gsum, u});
        }
        std::sort(cand.begin(), cand.end());
        for (const auto& c : cand) {
            if (w0 >= lo && w0 <= hi) break;
            best[c.second] = (char)(1 - from);
            w0 += from == 0 ? -1 : 1;
        }
    }
    return best;
}

// Recursive bisection of a vertex subset into k parts of at most B vertices;
// parts are numbered in the order of the recursion's leaves (side 0 first),
// so consecutive part ids are neighbours in the bisection tree.
struct Bisector {
    int n;
    const uint32_t *st, *num, *lst;
    int B;
    int* part;
    int next_part = 0;
    std::vector<int> loc;

    Bisector(int n_, const uint32_t* s_, const uint32_t* m_, const uint32_t* l_, int b_, int* p_)
        : n(n_), st(s_), num(m_), lst(l_), B(b_), part(p_), loc(n_, -1)
    {
    }

    void split(std::vector<int>& vs, int k)
    {
        const int m = (int)vs.size();
        if (m == 0) return;
        if (m <= B) {
            const int p = next_part++;
            for (int v : vs) part[v] = p;
            return;
        }
        k = std::max(k, (m + B - 1) / B);  // (the bounds keep m <= k B; this is a guard)
        const int k1 = k / 2, k2 = k - k1;
        // side 0 carries k1 of the k parts: any size that lets both sides'
        // parts fit a bank (the part count's slack of about one vertex per
        // part is what the cut may spend), refined toward its share
        const int target = (int)(((long)m * k1 + k / 2) / k);
        const int lo = std::max(1, m - k2 * B), hi = std::min(m - 1, k1 * B);
        for (int i = 0; i < m; ++i) loc[vs[i]] = i;
        WGraph g;
        g.n = m;
        g.vw.assign(m, 1);
        g.xadj.assign(m + 1, 0);
        for (int i = 0; i < m; ++i) {
            const int u = vs[i];
            for (uint32_t k = 0; k < num[u]; ++k) {
                const int j = loc[lst[st[u] + k]];
                if (j >= 0 && j != i) {
                    g.adj.push_back(j);
                    g.ew.push_back(1);
                }
            }
            g.xadj[i + 1] = (int)g.adj.size();
        }
        for (int v : vs) loc[v] = -1;
        const std::vector<char> side = bisect(g, lo, hi, target, (uint32_t)next_part * 7919u + (uint32_t)m);
        std::vector<int> a, b;
        for (int i = 0; i < m; ++i) (side[i] == 0 ? a : b).push_back(vs[i]);
        std::vector<int>().swap(vs);
        split(a, k1);
        split(b, k2);
    }
};

}  // namespace

extern "C" {

int sgipc_mas_partition_maps(const int* part_id, int n, int bank, int* map_size, int* partId_map_real,
                             int* real_map_partId)
{
    if (n <= 0) return SGIPC_EMPTY;
    if (!part_id || !map_size || bank < 1 || bank > 64) return SGIPC_BAD_ARG;
    int mx = -1;
    for (int i = 0; i < n; ++i) {
        if (part_id[i] < 0) return SGIPC_BAD_INDEX;
        mx = std::max(mx, part_id[i]);
    }
    if ((long)(mx + 1) * bank > (1L << 30)) return SGIPC_CAPACITY;
    std::vector<int> fill((size_t)mx + 1, 0);
    for (int i = 0; i < n; ++i)
        if (++fill[part_id[i]] > bank) return SGIPC_CAPACITY;  // a part larger than one bank
    *map_size = (mx + 1) * bank;
    if (!partId_map_real && !real_map_partId) return SGIPC_OK;  // size query
    if (!partId_map_real || !real_map_partId) return SGIPC_BAD_ARG;
    std::fill(partId_map_real, partId_map_real + *map_size, -1);
    std::fill(fill.begin(), fill.end(), 0);
    for (int i = 0; i < n; ++i) {
        const int slot = bank * part_id[i] + fill[part_id[i]]++;
        partId_map_real[slot] = i;
        real_map_partId[i] = slot;
    }
    return SGIPC_OK;
}

int sgipc_partition_vertices(int n, const uint32_t* nbr_start, const uint32_t* nbr_num, const uint32_t* nbr_list,
                             int bank, int* part_id, int* n_parts)
{
    if (n <= 0) return SGIPC_EMPTY;
    if (!nbr_start || !nbr_num || !nbr_list || !part_id || !n_parts || bank < 1 || bank > 64) return SGIPC_BAD_ARG;
    for (int u = 0; u < n; ++u)
        for (uint32_t k = 0; k < nbr_num[u]; ++k)
            if (nbr_list[nbr_start[u] + k] >= (uint32_t)n) return SGIPC_BAD_INDEX;
    Bisector bs{n, nbr_start, nbr_num, nbr_list, bank, part_id};
    std::vector<int> all(n);
    for (int i = 0; i < n; ++i) all[i] = i;
    // parts of bank - 1 vertices on average (the reference's METIS call
    // settles at about that fill: its part count grows until the largest fits)
    const int fill = bank > 1 ? bank - 1 : 1;
    const int k = (n + fill - 1) / fill;
    bs.split(all, k);
    *n_parts = bs.next_part;
    return SGIPC_OK;
}

}  // extern "C"
