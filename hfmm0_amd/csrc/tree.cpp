// Host-side octree and interaction lists.
//
// Semantics follow the reference's cubic octree (jaxfmm/trees/cubic.py:316-695) and
// list builders (jaxfmm/trees/lists.py:619-802): Morton codes at the finest level,
// bottom-up coarsening where a parent absorbs its children iff its count is at most
// N_max and the 2:1 balance check passes, automatic depth refinement, and the
// V / adaptive-U / W / X lists.  The data structure is different: per-level sorted
// code arrays searched by binary search instead of fixed-size padded buffers, and
// boxes stored in canonical order (level ascending, Morton ascending).
#include <algorithm>
#include <cmath>
#include <numeric>

#include "common.h"

namespace hfmm {

static inline uint64_t expand_bits(uint64_t v) {
    v = (v | (v << 32)) & 0x1F00000000FFFFull;
    v = (v | (v << 16)) & 0x1F0000FF0000FFull;
    v = (v | (v << 8)) & 0x100F00F00F00F00Full;
    v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

static inline uint64_t compact_bits(uint64_t v) {
    v &= 0x1249249249249249ull;
    v = (v | (v >> 2)) & 0x10C30C30C30C30C3ull;
    v = (v | (v >> 4)) & 0x100F00F00F00F00Full;
    v = (v | (v >> 8)) & 0x1F0000FF0000FFull;
    v = (v | (v >> 16)) & 0x1F00000000FFFFull;
    v = (v | (v >> 32)) & 0x1FFFFFull;
    return v;
}

uint64_t morton3(uint64_t x, uint64_t y, uint64_t z) {
    return expand_bits(x) | (expand_bits(y) << 1) | (expand_bits(z) << 2);
}

void morton_grid(uint64_t c, int& x, int& y, int& z) {
    x = (int)compact_bits(c);
    y = (int)compact_bits(c >> 1);
    z = (int)compact_bits(c >> 2);
}

static int lookup(const std::vector<uint64_t>& codes, uint64_t c) {
    auto it = std::lower_bound(codes.begin(), codes.end(), c);
    if (it == codes.end() || *it != c) return -1;
    return (int)(it - codes.begin());
}

int Tree::find(int lvl, uint64_t c) const {
    if (lvl < 1 || lvl > depth) return -1;
    auto b = code.begin() + level_start[lvl], e = code.begin() + level_start[lvl + 1];
    auto it = std::lower_bound(b, e, c);
    if (it == e || *it != c) return -1;
    return (int)(it - code.begin());
}

namespace {

struct LevelBoxes {
    std::vector<uint64_t> code;
    std::vector<int> count;
    std::vector<uint8_t> leaf;
};

const int kNeigh[26][3] = {
    {-1, -1, -1}, {-1, -1, 0}, {-1, -1, 1}, {-1, 0, -1}, {-1, 0, 0}, {-1, 0, 1}, {-1, 1, -1},
    {-1, 1, 0},   {-1, 1, 1},  {0, -1, -1}, {0, -1, 0},  {0, -1, 1}, {0, 0, -1}, {0, 0, 1},
    {0, 1, -1},   {0, 1, 0},   {0, 1, 1},   {1, -1, -1}, {1, -1, 0}, {1, -1, 1}, {1, 0, -1},
    {1, 0, 0},    {1, 0, 1},   {1, 1, -1},  {1, 1, 0},   {1, 1, 1}};

// child octant c of the neighbour in direction d touches the centre box
bool adj_child(const int* d, int c) {
    const int b[3] = {c & 1, (c >> 1) & 1, (c >> 2) & 1};
    for (int a = 0; a < 3; ++a) {
        if (d[a] == -1 && b[a] != 1) return false;
        if (d[a] == 1 && b[a] != 0) return false;
    }
    return true;
}

// status of the children of parent code pc at level lc: any non-leaf among masked octants?
bool masked_child_internal(const LevelBoxes& lv, uint64_t pc, const int* d) {
    for (int c = 0; c < 8; ++c) {
        if (!adj_child(d, c)) continue;
        int i = lookup(lv.code, pc * 8 + (uint64_t)c);
        if (i >= 0 && !lv.leaf[i]) return true;
    }
    return false;
}

struct FixedTree {
    std::vector<LevelBoxes> levels;   // index = level
    std::vector<uint64_t> sorted_codes;
    std::vector<int64_t> order;
    int maxleaf = 0;
};

FixedTree build_fixed(const double* xyz, int64_t n, int md, int n_max, const double* pmin, double extent) {
    FixedTree ft;
    const double grid = std::ldexp(1.0, md);
    const uint64_t gmax = (1ull << md) - 1;
    std::vector<uint64_t> pc(n);
    for (int64_t i = 0; i < n; ++i) {
        uint64_t g[3];
        for (int a = 0; a < 3; ++a) {
            double s = (xyz[3 * i + a] - pmin[a]) / extent * grid;   // cubic.py:390-391 order
            uint64_t u = s > 0 ? (uint64_t)s : 0;
            g[a] = std::min(u, gmax);
        }
        pc[i] = morton3(g[0], g[1], g[2]);
    }
    ft.order.resize(n);
    std::iota(ft.order.begin(), ft.order.end(), 0);
    std::stable_sort(ft.order.begin(), ft.order.end(), [&](int64_t a, int64_t b) { return pc[a] < pc[b]; });
    ft.sorted_codes.resize(n);
    for (int64_t i = 0; i < n; ++i) ft.sorted_codes[i] = pc[ft.order[i]];

    ft.levels.assign(md + 2, LevelBoxes());
    LevelBoxes& fin = ft.levels[md];
    for (int64_t i = 0; i < n; ++i) {
        if (i == 0 || ft.sorted_codes[i] != ft.sorted_codes[i - 1]) {
            fin.code.push_back(ft.sorted_codes[i]);
            fin.count.push_back(0);
            fin.leaf.push_back(1);
        }
        fin.count.back()++;
    }
    for (int L = md; L >= 2; --L) {
        LevelBoxes& cur = ft.levels[L];
        LevelBoxes par;
        std::vector<int> parent_of(cur.code.size());
        for (size_t i = 0; i < cur.code.size(); ++i) {
            uint64_t p = cur.code[i] >> 3;
            if (par.code.empty() || par.code.back() != p) {
                par.code.push_back(p);
                par.count.push_back(0);
                par.leaf.push_back(0);
            }
            par.count.back() += cur.count[i];
            parent_of[i] = (int)par.code.size() - 1;
        }
        const int Lp = L - 1;
        const int64_t size = 1ll << Lp;
        std::vector<uint8_t> merge(par.code.size());
        const uint64_t last_code = cur.code.back();     // last box of level L (quirk lookup)
        for (size_t b = 0; b < par.code.size(); ++b) {
            bool ok = par.count[b] <= n_max;
            if (ok && L != md) {                        // cubic.py:510-511: no check at the first step
                int gx, gy, gz;
                morton_grid(par.code[b], gx, gy, gz);
                for (int d = 0; d < 26 && ok; ++d) {
                    int64_t nx = gx + kNeigh[d][0], ny = gy + kNeigh[d][1], nz = gz + kNeigh[d][2];
                    if (nx < 0 || ny < 0 || nz < 0 || nx >= size || ny >= size || nz >= size) continue;
                    uint64_t nc = morton3(nx, ny, nz);
                    if (lookup(par.code, nc) >= 0) {
                        if (masked_child_internal(cur, nc, kNeigh[d])) ok = false;
                    } else if (nc == 0) {
                        // reference lookup of an absent code-0 neighbour: the children of the
                        // last level-L box are tested (cubic.py:466-476)
                        if (masked_child_internal(ft.levels[L + 1], last_code, kNeigh[d])) ok = false;
                    }
                }
            }
            merge[b] = ok;
        }
        LevelBoxes kept;
        for (size_t i = 0; i < cur.code.size(); ++i) {
            if (merge[parent_of[i]]) continue;
            kept.code.push_back(cur.code[i]);
            kept.count.push_back(cur.count[i]);
            kept.leaf.push_back(cur.leaf[i]);
        }
        cur = std::move(kept);
        par.leaf = merge;
        ft.levels[Lp] = std::move(par);
    }
    for (int L = 1; L <= md; ++L) {
        const LevelBoxes& lv = ft.levels[L];
        for (size_t i = 0; i < lv.code.size(); ++i)
            if (lv.leaf[i]) ft.maxleaf = std::max(ft.maxleaf, lv.count[i]);
    }
    return ft;
}

}  // namespace

Tree build_tree(const double* xyz, int64_t n, int n_max, int max_depth) {
    if (n <= 0) fail(-2, "positions must be a nonempty finite (N,3) array");
    Tree t;
    t.n_points = n;
    double pmax[3];
    for (int a = 0; a < 3; ++a) {
        t.pmin[a] = xyz[a];
        pmax[a] = xyz[a];
    }
    for (int64_t i = 1; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            t.pmin[a] = std::min(t.pmin[a], xyz[3 * i + a]);
            pmax[a] = std::max(pmax[a], xyz[3 * i + a]);
        }
    t.extent = std::max(pmax[0] - t.pmin[0], std::max(pmax[1] - t.pmin[1], pmax[2] - t.pmin[2]));
    if (!(t.extent > 0)) t.extent = 1.0;   // all points coincide: any root cube holds them

    const bool given = max_depth > 0;
    int md = given ? max_depth
                   : std::max(2, (int)std::ceil(std::log(std::max((double)n / std::max(n_max, 1), 1.0)) /
                                                std::log(8.0)));
    const int depth_cap = 10;
    FixedTree ft;
    while (true) {
        if (md > kMaxLevels - 2) fail(-2, "tree depth exceeds the supported maximum");
        ft = build_fixed(xyz, n, md, n_max, t.pmin, t.extent);
        if (given || ft.maxleaf <= 2 * n_max || md >= depth_cap) break;
        ++md;
    }
    t.depth = md;
    t.maxleaf = ft.maxleaf;
    t.level_start.assign(md + 2, 0);
    for (int L = 1; L <= md; ++L) {
        t.level_start[L] = (int)t.level.size();
        const LevelBoxes& lv = ft.levels[L];
        for (size_t i = 0; i < lv.code.size(); ++i) {
            t.level.push_back(L);
            t.code.push_back(lv.code[i]);
            t.count.push_back(lv.count[i]);
            t.is_leaf.push_back(lv.leaf[i]);
        }
    }
    t.level_start[md + 1] = (int)t.level.size();
    t.level_start[0] = 0;
    const int nb = t.n_boxes();
    t.parent.assign(nb, -1);
    t.pstart.assign(nb, 0);
    for (int b = 0; b < nb; ++b) {
        if (t.level[b] > 1) t.parent[b] = t.find(t.level[b] - 1, t.code[b] >> 3);
        uint64_t first = t.code[b] << (3 * (md - t.level[b]));
        t.pstart[b] = (int)(std::lower_bound(ft.sorted_codes.begin(), ft.sorted_codes.end(), first) -
                            ft.sorted_codes.begin());
    }
    t.perm = std::move(ft.order);
    return t;
}

static inline bool adj_coarse(int b, int c) { return (b <= 2 * c + 2) && (2 * c <= b + 1); }
static inline bool adj_fine(int b, int f) { return (2 * b <= f + 1) && (f <= 2 * b + 2); }

Lists build_lists(const Tree& t) {
    Lists ls;
    const int nb = t.n_boxes();
    for (int b = 0; b < nb; ++b) {
        const int L = t.level[b];
        const uint64_t c = t.code[b];
        int bx, by, bz;
        morton_grid(c, bx, by, bz);
        const bool leaf = t.is_leaf[b];
        const int px = bx >> 1, py = by >> 1, pz = bz >> 1;
        const int64_t psize = 1ll << (L - 1), size = 1ll << L;
        // V, coarser U, X: parent colleagues
        std::vector<std::pair<int, int>> v_here;
        for (int d = 0; d < 26; ++d) {
            const int64_t nx = px + kNeigh[d][0], ny = py + kNeigh[d][1], nz = pz + kNeigh[d][2];
            if (nx < 0 || ny < 0 || nz < 0 || nx >= psize || ny >= psize || nz >= psize) continue;
            const uint64_t nc = morton3(nx, ny, nz);
            for (int ch = 0; ch < 8; ++ch) {
                const uint64_t cc = nc * 8 + ch;
                if (cc == c) continue;
                const int s = t.find(L, cc);
                if (s < 0) continue;
                int cx, cy, cz;
                morton_grid(cc, cx, cy, cz);
                if (std::max(std::abs(cx - bx), std::max(std::abs(cy - by), std::abs(cz - bz))) > 1)
                    v_here.push_back({b, s});
            }
            if (L > 1) {
                const int s = t.find(L - 1, nc);
                if (s >= 0 && t.is_leaf[s]) {
                    const bool adj = adj_coarse(bx, (int)nx) && adj_coarse(by, (int)ny) && adj_coarse(bz, (int)nz);
                    if (adj && leaf) {
                        ls.u_tgt.push_back(b);
                        ls.u_src.push_back(s);
                    }
                    if (!adj) {
                        ls.x_tgt.push_back(b);
                        ls.x_src.push_back(s);
                    }
                }
            }
        }
        for (auto& pr : v_here) {
            ls.v_tgt.push_back(pr.first);
            ls.v_src.push_back(pr.second);
        }
        if (!leaf) continue;
        // same-level U, finer U, W: colleagues
        for (int d = 0; d < 26; ++d) {
            const int64_t mx = bx + kNeigh[d][0], my = by + kNeigh[d][1], mz = bz + kNeigh[d][2];
            if (mx < 0 || my < 0 || mz < 0 || mx >= size || my >= size || mz >= size) continue;
            const uint64_t mc = morton3(mx, my, mz);
            const int s = t.find(L, mc);
            if (s >= 0 && t.is_leaf[s]) {
                ls.u_tgt.push_back(b);
                ls.u_src.push_back(s);
            }
            if (L + 1 > t.depth) continue;
            for (int ch = 0; ch < 8; ++ch) {
                const uint64_t fc = mc * 8 + ch;
                const int f = t.find(L + 1, fc);
                if (f < 0) continue;
                int fx, fy, fz;
                morton_grid(fc, fx, fy, fz);
                const bool adj = adj_fine(bx, fx) && adj_fine(by, fy) && adj_fine(bz, fz);
                if (adj && t.is_leaf[f]) {
                    ls.u_tgt.push_back(b);
                    ls.u_src.push_back(f);
                }
                if (!adj) {
                    ls.w_tgt.push_back(b);
                    ls.w_src.push_back(f);
                }
            }
        }
    }
    auto sort_pairs = [](std::vector<int>& a, std::vector<int>& b) {
        std::vector<size_t> idx(a.size());
        std::iota(idx.begin(), idx.end(), 0);
        std::sort(idx.begin(), idx.end(), [&](size_t i, size_t j) {
            return a[i] != a[j] ? a[i] < a[j] : b[i] < b[j];
        });
        std::vector<int> a2(a.size()), b2(b.size());
        for (size_t i = 0; i < idx.size(); ++i) {
            a2[i] = a[idx[i]];
            b2[i] = b[idx[i]];
        }
        a.swap(a2);
        b.swap(b2);
    };
    sort_pairs(ls.v_tgt, ls.v_src);
    sort_pairs(ls.u_tgt, ls.u_src);
    sort_pairs(ls.w_tgt, ls.w_src);
    sort_pairs(ls.x_tgt, ls.x_src);
    return ls;
}

}  // namespace hfmm
