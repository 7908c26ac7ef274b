// sphk_tree.cpp -- host octree, walk groups and interaction-list builders
// (the roles FDPS plays for the reference; policies in include/sphk_tree.h).
//
// Morton keys: 21 bits per dimension over the cube that bounds all
// particles.  Nodes are stored breadth-first with contiguous children; the
// walk list of each group is produced by an explicit-stack traversal, one
// group per OpenMP iteration, and concatenated in group order so the output
// is independent of the thread count.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <new>
#include <vector>

#include "sphk.h"
#include "sphk_internal.hpp"
#include "sphk_tree.h"

namespace {

struct Box {
    double lo[3], hi[3];
    void empty()
    {
        for (int d = 0; d < 3; d++) {
            lo[d] = std::numeric_limits<double>::infinity();
            hi[d] = -std::numeric_limits<double>::infinity();
        }
    }
    void add(const double p[3], double r)
    {
        for (int d = 0; d < 3; d++) {
            lo[d] = std::min(lo[d], p[d] - r);
            hi[d] = std::max(hi[d], p[d] + r);
        }
    }
    void add(const Box &b)
    {
        for (int d = 0; d < 3; d++) {
            lo[d] = std::min(lo[d], b.lo[d]);
            hi[d] = std::max(hi[d], b.hi[d]);
        }
    }
    bool overlaps(const Box &b) const
    {
        for (int d = 0; d < 3; d++)
            if (b.hi[d] < lo[d] || b.lo[d] > hi[d]) return false;
        return true;
    }
    bool contains(const double p[3]) const
    {
        for (int d = 0; d < 3; d++)
            if (p[d] < lo[d] || p[d] > hi[d]) return false;
        return true;
    }
    double dist2(const double p[3]) const
    {
        double s = 0.;
        for (int d = 0; d < 3; d++) {
            double e = 0.;
            if (p[d] < lo[d]) e = lo[d] - p[d];
            else if (p[d] > hi[d]) e = p[d] - hi[d];
            s += e * e;
        }
        return s;
    }
};

struct Node {
    int32_t begin, end, child, nchild, level;
    Box box;   // tight box of positions
    Box obox;  // box of x_j +- h_j (symmetry lists)
    double len;
};

struct Lists {
    std::vector<int32_t> j_off, j_idx, s_off, s_idx;
};

uint64_t spread3(uint64_t v)
{
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
}

struct Frame {
    double org[3];
    double side;
};

Frame frame_of(int n, const double *x, const double *y, const double *z)
{
    Box b;
    b.empty();
    for (int i = 0; i < n; i++) {
        const double p[3] = {x[i], y[i], z[i]};
        b.add(p, 0.);
    }
    double side = 0.;
    for (int d = 0; d < 3; d++) side = std::max(side, b.hi[d] - b.lo[d]);
    if (!(side > 0.)) side = 1.;
    side *= 1. + 1e-9;
    Frame f;
    for (int d = 0; d < 3; d++) f.org[d] = b.lo[d];
    f.side = side;
    return f;
}

uint64_t key_of(const Frame &f, double x, double y, double z)
{
    const double s = double(1 << 21) / f.side;
    auto q = [&](double v, double o) {
        double t = (v - o) * s;
        if (!(t >= 0.)) t = 0.;
        uint64_t u = (uint64_t)t;
        return u > 0x1fffff ? (uint64_t)0x1fffff : u;
    };
    return spread3(q(x, f.org[0])) << 2 | spread3(q(y, f.org[1])) << 1 | spread3(q(z, f.org[2]));
}

}  // namespace

struct sphk_tree {
    int n = 0;
    std::vector<double> x, y, z;
    std::vector<uint64_t> key;
    std::vector<Node> node;
    std::vector<double> mm, mx, my, mz, me;  // node monopoles
    std::vector<int32_t> group;               // node index per walk (tree groups)
    std::vector<Box> gbox;                    // tight box of each walk's particles
    std::vector<int32_t> i_off;
    Lists lists[3];
    bool built[3] = {false, false, false};
};

using sphk::set_error;

extern "C" {

int sphk_morton_order(int n, const double *x, const double *y, const double *z, int32_t *perm)
{
    if (n < 0 || (n > 0 && (!x || !y || !z || !perm))) return set_error(SPHK_E_ARG, "morton: bad args");
    if (n == 0) return SPHK_OK;
    const Frame f = frame_of(n, x, y, z);
    std::vector<std::pair<uint64_t, int32_t>> kv(n);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) kv[i] = {key_of(f, x[i], y[i], z[i]), i};
    std::sort(kv.begin(), kv.end());
    for (int i = 0; i < n; i++) perm[i] = kv[i].second;
    return SPHK_OK;
}

int sphk_tree_build(int n, const double *x, const double *y, const double *z, const double *m,
                    const double *eps2, int n_leaf, int n_group, sphk_tree **out)
{
    if (!out) return set_error(SPHK_E_ARG, "tree: null output");
    *out = nullptr;
    if (n <= 0 || !x || !y || !z || !m) return set_error(SPHK_E_ARG, "tree: bad particle arrays");
    if (n_leaf < 1 || n_group < 1) return set_error(SPHK_E_ARG, "tree: n_leaf, n_group must be >= 1");
    sphk_tree *t = new (std::nothrow) sphk_tree();
    if (!t) return set_error(SPHK_E_CAPACITY, "tree: out of memory");
    t->n = n;
    t->x.assign(x, x + n);
    t->y.assign(y, y + n);
    t->z.assign(z, z + n);
    const Frame f = frame_of(n, x, y, z);
    t->key.resize(n);
    for (int i = 0; i < n; i++) {
        t->key[i] = key_of(f, x[i], y[i], z[i]);
        if (i && t->key[i] < t->key[i - 1]) {
            delete t;
            return set_error(SPHK_E_ARG, "tree: particles are not in Morton order (use sphk_morton_order)");
        }
    }
    // breadth-first build with contiguous children
    t->node.reserve(2 * (size_t)n / std::max(1, n_leaf) + 16);
    Node root{};
    root.begin = 0;
    root.end = n;
    root.child = -1;
    root.nchild = 0;
    root.level = 0;
    root.len = f.side;
    t->node.push_back(root);
    for (size_t a = 0; a < t->node.size(); a++) {
        Node nd = t->node[a];
        if (nd.end - nd.begin <= n_leaf || nd.level >= 21) continue;
        const int shift = 3 * (20 - nd.level);
        const int32_t first = (int32_t)t->node.size();
        int32_t b = nd.begin;
        int nch = 0;
        for (int o = 0; o < 8 && b < nd.end; o++) {
            auto it = std::partition_point(t->key.begin() + b, t->key.begin() + nd.end,
                                           [&](uint64_t k) { return (int)((k >> shift) & 7) <= o; });
            const int32_t e = (int32_t)(it - t->key.begin());
            if (e > b) {
                Node c{};
                c.begin = b;
                c.end = e;
                c.child = -1;
                c.nchild = 0;
                c.level = nd.level + 1;
                c.len = nd.len * 0.5;
                t->node.push_back(c);
                nch++;
            }
            b = e;
        }
        t->node[a].child = first;
        t->node[a].nchild = nch;
    }
    const size_t nn = t->node.size();
    t->mm.assign(nn, 0.);
    t->mx.assign(nn, 0.);
    t->my.assign(nn, 0.);
    t->mz.assign(nn, 0.);
    t->me.assign(nn, 0.);
    std::vector<double> sx(nn, 0.), sy(nn, 0.), sz(nn, 0.), se(nn, 0.);
    for (size_t a = nn; a-- > 0;) {
        Node &nd = t->node[a];
        nd.box.empty();
        double M = 0., X = 0., Y = 0., Z = 0., E = 0.;
        if (nd.child < 0) {
            for (int i = nd.begin; i < nd.end; i++) {
                const double p[3] = {x[i], y[i], z[i]};
                nd.box.add(p, 0.);
                const double e2 = eps2 ? eps2[i] : 0.;
                M += m[i];
                X += m[i] * x[i];
                Y += m[i] * y[i];
                Z += m[i] * z[i];
                E += m[i] * e2;
            }
        } else {
            for (int c = nd.child; c < nd.child + nd.nchild; c++) {
                nd.box.add(t->node[c].box);
                M += t->mm[c];
                X += sx[c];
                Y += sy[c];
                Z += sz[c];
                E += se[c];
            }
        }
        t->mm[a] = M;
        sx[a] = X;
        sy[a] = Y;
        sz[a] = Z;
        se[a] = E;
        if (M > 0.) {
            t->mx[a] = X / M;
            t->my[a] = Y / M;
            t->mz[a] = Z / M;
            t->me[a] = E / M;
        } else {
            t->mx[a] = 0.5 * (nd.box.lo[0] + nd.box.hi[0]);
            t->my[a] = 0.5 * (nd.box.lo[1] + nd.box.hi[1]);
            t->mz[a] = 0.5 * (nd.box.lo[2] + nd.box.hi[2]);
            t->me[a] = 0.;
        }
    }
    // groups: largest nodes with <= n_group particles, in Morton (DFS) order
    std::vector<int32_t> stack{0};
    while (!stack.empty()) {
        const int32_t a = stack.back();
        stack.pop_back();
        const Node &nd = t->node[a];
        if (nd.end - nd.begin <= n_group || nd.child < 0) {
            t->group.push_back(a);
        } else {
            for (int c = nd.child + nd.nchild - 1; c >= nd.child; c--) stack.push_back(c);
        }
    }
    t->i_off.resize(t->group.size() + 1);
    t->gbox.resize(t->group.size());
    for (size_t g = 0; g < t->group.size(); g++) {
        t->i_off[g] = t->node[t->group[g]].begin;
        t->gbox[g] = t->node[t->group[g]].box;
    }
    t->i_off[t->group.size()] = n;
    *out = t;
    return SPHK_OK;
}

int sphk_tree_group_fixed(sphk_tree *t, int size)
{
    if (!t || size < 1) return set_error(SPHK_E_ARG, "group_fixed: bad argument");
    const int n = t->n;
    const int ng = (n + size - 1) / size;
    t->group.assign(ng, -1);
    t->gbox.resize(ng);
    t->i_off.resize(ng + 1);
    for (int g = 0; g < ng; g++) {
        t->i_off[g] = g * size;
        Box b;
        b.empty();
        for (int i = g * size; i < std::min(n, (g + 1) * size); i++) {
            const double p[3] = {t->x[i], t->y[i], t->z[i]};
            b.add(p, 0.);
        }
        t->gbox[g] = b;
    }
    t->i_off[ng] = n;
    for (auto &b : t->built) b = false;
    return SPHK_OK;
}

void sphk_tree_free(sphk_tree *t) { delete t; }

int sphk_tree_n_walk(const sphk_tree *t) { return t ? (int)t->group.size() : -1; }
int sphk_tree_n_node(const sphk_tree *t) { return t ? (int)t->node.size() : -1; }
const int32_t *sphk_tree_i_off(const sphk_tree *t) { return t ? t->i_off.data() : nullptr; }

}  // extern "C"

namespace {

template <typename Body>
int build_lists(sphk_tree *t, Lists &L, bool with_s, Body body)
{
    const int ng = (int)t->group.size();
    std::vector<std::vector<int32_t>> jl(ng), sl(with_s ? ng : 0);
#pragma omp parallel for schedule(dynamic, 4)
    for (int g = 0; g < ng; g++) body(g, jl[g], with_s ? &sl[g] : nullptr);
    L.j_off.assign(ng + 1, 0);
    int64_t s = 0;
    for (int g = 0; g < ng; g++) {
        L.j_off[g] = (int32_t)s;
        s += (int64_t)jl[g].size();
    }
    if (s > std::numeric_limits<int32_t>::max()) return set_error(SPHK_E_CAPACITY, "list too long");
    L.j_off[ng] = (int32_t)s;
    L.j_idx.resize(s);
#pragma omp parallel for schedule(static)
    for (int g = 0; g < ng; g++)
        if (!jl[g].empty()) std::memcpy(&L.j_idx[L.j_off[g]], jl[g].data(), jl[g].size() * 4);
    L.s_off.clear();
    L.s_idx.clear();
    if (with_s) {
        L.s_off.assign(ng + 1, 0);
        s = 0;
        for (int g = 0; g < ng; g++) {
            L.s_off[g] = (int32_t)s;
            s += (int64_t)sl[g].size();
        }
        if (s > std::numeric_limits<int32_t>::max()) return set_error(SPHK_E_CAPACITY, "list too long");
        L.s_off[ng] = (int32_t)s;
        L.s_idx.resize(s);
#pragma omp parallel for schedule(static)
        for (int g = 0; g < ng; g++)
            if (!sl[g].empty()) std::memcpy(&L.s_idx[L.s_off[g]], sl[g].data(), sl[g].size() * 4);
    }
    return SPHK_OK;
}

}  // namespace

extern "C" {

int sphk_tree_list_gather(sphk_tree *t, const double *rs)
{
    if (!t || !rs) return set_error(SPHK_E_ARG, "gather: null argument");
    int rc = build_lists(t, t->lists[0], false, [&](int g, std::vector<int32_t> &out, std::vector<int32_t> *) {
        Box ob;
        ob.empty();
        for (int i = t->i_off[g]; i < t->i_off[g + 1]; i++) {
            const double p[3] = {t->x[i], t->y[i], t->z[i]};
            ob.add(p, rs[i]);
        }
        std::vector<int32_t> st{0};
        while (!st.empty()) {
            const Node &nd = t->node[st.back()];
            st.pop_back();
            if (!nd.box.overlaps(ob)) continue;
            if (nd.child < 0) {
                for (int j = nd.begin; j < nd.end; j++) {
                    const double p[3] = {t->x[j], t->y[j], t->z[j]};
                    if (!ob.contains(p)) continue;
                    // exact membership: j lies in the search sphere of some i
                    for (int i = t->i_off[g]; i < t->i_off[g + 1]; i++) {
                        const double dx = p[0] - t->x[i], dy = p[1] - t->y[i], dz = p[2] - t->z[i];
                        if (dx * dx + dy * dy + dz * dz < rs[i] * rs[i]) {
                            out.push_back(j);
                            break;
                        }
                    }
                }
            } else {
                for (int c = nd.child + nd.nchild - 1; c >= nd.child; c--) st.push_back(c);
            }
        }
    });
    t->built[0] = rc == SPHK_OK;
    return rc;
}

int sphk_tree_list_symmetry(sphk_tree *t, const double *h)
{
    if (!t || !h) return set_error(SPHK_E_ARG, "symmetry: null argument");
    // per-node box of x_j +- h_j, bottom-up
    for (size_t a = t->node.size(); a-- > 0;) {
        Node &nd = t->node[a];
        nd.obox.empty();
        if (nd.child < 0) {
            for (int j = nd.begin; j < nd.end; j++) {
                const double p[3] = {t->x[j], t->y[j], t->z[j]};
                nd.obox.add(p, h[j]);
            }
        } else {
            for (int c = nd.child; c < nd.child + nd.nchild; c++) nd.obox.add(t->node[c].obox);
        }
    }
    int rc = build_lists(t, t->lists[1], false, [&](int g, std::vector<int32_t> &out, std::vector<int32_t> *) {
        Box ob;
        ob.empty();
        for (int i = t->i_off[g]; i < t->i_off[g + 1]; i++) {
            const double p[3] = {t->x[i], t->y[i], t->z[i]};
            ob.add(p, h[i]);
        }
        const Box &ib = t->gbox[g];
        std::vector<int32_t> st{0};
        while (!st.empty()) {
            const Node &nd = t->node[st.back()];
            st.pop_back();
            if (!nd.box.overlaps(ob) && !nd.obox.overlaps(ib)) continue;
            if (nd.child < 0) {
                for (int j = nd.begin; j < nd.end; j++) {
                    const double p[3] = {t->x[j], t->y[j], t->z[j]};
                    Box jb;
                    jb.empty();
                    jb.add(p, h[j]);
                    if (!ob.contains(p) && !jb.overlaps(ib)) continue;
                    // exact membership: r_ij < max(H_i, H_j) for some i
                    for (int i = t->i_off[g]; i < t->i_off[g + 1]; i++) {
                        const double dx = p[0] - t->x[i], dy = p[1] - t->y[i], dz = p[2] - t->z[i];
                        const double hm = h[i] > h[j] ? h[i] : h[j];
                        if (dx * dx + dy * dy + dz * dz < hm * hm) {
                            out.push_back(j);
                            break;
                        }
                    }
                }
            } else {
                for (int c = nd.child + nd.nchild - 1; c >= nd.child; c--) st.push_back(c);
            }
        }
    });
    t->built[1] = rc == SPHK_OK;
    return rc;
}

int sphk_tree_list_long(sphk_tree *t, double theta)
{
    if (!t || !(theta >= 0.)) return set_error(SPHK_E_ARG, "long: bad argument");
    const double inv_t2 = theta > 0. ? 1. / (theta * theta) : std::numeric_limits<double>::infinity();
    int rc = build_lists(t, t->lists[2], true, [&](int g, std::vector<int32_t> &out, std::vector<int32_t> *sp) {
        const Box &ib = t->gbox[g];
        std::vector<int32_t> st{0};
        while (!st.empty()) {
            const int32_t a = st.back();
            st.pop_back();
            const Node &nd = t->node[a];
            if (t->mm[a] == 0.) continue;  // massless node
            const double c[3] = {t->mx[a], t->my[a], t->mz[a]};
            const double d2 = ib.dist2(c);
            if (d2 > nd.len * nd.len * inv_t2) {
                sp->push_back(a);
            } else if (nd.child < 0) {
                for (int j = nd.begin; j < nd.end; j++) out.push_back(j);
            } else {
                for (int cc = nd.child + nd.nchild - 1; cc >= nd.child; cc--) st.push_back(cc);
            }
        }
    });
    t->built[2] = rc == SPHK_OK;
    return rc;
}

int sphk_tree_list(const sphk_tree *t, int kind, const int32_t **j_off, const int32_t **j_idx,
                   const int32_t **s_off, const int32_t **s_idx)
{
    if (!t || kind < 0 || kind > 2) return set_error(SPHK_E_ARG, "list: bad argument");
    if (!t->built[kind]) return set_error(SPHK_E_STATE, "list kind %d not built", kind);
    const Lists &L = t->lists[kind];
    if (j_off) *j_off = L.j_off.data();
    if (j_idx) *j_idx = L.j_idx.data();
    if (s_off) *s_off = kind == 2 ? L.s_off.data() : nullptr;
    if (s_idx) *s_idx = kind == 2 ? L.s_idx.data() : nullptr;
    return SPHK_OK;
}

int sphk_tree_moments(const sphk_tree *t, const double **m, const double **x, const double **y,
                      const double **z, const double **eps2)
{
    if (!t) return set_error(SPHK_E_ARG, "moments: null tree");
    if (m) *m = t->mm.data();
    if (x) *x = t->mx.data();
    if (y) *y = t->my.data();
    if (z) *z = t->mz.data();
    if (eps2) *eps2 = t->me.data();
    return SPHK_OK;
}

}  // extern "C"
