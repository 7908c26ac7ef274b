// sphk_dtree.hip -- device-resident octree, walk groups and interaction lists.
//
// The host builder (sphk_tree.cpp) and this one implement the same
// policies; this one keeps everything in HBM:
//   1. Morton keys over the bounding cube (block reduction, one key per
//      thread); sphk_morton_order_dev sorts (key, index) with a radix sort.
//   2. The octree is built breadth-first, one kernel pair per level: count
//      the non-empty child octants of every frontier node (binary searches
//      in the sorted keys), exclusive scan, write the children.  Children of
//      one level are numbered in parent order and octant order, as the
//      host's breadth-first loop numbers them.
//   3. Node boxes and monopoles bottom-up, one kernel per level, each node
//      summing its particles (leaf) or children in index order with FP
//      contraction off, so the FP64 sums round as on the host.
//   4. Groups: a node is a walk if it holds <= n_group particles (or is a
//      leaf) and its parent holds more; walks are ordered by first particle.
//   5. Lists: the host's depth-first traversal of each walk, 8 lanes per
//      walk (the children of one opened node at a time, a stack of pending
//      entries in LDS), run twice: count, then fill at scanned offsets; the
//      output is the host's, element for element.  The short-range lists test
//      each reached leaf's particles for exact membership on the way.
#include <hip/hip_runtime.h>

#include <cstring>   // rocPRIM's iterator headers use memset without including it

#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <cstdint>
#include <limits>
#include <mutex>
#include <new>
#include <vector>

#include "sphk.h"
#include "sphk_adjoint.h"
#include "sphk_dtree.h"
#include "sphk_internal.hpp"

using sphk::set_error;

namespace {

constexpr int kMaxLevel = 21;
constexpr int kBlock = 256;

#define SPHK_TRY(expr)                                                                          \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) return set_error(SPHK_E_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// Device array with capacity; grow() keeps the contents.
template <typename T>
struct DArr {
    T *p = nullptr;
    size_t cap = 0;
    hipError_t reserve(size_t n, bool keep, hipStream_t st)
    {
        if (n <= cap) return hipSuccess;
        T *q = nullptr;
        hipError_t e = hipMalloc(&q, n * sizeof(T));
        if (e != hipSuccess) return e;
        if (keep && p && cap) e = hipMemcpyAsync(q, p, cap * sizeof(T), hipMemcpyDeviceToDevice, st);
        // queued work may still read the old block: wait before freeing it
        if (e == hipSuccess && p) e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            (void)hipFree(q);
            return e;
        }
        if (p) (void)hipFree(p);
        p = q;
        cap = n;
        return hipSuccess;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

struct DLists {
    DArr<int32_t> j_off, j_idx, s_off, s_idx;
    int64_t n_j = 0, n_s = 0;
    bool built = false;
    // built by one walk and not yet packed: the entries are still in the
    // rooms of the handle's scratch (walk g's from g * room_j in mscr, from
    // g * room_s in sscr for the long list's nodes); j_off / s_off are final
    bool in_rooms = false;
    int64_t room_j = 0, room_s = 0;
};

__device__ __forceinline__ uint64_t spread3(uint64_t v)
{
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
}

// ---- bounding cube and keys ------------------------------------------------
__global__ __launch_bounds__(kBlock) void bbox_partial_kernel(int n, const double *__restrict__ x,
                                                              const double *__restrict__ y,
                                                              const double *__restrict__ z,
                                                              double *__restrict__ part)
{
    __shared__ double s[6][kBlock];
    const double inf = __builtin_inf();
    double lo[3] = {inf, inf, inf}, hi[3] = {-inf, -inf, -inf};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const double p[3] = {x[i], y[i], z[i]};
        for (int d = 0; d < 3; d++) {
            lo[d] = fmin(lo[d], p[d]);
            hi[d] = fmax(hi[d], p[d]);
        }
    }
    for (int d = 0; d < 3; d++) {
        s[d][threadIdx.x] = lo[d];
        s[3 + d][threadIdx.x] = hi[d];
    }
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int d = 0; d < 3; d++) {
                s[d][threadIdx.x] = fmin(s[d][threadIdx.x], s[d][threadIdx.x + w]);
                s[3 + d][threadIdx.x] = fmax(s[3 + d][threadIdx.x], s[3 + d][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int d = 0; d < 6; d++) part[6 * blockIdx.x + d] = s[d][0];
}

// frame = {origin x, y, z, side}: the bounding cube, side enlarged by 1e-9
// (the host's frame_of).
// One block combines the partial boxes (min and max are exact: any order
// gives the same box).
__global__ __launch_bounds__(kBlock) void bbox_final_kernel(int nb, const double *__restrict__ part,
                                                            double *__restrict__ frame)
{
    __shared__ double s[6][kBlock];
    const double inf = __builtin_inf();
    double lo[3] = {inf, inf, inf}, hi[3] = {-inf, -inf, -inf};
    for (int b = threadIdx.x; b < nb; b += kBlock)
        for (int d = 0; d < 3; d++) {
            lo[d] = fmin(lo[d], part[6 * b + d]);
            hi[d] = fmax(hi[d], part[6 * b + 3 + d]);
        }
    for (int d = 0; d < 3; d++) {
        s[d][threadIdx.x] = lo[d];
        s[3 + d][threadIdx.x] = hi[d];
    }
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int d = 0; d < 3; d++) {
                s[d][threadIdx.x] = fmin(s[d][threadIdx.x], s[d][threadIdx.x + w]);
                s[3 + d][threadIdx.x] = fmax(s[3 + d][threadIdx.x], s[3 + d][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    for (int d = 0; d < 3; d++) {
        lo[d] = s[d][0];
        hi[d] = s[3 + d][0];
    }
    double side = 0.;
    for (int d = 0; d < 3; d++) side = fmax(side, hi[d] - lo[d]);
    if (!(side > 0.)) side = 1.;
    side *= 1. + 1e-9;
    frame[0] = lo[0];
    frame[1] = lo[1];
    frame[2] = lo[2];
    frame[3] = side;
}

__device__ __forceinline__ uint64_t quant(double v, double o, double s)
{
    double t = (v - o) * s;
    if (!(t >= 0.)) t = 0.;
    const uint64_t u = (uint64_t)t;
    return u > 0x1fffff ? (uint64_t)0x1fffff : u;
}

__global__ __launch_bounds__(kBlock) void keys_kernel(int n, const double *__restrict__ x,
                                                      const double *__restrict__ y,
                                                      const double *__restrict__ z,
                                                      const double *__restrict__ frame,
                                                      uint64_t *__restrict__ key, int32_t *__restrict__ idx)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double s = double(1 << 21) / frame[3];
    key[i] = spread3(quant(x[i], frame[0], s)) << 2 | spread3(quant(y[i], frame[1], s)) << 1 |
             spread3(quant(z[i], frame[2], s));
    if (idx) idx[i] = i;
}

__global__ __launch_bounds__(kBlock) void unsorted_kernel(int n, const uint64_t *__restrict__ key,
                                                          int *__restrict__ flag)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= 1 && i < n && key[i] < key[i - 1]) atomicOr(flag, 1);
}

// Sets up the frame and keys of n particles (queued on st).
int make_keys(int n, const double *x, const double *y, const double *z, hipStream_t st, DArr<double> &part,
              DArr<double> &frame, DArr<uint64_t> &key, int32_t *idx)
{
    const int nb = std::min(1024, (n + kBlock - 1) / kBlock);
    SPHK_TRY(part.reserve(6 * (size_t)nb, false, st));
    SPHK_TRY(frame.reserve(4, false, st));
    SPHK_TRY(key.reserve(n, false, st));
    hipLaunchKernelGGL(bbox_partial_kernel, dim3(nb), dim3(kBlock), 0, st, n, x, y, z, part.p);
    hipLaunchKernelGGL(bbox_final_kernel, dim3(1), dim3(kBlock), 0, st, nb, part.p, frame.p);
    hipLaunchKernelGGL(keys_kernel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, n, x, y, z,
                       frame.p, key.p, idx);
    SPHK_TRY(hipGetLastError());
    return SPHK_OK;
}

// ---- octree levels ---------------------------------------------------------
struct NodeArrays {
    int32_t *begin, *end, *child, *nchild, *level, *parent;
    double *len;
};

// First index in [b, e) whose octant digit exceeds o (keys sorted).
__device__ __forceinline__ int32_t octant_end(const uint64_t *__restrict__ key, int32_t b, int32_t e, int shift,
                                              int o)
{
    while (b < e) {
        const int32_t mid = b + ((e - b) >> 1);
        if ((int)((key[mid] >> shift) & 7) <= o)
            b = mid + 1;
        else
            e = mid;
    }
    return b;
}

template <bool WRITE>
__global__ __launch_bounds__(kBlock) void split_kernel(int fa, int nf, const uint64_t *__restrict__ key,
                                                       NodeArrays nd, int n_leaf, int64_t *__restrict__ cnt,
                                                       const int64_t *__restrict__ off, int base)
{
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= nf) return;
    const int a = fa + k;
    const int32_t nb = nd.begin[a], ne = nd.end[a], lv = nd.level[a];
    if (ne - nb <= n_leaf || lv >= kMaxLevel) {
        if (!WRITE) cnt[k] = 0;
        return;
    }
    const int shift = 3 * (20 - lv);
    int32_t b = nb;
    int nch = 0;
    const int first = WRITE ? base + (int)off[k] : 0;
    for (int o = 0; o < 8 && b < ne; o++) {
        const int32_t e = octant_end(key, b, ne, shift, o);
        if (e > b) {
            if (WRITE) {
                const int c = first + nch;
                nd.begin[c] = b;
                nd.end[c] = e;
                nd.child[c] = -1;
                nd.nchild[c] = 0;
                nd.level[c] = lv + 1;
                nd.parent[c] = a;
                nd.len[c] = nd.len[a] * 0.5;
            }
            nch++;
        }
        b = e;
    }
    if (WRITE) {
        nd.child[a] = first;
        nd.nchild[a] = nch;
    } else {
        cnt[k] = nch;
    }
}

__global__ void root_kernel(int n, NodeArrays nd, const double *__restrict__ frame)
{
    nd.begin[0] = 0;
    nd.end[0] = n;
    nd.child[0] = -1;
    nd.nchild[0] = 0;
    nd.level[0] = 0;
    nd.parent[0] = -1;
    nd.len[0] = frame[3];
}

// ---- boxes and monopoles, one level at a time (bottom-up) ------------------
struct Moments {
    double *lo, *hi;        // [3][cap] tight box
    double *mm, *mx, *my, *mz, *me;
    double *sx, *sy, *sz, *se;   // mass-weighted sums
    size_t cap;
};

__global__ __launch_bounds__(kBlock) void moments_kernel(int a0, int na, NodeArrays nd, Moments mo,
                                                         const double *__restrict__ x,
                                                         const double *__restrict__ y,
                                                         const double *__restrict__ z,
                                                         const double *__restrict__ m,
                                                         const double *__restrict__ eps2)
{
#pragma clang fp contract(off)
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= na) return;
    const int a = a0 + k;
    const size_t C = mo.cap;
    const double inf = __builtin_inf();
    double lo[3] = {inf, inf, inf}, hi[3] = {-inf, -inf, -inf};
    double M = 0., X = 0., Y = 0., Z = 0., E = 0.;
    if (nd.child[a] < 0) {
        for (int i = nd.begin[a]; i < nd.end[a]; i++) {
            const double p[3] = {x[i], y[i], z[i]};
            for (int d = 0; d < 3; d++) {
                lo[d] = fmin(lo[d], p[d]);
                hi[d] = fmax(hi[d], p[d]);
            }
            const double e2 = eps2 ? eps2[i] : 0.;
            M += m[i];
            X += m[i] * x[i];
            Y += m[i] * y[i];
            Z += m[i] * z[i];
            E += m[i] * e2;
        }
    } else {
        for (int c = nd.child[a]; c < nd.child[a] + nd.nchild[a]; c++) {
            for (int d = 0; d < 3; d++) {
                lo[d] = fmin(lo[d], mo.lo[d * C + c]);
                hi[d] = fmax(hi[d], mo.hi[d * C + c]);
            }
            M += mo.mm[c];
            X += mo.sx[c];
            Y += mo.sy[c];
            Z += mo.sz[c];
            E += mo.se[c];
        }
    }
    for (int d = 0; d < 3; d++) {
        mo.lo[d * C + a] = lo[d];
        mo.hi[d * C + a] = hi[d];
    }
    mo.mm[a] = M;
    mo.sx[a] = X;
    mo.sy[a] = Y;
    mo.sz[a] = Z;
    mo.se[a] = E;
    if (M > 0.) {
        mo.mx[a] = X / M;
        mo.my[a] = Y / M;
        mo.mz[a] = Z / M;
        mo.me[a] = E / M;
    } else {
        mo.mx[a] = 0.5 * (lo[0] + hi[0]);
        mo.my[a] = 0.5 * (lo[1] + hi[1]);
        mo.mz[a] = 0.5 * (lo[2] + hi[2]);
        mo.me[a] = 0.;
    }
}

// Box of x_j +- h_j per node (symmetry lists), bottom-up.
__global__ __launch_bounds__(kBlock) void obox_kernel(int a0, int na, NodeArrays nd, double *__restrict__ olo,
                                                      double *__restrict__ ohi, size_t C,
                                                      const double *__restrict__ x,
                                                      const double *__restrict__ y,
                                                      const double *__restrict__ z,
                                                      const double *__restrict__ h)
{
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= na) return;
    const int a = a0 + k;
    const double inf = __builtin_inf();
    double lo[3] = {inf, inf, inf}, hi[3] = {-inf, -inf, -inf};
    if (nd.child[a] < 0) {
        for (int i = nd.begin[a]; i < nd.end[a]; i++) {
            const double p[3] = {x[i], y[i], z[i]};
            for (int d = 0; d < 3; d++) {
                lo[d] = fmin(lo[d], p[d] - h[i]);
                hi[d] = fmax(hi[d], p[d] + h[i]);
            }
        }
    } else {
        for (int c = nd.child[a]; c < nd.child[a] + nd.nchild[a]; c++)
            for (int d = 0; d < 3; d++) {
                lo[d] = fmin(lo[d], olo[d * C + c]);
                hi[d] = fmax(hi[d], ohi[d * C + c]);
            }
    }
    for (int d = 0; d < 3; d++) {
        olo[d * C + a] = lo[d];
        ohi[d * C + a] = hi[d];
    }
}

// ---- groups ------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void fill_kernel(int n, int32_t *__restrict__ a, int32_t v)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ __launch_bounds__(kBlock) void group_mark_kernel(int nn, NodeArrays nd, int n_group,
                                                            int32_t *__restrict__ head)
{
    const int a = blockIdx.x * kBlock + threadIdx.x;
    if (a >= nn) return;
    const bool small = nd.end[a] - nd.begin[a] <= n_group || nd.child[a] < 0;
    const int p = nd.parent[a];
    const bool opened = p < 0 || nd.end[p] - nd.begin[p] > n_group;
    if (small && opened) head[nd.begin[a]] = a;
}

__global__ __launch_bounds__(kBlock) void head_flag_kernel(int n, const int32_t *__restrict__ head,
                                                           int64_t *__restrict__ flag)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) flag[i] = head[i] >= 0 ? 1 : 0;
}

__global__ __launch_bounds__(kBlock) void group_write_kernel(int n, const int32_t *__restrict__ head,
                                                             const int64_t *__restrict__ pos,
                                                             int32_t *__restrict__ group,
                                                             int32_t *__restrict__ i_off, int ng)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    if (head[i] >= 0) {
        group[pos[i]] = head[i];
        i_off[pos[i]] = i;
    }
    if (i == 0) i_off[ng] = n;
}

// ---- lists -------------------------------------------------------------------
// Node records for the list walks: the fields one visit reads, packed in
// one 64-byte record per node (one memory segment per visit instead of one
// per field array).  Values are copies of the node arrays, so the walks'
// comparisons and their results are unchanged.
struct alignas(64) BoxRec {     // gather and symmetry walks
    double lo[3], hi[3];        // tight box
    int32_t child, nchild, begin, end;
};
struct alignas(64) MacRec {     // long walk
    double mx, my, mz, mm, len;
    int32_t child, nchild, begin, end;
};
struct alignas(64) OboxRec {    // symmetry walk: box of the nodes' particle spheres
    double lo[3], hi[3];
};

__global__ __launch_bounds__(kBlock) void pack_nodes_kernel(int nn, NodeArrays nd, Moments mo,
                                                            BoxRec *__restrict__ br, MacRec *__restrict__ mr)
{
    const int a = blockIdx.x * kBlock + threadIdx.x;
    if (a >= nn) return;
    const size_t C = mo.cap;
    BoxRec b;
    for (int d = 0; d < 3; d++) {
        b.lo[d] = mo.lo[d * C + a];
        b.hi[d] = mo.hi[d * C + a];
    }
    b.child = nd.child[a];
    b.nchild = nd.nchild[a];
    b.begin = nd.begin[a];
    b.end = nd.end[a];
    br[a] = b;
    mr[a] = MacRec{mo.mx[a], mo.my[a], mo.mz[a], mo.mm[a], nd.len[a], b.child, b.nchild, b.begin, b.end};
}

__global__ __launch_bounds__(kBlock) void pack_obox_kernel(int nn, const double *__restrict__ olo,
                                                           const double *__restrict__ ohi, size_t C,
                                                           OboxRec *__restrict__ orr)
{
    const int a = blockIdx.x * kBlock + threadIdx.x;
    if (a >= nn) return;
    OboxRec o;
    for (int d = 0; d < 3; d++) {
        o.lo[d] = olo[d * C + a];
        o.hi[d] = ohi[d * C + a];
    }
    orr[a] = o;
}

struct TreeView {
    NodeArrays nd;
    Moments mo;
    const BoxRec *br;
    const MacRec *mr;
    const OboxRec *orr;
    const double *x, *y, *z;
    const int32_t *group, *i_off;
};

__device__ __forceinline__ bool box_overlaps(const double (&alo)[3], const double (&ahi)[3],
                                             const double (&blo)[3], const double (&bhi)[3])
{
    for (int d = 0; d < 3; d++)
        if (bhi[d] < alo[d] || blo[d] > ahi[d]) return false;
    return true;
}

// Box of the i search spheres of each walk (gather: rs, symmetry: h).
__global__ __launch_bounds__(kBlock) void walk_box_kernel(int ng, const int32_t *__restrict__ i_off,
                                                          const double *__restrict__ x,
                                                          const double *__restrict__ y,
                                                          const double *__restrict__ z,
                                                          const double *__restrict__ r, double *__restrict__ wb)
{
    const int g = blockIdx.x * kBlock + threadIdx.x;
    if (g >= ng) return;
    const double inf = __builtin_inf();
    double lo[3] = {inf, inf, inf}, hi[3] = {-inf, -inf, -inf};
    for (int i = i_off[g]; i < i_off[g + 1]; i++) {
        const double p[3] = {x[i], y[i], z[i]};
        for (int d = 0; d < 3; d++) {
            lo[d] = fmin(lo[d], p[d] - r[i]);
            hi[d] = fmax(hi[d], p[d] + r[i]);
        }
    }
    for (int d = 0; d < 3; d++) {
        wb[d * ng + g] = lo[d];
        wb[(3 + d) * ng + g] = hi[d];
    }
}

// What the depth-first walk does at node a: skip it (pruned, or massless for
// the long list), take it as one SPJ, emit its particles (leaf), or open it.
enum Visit { kSkip, kSpj, kLeaf, kOpen };

// Child range and particle range of the visited node, from its record.
struct NodeLinks {
    int32_t child, nchild, begin, end;
};

template <int MODE>
__device__ __forceinline__ Visit visit(const TreeView &t, int a, const double (&olo)[3],
                                       const double (&ohi)[3], const double (&glo)[3], const double (&ghi)[3],
                                       double inv_t2, NodeLinks &ln)
{
#pragma clang fp contract(off)
    if (MODE != 2) {
        const BoxRec b = t.br[a];
        OboxRec o{};
        if (MODE == 1) o = t.orr[a];   // loaded with b: one memory round trip per visit
        ln = NodeLinks{b.child, b.nchild, b.begin, b.end};
        const bool leaf = b.child < 0;
        if (MODE == 0) {
            if (!box_overlaps(olo, ohi, b.lo, b.hi)) return kSkip;
            return leaf ? kLeaf : kOpen;
        }
        if (!box_overlaps(olo, ohi, b.lo, b.hi)) {
            if (!box_overlaps(glo, ghi, o.lo, o.hi)) return kSkip;
        }
        return leaf ? kLeaf : kOpen;
    }
    const MacRec m = t.mr[a];
    ln = NodeLinks{m.child, m.nchild, m.begin, m.end};
    if (m.mm == 0.) return kSkip;
    const double c[3] = {m.mx, m.my, m.mz};
    double s = 0.;
    for (int d = 0; d < 3; d++) {
        double e = 0.;
        if (c[d] < glo[d])
            e = glo[d] - c[d];
        else if (c[d] > ghi[d])
            e = c[d] - ghi[d];
        s += e * e;
    }
    const double L = m.len;
    if (s > L * L * inv_t2) return kSpj;
    return m.child < 0 ? kLeaf : kOpen;
}

// ---- interaction lists: 8 lanes per walk -----------------------------------
// The host builder's depth-first walk of each group, 8 walks per wave and 8
// lanes per walk (an octree node has at most 8 children).  A walk keeps a
// stack of evaluated entries in LDS, the next one in preorder on top.  Each
// step evaluates the children of one opened node, a lane each: the finished
// children (SPJ, leaf) before the first child to open are emitted at once,
// the rest are pushed, first child on top.  Then the finished entries on top
// of the stack are popped and emitted, up to 8 at a time, down to the next
// node to open, whose children the next step evaluates.  Emitted in this
// order, the entries are the host's list element for element.  Each tree
// level leaves at most 7 siblings pending, so the stack never holds more
// than 7 (kMaxLevel + 1) + 8 entries (kListStack); walks run first with a
// stack of kListStackSmall entries (less LDS, more waves in flight) and,
// if one would overflow it, again with kListStack.
//
// Entry: word 0 the node (SPJ, and leaves in candidate-leaf mode), the first
// child (to open) or the first particle (leaf, particles listed); word 1 the
// type in bits 29-31 and the child or particle count below.
constexpr int kListWalks = 8;
constexpr int kListStack = 7 * (kMaxLevel + 1) + 8 + 2;   // 164 (even: 8-byte rows)
constexpr int kListStackSmall = 64;
constexpr uint32_t kTySpj = 1u << 29, kTyLeaf = 2u << 29, kTyOpen = 3u << 29, kTyMask = 7u << 29;
constexpr uint32_t kCountMask = (1u << 29) - 1;

// bits of the 8 lanes of this lane's walk
__device__ __forceinline__ unsigned walk_bits(bool p, int lane)
{
    return (unsigned)((__ballot(p) >> (lane & ~7)) & 0xffu);
}

__device__ __forceinline__ int walk_scan_incl(int v, int c)
{
    for (int d = 1; d < 8; d <<= 1) {
        const int u = __shfl_up(v, d, 8);
        if (c >= d) v += u;
    }
    return v;
}

// Exact membership of particle j of a reached leaf in a short-range list,
// as the host's leaf loop decides it: MODE 0 |x_j - x_i| < rs_i for some i of
// the walk; MODE 1 r_ij < max(H_i, H_j) for some i, after the box pre-tests
// (j inside the walk's search box, or j's sphere reaching the group's box).
// j's position is p (and its H hj in MODE 1).  The walk's i-particles
// (x, y, z, r) are read from si, the first nl of them staged in LDS.
constexpr int kWalkI = 256;

template <int MODE>
__device__ __forceinline__ bool member(const TreeView &t, const double *__restrict__ r, const double (&p)[3],
                                       double hj, int ib, int ni, int nl, const double4 *si,
                                       const double (&olo)[3], const double (&ohi)[3], const double (&glo)[3],
                                       const double (&ghi)[3])
{
#pragma clang fp contract(off)
    bool in = true, jov = MODE == 1;
    for (int d = 0; d < 3; d++) {
        in = in && !(p[d] < olo[d] || p[d] > ohi[d]);
        if (MODE == 1) jov = jov && !(ghi[d] < p[d] - hj || glo[d] > p[d] + hj);
    }
    if (!in && !jov) return false;
    auto hit = [&](double4 P) {
        const double dx = p[0] - P.x, dy = p[1] - P.y, dz = p[2] - P.z;
        const double rr = MODE == 1 ? (P.w > hj ? P.w : hj) : P.w;
        return dx * dx + dy * dy + dz * dz < rr * rr;
    };
    int q = 0;
    for (; q + 2 <= nl; q += 2) {   // LDS, two at a time
        const double4 P0 = si[q], P1 = si[q + 1];
        if (hit(P0) || hit(P1)) return true;
    }
    if (q < nl && hit(si[q])) return true;
    for (q = nl; q < ni; q++) {   // beyond the staged particles: from memory
        const int i = ib + q;
        if (hit(make_double4(t.x[i], t.y[i], t.z[i], r[i]))) return true;
    }
    return false;
}

// The particles of the leaves reached into jout (MODE 0 gather, 1 symmetry:
// the candidates of the short-range lists), and for MODE 2 (long) the SPJ
// nodes into sout.  FILL 0: the walk's counts go to cj/cs.  FILL 1: written
// from the walk's offsets oj/os.  FILL 2: written into rooms of `room`
// (jout) and `room_s` (sout) entries per walk, walk g's from g * room and
// g * room_s; short lists: the count to cj and the room's offset to cs;
// long list: the counts to cj and cs.  Bit 2 (4) of err is set if a walk
// does not fit its jout (sout) room: entries past it are not written and
// its count is the room's.
// Bit 1 of err is set if a stack of STACK entries would overflow.
template <int MODE, int FILL, int STACK>
__global__ __launch_bounds__(64) void walk_list_kernel(int ng, TreeView t, const double *__restrict__ wb,
                                                       const double *__restrict__ r, double inv_t2,
                                                       int64_t *__restrict__ cj, int64_t *__restrict__ cs,
                                                       const int64_t *__restrict__ oj, const int64_t *__restrict__ os,
                                                       int32_t *__restrict__ jout, int32_t *__restrict__ sout,
                                                       int *__restrict__ err, int64_t room = 0, int64_t room_s = 0)
{
#pragma clang fp contract(off)
    __shared__ uint2 stk[kListWalks][STACK];
    const int lane = threadIdx.x, sub = lane >> 3, c = lane & 7;
    const int g = blockIdx.x * kListWalks + sub;
    const bool live = g < ng;
    const unsigned below = (1u << c) - 1u, above = 0xffu & ~((2u << c) - 1u);
    const size_t C = t.mo.cap;
    double olo[3] = {0., 0., 0.}, ohi[3] = {0., 0., 0.}, glo[3] = {0., 0., 0.}, ghi[3] = {0., 0., 0.};
    int64_t nj = 0, ns = 0, j0 = 0, s0 = 0;
    if (live) {
        const int gn = t.group[g];
        for (int d = 0; d < 3; d++) {
            if (MODE != 2) {
                olo[d] = wb[d * ng + g];
                ohi[d] = wb[(3 + d) * ng + g];
            }
            glo[d] = t.mo.lo[d * C + gn];
            ghi[d] = t.mo.hi[d * C + gn];
        }
        if (FILL == 1) {
            j0 = oj[g];
            if (MODE == 2) s0 = os[g];
        }
        if (FILL == 2) {
            j0 = (int64_t)g * room;
            if (MODE == 2) s0 = (int64_t)g * room_s;
        }
    }
    // emit the entries (w0, w1) of the lanes with em set, in lane order
    auto emit = [&](bool em, uint32_t w0, uint32_t w1) {
        const uint32_t ty = w1 & kTyMask;
        if (MODE == 2) {
            const bool is_s = em && ty == kTySpj;
            const unsigned sb = walk_bits(is_s, lane);
            if (FILL == 1 && is_s) sout[s0 + ns + __popc(sb & below)] = (int32_t)w0;
        if (FILL == 2 && is_s && ns + __popc(sb & below) < room_s) sout[s0 + ns + __popc(sb & below)] = (int32_t)w0;
            ns += __popc(sb);
        }
        const bool is_l = em && ty == kTyLeaf;
        const int sz = is_l ? (int)(w1 & kCountMask) : 0;
        const int incl = walk_scan_incl(sz, c);
        if (FILL == 1 && is_l) {
            const int64_t p0 = j0 + nj + incl - sz;
            for (int q = 0; q < sz; q++) jout[p0 + q] = (int32_t)w0 + q;
        }
        if (FILL == 2 && is_l) {
            const int64_t p0 = nj + incl - sz;   // within the room
            for (int q = 0; q < sz; q++)
                if (p0 + q < room) jout[j0 + p0 + q] = (int32_t)w0 + q;
        }
        nj += __shfl(incl, 7, 8);
    };

    int sp = 0;                        // the walk's stack depth
    int first = 0, nch = live ? 1 : 0; // nodes evaluated by the next step: the root first
    bool bad = false;
    while (__ballot(nch > 0)) {
        // evaluate nodes first + c
        const bool ev = c < nch;
        uint32_t w0 = 0, w1 = 0;
        if (ev) {
            const int a = first + c;
            NodeLinks ln;
            const Visit v = visit<MODE>(t, a, olo, ohi, glo, ghi, inv_t2, ln);
            if (v == kSpj) {
                w0 = (uint32_t)a;
                w1 = kTySpj;
            } else if (v == kLeaf) {
                w0 = (uint32_t)ln.begin;
                w1 = kTyLeaf | (uint32_t)(ln.end - ln.begin);
            } else if (v == kOpen) {
                w0 = (uint32_t)ln.child;
                w1 = kTyOpen | (uint32_t)ln.nchild;
            }
        }
        const bool kept = ev && w1 != 0;
        const unsigned ob = walk_bits(kept && (w1 & kTyMask) == kTyOpen, lane);
        const int fo = ob ? __builtin_ctz(ob) : 8;
        emit(kept && c < fo, w0, w1);
        const bool push = kept && c >= fo;
        const unsigned pb = walk_bits(push, lane);
        if (sp + __popc(pb) > STACK) bad = true;
        if (push && !bad) stk[sub][sp + __popc(pb & above)] = make_uint2(w0, w1);
        if (!bad) sp += __popc(pb);
        // pop finished entries down to the next node to open
        nch = 0;
        while (sp > 0 && !bad) {
            const int avail = sp < 8 ? sp : 8;
            const uint2 e = c < avail ? stk[sub][sp - 1 - c] : make_uint2(0u, 0u);
            const unsigned ob2 = walk_bits(c < avail && (e.y & kTyMask) == kTyOpen, lane);
            const int k = ob2 ? __builtin_ctz(ob2) : avail;
            emit(c < k, e.x, e.y);
            if (ob2) {
                first = (int)__shfl((int)e.x, k, 8);
                nch = (int)(__shfl((int)e.y, k, 8) & kCountMask);
                sp -= k + 1;
                break;
            }
            sp -= k;
        }
    }
    if (bad && c == 0) atomicOr(err, 1);
    if (FILL == 0 && live && c == 0) {
        cj[g] = nj;
        if (MODE == 2) cs[g] = ns;
    }
    if (FILL == 2 && live && c == 0) {
        cj[g] = nj < room ? nj : room;   // what the room holds (all of it unless err bit 2)
        if (MODE == 2)
            cs[g] = ns < room_s ? ns : room_s;
        else
            cs[g] = j0;
        if (nj > room) atomicOr(err, 2);
        if (MODE == 2 && ns > room_s) atomicOr(err, 4);
    }
}

// Exact membership of each walk's candidates (the particles of the leaves its
// walk reached, in the host's order), one wave per walk: the first n_si
// (<= kWalkI) i-particles of the walk sit in LDS (dynamic, n_si double4), any
// beyond are read from memory; lane l tests candidate c0 + l, and the members
// are compacted in place (in order) at the front of the walk's stretch of
// scr; their number goes to cnt.  Candidates are taken kMemberU * 64 at a
// time: their indices and positions are all loaded before any is tested.
constexpr int kMemberU = 4;

template <int MODE>
__global__ __launch_bounds__(64) void member_kernel(int ng, TreeView t, const double *__restrict__ wb,
                                                    const double *__restrict__ r, const int64_t *__restrict__ coff,
                                                    const int64_t *__restrict__ ccnt, int32_t *__restrict__ scr,
                                                    int64_t *__restrict__ cnt, int n_si)
{
    extern __shared__ double4 si[];
    const int g = blockIdx.x;
    if (g >= ng) return;
    const int lane = threadIdx.x;
    const size_t C = t.mo.cap;
    const int ib = t.i_off[g], ni = t.i_off[g + 1] - ib;
    const int nl = ni < n_si ? ni : n_si;
    for (int q = lane; q < nl; q += 64)
        si[q] = make_double4(t.x[ib + q], t.y[ib + q], t.z[ib + q], r[ib + q]);
    const int gn = t.group[g];
    double olo[3], ohi[3], glo[3], ghi[3];
    for (int d = 0; d < 3; d++) {
        olo[d] = wb[d * ng + g];
        ohi[d] = wb[(3 + d) * ng + g];
        glo[d] = t.mo.lo[d * C + gn];
        ghi[d] = t.mo.hi[d * C + gn];
    }
    __syncthreads();
    int32_t *s = scr + coff[g];
    const int64_t nc = ccnt[g];
    int64_t count = 0;
    for (int64_t c0 = 0; c0 < nc; c0 += 64 * kMemberU) {
        int j[kMemberU];
        double p[kMemberU][3], hj[kMemberU];
#pragma unroll
        for (int u = 0; u < kMemberU; u++) {
            const int64_t e = c0 + 64 * u + lane;
            j[u] = e < nc ? s[e] : -1;
        }
#pragma unroll
        for (int u = 0; u < kMemberU; u++) {
            const int k = j[u] < 0 ? 0 : j[u];
            p[u][0] = t.x[k];
            p[u][1] = t.y[k];
            p[u][2] = t.z[k];
            hj[u] = MODE == 1 ? r[k] : 0.;
        }
        // every slot of this stretch is read: writes land at or before them
#pragma unroll
        for (int u = 0; u < kMemberU; u++) {
            const bool acc = j[u] >= 0 && member<MODE>(t, r, p[u], hj[u], ib, ni, nl, si, olo, ohi, glo, ghi);
            const uint64_t bal = __ballot(acc);
            if (acc) s[count + __popcll(bal & ((1ull << lane) - 1ull))] = j[u];
            count += __popcll(bal);
        }
    }
    if (lane == 0) cnt[g] = count;
}

// Members of each walk, from its stretch of the candidate scratch to its
// place in the list.
__global__ __launch_bounds__(64) void member_copy_kernel(int ng, const int64_t *__restrict__ soff,
                                                         const int32_t *__restrict__ j_off,
                                                         const int32_t *__restrict__ scr, int32_t *__restrict__ jout)
{
    const int g = blockIdx.x;
    if (g >= ng) return;
    const int64_t s0 = soff[g];
    const int b = j_off[g], n = j_off[g + 1] - b;
    for (int k = threadIdx.x; k < n; k += 64) jout[b + k] = scr[s0 + k];
}

// The long list's entries from their rooms (walk g's from g * room) to the
// walk offsets.
__global__ __launch_bounds__(64) void room_copy_kernel(int ng, int64_t room, const int32_t *__restrict__ off,
                                                       const int32_t *__restrict__ scr, int32_t *__restrict__ out)
{
    const int g = blockIdx.x;
    if (g >= ng) return;
    const int64_t s0 = (int64_t)g * room;
    const int b = off[g], n = off[g + 1] - b;
    for (int k = threadIdx.x; k < n; k += 64) out[b + k] = scr[s0 + k];
}

// Walk offsets from the scanned per-walk counts, and the total.
__global__ void offsets_kernel(int ng, const int64_t *__restrict__ scan, const int64_t *__restrict__ cnt,
                               int32_t *__restrict__ off, int64_t *__restrict__ total)
{
    const int g = blockIdx.x * kBlock + threadIdx.x;
    if (g < ng) off[g] = (int32_t)scan[g];
    if (g == ng - 1) {
        const int64_t t = scan[g] + cnt[g];
        *total = t;
        off[ng] = (int32_t)t;
    }
}

}  // namespace

constexpr int64_t kRoomStart = 2048;          // candidates (SPJ nodes) per walk
constexpr int64_t kRoomLongStart = 8192;      // long-list particles per walk
constexpr int64_t kRoomMax = 1 << 16;
constexpr int64_t kRoomMaxEntries = 1ll << 31;  // scratch entries (8 GiB)

struct sphk_dtree {
    int n = 0, n_leaf = 8, n_group = 16;
    hipStream_t st = nullptr;
    DArr<double> x, y, z, part, frame, len;
    DArr<uint64_t> key;
    DArr<int32_t> begin, end, child, nchild, level, parent;
    std::vector<int> level_start;   // node index of the first node of each level
    int n_node = 0;
    DArr<double> box_lo, box_hi, mm, mx, my, mz, me, sx, sy, sz, se, olo, ohi;
    DArr<int32_t> head, group, i_off, mscr;
    int n_walk = 0;
    DArr<double> wbox;
    DArr<BoxRec> brec;
    DArr<MacRec> mrec;
    DArr<OboxRec> orec;
    DArr<int64_t> i64a, i64b, i64c, i64d;
    DArr<int64_t> total;
    DArr<int> err;
    DArr<char> scratch;
    DLists lists[3];
    // per-walk rooms of the one-walk list builds (gather and symmetry
    // candidates; long-list EPJ and SPJ entries), doubled after a build that
    // overflowed them
    int64_t room[2] = {kRoomStart, kRoomStart};
    int64_t room_long[2] = {kRoomLongStart, kRoomStart};
    DArr<int32_t> sscr;   // SPJ room of the long list

    NodeArrays nodes() { return NodeArrays{begin.p, end.p, child.p, nchild.p, level.p, parent.p, len.p}; }
    Moments moments()
    {
        return Moments{box_lo.p, box_hi.p, mm.p, mx.p, my.p, mz.p, me.p, sx.p, sy.p, sz.p, se.p, box_lo.cap / 3};
    }
    void release()
    {
        for (auto *a : {&x, &y, &z, &part, &frame, &len, &box_lo, &box_hi, &mm, &mx, &my, &mz, &me, &sx, &sy,
                        &sz, &se, &olo, &ohi})
            a->release();
        key.release();
        for (auto *a : {&begin, &end, &child, &nchild, &level, &parent, &head, &group, &i_off, &mscr})
            a->release();
        wbox.release();
        brec.release();
        mrec.release();
        orec.release();
        for (auto *a : {&i64a, &i64b, &i64c, &i64d, &total}) a->release();
        err.release();
        scratch.release();
        sscr.release();
        for (auto &L : lists) {
            L.j_off.release();
            L.j_idx.release();
            L.s_off.release();
            L.s_idx.release();
        }
    }
};

namespace {

int reserve_nodes(sphk_dtree *t, size_t cap)
{
    if (cap <= t->begin.cap) return SPHK_OK;
    for (auto *a : {&t->begin, &t->end, &t->child, &t->nchild, &t->level, &t->parent})
        SPHK_TRY(a->reserve(cap, true, t->st));
    SPHK_TRY(t->len.reserve(cap, true, t->st));
    return SPHK_OK;
}

// exclusive scan of n int64 counts (queued on st)
int scan64(sphk_dtree *t, const int64_t *in, int64_t *out, int n)
{
    size_t bytes = 0;
    SPHK_TRY(rocprim::exclusive_scan(nullptr, bytes, in, out, (int64_t)0, (size_t)n, rocprim::plus<int64_t>(),
                                     t->st));
    SPHK_TRY(t->scratch.reserve(bytes + 16, false, t->st));
    SPHK_TRY(rocprim::exclusive_scan(t->scratch.p, bytes, in, out, (int64_t)0, (size_t)n,
                                     rocprim::plus<int64_t>(), t->st));
    return SPHK_OK;
}

template <typename T>
int read1(hipStream_t st, const T *dev, T *host)
{
    SPHK_TRY(hipMemcpyAsync(host, dev, sizeof(T), hipMemcpyDeviceToHost, st));
    SPHK_TRY(hipStreamSynchronize(st));
    return SPHK_OK;
}

int grid(int n) { return (n + kBlock - 1) / kBlock; }

// Both totals of a list (two int64 written by offsets_kernel) and the list
// walks' overflow flag, in one host read.
int read_totals(sphk_dtree *t, int64_t (&tot)[2], int &overflow)
{
    overflow = 0;
    SPHK_TRY(hipMemcpyAsync(tot, t->total.p, 2 * sizeof(int64_t), hipMemcpyDeviceToHost, t->st));
    SPHK_TRY(hipMemcpyAsync(&overflow, t->err.p, sizeof(int), hipMemcpyDeviceToHost, t->st));
    SPHK_TRY(hipStreamSynchronize(t->st));
    return SPHK_OK;
}

// Per-walk counts -> offsets and totals: counts a (and b) are scanned into
// c (and d); off_a / off_b receive the walk offsets.
int walk_offsets(sphk_dtree *t, int ng, const int64_t *a, int64_t *c, int32_t *off_a, const int64_t *b,
                 int64_t *d, int32_t *off_b)
{
    if (int rc = scan64(t, a, c, ng)) return rc;
    hipLaunchKernelGGL(offsets_kernel, dim3(grid(ng)), dim3(kBlock), 0, t->st, ng, c, a, off_a, t->total.p);
    if (b) {
        if (int rc = scan64(t, b, d, ng)) return rc;
        hipLaunchKernelGGL(offsets_kernel, dim3(grid(ng)), dim3(kBlock), 0, t->st, ng, d, b, off_b, t->total.p + 1);
    }
    SPHK_TRY(hipGetLastError());
    return SPHK_OK;
}

// A list still in the rooms of the scratch: packed to its offsets, into the
// caller's arrays (j_idx / s_idx non-NULL; the list stays in the rooms) or
// into the handle's own (the list leaves the rooms).
int pack_list(sphk_dtree *t, int kind, int32_t *j_idx, int32_t *s_idx)
{
    DLists &L = t->lists[kind];
    if (!L.built || !L.in_rooms) return SPHK_OK;
    const int ng = t->n_walk;
    const bool own = j_idx == nullptr;
    if (own) {
        SPHK_TRY(L.j_idx.reserve(std::max<int64_t>(L.n_j, 1), false, t->st));
        j_idx = L.j_idx.p;
        if (kind == 2) {
            SPHK_TRY(L.s_idx.reserve(std::max<int64_t>(L.n_s, 1), false, t->st));
            s_idx = L.s_idx.p;
        }
    }
    hipLaunchKernelGGL(room_copy_kernel, dim3(ng), dim3(64), 0, t->st, ng, L.room_j, L.j_off.p, t->mscr.p, j_idx);
    if (kind == 2 && s_idx)
        hipLaunchKernelGGL(room_copy_kernel, dim3(ng), dim3(64), 0, t->st, ng, L.room_s, L.s_off.p, t->sscr.p,
                           s_idx);
    SPHK_TRY(hipGetLastError());
    if (own) L.in_rooms = false;
    return SPHK_OK;
}

// Interaction lists of every walk.  Long list: the tree walk (count, scan,
// fill) gives each walk's EPJ particles and SPJ nodes.  Short-range lists
// (gather: r = rs, symmetry: r = H): the tree walk (count, scan, fill)
// writes the particles of the leaves reached into the walk's stretch of a
// scratch array, member_kernel keeps the members there, and they are copied
// to their offsets in the list.
template <int MODE>
int build_list(sphk_dtree *t, const double *r, double inv_t2)
{
    // lists still in the rooms of the scratch this build reuses: packed first
    for (int k = 0; k < 3; k++)
        if (int rc = pack_list(t, k, nullptr, nullptr)) return rc;
    DLists &L = t->lists[MODE];
    L.built = false;
    L.in_rooms = false;
    const int ng = t->n_walk;
    // i-particles staged in LDS by the member test: a walk holds <= n_group
    // particles unless it is a leaf (<= n_leaf; more only at the deepest level)
    const int n_si = std::min(kWalkI, std::max(t->n_group, t->n_leaf));
    SPHK_TRY(t->i64a.reserve(ng, false, t->st));   // EPJ counts per walk
    SPHK_TRY(t->i64b.reserve(ng, false, t->st));   // SPJ counts
    SPHK_TRY(t->i64c.reserve(ng, false, t->st));   // EPJ offsets
    SPHK_TRY(t->i64d.reserve(ng, false, t->st));   // SPJ offsets
    SPHK_TRY(t->total.reserve(2, false, t->st));
    SPHK_TRY(L.j_off.reserve(ng + 1, false, t->st));
    if (MODE == 2) SPHK_TRY(L.s_off.reserve(ng + 1, false, t->st));
    SPHK_TRY(hipMemsetAsync(t->err.p, 0, sizeof(int), t->st));
    if (MODE != 2) {
        SPHK_TRY(t->wbox.reserve(6 * (size_t)ng, false, t->st));
        hipLaunchKernelGGL(walk_box_kernel, dim3(grid(ng)), dim3(kBlock), 0, t->st, ng, t->i_off.p, t->x.p,
                           t->y.p, t->z.p, r, t->wbox.p);
    }
    TreeView v{t->nodes(), t->moments(), t->brec.p, t->mrec.p, t->orec.p, t->x.p, t->y.p, t->z.p, t->group.p,
               t->i_off.p};
    const double *wb = MODE == 2 ? nullptr : t->wbox.p;
    const dim3 lg((unsigned)((ng + kListWalks - 1) / kListWalks)), lb(64);
    if constexpr (MODE != 2) {
        // one walk into a room per walk: no count walk, one host read.  If a
        // walk overflows its room or the small stack, the counted build below
        // runs instead (and later builds get twice the room).
        int64_t &room = t->room[MODE];
        if ((int64_t)ng * room <= kRoomMaxEntries) {
            SPHK_TRY(t->mscr.reserve((size_t)ng * (size_t)room, false, t->st));
            hipLaunchKernelGGL((walk_list_kernel<MODE, 2, kListStackSmall>), lg, lb, 0, t->st, ng, v, wb, r, inv_t2,
                               t->i64a.p, t->i64c.p, nullptr, nullptr, t->mscr.p, nullptr, t->err.p, room);
            hipLaunchKernelGGL((member_kernel<MODE>), dim3(ng), dim3(64), n_si * sizeof(double4), t->st, ng, v, wb,
                               r, t->i64c.p, t->i64a.p, t->mscr.p, t->i64b.p, n_si);
            SPHK_TRY(hipGetLastError());
            if (int rc = walk_offsets(t, ng, t->i64b.p, t->i64d.p, L.j_off.p, nullptr, nullptr, nullptr)) return rc;
            int64_t tot1[2] = {0, 0};
            int flags = 0;
            if (int rc = read_totals(t, tot1, flags)) return rc;
            if (flags == 0) {
                if (tot1[0] > std::numeric_limits<int32_t>::max())
                    return set_error(SPHK_E_CAPACITY, "dtree: list longer than 2^31 entries");
                // the members stay in their rooms until the list is read
                L.n_j = tot1[0];
                L.n_s = 0;
                L.in_rooms = true;
                L.room_j = room;
                L.room_s = 0;
                L.built = true;
                return SPHK_OK;
            }
            if (flags & 2) room = std::min<int64_t>(room * 2, kRoomMax);
            SPHK_TRY(hipMemsetAsync(t->err.p, 0, sizeof(int), t->st));
        }
    }
    if constexpr (MODE == 2) {
        // one walk into rooms per walk (EPJ and SPJ), then the entries packed
        // to the walk offsets: no count walk, one host read.  The counted
        // build below if a walk overflows a room or the small stack (and later
        // builds get twice the room that overflowed).
        int64_t &rj = t->room_long[0], &rsp = t->room_long[1];
        if ((int64_t)ng * rj <= kRoomMaxEntries && (int64_t)ng * rsp <= kRoomMaxEntries) {
            SPHK_TRY(t->mscr.reserve((size_t)ng * (size_t)rj, false, t->st));
            SPHK_TRY(t->sscr.reserve((size_t)ng * (size_t)rsp, false, t->st));
            hipLaunchKernelGGL((walk_list_kernel<2, 2, kListStackSmall>), lg, lb, 0, t->st, ng, v, wb, r, inv_t2,
                               t->i64a.p, t->i64b.p, nullptr, nullptr, t->mscr.p, t->sscr.p, t->err.p, rj, rsp);
            SPHK_TRY(hipGetLastError());
            if (int rc = walk_offsets(t, ng, t->i64a.p, t->i64c.p, L.j_off.p, t->i64b.p, t->i64d.p, L.s_off.p))
                return rc;
            int64_t tot1[2] = {0, 0};
            int flags = 0;
            if (int rc = read_totals(t, tot1, flags)) return rc;
            if (flags == 0) {
                if (tot1[0] > std::numeric_limits<int32_t>::max() || tot1[1] > std::numeric_limits<int32_t>::max())
                    return set_error(SPHK_E_CAPACITY, "dtree: list longer than 2^31 entries");
                // the entries stay in their rooms until the list is read
                L.n_j = tot1[0];
                L.n_s = tot1[1];
                L.in_rooms = true;
                L.room_j = rj;
                L.room_s = rsp;
                L.built = true;
                return SPHK_OK;
            }
            if (flags & 2) rj = std::min<int64_t>(rj * 2, kRoomMax);
            if (flags & 4) rsp = std::min<int64_t>(rsp * 2, kRoomMax);
            SPHK_TRY(hipMemsetAsync(t->err.p, 0, sizeof(int), t->st));
        }
    }
    // count walk with the small stack, and with the worst-case one if a walk overflowed it
    int64_t tot[2] = {0, 0};
    int overflow = 0;
    bool big = false;
    for (;;) {
        if (big)
            hipLaunchKernelGGL((walk_list_kernel<MODE, 0, kListStack>), lg, lb, 0, t->st, ng, v, wb, r, inv_t2,
                               t->i64a.p, t->i64b.p, nullptr, nullptr, nullptr, nullptr, t->err.p);
        else
            hipLaunchKernelGGL((walk_list_kernel<MODE, 0, kListStackSmall>), lg, lb, 0, t->st, ng, v, wb, r,
                               inv_t2, t->i64a.p, t->i64b.p, nullptr, nullptr, nullptr, nullptr, t->err.p);
        SPHK_TRY(hipGetLastError());
        if (int rc = walk_offsets(t, ng, t->i64a.p, t->i64c.p, L.j_off.p, MODE == 2 ? t->i64b.p : nullptr,
                                  t->i64d.p, L.s_off.p))
            return rc;
        if (int rc = read_totals(t, tot, overflow)) return rc;
        if (!overflow) break;
        if (big) return set_error(SPHK_E_STATE, "dtree: list walk stack overflow (corrupt tree)");
        big = true;
        SPHK_TRY(hipMemsetAsync(t->err.p, 0, sizeof(int), t->st));
    }
    if (MODE != 2) tot[1] = 0;
    if (tot[0] > std::numeric_limits<int32_t>::max() || tot[1] > std::numeric_limits<int32_t>::max())
        return set_error(SPHK_E_CAPACITY, "dtree: list longer than 2^31 entries");
    if constexpr (MODE == 2) {
        SPHK_TRY(L.j_idx.reserve(std::max<int64_t>(tot[0], 1), false, t->st));
        SPHK_TRY(L.s_idx.reserve(std::max<int64_t>(tot[1], 1), false, t->st));
        if (big)
            hipLaunchKernelGGL((walk_list_kernel<MODE, 1, kListStack>), lg, lb, 0, t->st, ng, v, wb, r, inv_t2,
                               nullptr, nullptr, t->i64c.p, t->i64d.p, L.j_idx.p, L.s_idx.p, t->err.p);
        else
            hipLaunchKernelGGL((walk_list_kernel<MODE, 1, kListStackSmall>), lg, lb, 0, t->st, ng, v, wb, r,
                               inv_t2, nullptr, nullptr, t->i64c.p, t->i64d.p, L.j_idx.p, L.s_idx.p, t->err.p);
        SPHK_TRY(hipGetLastError());
    } else {
        // tot[0] candidates into the scratch; the members compacted in place,
        // their counts to i64b
        SPHK_TRY(t->mscr.reserve(std::max<int64_t>(tot[0], 1), false, t->st));
        if (big)
            hipLaunchKernelGGL((walk_list_kernel<MODE, 1, kListStack>), lg, lb, 0, t->st, ng, v, wb, r, inv_t2,
                               nullptr, nullptr, t->i64c.p, nullptr, t->mscr.p, nullptr, t->err.p);
        else
            hipLaunchKernelGGL((walk_list_kernel<MODE, 1, kListStackSmall>), lg, lb, 0, t->st, ng, v, wb, r,
                               inv_t2, nullptr, nullptr, t->i64c.p, nullptr, t->mscr.p, nullptr, t->err.p);
        hipLaunchKernelGGL((member_kernel<MODE>), dim3(ng), dim3(64), n_si * sizeof(double4), t->st, ng, v, wb, r,
                           t->i64c.p, t->i64a.p, t->mscr.p, t->i64b.p, n_si);
        SPHK_TRY(hipGetLastError());
        if (int rc = walk_offsets(t, ng, t->i64b.p, t->i64d.p, L.j_off.p, nullptr, nullptr, nullptr)) return rc;
        if (int rc = read_totals(t, tot, overflow)) return rc;
        if (overflow) return set_error(SPHK_E_STATE, "dtree: list walk stack overflow in the fill walk");
        tot[1] = 0;
        if (tot[0] > std::numeric_limits<int32_t>::max())
            return set_error(SPHK_E_CAPACITY, "dtree: list longer than 2^31 entries");
        SPHK_TRY(L.j_idx.reserve(std::max<int64_t>(tot[0], 1), false, t->st));
        hipLaunchKernelGGL(member_copy_kernel, dim3(ng), dim3(64), 0, t->st, ng, t->i64c.p, L.j_off.p, t->mscr.p,
                           L.j_idx.p);
        SPHK_TRY(hipGetLastError());
    }
    L.n_j = tot[0];
    L.n_s = tot[1];
    L.built = true;
    return SPHK_OK;
}

// Workspace of sphk_morton_order_dev, one per device, grown as needed and
// kept until the process ends (no allocation once it has reached the run's
// size).  Calls on one device take it in turn; each call's stream first
// waits for the previous call's work (an event), so no host wait is needed.
struct MortonWork {
    std::mutex mu;
    DArr<double> part, frame;
    DArr<uint64_t> key, key2;
    DArr<int32_t> idx;
    DArr<char> tmp;
    hipEvent_t done = nullptr;
};
constexpr int kMortonDevices = 16;
MortonWork g_morton[kMortonDevices];

}  // namespace

extern "C" {

int sphk_morton_order_dev(int n, const double *x, const double *y, const double *z, int32_t *perm, void *stream)
{
    if (n < 0 || (n > 0 && (!x || !y || !z || !perm))) return set_error(SPHK_E_ARG, "morton_dev: bad args");
    if (n == 0) return SPHK_OK;
    hipStream_t st = (hipStream_t)stream;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMortonDevices)
        return set_error(SPHK_E_DEVICE, "morton_dev: no usable current device");
    MortonWork &w = g_morton[dev];
    std::lock_guard<std::mutex> lock(w.mu);
    // the previous call's kernels (maybe on another stream) finish before the workspace is reused
    if (!w.done) {
        if (hipEventCreateWithFlags(&w.done, hipEventDisableTiming) != hipSuccess) {
            w.done = nullptr;
            return set_error(SPHK_E_DEVICE, "morton_dev: event creation failed");
        }
    } else {
        SPHK_TRY(hipStreamWaitEvent(st, w.done, 0));
    }
    SPHK_TRY(w.idx.reserve(n, false, st));
    SPHK_TRY(w.key2.reserve(n, false, st));
    if (int rc = make_keys(n, x, y, z, st, w.part, w.frame, w.key, w.idx.p)) return rc;
    size_t bytes = 0;
    SPHK_TRY(rocprim::radix_sort_pairs(nullptr, bytes, w.key.p, w.key2.p, w.idx.p, perm, (size_t)n, 0, 63, st));
    SPHK_TRY(w.tmp.reserve(bytes + 16, false, st));
    SPHK_TRY(rocprim::radix_sort_pairs(w.tmp.p, bytes, w.key.p, w.key2.p, w.idx.p, perm, (size_t)n, 0, 63, st));
    SPHK_TRY(hipEventRecord(w.done, st));
    return SPHK_OK;
}

namespace {

// The build proper, into t's buffers (grown as needed).  On failure t holds
// no tree; its buffers stay allocated for a later build or sphk_dtree_free.
int build_tree(sphk_dtree *t, int n, const double *x, const double *y, const double *z, const double *m,
               const double *eps2, int n_leaf, int n_group, hipStream_t stream)
{
    // a previous user of the buffers may still be queued on the old stream
    if (t->st != stream && hipStreamSynchronize(t->st) != hipSuccess)
        return set_error(SPHK_E_DEVICE, "dtree: stream synchronisation failed");
    t->n = n;
    t->n_leaf = n_leaf;
    t->n_group = n_group;
    t->st = stream;
    t->n_walk = 0;
    t->n_node = 0;
    for (auto &L : t->lists) {
        L.built = false;
        L.in_rooms = false;
    }
    auto fail = [&](int rc) {
        t->n_walk = 0;
        t->n_node = 0;
        return rc;
    };
    hipStream_t st = t->st;
    // own copies of the positions (list builds read them later)
    for (auto pr : {std::make_pair(&t->x, x), std::make_pair(&t->y, y), std::make_pair(&t->z, z)}) {
        if (pr.first->reserve(n, false, st) != hipSuccess ||
            hipMemcpyAsync(pr.first->p, pr.second, n * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess)
            return fail(set_error(SPHK_E_DEVICE, "dtree: position copy failed"));
    }
    if (int rc = make_keys(n, t->x.p, t->y.p, t->z.p, st, t->part, t->frame, t->key, nullptr)) return fail(rc);
    if (t->err.reserve(1, false, st) != hipSuccess || hipMemsetAsync(t->err.p, 0, sizeof(int), st) != hipSuccess)
        return fail(set_error(SPHK_E_DEVICE, "dtree: allocation failed"));
    hipLaunchKernelGGL(unsorted_kernel, dim3(grid(n)), dim3(kBlock), 0, st, n, t->key.p, t->err.p);
    int unsorted = 0;
    if (int rc = read1(st, t->err.p, &unsorted)) return fail(rc);
    if (unsorted)
        return fail(set_error(SPHK_E_ARG, "dtree: particles are not in Morton order (use sphk_morton_order_dev)"));

    // root, then one level at a time
    if (int rc = reserve_nodes(t, 2 * (size_t)n / std::max(1, n_leaf) + 16)) return fail(rc);
    hipLaunchKernelGGL(root_kernel, dim3(1), dim3(1), 0, st, n, t->nodes(), t->frame.p);
    t->level_start = {0, 1};
    int fa = 0, fb = 1;
    while (fa < fb) {
        const int nf = fb - fa;
        if (t->i64a.reserve(nf, false, st) != hipSuccess || t->i64c.reserve(nf, false, st) != hipSuccess ||
            t->total.reserve(2, false, st) != hipSuccess)
            return fail(set_error(SPHK_E_DEVICE, "dtree: allocation failed"));
        hipLaunchKernelGGL((split_kernel<false>), dim3(grid(nf)), dim3(kBlock), 0, st, fa, nf, t->key.p,
                           t->nodes(), n_leaf, t->i64a.p, nullptr, 0);
        if (int rc = scan64(t, t->i64a.p, t->i64c.p, nf)) return fail(rc);
        int64_t last[2];
        if (hipMemcpyAsync(&last[0], t->i64c.p + nf - 1, sizeof(int64_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(&last[1], t->i64a.p + nf - 1, sizeof(int64_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return fail(set_error(SPHK_E_DEVICE, "dtree: level size read failed"));
        const int64_t nnew = last[0] + last[1];
        if (nnew == 0) break;
        if (fb + nnew >= (1 << 29))   // list walks pack node indices into 29 bits
            return fail(set_error(SPHK_E_CAPACITY, "dtree: too many nodes"));
        if ((size_t)(fb + nnew) > t->begin.cap)
            if (int rc = reserve_nodes(t, std::max<size_t>((size_t)(fb + nnew), t->begin.cap * 3 / 2))) return fail(rc);
        hipLaunchKernelGGL((split_kernel<true>), dim3(grid(nf)), dim3(kBlock), 0, st, fa, nf, t->key.p, t->nodes(),
                           n_leaf, nullptr, t->i64c.p, fb);
        fa = fb;
        fb = fb + (int)nnew;
        t->level_start.push_back(fb);
    }
    t->n_node = fb;
    const size_t nn = (size_t)fb;

    // boxes and monopoles, deepest level first
    for (auto *a : {&t->mm, &t->mx, &t->my, &t->mz, &t->me, &t->sx, &t->sy, &t->sz, &t->se})
        if (a->reserve(nn, false, st) != hipSuccess) return fail(set_error(SPHK_E_DEVICE, "dtree: allocation failed"));
    if (t->box_lo.reserve(3 * nn, false, st) != hipSuccess || t->box_hi.reserve(3 * nn, false, st) != hipSuccess)
        return fail(set_error(SPHK_E_DEVICE, "dtree: allocation failed"));
    // the box arrays are [3][cap]; cap = box_lo.cap / 3 must match box_hi's
    if (t->box_lo.cap != t->box_hi.cap) return fail(set_error(SPHK_E_STATE, "dtree: box capacity mismatch"));
    for (int L = (int)t->level_start.size() - 2; L >= 0; L--) {
        const int a0 = t->level_start[L], na = t->level_start[L + 1] - a0;
        hipLaunchKernelGGL(moments_kernel, dim3(grid(na)), dim3(kBlock), 0, st, a0, na, t->nodes(), t->moments(),
                           t->x.p, t->y.p, t->z.p, m, eps2);
    }
    if (hipGetLastError() != hipSuccess) return fail(set_error(SPHK_E_DEVICE, "dtree: moments launch failed"));
    if (t->brec.reserve(nn, false, st) != hipSuccess || t->mrec.reserve(nn, false, st) != hipSuccess)
        return fail(set_error(SPHK_E_DEVICE, "dtree: allocation failed"));
    hipLaunchKernelGGL(pack_nodes_kernel, dim3(grid((int)nn)), dim3(kBlock), 0, st, (int)nn, t->nodes(),
                       t->moments(), t->brec.p, t->mrec.p);

    // groups
    if (t->head.reserve(n, false, st) != hipSuccess || t->i64a.reserve(n, false, st) != hipSuccess ||
        t->i64c.reserve(n, false, st) != hipSuccess)
        return fail(set_error(SPHK_E_DEVICE, "dtree: allocation failed"));
    hipLaunchKernelGGL(fill_kernel, dim3(grid(n)), dim3(kBlock), 0, st, n, t->head.p, -1);
    hipLaunchKernelGGL(group_mark_kernel, dim3(grid((int)nn)), dim3(kBlock), 0, st, (int)nn, t->nodes(), n_group,
                       t->head.p);
    hipLaunchKernelGGL(head_flag_kernel, dim3(grid(n)), dim3(kBlock), 0, st, n, t->head.p, t->i64a.p);
    if (int rc = scan64(t, t->i64a.p, t->i64c.p, n)) return fail(rc);
    int64_t last[2];
    if (hipMemcpyAsync(&last[0], t->i64c.p + n - 1, sizeof(int64_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(&last[1], t->i64a.p + n - 1, sizeof(int64_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return fail(set_error(SPHK_E_DEVICE, "dtree: group count read failed"));
    const int ng = (int)(last[0] + last[1]);
    if (t->group.reserve(ng, false, st) != hipSuccess || t->i_off.reserve(ng + 1, false, st) != hipSuccess)
        return fail(set_error(SPHK_E_DEVICE, "dtree: allocation failed"));
    hipLaunchKernelGGL(group_write_kernel, dim3(grid(n)), dim3(kBlock), 0, st, n, t->head.p, t->i64c.p, t->group.p,
                       t->i_off.p, ng);
    if (hipGetLastError() != hipSuccess) return fail(set_error(SPHK_E_DEVICE, "dtree: build failed"));
    t->n_walk = ng;
    return SPHK_OK;
}

int check_build_args(int n, const double *x, const double *y, const double *z, const double *m, int n_leaf,
                     int n_group)
{
    if (n <= 0 || !x || !y || !z || !m) return set_error(SPHK_E_ARG, "dtree: bad particle arrays");
    // list walks pack node and particle indices into 29 bits
    if (n >= (1 << 29)) return set_error(SPHK_E_CAPACITY, "dtree: at most 2^29 - 1 particles per tree");
    if (n_leaf < 1 || n_group < 1) return set_error(SPHK_E_ARG, "dtree: n_leaf, n_group must be >= 1");
    return SPHK_OK;
}

}  // namespace

int sphk_dtree_build(int n, const double *x, const double *y, const double *z, const double *m,
                     const double *eps2, int n_leaf, int n_group, void *stream, sphk_dtree **out)
{
    if (!out) return set_error(SPHK_E_ARG, "dtree: null output");
    *out = nullptr;
    if (int rc = check_build_args(n, x, y, z, m, n_leaf, n_group)) return rc;
    sphk_dtree *t = new (std::nothrow) sphk_dtree();
    if (!t) return set_error(SPHK_E_CAPACITY, "dtree: out of memory");
    t->st = (hipStream_t)stream;
    if (int rc = build_tree(t, n, x, y, z, m, eps2, n_leaf, n_group, (hipStream_t)stream)) {
        t->release();
        delete t;
        return rc;
    }
    *out = t;
    return SPHK_OK;
}

int sphk_dtree_rebuild(sphk_dtree *t, int n, const double *x, const double *y, const double *z,
                       const double *m, const double *eps2, int n_leaf, int n_group, void *stream)
{
    if (!t) return set_error(SPHK_E_ARG, "dtree: null handle");
    if (int rc = check_build_args(n, x, y, z, m, n_leaf, n_group)) return rc;
    return build_tree(t, n, x, y, z, m, eps2, n_leaf, n_group, (hipStream_t)stream);
}

void sphk_dtree_free(sphk_dtree *t)
{
    if (!t) return;
    (void)hipStreamSynchronize(t->st);
    t->release();
    delete t;
}

int sphk_dtree_n_walk(const sphk_dtree *t) { return t ? t->n_walk : -1; }
int sphk_dtree_n_node(const sphk_dtree *t) { return t ? t->n_node : -1; }
const int32_t *sphk_dtree_i_off(const sphk_dtree *t) { return t ? t->i_off.p : nullptr; }

int sphk_dtree_list_gather(sphk_dtree *t, const double *rs)
{
    if (!t || !rs) return set_error(SPHK_E_ARG, "dtree gather: null argument");
    if (t->n_walk <= 0) return set_error(SPHK_E_STATE, "dtree gather: no tree built");
    return build_list<0>(t, rs, 0.);
}

int sphk_dtree_list_symmetry(sphk_dtree *t, const double *h)
{
    if (!t || !h) return set_error(SPHK_E_ARG, "dtree symmetry: null argument");
    if (t->n_walk <= 0) return set_error(SPHK_E_STATE, "dtree symmetry: no tree built");
    const size_t nn = (size_t)t->n_node;
    if (t->olo.reserve(3 * nn, false, t->st) != hipSuccess || t->ohi.reserve(3 * nn, false, t->st) != hipSuccess)
        return set_error(SPHK_E_DEVICE, "dtree symmetry: allocation failed");
    // the obox arrays use the same [3][cap] stride as the tight boxes
    const size_t C = t->box_lo.cap / 3;
    if (t->olo.cap < 3 * C || t->ohi.cap < 3 * C) {
        t->olo.release();
        t->ohi.release();
        if (t->olo.reserve(3 * C, false, t->st) != hipSuccess || t->ohi.reserve(3 * C, false, t->st) != hipSuccess)
            return set_error(SPHK_E_DEVICE, "dtree symmetry: allocation failed");
    }
    for (int L = (int)t->level_start.size() - 2; L >= 0; L--) {
        const int a0 = t->level_start[L], na = t->level_start[L + 1] - a0;
        hipLaunchKernelGGL(obox_kernel, dim3(grid(na)), dim3(kBlock), 0, t->st, a0, na, t->nodes(), t->olo.p,
                           t->ohi.p, C, t->x.p, t->y.p, t->z.p, h);
    }
    if (t->orec.reserve(nn, false, t->st) != hipSuccess)
        return set_error(SPHK_E_DEVICE, "dtree symmetry: allocation failed");
    hipLaunchKernelGGL(pack_obox_kernel, dim3(grid((int)nn)), dim3(kBlock), 0, t->st, (int)nn, t->olo.p, t->ohi.p, C,
                       t->orec.p);
    SPHK_TRY(hipGetLastError());
    return build_list<1>(t, h, 0.);
}

int sphk_dtree_list_long(sphk_dtree *t, double theta)
{
    if (!t || !(theta >= 0.)) return set_error(SPHK_E_ARG, "dtree long: bad argument");
    if (t->n_walk <= 0) return set_error(SPHK_E_STATE, "dtree long: no tree built");
    const double inv_t2 = theta > 0. ? 1. / (theta * theta) : std::numeric_limits<double>::infinity();
    return build_list<2>(t, nullptr, inv_t2);
}

int sphk_dtree_list(sphk_dtree *t, int kind, const int32_t **j_off, const int32_t **j_idx,
                    const int32_t **s_off, const int32_t **s_idx, int64_t *n_j, int64_t *n_s)
{
    if (!t || kind < 0 || kind > 2) return set_error(SPHK_E_ARG, "dtree list: bad argument");
    DLists &L = t->lists[kind];
    if (!L.built) return set_error(SPHK_E_STATE, "dtree list kind %d not built", kind);
    if (int rc = pack_list(t, kind, nullptr, nullptr)) return rc;
    if (j_off) *j_off = L.j_off.p;
    if (j_idx) *j_idx = L.j_idx.p;
    if (s_off) *s_off = kind == 2 ? L.s_off.p : nullptr;
    if (s_idx) *s_idx = kind == 2 ? L.s_idx.p : nullptr;
    if (n_j) *n_j = L.n_j;
    if (n_s) *n_s = L.n_s;
    return SPHK_OK;
}

int sphk_dtree_list_size(const sphk_dtree *t, int kind, int64_t *n_j, int64_t *n_s)
{
    if (!t || kind < 0 || kind > 2) return set_error(SPHK_E_ARG, "dtree list size: bad argument");
    const DLists &L = t->lists[kind];
    if (!L.built) return set_error(SPHK_E_STATE, "dtree list kind %d not built", kind);
    if (n_j) *n_j = L.n_j;
    if (n_s) *n_s = L.n_s;
    return SPHK_OK;
}

int sphk_dtree_list_copy(sphk_dtree *t, int kind, int32_t *j_off, int32_t *j_idx, int32_t *s_off, int32_t *s_idx)
{
    if (!t || kind < 0 || kind > 2 || !j_off || (!j_idx && t->lists[kind].n_j > 0))
        return set_error(SPHK_E_ARG, "dtree list copy: bad argument");
    DLists &L = t->lists[kind];
    if (!L.built) return set_error(SPHK_E_STATE, "dtree list kind %d not built", kind);
    if (kind == 2 && (!s_off || (!s_idx && L.n_s > 0)))
        return set_error(SPHK_E_ARG, "dtree list copy: null SPJ array");
    const size_t no = (size_t)t->n_walk + 1;
    SPHK_TRY(hipMemcpyAsync(j_off, L.j_off.p, no * sizeof(int32_t), hipMemcpyDeviceToDevice, t->st));
    if (kind == 2) SPHK_TRY(hipMemcpyAsync(s_off, L.s_off.p, no * sizeof(int32_t), hipMemcpyDeviceToDevice, t->st));
    if (L.in_rooms) return pack_list(t, kind, j_idx, kind == 2 ? s_idx : nullptr);
    if (L.n_j > 0)
        SPHK_TRY(hipMemcpyAsync(j_idx, L.j_idx.p, (size_t)L.n_j * sizeof(int32_t), hipMemcpyDeviceToDevice, t->st));
    if (kind == 2 && L.n_s > 0)
        SPHK_TRY(hipMemcpyAsync(s_idx, L.s_idx.p, (size_t)L.n_s * sizeof(int32_t), hipMemcpyDeviceToDevice, t->st));
    return SPHK_OK;
}

int sphk_dtree_moments(const sphk_dtree *t, const double **m, const double **x, const double **y,
                       const double **z, const double **eps2)
{
    if (!t) return set_error(SPHK_E_ARG, "dtree moments: null tree");
    if (m) *m = t->mm.p;
    if (x) *x = t->mx.p;
    if (y) *y = t->my.p;
    if (z) *z = t->mz.p;
    if (eps2) *eps2 = t->me.p;
    return SPHK_OK;
}

int sphk_dtree_nodes(const sphk_dtree *t, const int32_t **begin, const int32_t **end, const int32_t **child,
                     const int32_t **nchild, const int32_t **level, const double **len)
{
    if (!t) return set_error(SPHK_E_ARG, "dtree nodes: null tree");
    if (begin) *begin = t->begin.p;
    if (end) *end = t->end.p;
    if (child) *child = t->child.p;
    if (nchild) *nchild = t->nchild.p;
    if (level) *level = t->level.p;
    if (len) *len = t->len.p;
    return SPHK_OK;
}

}  // extern "C"

namespace {

// cotangent per unit mass of each node's centre of mass, summed over the node
// and its ancestors (at most 21 levels up)
__global__ __launch_bounds__(kBlock) void spread_node_kernel(int nn, NodeArrays nd, const double *__restrict__ M,
                                                             const double *__restrict__ nx,
                                                             const double *__restrict__ ny,
                                                             const double *__restrict__ nz, double *__restrict__ cx,
                                                             double *__restrict__ cy, double *__restrict__ cz)
{
    const int a = blockIdx.x * kBlock + threadIdx.x;
    if (a >= nn) return;
    double s[3] = {0., 0., 0.};
    for (int b = a; b >= 0; b = nd.parent[b])
        if (M[b] > 0.) {
            s[0] += nx[b] / M[b];
            s[1] += ny[b] / M[b];
            s[2] += nz[b] / M[b];
        }
    cx[a] = s[0];
    cy[a] = s[1];
    cz[a] = s[2];
}

__global__ __launch_bounds__(kBlock) void spread_leaf_kernel(int nn, NodeArrays nd, const double *__restrict__ m,
                                                             const double *__restrict__ cx,
                                                             const double *__restrict__ cy,
                                                             const double *__restrict__ cz, double *gx, double *gy,
                                                             double *gz)
{
    const int a = blockIdx.x * kBlock + threadIdx.x;
    if (a >= nn || nd.child[a] >= 0) return;
    for (int k = nd.begin[a]; k < nd.end[a]; k++) {
        gx[k] += m[k] * cx[a];
        gy[k] += m[k] * cy[a];
        gz[k] += m[k] * cz[a];
    }
}

}  // namespace

extern "C" int sphk_dtree_spread_dev(const sphk_dtree *t, const double *m, const double *nx, const double *ny,
                                     const double *nz, double *gx, double *gy, double *gz)
{
    if (!t || !m || !nx || !ny || !nz || !gx || !gy || !gz) return set_error(SPHK_E_ARG, "dtree spread: null argument");
    sphk_dtree *tt = const_cast<sphk_dtree *>(t);
    const int nn = t->n_node;
    // three per-node sums, kept in the SPJ scratch of the tree
    SPHK_TRY(tt->sx.reserve(nn, false, t->st));
    SPHK_TRY(tt->sy.reserve(nn, false, t->st));
    SPHK_TRY(tt->sz.reserve(nn, false, t->st));
    hipLaunchKernelGGL(spread_node_kernel, dim3(grid(nn)), dim3(kBlock), 0, t->st, nn, tt->nodes(), t->mm.p, nx, ny,
                       nz, tt->sx.p, tt->sy.p, tt->sz.p);
    hipLaunchKernelGGL(spread_leaf_kernel, dim3(grid(nn)), dim3(kBlock), 0, t->st, nn, tt->nodes(), m, tt->sx.p,
                       tt->sy.p, tt->sz.p, gx, gy, gz);
    SPHK_TRY(hipGetLastError());
    return SPHK_OK;
}

namespace {

// ---- reorder and input comparison (host glue of the rebuilds) ----------------
constexpr int kPermCols = 64;
struct PermCols {
    char *col[kPermCols];
    char *tmp[kPermCols];
    int elem[kPermCols];
};

template <typename T>
__device__ __forceinline__ void gather_col(int n, const int32_t *__restrict__ perm, const char *src, char *dst)
{
    const T *s = reinterpret_cast<const T *>(src);
    T *d = reinterpret_cast<T *>(dst);
    for (int k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock) d[k] = s[perm ? perm[k] : k];
}

// tmp_c[k] = col_c[perm[k]], or with perm == nullptr col_c[k] = tmp_c[k];
// grid row c handles column c
__global__ __launch_bounds__(kBlock) void permute_kernel(int n, const int32_t *__restrict__ perm, PermCols pc)
{
    const int c = blockIdx.y;
    const char *src = perm ? pc.col[c] : pc.tmp[c];
    char *dst = perm ? pc.tmp[c] : pc.col[c];
    if (!perm && ((uintptr_t)dst % 16) == 0) {
        // copy back 16 bytes per thread (the scratch is 256-byte aligned)
        const size_t bytes = (size_t)n * pc.elem[c];
        const size_t n16 = bytes / 16;
        const uint4 *s4 = reinterpret_cast<const uint4 *>(src);
        uint4 *d4 = reinterpret_cast<uint4 *>(dst);
        const size_t stride = (size_t)gridDim.x * kBlock;
        for (size_t k = blockIdx.x * kBlock + threadIdx.x; k < n16; k += stride) d4[k] = s4[k];
        for (size_t k = n16 * 16 + blockIdx.x * kBlock + threadIdx.x; k < bytes; k += stride) dst[k] = src[k];
        return;
    }
    // gather, or a copy back by elements into a column aligned to less than 16 bytes
    switch (pc.elem[c]) {
    case 8: gather_col<uint64_t>(n, perm, src, dst); break;
    case 4: gather_col<uint32_t>(n, perm, src, dst); break;
    case 2: gather_col<uint16_t>(n, perm, src, dst); break;
    default: gather_col<uint8_t>(n, perm, src, dst); break;
    }
}

constexpr int kEqualArrays = 8;
struct EqualArrays {
    const uint64_t *a[kEqualArrays];
    const uint64_t *b[kEqualArrays];
};

__global__ __launch_bounds__(kBlock) void equal_kernel(int n, EqualArrays ea, int32_t *__restrict__ flag)
{
    const int c = blockIdx.y;
    bool diff = false;
    for (int k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock)
        diff = diff || ea.a[c][k] != ea.b[c][k];
    if (__any(diff) && threadIdx.x % 64 == 0) *flag = 1;   // every writer stores 1
}

__global__ void flag_zero_kernel(int32_t *flag) { *flag = 0; }

size_t permute_work(int n, int count, const int *elem)
{
    size_t w = 0;
    for (int c = 0; c < count; c++) w += (((size_t)n * (size_t)elem[c]) + 255) / 256 * 256;
    return w;
}

}  // namespace

extern "C" {

size_t sphk_permute_work_bytes(int n, int count, const int *elem)
{
    if (n <= 0 || count <= 0 || count > kPermCols || !elem) return 0;
    return permute_work(n, count, elem);
}

int sphk_permute_dev(int n, const int32_t *perm, int count, void *const *cols, const int *elem, void *work,
                     size_t work_bytes, void *stream)
{
    if (n < 0 || count < 0 || count > kPermCols) return set_error(SPHK_E_ARG, "permute_dev: bad sizes");
    if (n == 0 || count == 0) return SPHK_OK;
    if (!perm || !cols || !elem || !work) return set_error(SPHK_E_ARG, "permute_dev: null pointer");
    PermCols pc{};
    size_t off = 0;
    for (int c = 0; c < count; c++) {
        if (!cols[c]) return set_error(SPHK_E_ARG, "permute_dev: null column %d", c);
        if (elem[c] != 1 && elem[c] != 2 && elem[c] != 4 && elem[c] != 8)
            return set_error(SPHK_E_ARG, "permute_dev: element size %d of column %d", elem[c], c);
        if (((uintptr_t)cols[c]) % (uintptr_t)elem[c])
            return set_error(SPHK_E_ARG, "permute_dev: column %d misaligned", c);
        pc.col[c] = static_cast<char *>(cols[c]);
        pc.tmp[c] = static_cast<char *>(work) + off;
        pc.elem[c] = elem[c];
        off += (((size_t)n * (size_t)elem[c]) + 255) / 256 * 256;
    }
    if (work_bytes < off) return set_error(SPHK_E_CAPACITY, "permute_dev: work of %zu bytes, %zu needed", work_bytes, off);
    if (((uintptr_t)work) % 256) return set_error(SPHK_E_ARG, "permute_dev: work not 256-byte aligned");
    hipStream_t st = (hipStream_t)stream;
    const dim3 grid(std::min(1024, (n + kBlock - 1) / kBlock), count);
    hipLaunchKernelGGL(permute_kernel, grid, dim3(kBlock), 0, st, n, perm, pc);
    hipLaunchKernelGGL(permute_kernel, grid, dim3(kBlock), 0, st, n, nullptr, pc);
    SPHK_TRY(hipGetLastError());
    return SPHK_OK;
}

int sphk_equal_dev(int n, int count, const double *const *a, const double *const *b, int32_t *flag, void *stream)
{
    if (n < 0 || count < 0 || count > kEqualArrays) return set_error(SPHK_E_ARG, "equal_dev: bad sizes");
    if (!flag) return set_error(SPHK_E_ARG, "equal_dev: null flag");
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(flag_zero_kernel, dim3(1), dim3(1), 0, st, flag);
    if (n > 0 && count > 0) {
        if (!a || !b) return set_error(SPHK_E_ARG, "equal_dev: null array list");
        EqualArrays ea{};
        for (int c = 0; c < count; c++) {
            if (!a[c] || !b[c]) return set_error(SPHK_E_ARG, "equal_dev: null array %d", c);
            ea.a[c] = reinterpret_cast<const uint64_t *>(a[c]);
            ea.b[c] = reinterpret_cast<const uint64_t *>(b[c]);
        }
        const dim3 grid(std::min(512, (n + kBlock - 1) / kBlock), count);
        hipLaunchKernelGGL(equal_kernel, grid, dim3(kBlock), 0, st, n, ea, flag);
    }
    SPHK_TRY(hipGetLastError());
    return SPHK_OK;
}

int sphk_copy_dev(void *dst, const void *src, size_t bytes, void *stream)
{
    if (bytes == 0) return SPHK_OK;
    if (!dst || !src) return set_error(SPHK_E_ARG, "copy_dev: null pointer");
    SPHK_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return SPHK_OK;
}

}  // extern "C"
