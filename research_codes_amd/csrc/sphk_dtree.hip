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
//   5. Lists: the host's depth-first traversal of each walk is split at a
//      frontier of subtrees (one tree level plus the leaves above it), one
//      thread per (walk, subtree), run twice: count, then fill at scanned
//      offsets; the output is the host's, element for element.
#include <hip/hip_runtime.h>

#include <cstring>   // rocPRIM's iterator headers use memset without including it

#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <cstdint>
#include <limits>
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
// `next` is the node the depth-first walk visits after a's subtree (its
// next sibling, else its parent's next, -1 after the last): the walks need
// no stack.
struct alignas(64) BoxRec {     // gather and symmetry walks
    double lo[3], hi[3];        // tight box
    int32_t child, next, begin, end;
};
struct alignas(64) MacRec {     // long walk
    double mx, my, mz, mm, len;
    int32_t child, next, begin, end;
};
struct alignas(64) OboxRec {    // symmetry walk: box of the nodes' particle spheres
    double lo[3], hi[3];
};

// Skip pointers of one level, parents' already set (levels top-down).
__global__ __launch_bounds__(kBlock) void next_kernel(int a0, int na, NodeArrays nd, int32_t *__restrict__ nxt)
{
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= na) return;
    const int a = a0 + k;
    const int p = nd.parent[a];
    if (p < 0) {
        nxt[a] = -1;
        return;
    }
    nxt[a] = a < nd.child[p] + nd.nchild[p] - 1 ? a + 1 : nxt[p];
}

__global__ __launch_bounds__(kBlock) void pack_nodes_kernel(int nn, NodeArrays nd, Moments mo,
                                                            const int32_t *__restrict__ nxt,
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
    b.next = nxt[a];
    b.begin = nd.begin[a];
    b.end = nd.end[a];
    br[a] = b;
    mr[a] = MacRec{mo.mx[a], mo.my[a], mo.mz[a], mo.mm[a], nd.len[a], b.child, b.next, b.begin, b.end};
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
    int32_t child, next, begin, end;
};

template <int MODE>
__device__ __forceinline__ Visit visit(const TreeView &t, int a, const double (&olo)[3],
                                       const double (&ohi)[3], const double (&glo)[3], const double (&ghi)[3],
                                       double inv_t2, NodeLinks &ln)
{
#pragma clang fp contract(off)
    if (MODE != 2) {
        const BoxRec b = t.br[a];
        ln = NodeLinks{b.child, b.next, b.begin, b.end};
        const bool leaf = b.child < 0;
        if (MODE == 0) {
            if (!box_overlaps(olo, ohi, b.lo, b.hi)) return kSkip;
            return leaf ? kLeaf : kOpen;
        }
        if (!box_overlaps(olo, ohi, b.lo, b.hi)) {
            const OboxRec o = t.orr[a];
            if (!box_overlaps(glo, ghi, o.lo, o.hi)) return kSkip;
        }
        return leaf ? kLeaf : kOpen;
    }
    const MacRec m = t.mr[a];
    ln = NodeLinks{m.child, m.next, m.begin, m.end};
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

// One lane's output stream of int32 entries, written in aligned 16-byte
// pieces: entries before the first multiple-of-4 position go out one by one,
// then four at a time from registers, the rest at flush().  The long list's
// entries arrive one per tree visit; single 4-byte stores from 64 walks at
// once leave partly written cache lines that the memory side completes one
// sector per entry.
struct Out4 {
    int32_t *out;
    int64_t pos;   // next position; entries [pos - n, pos) wait in b0..b3
    int n;
    int32_t b0, b1, b2, b3;
    __device__ __forceinline__ void put(int32_t v)
    {
        if (n == 0 && (pos & 3) != 0) {
            out[pos++] = v;
            return;
        }
        b0 = n == 0 ? v : b0;
        b1 = n == 1 ? v : b1;
        b2 = n == 2 ? v : b2;
        b3 = n == 3 ? v : b3;
        n++;
        pos++;
        if (n == 4) {
            *reinterpret_cast<int4 *>(out + pos - 4) = make_int4(b0, b1, b2, b3);
            n = 0;
        }
    }
    __device__ __forceinline__ void flush()
    {
        const int64_t p = pos - n;
        if (n > 0) out[p] = b0;
        if (n > 1) out[p + 1] = b1;
        if (n > 2) out[p + 2] = b2;
        n = 0;
    }
};

// The host builder's depth-first walk of one group, split at the frontier:
// frontier node f (the nodes of one level plus the leaves above it, in
// particle order) gets its own thread per walk.  The thread first replays
// the walk's decisions on the ancestors of f, root first: if one is skipped
// nothing is emitted; if one is taken as an SPJ it is emitted by the thread
// of its first frontier node only.  Otherwise it walks the subtree of f.
// Concatenated in frontier order, the threads' outputs are the host's list.
// MODE 0 gather, 1 symmetry, 2 long; with FILL entries are written from the
// scanned offsets, else counted.
template <int MODE, bool FILL, bool LEAVES = false>
__global__ __launch_bounds__(64) void list_kernel(int ng, int nf, TreeView t, const int32_t *__restrict__ front,
                                                  const double *__restrict__ wb, const double *__restrict__ r,
                                                  double inv_t2, int64_t *__restrict__ cj,
                                                  int64_t *__restrict__ cs, const int64_t *__restrict__ oj,
                                                  const int64_t *__restrict__ os, int32_t *__restrict__ jout,
                                                  int32_t *__restrict__ sout)
{
#pragma clang fp contract(off)
    // neighbouring lanes take neighbouring walks under the same frontier
    // node: their traversals follow nearly the same path
    const int64_t t0 = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (t0 >= (int64_t)ng * nf) return;
    const int g = (int)(t0 % ng), k = (int)(t0 / ng);
    const int64_t tid = (int64_t)g * nf + k;   // slot of (walk, frontier node) in the count arrays
    const int f = front[k];
    const size_t C = t.mo.cap;
    const int ib = t.i_off[g], ie = t.i_off[g + 1];
    const int gn = t.group[g];
    double olo[3] = {0., 0., 0.}, ohi[3] = {0., 0., 0.}, glo[3], ghi[3];
    for (int d = 0; d < 3; d++) {
        if (MODE != 2) {
            olo[d] = wb[d * ng + g];
            ohi[d] = wb[(3 + d) * ng + g];
        }
        glo[d] = t.mo.lo[d * C + gn];
        ghi[d] = t.mo.hi[d * C + gn];
    }
    int64_t nj = 0, ns = 0;
    const int64_t j0 = FILL ? oj[tid] : 0;
    const int64_t s0 = (FILL && MODE == 2) ? os[tid] : 0;
    // long list: both streams through write-combining buffers
    Out4 jw{jout, j0, 0, 0, 0, 0, 0}, sw{sout, s0, 0, 0, 0, 0, 0};

    // ancestors of f, root first (the frontier is a few levels deep)
    const int lf = t.nd.level[f];
    bool reach = true;
    NodeLinks ln;
    for (int L = 0; L < lf && reach; L++) {
        int a = f;
        for (int u = lf; u > L; u--) a = t.nd.parent[a];
        const Visit v = visit<MODE>(t, a, olo, ohi, glo, ghi, inv_t2, ln);
        if (v == kOpen) continue;
        reach = false;
        if (v == kSpj && ln.begin == t.nd.begin[f]) {
            if (FILL) sw.put(a);
            ns++;
        }
    }
    if (reach) {
        // preorder over f's subtree by skip pointers: open = first child,
        // otherwise continue at next (f's own next ends the subtree)
        const int stop = MODE == 2 ? t.mr[f].next : t.br[f].next;
        for (int nxt = f; nxt != stop;) {
            const int a = nxt;
            const Visit v = visit<MODE>(t, a, olo, ohi, glo, ghi, inv_t2, ln);
            if (v == kOpen) {
                nxt = ln.child;
                continue;
            }
            nxt = ln.next;
            if (v == kSkip) continue;
            if (v == kSpj) {
                if (FILL) sw.put(a);
                ns++;
                continue;
            }
            if (LEAVES) {   // candidate leaf; its particles are filtered by member_kernel
                if (FILL) jout[j0 + nj] = a;
                nj++;
                continue;
            }
            // leaf: its particles, filtered by exact membership for the short lists
            for (int j = ln.begin; j < ln.end; j++) {
                if (MODE == 2) {
                    if (FILL) jw.put(j);
                    nj++;
                    continue;
                }
                const double p[3] = {t.x[j], t.y[j], t.z[j]};
                const double hj = MODE == 1 ? r[j] : 0.;
                bool in = true, jov = MODE == 1;
                for (int d = 0; d < 3; d++) {
                    in = in && !(p[d] < olo[d] || p[d] > ohi[d]);
                    if (MODE == 1) jov = jov && !(ghi[d] < p[d] - hj || glo[d] > p[d] + hj);
                }
                if (!in && !jov) continue;
                for (int i = ib; i < ie; i++) {
                    const double dx = p[0] - t.x[i], dy = p[1] - t.y[i], dz = p[2] - t.z[i];
                    const double rr = MODE == 1 ? (r[i] > hj ? r[i] : hj) : r[i];
                    if (dx * dx + dy * dy + dz * dz < rr * rr) {
                        if (FILL) jout[j0 + nj] = j;
                        nj++;
                        break;
                    }
                }
            }
        }
    }
    if (FILL && MODE == 2) {
        jw.flush();
        sw.flush();
    }
    if (!FILL) {
        cj[tid] = nj;
        if (MODE == 2) cs[tid] = ns;
    }
}

// Exact membership of the particles of each walk's candidate leaves, one
// wave per walk: the walk's i-particles sit in LDS, lane l tests candidate
// c0 + l of the walk's leaves laid end to end (the host's order), and a
// ballot orders the accepted ones.  MODE 0: |x_j - x_i| < rs_i for some i;
// MODE 1: r_ij < max(H_i, H_j) for some i (after the same box pre-tests as
// the host's leaf loop).  The accepted j are written from off[g] (room for
// all of the walk's candidates) and their number to cnt[g].
constexpr int kMemberCap = 256;   // i-particles of a walk held in LDS

template <int MODE>
__global__ __launch_bounds__(64) void member_kernel(int ng, TreeView t, const int32_t *__restrict__ lo,
                                                    const int32_t *__restrict__ li, const double *__restrict__ wb,
                                                    const double *__restrict__ r, int64_t *__restrict__ cnt,
                                                    const int64_t *__restrict__ off, int32_t *__restrict__ jout)
{
#pragma clang fp contract(off)
    __shared__ double si[4][kMemberCap];
    __shared__ int32_t lbeg[64], lpre[65];
    const int g = blockIdx.x;
    if (g >= ng) return;
    const int lane = threadIdx.x;
    const size_t C = t.mo.cap;
    const int ib = t.i_off[g], ie = t.i_off[g + 1];
    const int ni = ie - ib;
    const bool in_lds = ni <= kMemberCap;
    if (in_lds)
        for (int q = lane; q < ni; q += 64) {
            si[0][q] = t.x[ib + q];
            si[1][q] = t.y[ib + q];
            si[2][q] = t.z[ib + q];
            si[3][q] = r[ib + q];
        }
    const int gn = t.group[g];
    double olo[3], ohi[3], glo[3], ghi[3];
    for (int d = 0; d < 3; d++) {
        olo[d] = wb[d * ng + g];
        ohi[d] = wb[(3 + d) * ng + g];
        glo[d] = t.mo.lo[d * C + gn];
        ghi[d] = t.mo.hi[d * C + gn];
    }
    int64_t count = 0;
    const int64_t base = off[g];
    const int lb = lo[g], le = lo[g + 1];
    for (int l0 = lb; l0 < le; l0 += 64) {
        const int nl = min(64, le - l0);
        __syncthreads();
        int sz = 0;
        if (lane < nl) {
            const int a = li[l0 + lane];
            lbeg[lane] = t.nd.begin[a];
            sz = t.nd.end[a] - t.nd.begin[a];
        }
        // inclusive scan of the leaf sizes
        int incl = sz;
        for (int d = 1; d < 64; d <<= 1) {
            const int v = __shfl_up(incl, d, 64);
            if (lane >= d) incl += v;
        }
        lpre[lane + 1] = incl;
        if (lane == 0) lpre[0] = 0;
        __syncthreads();
        const int total = lpre[nl];
        for (int c0 = 0; c0 < total; c0 += 64) {
            const int c = c0 + lane;
            bool pass = false;
            int j = -1;
            if (c < total) {
                int lo_ = 0, hi_ = nl - 1;   // leaf q with lpre[q] <= c < lpre[q+1]
                while (lo_ < hi_) {
                    const int mid = (lo_ + hi_ + 1) >> 1;
                    if (lpre[mid] <= c)
                        lo_ = mid;
                    else
                        hi_ = mid - 1;
                }
                j = lbeg[lo_] + (c - lpre[lo_]);
                const double p[3] = {t.x[j], t.y[j], t.z[j]};
                const double hj = MODE == 1 ? r[j] : 0.;
                bool in = true, jov = MODE == 1;
                for (int d = 0; d < 3; d++) {
                    in = in && !(p[d] < olo[d] || p[d] > ohi[d]);
                    if (MODE == 1) jov = jov && !(ghi[d] < p[d] - hj || glo[d] > p[d] + hj);
                }
                if (in || jov)
                    for (int q = 0; q < ni; q++) {
                        const double xi = in_lds ? si[0][q] : t.x[ib + q];
                        const double yi = in_lds ? si[1][q] : t.y[ib + q];
                        const double zi = in_lds ? si[2][q] : t.z[ib + q];
                        const double ri = in_lds ? si[3][q] : r[ib + q];
                        const double dx = p[0] - xi, dy = p[1] - yi, dz = p[2] - zi;
                        const double rr = MODE == 1 ? (ri > hj ? ri : hj) : ri;
                        if (dx * dx + dy * dy + dz * dz < rr * rr) {
                            pass = true;
                            break;
                        }
                    }
            }
            const uint64_t bal = __ballot(pass);
            if (pass) jout[base + count + __popcll(bal & ((1ull << lane) - 1ull))] = j;
            count += __popcll(bal);
        }
    }
    if (lane == 0) cnt[g] = count;
}

// Particles in each walk's candidate leaves (the room member_kernel may
// need for the walk's accepted particles), one wave per walk.
__global__ __launch_bounds__(64) void cand_total_kernel(int ng, const int32_t *__restrict__ lo,
                                                        const int32_t *__restrict__ li, NodeArrays nd,
                                                        int64_t *__restrict__ tot)
{
    const int g = blockIdx.x;
    if (g >= ng) return;
    int s = 0;
    for (int l = lo[g] + (int)threadIdx.x; l < lo[g + 1]; l += 64) {
        const int a = li[l];
        s += nd.end[a] - nd.begin[a];
    }
    for (int d = 32; d > 0; d >>= 1) s += __shfl_xor(s, d, 64);
    if (threadIdx.x == 0) tot[g] = s;
}

// Accepted particles of each walk, from its stretch of the member scratch
// to its place in the list.
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

// Frontier nodes: the nodes of level L and the leaves above it.
__global__ __launch_bounds__(kBlock) void front_mark_kernel(int nn, NodeArrays nd, int L, int32_t *__restrict__ head)
{
    const int a = blockIdx.x * kBlock + threadIdx.x;
    if (a >= nn) return;
    const int lv = nd.level[a];
    if (lv == L || (lv < L && nd.child[a] < 0)) head[nd.begin[a]] = a;
}

__global__ __launch_bounds__(kBlock) void front_write_kernel(int n, const int32_t *__restrict__ head,
                                                             const int64_t *__restrict__ pos,
                                                             int32_t *__restrict__ front)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n && head[i] >= 0) front[pos[i]] = head[i];
}

// Walk offsets from the scanned (walk, frontier) counts: off[g] = scan[g nf].
__global__ void offsets_kernel(int ng, int nf, const int64_t *__restrict__ scan, const int64_t *__restrict__ cnt,
                               int32_t *__restrict__ off, int64_t *__restrict__ total)
{
    const int g = blockIdx.x * kBlock + threadIdx.x;
    if (g < ng) off[g] = (int32_t)scan[(int64_t)g * nf];
    if (g == ng - 1) {
        const int64_t last = (int64_t)ng * nf - 1;
        const int64_t t = scan[last] + cnt[last];
        *total = t;
        off[ng] = (int32_t)t;
    }
}

}  // namespace

struct sphk_dtree {
    int n = 0, n_leaf = 8, n_group = 16;
    hipStream_t st = nullptr;
    DArr<double> x, y, z, part, frame, len;
    DArr<uint64_t> key;
    DArr<int32_t> begin, end, child, nchild, level, parent;
    std::vector<int> level_start;   // node index of the first node of each level
    int n_node = 0;
    DArr<double> box_lo, box_hi, mm, mx, my, mz, me, sx, sy, sz, se, olo, ohi;
    DArr<int32_t> head, group, i_off, front, leaf_off, leaf_idx, mscr, nextp;
    int n_walk = 0, n_front = 0;
    DArr<double> wbox;
    DArr<BoxRec> brec;
    DArr<MacRec> mrec;
    DArr<OboxRec> orec;
    DArr<int64_t> i64a, i64b, i64c, i64d;
    DArr<int64_t> total;
    DArr<int> err;
    DArr<char> scratch;
    DLists lists[3];

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
        for (auto *a : {&begin, &end, &child, &nchild, &level, &parent, &head, &group, &i_off, &front, &leaf_off,
                        &leaf_idx, &mscr, &nextp})
            a->release();
        wbox.release();
        brec.release();
        mrec.release();
        orec.release();
        for (auto *a : {&i64a, &i64b, &i64c, &i64d, &total}) a->release();
        err.release();
        scratch.release();
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

template <int MODE>
int build_list_two_stage(sphk_dtree *t, const double *r);

template <int MODE>
int build_list(sphk_dtree *t, const double *r, double inv_t2)
{
    if (MODE != 2) return build_list_two_stage<MODE>(t, r);
    DLists &L = t->lists[MODE];
    L.built = false;
    const int ng = t->n_walk, nf = t->n_front;
    const int64_t nt = (int64_t)ng * nf;
    SPHK_TRY(t->i64a.reserve(nt, false, t->st));   // EPJ counts per (walk, frontier node)
    SPHK_TRY(t->i64b.reserve(nt, false, t->st));   // SPJ counts
    SPHK_TRY(t->i64c.reserve(nt, false, t->st));   // EPJ offsets
    SPHK_TRY(t->i64d.reserve(nt, false, t->st));   // SPJ offsets
    SPHK_TRY(t->total.reserve(2, false, t->st));
    SPHK_TRY(L.j_off.reserve(ng + 1, false, t->st));
    if (MODE == 2) SPHK_TRY(L.s_off.reserve(ng + 1, false, t->st));
    if (MODE != 2) {
        SPHK_TRY(t->wbox.reserve(6 * (size_t)ng, false, t->st));
        hipLaunchKernelGGL(walk_box_kernel, dim3(grid(ng)), dim3(kBlock), 0, t->st, ng, t->i_off.p, t->x.p,
                           t->y.p, t->z.p, r, t->wbox.p);
    }
    TreeView v{t->nodes(), t->moments(), t->brec.p, t->mrec.p, t->orec.p, t->x.p, t->y.p, t->z.p, t->group.p,
               t->i_off.p};
    const dim3 lg((unsigned)((nt + 63) / 64)), lb(64);
    hipLaunchKernelGGL((list_kernel<MODE, false>), lg, lb, 0, t->st, ng, nf, v, t->front.p, t->wbox.p, r, inv_t2,
                       t->i64a.p, t->i64b.p, nullptr, nullptr, nullptr, nullptr);
    SPHK_TRY(hipGetLastError());
    int64_t tot[2] = {0, 0};
    if (nt > std::numeric_limits<int>::max()) return set_error(SPHK_E_CAPACITY, "dtree: too many walk subtrees");
    if (int rc = scan64(t, t->i64a.p, t->i64c.p, (int)nt)) return rc;
    hipLaunchKernelGGL(offsets_kernel, dim3(grid(ng)), dim3(kBlock), 0, t->st, ng, nf, t->i64c.p, t->i64a.p,
                       L.j_off.p, t->total.p);
    if (int rc = read1(t->st, t->total.p, &tot[0])) return rc;
    if (MODE == 2) {
        if (int rc = scan64(t, t->i64b.p, t->i64d.p, (int)nt)) return rc;
        hipLaunchKernelGGL(offsets_kernel, dim3(grid(ng)), dim3(kBlock), 0, t->st, ng, nf, t->i64d.p, t->i64b.p,
                           L.s_off.p, t->total.p + 1);
        if (int rc = read1(t->st, t->total.p + 1, &tot[1])) return rc;
    }
    if (tot[0] > std::numeric_limits<int32_t>::max() || tot[1] > std::numeric_limits<int32_t>::max())
        return set_error(SPHK_E_CAPACITY, "dtree: list longer than 2^31 entries");
    SPHK_TRY(L.j_idx.reserve(std::max<int64_t>(tot[0], 1), false, t->st));
    if (MODE == 2) SPHK_TRY(L.s_idx.reserve(std::max<int64_t>(tot[1], 1), false, t->st));
    hipLaunchKernelGGL((list_kernel<MODE, true>), lg, lb, 0, t->st, ng, nf, v, t->front.p, t->wbox.p, r, inv_t2,
                       nullptr, nullptr, t->i64c.p, MODE == 2 ? t->i64d.p : nullptr, L.j_idx.p,
                       MODE == 2 ? L.s_idx.p : nullptr);
    SPHK_TRY(hipGetLastError());
    L.n_j = tot[0];
    L.n_s = MODE == 2 ? tot[1] : 0;
    L.built = true;
    return SPHK_OK;
}

// Short-range lists in two stages: the split tree walk lists each walk's
// candidate leaves (count, scan, fill), then member_kernel filters their
// particles (count, scan, fill).  Same output as the one-stage walk.
template <int MODE>
int build_list_two_stage(sphk_dtree *t, const double *r)
{
    DLists &L = t->lists[MODE];
    L.built = false;
    const int ng = t->n_walk, nf = t->n_front;
    const int64_t nt = (int64_t)ng * nf;
    if (nt > std::numeric_limits<int>::max()) return set_error(SPHK_E_CAPACITY, "dtree: too many walk subtrees");
    SPHK_TRY(t->i64a.reserve(nt, false, t->st));   // leaf counts per (walk, frontier node)
    SPHK_TRY(t->i64c.reserve(nt, false, t->st));   // their scan
    SPHK_TRY(t->total.reserve(2, false, t->st));
    SPHK_TRY(t->leaf_off.reserve(ng + 1, false, t->st));
    SPHK_TRY(L.j_off.reserve(ng + 1, false, t->st));
    SPHK_TRY(t->wbox.reserve(6 * (size_t)ng, false, t->st));
    hipLaunchKernelGGL(walk_box_kernel, dim3(grid(ng)), dim3(kBlock), 0, t->st, ng, t->i_off.p, t->x.p, t->y.p,
                       t->z.p, r, t->wbox.p);
    TreeView v{t->nodes(), t->moments(), t->brec.p, t->mrec.p, t->orec.p, t->x.p, t->y.p, t->z.p, t->group.p,
               t->i_off.p};
    const dim3 lg((unsigned)((nt + 63) / 64)), lb(64);
    hipLaunchKernelGGL((list_kernel<MODE, false, true>), lg, lb, 0, t->st, ng, nf, v, t->front.p, t->wbox.p, r, 0.,
                       t->i64a.p, nullptr, nullptr, nullptr, nullptr, nullptr);
    SPHK_TRY(hipGetLastError());
    if (int rc = scan64(t, t->i64a.p, t->i64c.p, (int)nt)) return rc;
    hipLaunchKernelGGL(offsets_kernel, dim3(grid(ng)), dim3(kBlock), 0, t->st, ng, nf, t->i64c.p, t->i64a.p,
                       t->leaf_off.p, t->total.p);
    int64_t n_leaves = 0;
    if (int rc = read1(t->st, t->total.p, &n_leaves)) return rc;
    if (n_leaves > std::numeric_limits<int32_t>::max()) return set_error(SPHK_E_CAPACITY, "dtree: too many leaves");
    SPHK_TRY(t->leaf_idx.reserve(std::max<int64_t>(n_leaves, 1), false, t->st));
    hipLaunchKernelGGL((list_kernel<MODE, true, true>), lg, lb, 0, t->st, ng, nf, v, t->front.p, t->wbox.p, r, 0.,
                       nullptr, nullptr, t->i64c.p, nullptr, t->leaf_idx.p, nullptr);
    SPHK_TRY(hipGetLastError());
    // stage 2: exact membership, one wave per walk, in one pass: the accepted
    // particles go to the walk's stretch of a scratch array (room for all its
    // candidates), then are copied to their offsets in the list
    SPHK_TRY(t->i64b.reserve(ng, false, t->st));
    SPHK_TRY(t->i64d.reserve(ng + 1, false, t->st));
    hipLaunchKernelGGL(cand_total_kernel, dim3(ng), dim3(64), 0, t->st, ng, t->leaf_off.p, t->leaf_idx.p,
                       t->nodes(), t->i64b.p);
    if (int rc = scan64(t, t->i64b.p, t->i64d.p, ng)) return rc;
    int64_t n_cand = 0, last[2];
    if (hipMemcpyAsync(&last[0], t->i64d.p + ng - 1, sizeof(int64_t), hipMemcpyDeviceToHost, t->st) != hipSuccess ||
        hipMemcpyAsync(&last[1], t->i64b.p + ng - 1, sizeof(int64_t), hipMemcpyDeviceToHost, t->st) != hipSuccess ||
        hipStreamSynchronize(t->st) != hipSuccess)
        return set_error(SPHK_E_DEVICE, "dtree: candidate count read failed");
    n_cand = last[0] + last[1];
    SPHK_TRY(t->mscr.reserve(std::max<int64_t>(n_cand, 1), false, t->st));
    hipLaunchKernelGGL((member_kernel<MODE>), dim3(ng), dim3(64), 0, t->st, ng, v, t->leaf_off.p, t->leaf_idx.p,
                       t->wbox.p, r, t->i64b.p, t->i64d.p, t->mscr.p);
    SPHK_TRY(hipGetLastError());
    // i64b now holds the accepted counts; their scan gives the list offsets
    SPHK_TRY(t->i64c.reserve(ng, false, t->st));
    if (int rc = scan64(t, t->i64b.p, t->i64c.p, ng)) return rc;
    hipLaunchKernelGGL(offsets_kernel, dim3(grid(ng)), dim3(kBlock), 0, t->st, ng, 1, t->i64c.p, t->i64b.p,
                       L.j_off.p, t->total.p);
    int64_t tot = 0;
    if (int rc = read1(t->st, t->total.p, &tot)) return rc;
    if (tot > std::numeric_limits<int32_t>::max()) return set_error(SPHK_E_CAPACITY, "dtree: list longer than 2^31 entries");
    SPHK_TRY(L.j_idx.reserve(std::max<int64_t>(tot, 1), false, t->st));
    hipLaunchKernelGGL(member_copy_kernel, dim3(ng), dim3(64), 0, t->st, ng, t->i64d.p, L.j_off.p, t->mscr.p,
                       L.j_idx.p);
    SPHK_TRY(hipGetLastError());
    L.n_j = tot;
    L.n_s = 0;
    L.built = true;
    return SPHK_OK;
}

}  // namespace

extern "C" {

int sphk_morton_order_dev(int n, const double *x, const double *y, const double *z, int32_t *perm, void *stream)
{
    if (n < 0 || (n > 0 && (!x || !y || !z || !perm))) return set_error(SPHK_E_ARG, "morton_dev: bad args");
    if (n == 0) return SPHK_OK;
    hipStream_t st = (hipStream_t)stream;
    DArr<double> part, frame;
    DArr<uint64_t> key, key2;
    DArr<int32_t> idx;
    DArr<char> tmp;
    int rc = SPHK_OK;
    do {
        if (idx.reserve(n, false, st) != hipSuccess || key2.reserve(n, false, st) != hipSuccess) {
            rc = set_error(SPHK_E_DEVICE, "morton_dev: allocation failed");
            break;
        }
        if ((rc = make_keys(n, x, y, z, st, part, frame, key, idx.p))) break;
        size_t bytes = 0;
        hipError_t e = rocprim::radix_sort_pairs(nullptr, bytes, key.p, key2.p, idx.p, perm, (size_t)n, 0, 63, st);
        if (e == hipSuccess) e = tmp.reserve(bytes + 16, false, st);
        if (e == hipSuccess)
            e = rocprim::radix_sort_pairs(tmp.p, bytes, key.p, key2.p, idx.p, perm, (size_t)n, 0, 63, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) rc = set_error(SPHK_E_DEVICE, "morton_dev: %s", hipGetErrorString(e));
    } while (0);
    part.release();
    frame.release();
    key.release();
    key2.release();
    idx.release();
    tmp.release();
    return rc;
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
    t->n_front = 0;
    for (auto &L : t->lists) L.built = false;
    auto fail = [&](int rc) {
        t->n_walk = 0;
        t->n_node = 0;
        t->n_front = 0;
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
        if (fb + nnew > std::numeric_limits<int32_t>::max())
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
    if (t->brec.reserve(nn, false, st) != hipSuccess || t->mrec.reserve(nn, false, st) != hipSuccess ||
        t->nextp.reserve(nn, false, st) != hipSuccess)
        return fail(set_error(SPHK_E_DEVICE, "dtree: allocation failed"));
    for (int L = 0; L + 1 < (int)t->level_start.size(); L++) {
        const int a0 = t->level_start[L], na = t->level_start[L + 1] - a0;
        hipLaunchKernelGGL(next_kernel, dim3(grid(na)), dim3(kBlock), 0, st, a0, na, t->nodes(), t->nextp.p);
    }
    hipLaunchKernelGGL(pack_nodes_kernel, dim3(grid((int)nn)), dim3(kBlock), 0, st, (int)nn, t->nodes(),
                       t->moments(), t->nextp.p, t->brec.p, t->mrec.p);

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
    if (hipStreamSynchronize(st) != hipSuccess || hipGetLastError() != hipSuccess)
        return fail(set_error(SPHK_E_DEVICE, "dtree: build failed"));
    t->n_walk = ng;

    // frontier for the split list walks: the first level with >= 64 nodes
    // (or the deepest), plus the leaves above it
    int L = (int)t->level_start.size() - 2;
    for (int q = 1; q + 1 < (int)t->level_start.size(); q++)
        if (t->level_start[q + 1] - t->level_start[q] >= 64) {
            L = q;
            break;
        }
    hipLaunchKernelGGL(fill_kernel, dim3(grid(n)), dim3(kBlock), 0, st, n, t->head.p, -1);
    hipLaunchKernelGGL(front_mark_kernel, dim3(grid((int)nn)), dim3(kBlock), 0, st, (int)nn, t->nodes(), L,
                       t->head.p);
    hipLaunchKernelGGL(head_flag_kernel, dim3(grid(n)), dim3(kBlock), 0, st, n, t->head.p, t->i64a.p);
    if (int rc = scan64(t, t->i64a.p, t->i64c.p, n)) return fail(rc);
    if (hipMemcpyAsync(&last[0], t->i64c.p + n - 1, sizeof(int64_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(&last[1], t->i64a.p + n - 1, sizeof(int64_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return fail(set_error(SPHK_E_DEVICE, "dtree: frontier count read failed"));
    t->n_front = (int)(last[0] + last[1]);
    if (t->front.reserve(t->n_front, false, st) != hipSuccess)
        return fail(set_error(SPHK_E_DEVICE, "dtree: allocation failed"));
    hipLaunchKernelGGL(front_write_kernel, dim3(grid(n)), dim3(kBlock), 0, st, n, t->head.p, t->i64c.p, t->front.p);
    if (hipStreamSynchronize(st) != hipSuccess || hipGetLastError() != hipSuccess)
        return fail(set_error(SPHK_E_DEVICE, "dtree: frontier failed"));
    return SPHK_OK;
}

int check_build_args(int n, const double *x, const double *y, const double *z, const double *m, int n_leaf,
                     int n_group)
{
    if (n <= 0 || !x || !y || !z || !m) return set_error(SPHK_E_ARG, "dtree: bad particle arrays");
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

int sphk_dtree_list(const sphk_dtree *t, int kind, const int32_t **j_off, const int32_t **j_idx,
                    const int32_t **s_off, const int32_t **s_idx, int64_t *n_j, int64_t *n_s)
{
    if (!t || kind < 0 || kind > 2) return set_error(SPHK_E_ARG, "dtree list: bad argument");
    const DLists &L = t->lists[kind];
    if (!L.built) return set_error(SPHK_E_STATE, "dtree list kind %d not built", kind);
    if (j_off) *j_off = L.j_off.p;
    if (j_idx) *j_idx = L.j_idx.p;
    if (s_off) *s_off = kind == 2 ? L.s_off.p : nullptr;
    if (s_idx) *s_idx = kind == 2 ? L.s_idx.p : nullptr;
    if (n_j) *n_j = L.n_j;
    if (n_s) *n_s = L.n_s;
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

extern "C" {

int sphk_copy_dev(void *dst, const void *src, size_t bytes, void *stream)
{
    if (bytes == 0) return SPHK_OK;
    if (!dst || !src) return set_error(SPHK_E_ARG, "copy_dev: null pointer");
    SPHK_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return SPHK_OK;
}

}  // extern "C"
