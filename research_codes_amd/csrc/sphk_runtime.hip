// sphk_runtime.hip -- context, buffers and the plain-C entry points of
// libsphk.so (declared in include/sphk.h).
//
// Host dispatch path (the reference's Dispatch/Retrieve model,
// sph_yamauchi/run/hdr_densityGPU.cu:496-621): the caller's packed SoA host
// arrays are gathered into one pinned staging block, moved with ONE async
// H2D copy into one device block per tag, the kernel is launched on the
// context stream, and retrieve brings all outputs back with one D2H copy.
// Device-resident path: the caller's device arrays are used in place.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "sphk.h"
#include "sphk_internal.hpp"

namespace sphk {

static thread_local char g_err[512] = "";

int set_error(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

#define SPHK_HIP(call)                                                                       \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return set_error(SPHK_E_DEVICE, "%s: %s (%s:%d)", #call, hipGetErrorString(e_),  \
                             __FILE__, __LINE__);                                            \
    } while (0)

static size_t align256(size_t n) { return (n + 255) & ~size_t(255); }

struct DevBlock {
    char *p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t n)
    {
        if (n <= cap) return hipSuccess;
        if (p) {
            hipError_t e = hipFree(p);
            if (e != hipSuccess) return e;
            p = nullptr;
            cap = 0;
        }
        size_t c = n + n / 4;
        hipError_t e = hipMalloc(&p, c);
        if (e == hipSuccess) cap = c;
        return e;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

struct HostBlock {
    char *p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t n)
    {
        if (n <= cap) return hipSuccess;
        if (p) {
            hipError_t e = hipHostFree(p);
            if (e != hipSuccess) return e;
            p = nullptr;
            cap = 0;
        }
        size_t c = n + n / 4;
        hipError_t e = hipHostMalloc(&p, c, hipHostMallocDefault);
        if (e == hipSuccess) cap = c;
        return e;
    }
    void release()
    {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
};

// Field tables: every src/out struct is a sequence of pointer members; these
// give each member's element size in bytes.
static const size_t kDensSrcSz[] = {4, 8, 8, 8, 8, 8, 8, 8, 8, 8};
static const size_t kDensOutSz[] = {8, 8, 8, 8, 8, 8, 1};
static const size_t kHydroSrcSz[] = {4, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8};
static const size_t kHydroOutSz[] = {8, 8, 8, 8, 8, 8, 8, 8, 8};
static const size_t kGravSrcSz[] = {8, 8, 8, 8, 8};
static const size_t kGravOutSz[] = {8, 8, 8, 8, 8};
static_assert(sizeof(sphk_dens_src) == sizeof(kDensSrcSz) / sizeof(size_t) * sizeof(void *), "");
static_assert(sizeof(sphk_dens_out) == sizeof(kDensOutSz) / sizeof(size_t) * sizeof(void *), "");
static_assert(sizeof(sphk_hydro_src) == sizeof(kHydroSrcSz) / sizeof(size_t) * sizeof(void *), "");
static_assert(sizeof(sphk_hydro_out) == sizeof(kHydroOutSz) / sizeof(size_t) * sizeof(void *), "");
static_assert(sizeof(sphk_grav_src) == sizeof(kGravSrcSz) / sizeof(size_t) * sizeof(void *), "");
static_assert(sizeof(sphk_grav_out) == sizeof(kGravOutSz) / sizeof(size_t) * sizeof(void *), "");

template <typename S>
static void **fields(S &s) { return reinterpret_cast<void **>(&s); }
template <typename S>
static void *const *fields(const S &s) { return reinterpret_cast<void *const *>(&s); }

enum Kind { kNone = -1, kDensity = 0, kHydro = 1, kGravity = 2 };

struct TagSlot {
    int kind = kNone;
    int n_walk = 0;
    size_t ni_tot = 0;
    DevBlock din, dout;
    HostBlock hin, hout;
    // device-side views of the output block, in field order
    std::vector<size_t> out_off;
    const size_t *out_sz = nullptr;
    int n_out = 0;
};

struct TimeRec {
    int which;
    hipEvent_t a, b;
};

}  // namespace sphk

// Walk queues of persistent launches: each launch takes the next of
// kQueueSlots zeroed queues (the launch leaves it zeroed), so up to that many
// persistent launches of one context may be in flight at once.
constexpr int kQueueSlots = 64;

struct sphk_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    sphk_params params{1.6, 1.936492, 1e30, 0, 0, 0, 0};
    int *queue = nullptr;          // kQueueSlots * kQueueInts ints on the device
    unsigned queue_next = 0;
    sphk::TagSlot tags[SPHK_TAG_MAX];
    bool timing = false;
    std::vector<sphk::TimeRec> pending;
    double total_ms[3] = {0, 0, 0};
    int count[3] = {0, 0, 0};
    std::mutex mu;
};

using namespace sphk;

namespace {

struct Timer {
    sphk_ctx *ctx;
    int which;
    hipStream_t st;
    hipEvent_t a = nullptr, b = nullptr;
    Timer(sphk_ctx *c, int w, hipStream_t s) : ctx(c), which(w), st(s)
    {
        if (ctx->timing) {
            if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
                a = b = nullptr;
                return;
            }
            (void)hipEventRecord(a, st);
        }
    }
    void stop()
    {
        if (a && b) {
            (void)hipEventRecord(b, st);
            std::lock_guard<std::mutex> g(ctx->mu);
            ctx->pending.push_back({which, a, b});
            a = b = nullptr;
        }
    }
    ~Timer()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

// Device-resident launches run on exactly the stream the caller names; NULL is
// the HIP null stream (what a caller's default stream, e.g. torch's, maps to).
hipStream_t pick_stream(sphk_ctx *, void *stream) { return reinterpret_cast<hipStream_t>(stream); }

int check_ctx(sphk_ctx *ctx)
{
    if (!ctx) return set_error(SPHK_E_ARG, "null context");
    return SPHK_OK;
}

int check_precision(int precision)
{
    if (precision != SPHK_FP64 && precision != SPHK_MIXED)
        return set_error(SPHK_E_ARG, "precision must be SPHK_FP64 or SPHK_MIXED (got %d)", precision);
    return SPHK_OK;
}

int check_tag(sphk_ctx *ctx, int tag)
{
    if (tag < 0 || tag >= SPHK_TAG_MAX)
        return set_error(SPHK_E_STATE, "tag %d outside [0,%d)", tag, SPHK_TAG_MAX);
    (void)ctx;
    return SPHK_OK;
}

// Prefix sums of per-walk counts; validates sizes against the limits.
int prefix(int n_walk, const int32_t *cnt, std::vector<int32_t> &off, size_t limit, const char *what)
{
    off.assign(n_walk + 1, 0);
    int64_t s = 0;
    for (int w = 0; w < n_walk; w++) {
        if (cnt[w] < 0) return set_error(SPHK_E_ARG, "%s[%d] = %d < 0", what, w, cnt[w]);
        off[w] = (int32_t)s;
        s += cnt[w];
        if ((size_t)s > limit)
            return set_error(SPHK_E_CAPACITY, "sum of %s exceeds %zu", what, limit);
    }
    off[n_walk] = (int32_t)s;
    return SPHK_OK;
}

// Stage: offsets arrays followed by each non-null field of each src struct
// (caller: the field is one of the caller's arrays).
struct StageItem {
    const void *src;
    size_t bytes;
    size_t off;
    bool caller = false;
};

// One device block per tag holds the items.  They go through the tag's
// pinned staging block in one transfer, or, for the caller's arrays when the
// caller declared them pinned (params.pinned_src), straight from them.
int stage_and_copy(sphk_ctx *ctx, TagSlot &slot, std::vector<StageItem> &items)
{
    const bool direct = ctx->params.pinned_src != 0;
    size_t total = 0, staged = 0;
    for (auto &it : items) {
        it.off = total;
        total += align256(it.bytes);
        if (!(direct && it.caller)) staged = total;   // staged items lead the block
    }
    SPHK_HIP(hipStreamSynchronize(ctx->stream));  // staging block may still be in flight
    SPHK_HIP(slot.hin.ensure(total));
    SPHK_HIP(slot.din.ensure(total));
    for (auto &it : items)
        if (it.bytes && !(direct && it.caller)) std::memcpy(slot.hin.p + it.off, it.src, it.bytes);
    if (staged) SPHK_HIP(hipMemcpyAsync(slot.din.p, slot.hin.p, staged, hipMemcpyHostToDevice, ctx->stream));
    if (direct)
        for (auto &it : items)
            if (it.caller && it.bytes)
                SPHK_HIP(hipMemcpyAsync(slot.din.p + it.off, it.src, it.bytes, hipMemcpyHostToDevice, ctx->stream));
    return SPHK_OK;
}

template <typename S>
int add_fields(std::vector<StageItem> &items, const S &s, const size_t *sz, int nf, size_t rows,
               std::vector<int> &slot_of)
{
    void *const *f = fields(s);
    for (int k = 0; k < nf; k++) {
        if (f[k]) {
            slot_of.push_back((int)items.size());
            items.push_back({f[k], sz[k] * rows, 0, true});
        } else {
            slot_of.push_back(-1);
        }
    }
    return SPHK_OK;
}

template <typename S>
S device_view(const TagSlot &slot, const std::vector<StageItem> &items, const std::vector<int> &slot_of,
              int first)
{
    S d;
    void **f = fields(d);
    for (size_t k = 0; k < sizeof(S) / sizeof(void *); k++) {
        int it = slot_of[first + k];
        f[k] = it >= 0 ? (void *)(slot.din.p + items[it].off) : nullptr;
    }
    return d;
}

template <typename O>
int prepare_out(TagSlot &slot, const size_t *sz, int nf, size_t rows, O &dview)
{
    size_t total = 0;
    slot.out_off.assign(nf, 0);
    for (int k = 0; k < nf; k++) {
        slot.out_off[k] = total;
        total += align256(sz[k] * (rows ? rows : 1));
    }
    SPHK_HIP(slot.dout.ensure(total));
    SPHK_HIP(slot.hout.ensure(total));
    slot.out_sz = sz;
    slot.n_out = nf;
    void **f = fields(dview);
    for (int k = 0; k < nf; k++) f[k] = slot.dout.p + slot.out_off[k];
    return SPHK_OK;
}

template <typename O>
int retrieve_common(sphk_ctx *ctx, int tag, int kind, int n_walk, const int32_t *nip, O &out)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if ((rc = check_tag(ctx, tag))) return rc;
    TagSlot &slot = ctx->tags[tag];
    if (slot.kind != kind)
        return set_error(SPHK_E_STATE, "retrieve on tag %d without a matching dispatch", tag);
    if (n_walk != slot.n_walk)
        return set_error(SPHK_E_ARG, "retrieve n_walk %d != dispatched %d", n_walk, slot.n_walk);
    size_t ni = 0;
    for (int w = 0; w < n_walk; w++) ni += (size_t)nip[w];
    if (ni != slot.ni_tot) return set_error(SPHK_E_ARG, "retrieve sum(ni) %zu != %zu", ni, slot.ni_tot);
    size_t total = 0;
    for (int k = 0; k < slot.n_out; k++) total = slot.out_off[k] + align256(slot.out_sz[k] * (ni ? ni : 1));
    SPHK_HIP(hipMemcpyAsync(slot.hout.p, slot.dout.p, total, hipMemcpyDeviceToHost, ctx->stream));
    SPHK_HIP(hipStreamSynchronize(ctx->stream));
    void **f = fields(out);
    for (int k = 0; k < slot.n_out; k++)
        if (f[k]) std::memcpy(f[k], slot.hout.p + slot.out_off[k], slot.out_sz[k] * ni);
    slot.kind = kNone;
    return SPHK_OK;
}

}  // namespace

extern "C" {

const char *sphk_last_error(void) { return g_err; }

int sphk_init(int my_proc, sphk_ctx **out)
{
    if (!out) return set_error(SPHK_E_ARG, "null output pointer");
    *out = nullptr;
    int ndev = 0;
    SPHK_HIP(hipGetDeviceCount(&ndev));
    if (ndev <= 0) return set_error(SPHK_E_DEVICE, "no HIP device");
    const int dev = (my_proc < 0 ? 0 : my_proc) % ndev;  // gpu_info.cu:20-25
    SPHK_HIP(hipSetDevice(dev));
    sphk_ctx *ctx = new sphk_ctx();
    ctx->device = dev;
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        return set_error(SPHK_E_DEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    const size_t qbytes = (size_t)kQueueSlots * kQueueInts * sizeof(int);
    e = hipMalloc(&ctx->queue, qbytes);
    if (e == hipSuccess) e = hipMemset(ctx->queue, 0, qbytes);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        if (ctx->queue) (void)hipFree(ctx->queue);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return set_error(SPHK_E_DEVICE, "walk queues: %s", hipGetErrorString(e));
    }
    if (getenv("SPHK_VERBOSE")) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess)
            fprintf(stderr, "sphk: %d device(s); using %d: %s (%s)\n", ndev, dev, prop.name,
                    prop.gcnArchName);
    }
    *out = ctx;
    return SPHK_OK;
}

int sphk_finalize(sphk_ctx *ctx)
{
    if (!ctx) return SPHK_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &r : ctx->pending) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (auto &t : ctx->tags) {
        t.din.release();
        t.dout.release();
        t.hin.release();
        t.hout.release();
    }
    if (ctx->queue) (void)hipFree(ctx->queue);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return SPHK_OK;
}

int sphk_set_params(sphk_ctx *ctx, const sphk_params *p)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!p) return set_error(SPHK_E_ARG, "null params");
    if (!(p->eta > 0) || !(p->ksrh > 0) || !(p->hmax > 0))
        return set_error(SPHK_E_ARG, "params must be positive (eta %g ksrh %g hmax %g)", p->eta,
                         p->ksrh, p->hmax);
    if (p->list_cap != 0 && p->list_cap != 192 && p->list_cap != 256)
        return set_error(SPHK_E_ARG, "list_cap must be 0, 192 or 256 (got %d)", p->list_cap);
    if (p->persistent != 0 && p->persistent != 1)
        return set_error(SPHK_E_ARG, "persistent must be 0 or 1 (got %d)", p->persistent);
    if (p->pinned_src != 0 && p->pinned_src != 1)
        return set_error(SPHK_E_ARG, "pinned_src must be 0 or 1 (got %d)", p->pinned_src);
    ctx->params = *p;
    return SPHK_OK;
}

int sphk_get_params(const sphk_ctx *ctx, sphk_params *p)
{
    if (!ctx || !p) return set_error(SPHK_E_ARG, "null argument");
    *p = ctx->params;
    return SPHK_OK;
}

int sphk_device(const sphk_ctx *ctx) { return ctx ? ctx->device : -1; }
void *sphk_stream(const sphk_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int sphk_synchronize(sphk_ctx *ctx)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    SPHK_HIP(hipStreamSynchronize(ctx->stream));
    return SPHK_OK;
}

// ---- device-resident ----------------------------------------------------

namespace {
// The walk queue of the next launch: nullptr for one workgroup per walk.
int *walk_queue(sphk_ctx *ctx)
{
    if (!ctx->params.persistent) return nullptr;
    std::lock_guard<std::mutex> lock(ctx->mu);
    const unsigned slot = ctx->queue_next++ % kQueueSlots;
    return ctx->queue + (size_t)slot * kQueueInts;
}
}  // namespace

int sphk_density_dev(sphk_ctx *ctx, int n_walk, const int32_t *i_off, const int32_t *j_off,
                     const int32_t *j_idx, sphk_dens_src pi, sphk_dens_src pj, sphk_dens_out out,
                     int precision, void *stream)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if ((rc = check_precision(precision))) return rc;
    if (n_walk < 0 || (n_walk > 0 && (!i_off || !j_off)))
        return set_error(SPHK_E_ARG, "density: bad walk table");
    if (n_walk == 0) return SPHK_OK;
    if (!pi.id || !pi.m || !pi.x || !pi.y || !pi.z || !pi.vx || !pi.vy || !pi.vz || !pi.ksr ||
        !pi.rs || !pj.id || !pj.m || !pj.x || !pj.y || !pj.z || !pj.vx || !pj.vy || !pj.vz ||
        !out.dens || !out.divv || !out.rotv || !out.grdh || !out.ksr || !out.np || !out.itr)
        return set_error(SPHK_E_ARG, "density: null field pointer");
    hipStream_t st = pick_stream(ctx, stream);
    Timer t(ctx, 0, st);
    SPHK_HIP(launch_density(st, n_walk, i_off, j_off, j_idx, pi, pj, out,
                            ctx->params.eta * ctx->params.ksrh, ctx->params.hmax,
                            precision == SPHK_MIXED, ctx->params.list_cap == 192 ? 192 : 256,
                            walk_queue(ctx)));
    t.stop();
    return SPHK_OK;
}

int sphk_hydro_dev(sphk_ctx *ctx, int n_walk, const int32_t *i_off, const int32_t *j_off,
                   const int32_t *j_idx, sphk_hydro_src pi, sphk_hydro_src pj, sphk_hydro_out out,
                   int precision, void *stream)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if ((rc = check_precision(precision))) return rc;
    if (n_walk < 0 || (n_walk > 0 && (!i_off || !j_off)))
        return set_error(SPHK_E_ARG, "hydro: bad walk table");
    if (n_walk == 0) return SPHK_OK;
    void *const *fi = fields(pi);
    void *const *fj = fields(pj);
    for (int k = 0; k < 18; k++) {
        if (!fj[k]) return set_error(SPHK_E_ARG, "hydro: null j field %d", k);
        if (k != 1 && !fi[k]) return set_error(SPHK_E_ARG, "hydro: null i field %d", k);
    }
    void *const *fo = fields(out);
    for (int k = 0; k < 9; k++)
        if (!fo[k]) return set_error(SPHK_E_ARG, "hydro: null output field %d", k);
    hipStream_t st = pick_stream(ctx, stream);
    Timer t(ctx, 1, st);
    SPHK_HIP(launch_hydro(st, n_walk, i_off, j_off, j_idx, pi, pj, out, precision == SPHK_MIXED,
                          walk_queue(ctx)));
    t.stop();
    return SPHK_OK;
}

int sphk_gravity_dev(sphk_ctx *ctx, int n_walk, const int32_t *i_off, const int32_t *j_off,
                     const int32_t *j_idx, sphk_grav_src pj, const int32_t *s_off,
                     const int32_t *s_idx, sphk_grav_src ps, sphk_grav_src pi, sphk_grav_out out,
                     int precision, void *stream)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if ((rc = check_precision(precision))) return rc;
    if (n_walk < 0 || (n_walk > 0 && (!i_off || !j_off)))
        return set_error(SPHK_E_ARG, "gravity: bad walk table");
    if (n_walk == 0) return SPHK_OK;
    if (!pi.x || !pi.y || !pi.z || !pi.eps2 || !pj.m || !pj.x || !pj.y || !pj.z || !pj.eps2)
        return set_error(SPHK_E_ARG, "gravity: null field pointer");
    if (s_off && (!ps.m || !ps.x || !ps.y || !ps.z || !ps.eps2))
        return set_error(SPHK_E_ARG, "gravity: null SPJ field pointer");
    if (!out.ax || !out.ay || !out.az || !out.pot)   // out.eta may be NULL: not formed
        return set_error(SPHK_E_ARG, "gravity: null output pointer");
    hipStream_t st = pick_stream(ctx, stream);
    Timer t(ctx, 2, st);
    SPHK_HIP(launch_gravity(st, n_walk, i_off, j_off, j_idx, pj, s_off, s_idx, ps, pi, out,
                            precision == SPHK_MIXED, ctx->params.uniform_softening != 0, walk_queue(ctx)));
    t.stop();
    return SPHK_OK;
}

// ---- host dispatch / retrieve -------------------------------------------

int sphk_density_dispatch(sphk_ctx *ctx, int tag, int n_walk, const int32_t *nip,
                          const int32_t *njp, sphk_dens_src pi, sphk_dens_src pj, int precision)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if ((rc = check_precision(precision))) return rc;
    if ((rc = check_tag(ctx, tag))) return rc;
    if (n_walk < 0 || n_walk > SPHK_WALK_LIMIT)
        return set_error(SPHK_E_CAPACITY, "n_walk %d outside [0,%d]", n_walk, SPHK_WALK_LIMIT);
    if (n_walk > 0 && (!nip || !njp)) return set_error(SPHK_E_ARG, "null nip/njp");
    std::vector<int32_t> ioff, joff;
    if ((rc = prefix(n_walk, nip, ioff, SPHK_NI_LIMIT, "nip"))) return rc;
    if ((rc = prefix(n_walk, njp, joff, SPHK_NJ_LIMIT, "njp"))) return rc;
    const size_t ni = ioff[n_walk], nj = joff[n_walk];
    TagSlot &slot = ctx->tags[tag];
    std::vector<StageItem> items;
    std::vector<int> slot_of;
    items.push_back({ioff.data(), ioff.size() * 4, 0});
    items.push_back({joff.data(), joff.size() * 4, 0});
    add_fields(items, pi, kDensSrcSz, 10, ni, slot_of);
    add_fields(items, pj, kDensSrcSz, 10, nj, slot_of);
    if ((rc = stage_and_copy(ctx, slot, items))) return rc;
    sphk_dens_src di = device_view<sphk_dens_src>(slot, items, slot_of, 0);
    sphk_dens_src dj = device_view<sphk_dens_src>(slot, items, slot_of, 10);
    sphk_dens_out dout;
    if ((rc = prepare_out(slot, kDensOutSz, 7, ni, dout))) return rc;
    slot.kind = kDensity;
    slot.n_walk = n_walk;
    slot.ni_tot = ni;
    return sphk_density_dev(ctx, n_walk, (const int32_t *)(slot.din.p + items[0].off),
                            (const int32_t *)(slot.din.p + items[1].off), nullptr, di, dj, dout,
                            precision, (void *)ctx->stream);
}

int sphk_density_retrieve(sphk_ctx *ctx, int tag, int n_walk, const int32_t *nip, sphk_dens_out out)
{
    return retrieve_common(ctx, tag, kDensity, n_walk, nip, out);
}

int sphk_hydro_dispatch(sphk_ctx *ctx, int tag, int n_walk, const int32_t *nip, const int32_t *njp,
                        sphk_hydro_src pi, sphk_hydro_src pj, int precision)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if ((rc = check_precision(precision))) return rc;
    if ((rc = check_tag(ctx, tag))) return rc;
    if (n_walk < 0 || n_walk > SPHK_WALK_LIMIT)
        return set_error(SPHK_E_CAPACITY, "n_walk %d outside [0,%d]", n_walk, SPHK_WALK_LIMIT);
    if (n_walk > 0 && (!nip || !njp)) return set_error(SPHK_E_ARG, "null nip/njp");
    std::vector<int32_t> ioff, joff;
    if ((rc = prefix(n_walk, nip, ioff, SPHK_NI_LIMIT, "nip"))) return rc;
    if ((rc = prefix(n_walk, njp, joff, SPHK_NJ_LIMIT, "njp"))) return rc;
    const size_t ni = ioff[n_walk], nj = joff[n_walk];
    TagSlot &slot = ctx->tags[tag];
    std::vector<StageItem> items;
    std::vector<int> slot_of;
    items.push_back({ioff.data(), ioff.size() * 4, 0});
    items.push_back({joff.data(), joff.size() * 4, 0});
    add_fields(items, pi, kHydroSrcSz, 18, ni, slot_of);
    add_fields(items, pj, kHydroSrcSz, 18, nj, slot_of);
    if ((rc = stage_and_copy(ctx, slot, items))) return rc;
    sphk_hydro_src di = device_view<sphk_hydro_src>(slot, items, slot_of, 0);
    sphk_hydro_src dj = device_view<sphk_hydro_src>(slot, items, slot_of, 18);
    sphk_hydro_out dout;
    if ((rc = prepare_out(slot, kHydroOutSz, 9, ni, dout))) return rc;
    slot.kind = kHydro;
    slot.n_walk = n_walk;
    slot.ni_tot = ni;
    return sphk_hydro_dev(ctx, n_walk, (const int32_t *)(slot.din.p + items[0].off),
                          (const int32_t *)(slot.din.p + items[1].off), nullptr, di, dj, dout,
                          precision, (void *)ctx->stream);
}

int sphk_hydro_retrieve(sphk_ctx *ctx, int tag, int n_walk, const int32_t *nip, sphk_hydro_out out)
{
    return retrieve_common(ctx, tag, kHydro, n_walk, nip, out);
}

int sphk_gravity_dispatch(sphk_ctx *ctx, int tag, int n_walk, const int32_t *nip,
                          const int32_t *njp, const int32_t *nsp, sphk_grav_src pi,
                          sphk_grav_src pj, sphk_grav_src ps, int precision)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if ((rc = check_precision(precision))) return rc;
    if ((rc = check_tag(ctx, tag))) return rc;
    if (n_walk < 0 || n_walk > SPHK_WALK_LIMIT)
        return set_error(SPHK_E_CAPACITY, "n_walk %d outside [0,%d]", n_walk, SPHK_WALK_LIMIT);
    if (n_walk > 0 && (!nip || !njp)) return set_error(SPHK_E_ARG, "null nip/njp");
    std::vector<int32_t> ioff, joff, soff;
    if ((rc = prefix(n_walk, nip, ioff, SPHK_NI_LIMIT, "nip"))) return rc;
    if ((rc = prefix(n_walk, njp, joff, SPHK_NJ_LIMIT, "njp"))) return rc;
    if (nsp && (rc = prefix(n_walk, nsp, soff, SPHK_NJ_LIMIT, "nsp"))) return rc;
    if (nsp && soff[n_walk] == 0) nsp = nullptr;  // no SPJ at all: same as no list
    const size_t ni = ioff[n_walk], nj = joff[n_walk], ns = nsp ? soff[n_walk] : 0;
    if ((size_t)nj + ns > (size_t)SPHK_NJ_LIMIT)
        return set_error(SPHK_E_CAPACITY, "sum(njp)+sum(nsp) exceeds %d", SPHK_NJ_LIMIT);
    TagSlot &slot = ctx->tags[tag];
    std::vector<StageItem> items;
    std::vector<int> slot_of;
    items.push_back({ioff.data(), ioff.size() * 4, 0});
    items.push_back({joff.data(), joff.size() * 4, 0});
    items.push_back({nsp ? soff.data() : nullptr, nsp ? soff.size() * 4 : 0, 0});
    add_fields(items, pi, kGravSrcSz, 5, ni, slot_of);
    add_fields(items, pj, kGravSrcSz, 5, nj, slot_of);
    sphk_grav_src ps_use = nsp ? ps : sphk_grav_src{nullptr, nullptr, nullptr, nullptr, nullptr};
    add_fields(items, ps_use, kGravSrcSz, 5, ns, slot_of);
    if ((rc = stage_and_copy(ctx, slot, items))) return rc;
    sphk_grav_src di = device_view<sphk_grav_src>(slot, items, slot_of, 0);
    sphk_grav_src dj = device_view<sphk_grav_src>(slot, items, slot_of, 5);
    sphk_grav_src ds = device_view<sphk_grav_src>(slot, items, slot_of, 10);
    sphk_grav_out dout;
    if ((rc = prepare_out(slot, kGravOutSz, 5, ni, dout))) return rc;
    slot.kind = kGravity;
    slot.n_walk = n_walk;
    slot.ni_tot = ni;
    return sphk_gravity_dev(ctx, n_walk, (const int32_t *)(slot.din.p + items[0].off),
                            (const int32_t *)(slot.din.p + items[1].off), nullptr, dj,
                            nsp ? (const int32_t *)(slot.din.p + items[2].off) : nullptr, nullptr,
                            ds, di, dout, precision, (void *)ctx->stream);
}

int sphk_gravity_retrieve(sphk_ctx *ctx, int tag, int n_walk, const int32_t *nip, sphk_grav_out out)
{
    return retrieve_common(ctx, tag, kGravity, n_walk, nip, out);
}

// ---- timing -------------------------------------------------------------

int sphk_timing(sphk_ctx *ctx, int enable)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    ctx->timing = enable != 0;
    return SPHK_OK;
}

static int harvest(sphk_ctx *ctx)
{
    std::lock_guard<std::mutex> g(ctx->mu);
    for (auto &r : ctx->pending) {
        SPHK_HIP(hipEventSynchronize(r.b));
        float ms = 0.f;
        SPHK_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        ctx->total_ms[r.which] += ms;
        ctx->count[r.which] += 1;
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    ctx->pending.clear();
    return SPHK_OK;
}

int sphk_kernel_time(sphk_ctx *ctx, int which, double *total_ms, int *count)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (which < 0 || which > 2 || !total_ms || !count) return set_error(SPHK_E_ARG, "bad argument");
    if ((rc = harvest(ctx))) return rc;
    *total_ms = ctx->total_ms[which];
    *count = ctx->count[which];
    return SPHK_OK;
}

int sphk_reset_kernel_time(sphk_ctx *ctx)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if ((rc = harvest(ctx))) return rc;
    for (int k = 0; k < 3; k++) {
        ctx->total_ms[k] = 0;
        ctx->count[k] = 0;
    }
    return SPHK_OK;
}

}  // extern "C"
