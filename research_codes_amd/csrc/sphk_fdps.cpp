// sphk_fdps.cpp -- the reference's six multiwalk functions and
// setup_gpu/reset_gpu (sph_yamauchi/run/use_gpu.hpp:16-64) over the plain-C
// ABI of libsphk.so.
//
// Dispatch copies each walk's AoS EPI/EPJ arrays (owned by FDPS, valid only
// during the call) into per-tag structure-of-arrays host buffers in pinned
// memory, walks in parallel (OpenMP), and calls sphk_*_dispatch, which
// copies them to the device straight from those buffers (the context is set
// to pinned_src; FDPS retrieves a tag before it dispatches on it again, so
// the buffers stay unchanged while a copy is in flight) and launches
// asynchronously.  Retrieve waits for
// the result and scatters it back into FDPS's per-walk Force arrays in walk
// order (as hdr_densityGPU.cu:594-621 does).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "sphk.h"
#include "sphk_fdps.hpp"
#include "sphk_fdps_types.hpp"

namespace RunParameter {
// Defined by the caller's program in the reference (hdr_run.hpp:24); weak so
// that libsphk.so loads in programs that do not define it.
extern double KernelSupportRadiusMaximum __attribute__((weak));
}  // namespace RunParameter

namespace {

sphk_ctx *g_ctx = nullptr;
int g_precision = SPHK_MIXED;
std::mutex g_mu;

void die(const char *where, int rc)
{
    std::fprintf(stderr, "sphk: %s failed (%d): %s\n", where, rc, sphk_last_error());
    std::abort();
}

void need_ctx()
{
    if (!g_ctx) {
        std::fprintf(stderr, "sphk: setup_gpu() has not been called\n");
        std::abort();
    }
}

// Host memory for the packed columns: pinned, so the dispatch copies them to
// the device directly.
template <typename T>
struct PinnedAlloc {
    using value_type = T;
    PinnedAlloc() = default;
    template <typename U>
    PinnedAlloc(const PinnedAlloc<U> &)
    {
    }
    T *allocate(size_t n)
    {
        void *p = nullptr;
        if (n > SIZE_MAX / sizeof(T) ||
            hipHostMalloc(&p, (n ? n : 1) * sizeof(T), hipHostMallocDefault) != hipSuccess)
            throw std::bad_alloc();
        return static_cast<T *>(p);
    }
    void deallocate(T *p, size_t)
    {
        if (p) (void)hipHostFree(p);
    }
};
template <typename T, typename U>
bool operator==(const PinnedAlloc<T> &, const PinnedAlloc<U> &)
{
    return true;
}
template <typename T, typename U>
bool operator!=(const PinnedAlloc<T> &, const PinnedAlloc<U> &)
{
    return false;
}
template <typename T>
using PinnedVec = std::vector<T, PinnedAlloc<T>>;

struct Cols {
    std::vector<PinnedVec<double>> f;
    PinnedVec<int32_t> id;
    void resize(size_t nf, size_t n)
    {
        f.resize(nf);
        for (auto &v : f) v.resize(n);
        id.resize(n);
    }
    const double *p(int k) const { return f[k].data(); }
};

struct TagBuf {
    Cols di, dj;   // density
    Cols hi, hj;   // hydro
    Cols gi, gj, gs;  // gravity
    std::vector<double> o[9];
    std::vector<int64_t> onp;
    std::vector<uint8_t> oitr;
};
// Never destroyed: its pinned buffers are freed by reset_gpu, or left to
// the process's end (not freed after the HIP runtime has shut down).
TagBuf *g_tag = new TagBuf[SPHK_TAG_MAX];

TagBuf &tagbuf(int tag)
{
    if (tag < 0 || tag >= SPHK_TAG_MAX) {
        std::fprintf(stderr, "sphk: tag %d outside [0,%d)\n", tag, SPHK_TAG_MAX);
        std::abort();
    }
    return g_tag[tag];
}

// Start of each walk's rows in the packed arrays (exclusive prefix sum).
std::vector<size_t> starts(int n_walk, const int *cnt)
{
    std::vector<size_t> s(n_walk + 1, 0);
    for (int w = 0; w < n_walk; w++) s[w + 1] = s[w] + (size_t)cnt[w];
    return s;
}

}  // namespace

extern "C" sphk_ctx *sphk_fdps_context(void) { return g_ctx; }

// gpu_info.cu:6-34
void setup_gpu(int my_proc)
{
    std::lock_guard<std::mutex> g(g_mu);
    if (g_ctx) return;
    int rc = sphk_init(my_proc, &g_ctx);
    if (rc) die("setup_gpu", rc);
    // The tag buffers are pinned: copied to the device directly, unless
    // SPHK_STAGED=1 asks for the context's staging blocks.
    const char *st = std::getenv("SPHK_STAGED");
    sphk_params prm;
    sphk_get_params(g_ctx, &prm);
    prm.pinned_src = (st && std::strcmp(st, "1") == 0) ? 0 : 1;
    if ((rc = sphk_set_params(g_ctx, &prm))) die("setup_gpu", rc);
    const char *pr = std::getenv("SPHK_PRECISION");
    g_precision = (pr && std::strcmp(pr, "fp64") == 0) ? SPHK_FP64 : SPHK_MIXED;
    if (my_proc == 0 && std::getenv("SPHK_VERBOSE"))
        std::fprintf(stderr, "sphk: device %d, precision %s\n", sphk_device(g_ctx),
                     g_precision == SPHK_FP64 ? "fp64" : "mixed");
}

// gpu_info.cu:36-43
void reset_gpu(int)
{
    std::lock_guard<std::mutex> g(g_mu);
    if (g_ctx) sphk_finalize(g_ctx);
    g_ctx = nullptr;
    for (int t = 0; t < SPHK_TAG_MAX; t++) g_tag[t] = TagBuf();
}

int DispatchKernel_Dens(const int tag, const int n_walk, const DensityEPI **epi, const int *nip,
                        const DensityEPJ **epj, const int *njp)
{
    need_ctx();
    TagBuf &b = tagbuf(tag);
    if (&RunParameter::KernelSupportRadiusMaximum) {
        sphk_params prm;
        sphk_get_params(g_ctx, &prm);
        prm.hmax = RunParameter::KernelSupportRadiusMaximum;
        int rc = sphk_set_params(g_ctx, &prm);
        if (rc) die("DispatchKernel_Dens (hmax)", rc);
    }
    const std::vector<size_t> si = starts(n_walk, nip), sj = starts(n_walk, njp);
    const size_t ni = si[n_walk], nj = sj[n_walk];
    b.di.resize(9, ni);
    b.dj.resize(7, nj);
    // walks are packed in parallel (the reference packs with OpenMP,
    // hdr_densityGPU.cu:537)
#pragma omp parallel for schedule(dynamic, 4)
    for (int w = 0; w < n_walk; w++) {
        size_t a = si[w], c = sj[w];
        for (int i = 0; i < nip[w]; i++, a++) {
            const DensityEPI &e = epi[w][i];
            b.di.id[a] = e.id;
            b.di.f[0][a] = e.mass;
            b.di.f[1][a] = e.pos.x;
            b.di.f[2][a] = e.pos.y;
            b.di.f[3][a] = e.pos.z;
            b.di.f[4][a] = e.vel.x;
            b.di.f[5][a] = e.vel.y;
            b.di.f[6][a] = e.vel.z;
            b.di.f[7][a] = e.ksr;
            b.di.f[8][a] = e.rs;
        }
        for (int j = 0; j < njp[w]; j++, c++) {
            const DensityEPJ &e = epj[w][j];
            b.dj.id[c] = e.id;
            b.dj.f[0][c] = e.mass;
            b.dj.f[1][c] = e.pos.x;
            b.dj.f[2][c] = e.pos.y;
            b.dj.f[3][c] = e.pos.z;
            b.dj.f[4][c] = e.vel.x;
            b.dj.f[5][c] = e.vel.y;
            b.dj.f[6][c] = e.vel.z;
        }
    }
    const sphk_dens_src pi{b.di.id.data(), b.di.p(0), b.di.p(1), b.di.p(2), b.di.p(3),
                           b.di.p(4),      b.di.p(5), b.di.p(6), b.di.p(7), b.di.p(8)};
    const sphk_dens_src pj{b.dj.id.data(), b.dj.p(0), b.dj.p(1), b.dj.p(2), b.dj.p(3),
                           b.dj.p(4),      b.dj.p(5), b.dj.p(6), nullptr,   nullptr};
    int rc = sphk_density_dispatch(g_ctx, tag, n_walk, nip, njp, pi, pj, g_precision);
    if (rc) die("DispatchKernel_Dens", rc);
    return 0;
}

int RetrieveKernel_Dens(const int tag, const int n_walk, const int *ni, Density **density)
{
    need_ctx();
    TagBuf &b = tagbuf(tag);
    const std::vector<size_t> si = starts(n_walk, ni);
    const size_t n = si[n_walk];
    for (int k = 0; k < 5; k++) b.o[k].resize(n);
    b.onp.resize(n);
    b.oitr.resize(n);
    sphk_dens_out o{b.o[0].data(), b.o[1].data(), b.o[2].data(), b.o[3].data(), b.o[4].data(),
                    b.onp.data(),  b.oitr.data()};
    int rc = sphk_density_retrieve(g_ctx, tag, n_walk, ni, o);
    if (rc) die("RetrieveKernel_Dens", rc);
#pragma omp parallel for schedule(dynamic, 4)
    for (int w = 0; w < n_walk; w++) {
        size_t a = si[w];
        for (int i = 0; i < ni[w]; i++, a++) {
            Density &d = density[w][i];
            d.dens = b.o[0][a];
            d.divv = b.o[1][a];
            d.rotv = b.o[2][a];
            d.grdh = b.o[3][a];
            d.ksr = b.o[4][a];
            d.np = b.onp[a];
            d.itr = b.oitr[a] != 0;
        }
    }
    return 0;
}

namespace {
// hydro field order of sphk_hydro_src after id: m x y z vx vy vz hinv pres
// presc bswt dens vsnd alph alphu thrm eta
template <typename E>
void put_hydro(Cols &c, size_t a, const E &e, double mass)
{
    c.id[a] = (int32_t)e.id;
    const double v[17] = {mass,   e.pos.x, e.pos.y, e.pos.z, e.vel.x, e.vel.y, e.vel.z, e.hinv, e.pres,
                          e.presc, e.bswt, e.dens,  e.vsnd,  e.alph,  e.alphu, e.thrm,  e.eta};
    for (int k = 0; k < 17; k++) c.f[k][a] = v[k];
}

sphk_hydro_src hydro_view(const Cols &c)
{
    return sphk_hydro_src{c.id.data(), c.p(0),  c.p(1),  c.p(2),  c.p(3),  c.p(4),
                          c.p(5),      c.p(6),  c.p(7),  c.p(8),  c.p(9),  c.p(10),
                          c.p(11),     c.p(12), c.p(13), c.p(14), c.p(15), c.p(16)};
}
}  // namespace

int DispatchKernel_HydroForce(const int tag, const int n_walk, const HydroEPI **epi, const int *nip,
                              const HydroEPJ **epj, const int *njp)
{
    need_ctx();
    TagBuf &b = tagbuf(tag);
    const std::vector<size_t> si = starts(n_walk, nip), sj = starts(n_walk, njp);
    b.hi.resize(17, si[n_walk]);
    b.hj.resize(17, sj[n_walk]);
#pragma omp parallel for schedule(dynamic, 4)
    for (int w = 0; w < n_walk; w++) {
        size_t a = si[w], c = sj[w];
        for (int i = 0; i < nip[w]; i++, a++) put_hydro(b.hi, a, epi[w][i], 0.0);
        for (int j = 0; j < njp[w]; j++, c++) put_hydro(b.hj, c, epj[w][j], epj[w][j].mass);
    }
    int rc = sphk_hydro_dispatch(g_ctx, tag, n_walk, nip, njp, hydro_view(b.hi), hydro_view(b.hj),
                                 g_precision);
    if (rc) die("DispatchKernel_HydroForce", rc);
    return 0;
}

int RetrieveKernel_HydroForce(const int tag, const int n_walk, const int *ni, Hydro **hydro)
{
    need_ctx();
    TagBuf &b = tagbuf(tag);
    const std::vector<size_t> si = starts(n_walk, ni);
    const size_t n = si[n_walk];
    for (int k = 0; k < 9; k++) b.o[k].resize(n);
    sphk_hydro_out o{b.o[0].data(), b.o[1].data(), b.o[2].data(), b.o[3].data(), b.o[4].data(),
                     b.o[5].data(), b.o[6].data(), b.o[7].data(), b.o[8].data()};
    int rc = sphk_hydro_retrieve(g_ctx, tag, n_walk, ni, o);
    if (rc) die("RetrieveKernel_HydroForce", rc);
#pragma omp parallel for schedule(dynamic, 4)
    for (int w = 0; w < n_walk; w++) {
        size_t a = si[w];
        for (int i = 0; i < ni[w]; i++, a++) {
            Hydro &h = hydro[w][i];
            h.acch = {b.o[0][a], b.o[1][a], b.o[2][a]};
            h.accg = {b.o[3][a], b.o[4][a], b.o[5][a]};
            h.udot = b.o[6][a];
            h.vsmx = b.o[7][a];
            h.diffu = b.o[8][a];
        }
    }
    return 0;
}

int DispatchKernel_Gravity(const int tag, const int n_walk, const GravityEPI **epi, const int *nip,
                           const GravityEPJ **epj, const int *njp,
                           const ParticleSimulator::GravitySPJ **spj, const int *nsp)
{
    need_ctx();
    TagBuf &b = tagbuf(tag);
    const std::vector<size_t> si = starts(n_walk, nip), sj = starts(n_walk, njp);
    const std::vector<size_t> ss = nsp ? starts(n_walk, nsp) : std::vector<size_t>(n_walk + 1, 0);
    b.gi.resize(5, si[n_walk]);
    b.gj.resize(5, sj[n_walk]);
    b.gs.resize(5, ss[n_walk]);
#pragma omp parallel for schedule(dynamic, 4)
    for (int w = 0; w < n_walk; w++) {
        size_t a = si[w], c = sj[w], d = ss[w];
        for (int i = 0; i < nip[w]; i++, a++) {
            const GravityEPI &e = epi[w][i];
            b.gi.f[0][a] = 0.0;
            b.gi.f[1][a] = e.pos.x;
            b.gi.f[2][a] = e.pos.y;
            b.gi.f[3][a] = e.pos.z;
            b.gi.f[4][a] = e.eps2;
        }
        for (int j = 0; j < njp[w]; j++, c++) {
            const GravityEPJ &e = epj[w][j];
            b.gj.f[0][c] = e.mass;
            b.gj.f[1][c] = e.pos.x;
            b.gj.f[2][c] = e.pos.y;
            b.gj.f[3][c] = e.pos.z;
            b.gj.f[4][c] = e.eps2;
        }
        if (nsp)
            for (int s = 0; s < nsp[w]; s++, d++) {
                const ParticleSimulator::GravitySPJ &e = spj[w][s];
                b.gs.f[0][d] = e.mass;
                b.gs.f[1][d] = e.pos.x;
                b.gs.f[2][d] = e.pos.y;
                b.gs.f[3][d] = e.pos.z;
                b.gs.f[4][d] = e.eps2;
            }
    }
    const sphk_grav_src pi{b.gi.p(0), b.gi.p(1), b.gi.p(2), b.gi.p(3), b.gi.p(4)};
    const sphk_grav_src pj{b.gj.p(0), b.gj.p(1), b.gj.p(2), b.gj.p(3), b.gj.p(4)};
    const sphk_grav_src ps{b.gs.p(0), b.gs.p(1), b.gs.p(2), b.gs.p(3), b.gs.p(4)};
    int rc = sphk_gravity_dispatch(g_ctx, tag, n_walk, nip, njp, nsp, pi, pj, ps, g_precision);
    if (rc) die("DispatchKernel_Gravity", rc);
    return 0;
}

int RetrieveKernel_Gravity(const int tag, const int n_walk, const int *ni, Gravity **gravity)
{
    need_ctx();
    TagBuf &b = tagbuf(tag);
    const std::vector<size_t> si = starts(n_walk, ni);
    const size_t n = si[n_walk];
    for (int k = 0; k < 5; k++) b.o[k].resize(n);
    sphk_grav_out o{b.o[0].data(), b.o[1].data(), b.o[2].data(), b.o[3].data(), b.o[4].data()};
    int rc = sphk_gravity_retrieve(g_ctx, tag, n_walk, ni, o);
    if (rc) die("RetrieveKernel_Gravity", rc);
#pragma omp parallel for schedule(dynamic, 4)
    for (int w = 0; w < n_walk; w++) {
        size_t a = si[w];
        for (int i = 0; i < ni[w]; i++, a++) {
            Gravity &g = gravity[w][i];
            g.acc = {b.o[0][a], b.o[1][a], b.o[2][a]};
            g.pot = b.o[3][a];
            g.eta = b.o[4][a];
        }
    }
    return 0;
}
