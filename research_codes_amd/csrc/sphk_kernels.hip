// sphk_kernels.hip -- gfx950 kernels for the SPH interaction functors.
//
// Launch shape (all three kernels): one 64-lane workgroup (one wavefront)
// per walk.  A walk's i-particles are processed in chunks of <= 64.  Inside a
// chunk of ni particles the wave is split into S = floor(64/ni) j-slices:
// lane l holds i = l mod ni and slice l / ni (lanes beyond S*ni idle), so a
// walk of 5 particles keeps 60 of 64 lanes busy.  j-entries are staged
// through LDS in tiles of 64 records: lane k gathers entry k through j_idx
// (coalesced where the list is a run of consecutive particles) and the tile
// is padded with inert records, so the sweep k = slice, slice+S, ... < 64
// has no bounds checks.  Records are packed as double4 / R4 so a lane reads
// an entry with 128-bit LDS loads.  Per-i partials are combined across
// slices through LDS in a fixed order (bitwise reproducible).
//
// Precision.  Positions are FP64; every pair displacement is formed in FP64.
// MIXED=false evaluates the pair term in FP64; MIXED=true casts the
// displacement to FP32, evaluates the pair term in FP32, sums a tile's
// partials in FP32 and folds them into FP64 per-i totals.  The three
// smoothing-length passes of the density functor are FP64 in both modes.

#include "sphk_device.hpp"
#include "sphk_internal.hpp"

namespace sphk {

template <bool MIXED>
using real_t = typename std::conditional<MIXED, float, double>::type;

template <typename R>
struct alignas(4 * sizeof(R)) R4 {
    R a, b, c, d;
};

// Inert padding record: far away (q >> 1 for any realistic H), zero mass,
// id -1, very negative sound speed (cannot raise vsmx).
constexpr double kFar = 1e10;

// Walk of a workgroup.  Workgroups of a launch are dispatched round-robin
// over the 8 XCDs (block b runs on XCD b mod 8), each XCD with its own L2.
// Walks are Morton-ordered, so neighbouring walks share most of their
// j-records.  Chunks of kXcdChunk consecutive walks are dealt to the XCDs in
// turn: an XCD works through whole chunks (its L2 serves the shared
// records) while every XCD still sees a mix of dense and sparse regions
// (one contiguous range per XCD measured 20% slower from load imbalance).
// Walks past the last whole round of chunks keep their own index.
constexpr int kXcd = 8;
constexpr int kXcdChunk = 64;
static_assert(kQueueInts == (kXcd + 1) * kQueueStride, "walk queue layout");

__device__ __forceinline__ int xcd_walk(int b, int n_walk)
{
    const int n_full = n_walk / (kXcd * kXcdChunk) * (kXcd * kXcdChunk);
    if (b >= n_full) return b;
    const int x = b % kXcd, k = b / kXcd;
    return ((k / kXcdChunk) * kXcd + x) * kXcdChunk + k % kXcdChunk;
}

// Walk scheduling.  Ordinary launch: one workgroup per walk (grid = n_walk),
// workgroup b takes walk xcd_walk(b).  Persistent launch (a grid sized to the
// resident capacity of the chip): every workgroup repeatedly claims the next
// virtual index v = x + 8 k of its own XCD x (counter k of that XCD's queue)
// and runs walk xcd_walk(v), so an XCD works through the same walks as in
// the ordinary launch; it stops at the first v >= n_walk.  The last
// workgroup to finish zeroes the queue for its next use.  Queue layout:
// kXcd claim counters and one exit counter, kQueueStride ints apart.
template <bool PERSIST, typename F>
__device__ __forceinline__ void run_walks(int n_walk, int *queue, F &&walk)
{
    if (!PERSIST) {
        if ((int)blockIdx.x < n_walk) walk(xcd_walk(blockIdx.x, n_walk), (int)threadIdx.x);
        return;
    }
    const int x = (int)(blockIdx.x % kXcd);
    int *ctr = queue + x * kQueueStride;
    for (;;) {
        int k = 0;
        if (threadIdx.x == 0) k = atomicAdd(ctr, 1);
        const int v = x + kXcd * __shfl(k, 0, kWave);
        if (v >= n_walk) break;
        walk(xcd_walk(v, n_walk), (int)threadIdx.x);
        __syncthreads();   // LDS of this walk no longer read
    }
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(queue + kXcd * kQueueStride, 1) == (int)gridDim.x - 1) {
            for (int c = 0; c < kXcd; c++) atomicExch(queue + c * kQueueStride, 0);
            atomicExch(queue + kXcd * kQueueStride, 0);
        }
    }
}

// Lane map of one i-chunk: S = floor(64/ni) slices.
struct Map {
    int ni, S, il, slice;
    bool lead;  // slice 0 of a real i: owns the reduced result
    __device__ __forceinline__ Map(int ni_, int lane)
    {
        ni = ni_;
        S = kWave / ni;
        il = lane % ni;
        slice = lane / ni;
        if (slice >= S) slice = kTile;  // idle lane: its sweep is empty
        lead = lane < ni;
    }
};

// Sum (or max) of per-lane values over the slices of each i, through LDS.
// On return the lead lane of each i holds the result.  red has 64*NQ slots.
template <int NQ>
__device__ __forceinline__ void slice_reduce(double (&v)[NQ], double *red, const Map &m, int lane,
                                             int max_mask = 0)
{
    __syncthreads();
    for (int q = 0; q < NQ; q++) red[q * kWave + lane] = v[q];
    __syncthreads();
    if (m.lead) {
        for (int q = 0; q < NQ; q++) {
            double a = red[q * kWave + m.il];
            const bool is_max = (max_mask >> q) & 1;
            for (int t = 1; t < m.S; t++) {
                const double b = red[q * kWave + m.il + m.ni * t];
                a = is_max ? (a > b ? a : b) : a + b;
            }
            v[q] = a;
        }
    }
}

// ---------------------------------------------------------------------------
// Density: scalar calcDensity, sph_yamauchi/run/hdr_density.hpp:254-363.
// ---------------------------------------------------------------------------

// q = sqrt(r2) * hinv with the reference's FP64 rounding sequence (separate
// IEEE operations, correctly rounded sqrt): decides q < 1 for pairs within a
// few ulps of the kernel edge.
__device__ __attribute__((noinline)) double exact_q(double dx, double dy, double dz, double hinv)
{
#pragma clang fp contract(off)
    const double r2 = dx * dx + dy * dy + dz * dz;
    return __builtin_sqrt(r2) * hinv;
}

// r = sqrt(r2) with the reference's FP64 rounding sequence: q = r * (1/H)
// is then the reference's q bit for bit.
__device__ __forceinline__ double exact_r(double dx, double dy, double dz)
{
#pragma clang fp contract(off)
    const double r2 = dx * dx + dy * dy + dz * dz;
    return __builtin_sqrt(r2);
}

struct Pass1 {
    double rh;
    int nj;
};

__device__ __forceinline__ void dens_pass1_pair(const double4 &j, double px, double py, double pz,
                                                double hi, double hi3, Pass1 &a)
{
    const double dx = px - j.x;
    const double dy = py - j.y;
    const double dz = pz - j.z;
    const double r2 = dx * dx + dy * dy + dz * dz;
    double q = r2 * rsqrt_f64(r2 > 1e-300 ? r2 : 1e-300) * hi;
    if (__builtin_fabs(q - 1.) < 1e-13) q = exact_q(dx, dy, dz, hi);
    a.rh += j.w * hi3 * kernel0th(q);
    a.nj += (q < 1.) ? 1 : 0;
}

template <typename R>
struct Pass2 {
    R g, dv, rx, ry, rz;
};

// The id of a j-record travels in the fourth slot of its velocity record:
// bit pattern in FP32, exact value in FP64.
__device__ __forceinline__ float pack_id(int32_t id, float) { return __int_as_float(id); }
__device__ __forceinline__ double pack_id(int32_t id, double) { return (double)id; }
__device__ __forceinline__ int32_t unpack_id(float v) { return __float_as_int(v); }
__device__ __forceinline__ int32_t unpack_id(double v) { return (int32_t)v; }

template <typename R>
__device__ __forceinline__ void dens_pass2_pair(const double4 &j, const R4<R> &v, double px, double py,
                                                double pz, R vx, R vy, R vz, int32_t id_i, R hi, R hi4,
                                                Pass2<R> &a)
{
    const R dpx = (R)(px - j.x);
    const R dpy = (R)(py - j.y);
    const R dpz = (R)(pz - j.z);
    const R dvx = vx - v.a;
    const R dvy = vy - v.b;
    const R dvz = vz - v.c;
    const R r2 = dpx * dpx + dpy * dpy + dpz * dpz;
    const R ri = (id_i != unpack_id(v.d)) ? rsqrt_r(r2) : R(0);
    const R q = r2 * ri * hi;
    // with t = (1-q)_+: W' = -5 C1 q t^3, 3W + qW' = C0 t^3 (3 + 9q - 32q^2),
    // and W'(q) / r = -5 C1 h t^3 (the self entry has d = 0: no div/curl term)
    R t = R(1) - q;
    t = t > R(0) ? t : R(0);
    const R t3 = t * t * t;
    const R mh = (R)j.w * hi4;
    a.g -= R(kCeff0) * t3 * (R(3) + q * (R(9) - R(32) * q)) * mh;
    const R dw = R(-5. * kCeff1) * mh * hi * t3;
    const R dwx = dw * dpx, dwy = dw * dpy, dwz = dw * dpz;
    a.dv -= dvx * dwx + dvy * dwy + dvz * dwz;
    a.rx += dvy * dwz - dvz * dwy;
    a.ry += dvz * dwx - dvx * dwz;
    a.rz += dvx * dwy - dvy * dwx;
}

// hdr_density.hpp:351-362: Omega^-1, div v and |rot v| from the grad-h sums.
__device__ __forceinline__ void dens_finish(const sphk_dens_out &out, int i, const double (&acc)[5],
                                            double dens, long long np, double hs, double rs)
{
    const double grdh = acc[0];
    double divv = acc[1];
    const double rot2 = acc[2] * acc[2] + acc[3] * acc[3] + acc[4] * acc[4];
    const double deni = 1. / dens;
    const double omgi = 1. / (1. + hs * deni * grdh / 3.);
    double rotv = rot2 != 0. ? __builtin_sqrt(rot2) : 0.;
    rotv *= deni * omgi;
    divv *= deni * omgi;
    out.dens[i] = dens;
    out.np[i] = np;
    out.ksr[i] = hs;
    out.itr[i] = (hs > rs) ? 1 : 0;
    out.divv[i] = divv;
    out.rotv[i] = rotv;
    out.grdh[i] = omgi;
}

// Wave-wide inclusive prefix sum and maximum.
__device__ __forceinline__ int wave_scan_incl(int v, int lane)
{
    for (int d = 1; d < kWave; d <<= 1) {
        const int t = __shfl_up(v, d, kWave);
        if (lane >= d) v += t;
    }
    return v;
}

__device__ __forceinline__ int wave_max(int v)
{
    for (int d = 1; d < kWave; d <<= 1) {
        const int t = __shfl_xor(v, d, kWave);
        v = t > v ? t : v;
    }
    return v;
}

// New H from rho: hdr_density.hpp:296-300 (eta * kappa * (m / rho)^(1/3), capped).
__device__ __forceinline__ double dens_next_h(double eta_ksrh, double m, double dens, double hmax)
{
    const double h = eta_ksrh * cbrt_f64(m / dens);
    return h < hmax ? h : hmax;
}

// Per-i sums of the five grad-h partials over the slices of a full sweep.
// The mixed kernel's reduction buffer holds two FP64 quantities, so the five
// are combined two, two and one at a time.
template <bool MIXED>
__device__ __forceinline__ void dens_reduce_grad(double (&acc)[5], double *red, const Map &M, int lane)
{
    if (!MIXED) {
        slice_reduce<5>(acc, red, M, lane);
        return;
    }
    double a[2] = {acc[0], acc[1]};
    slice_reduce<2>(a, red, M, lane);
    double b[2] = {acc[2], acc[3]};
    slice_reduce<2>(b, red, M, lane);
    double c[1] = {acc[4]};
    slice_reduce<1>(c, red, M, lane);
    acc[0] = a[0];
    acc[1] = a[1];
    acc[2] = b[0];
    acc[3] = b[1];
    acc[4] = c[0];
}

// Walks whose j-list holds at most kDensCap records keep the whole list in
// LDS for the four sweeps (three h-passes, grad-h pass); longer lists stream
// 64-record tiles.
//
// Resident lists are compacted per i-particle.  One FP64 sweep marks the
// records within R_i = max(rs_i, H_0,i) (times 1 + 1e-12) of i; every record
// outside contributes exactly zero to all four sweeps as long as H_k <= R_i,
// so the sweeps visit only the marked records (positions in LDS, one byte
// each).  The i-particles whose lists fit in kDensComp entries are taken
// together; the lanes are shared out so that every taken i needs at most
// T = ceil(C / (64 - n)) steps (C entries over n particles).  If an H_k of
// the iteration exceeds R_i the chunk continues with full sweeps.
constexpr int kDensCap = 256;        // largest resident capacity (instantiated: 192, 256)
constexpr int kDensComp = 1728;
constexpr int kDensR = 12;           // members per lane held in registers for the h-passes
static_assert(kDensCap <= 256, "compacted positions are stored as bytes");

template <bool MIXED>
__device__ __forceinline__ void dens_stage(int e, int t0_abs, int nj_end, const int32_t *__restrict__ j_idx,
                                           const sphk_dens_src &pj, double4 *sp, R4<real_t<MIXED>> *sv,
                                           bool with_v)
{
    using R = real_t<MIXED>;
    const int k = t0_abs + e;
    if (k < nj_end) {
        const int j = j_idx ? j_idx[k] : k;
        sp[e] = make_double4(pj.x[j], pj.y[j], pj.z[j], pj.m[j]);
        if (with_v) sv[e] = R4<R>{(R)pj.vx[j], (R)pj.vy[j], (R)pj.vz[j], pack_id(pj.id[j], R(0))};
    } else {
        sp[e] = make_double4(kFar, kFar, kFar, 0.);
        if (with_v) sv[e] = R4<R>{R(0), R(0), R(0), pack_id(-1, R(0))};
    }
}

// The i-particle data one lane works with.
struct DensI {
    double x, y, z, m, ksr, rs, vx, vy, vz;
    int32_t id;
};

__device__ __forceinline__ DensI dens_load_i(const sphk_dens_src &pi, int i)
{
    return DensI{pi.x[i], pi.y[i], pi.z[i], pi.m[i], pi.ksr[i], pi.rs[i], pi.vx[i], pi.vy[i], pi.vz[i],
                 pi.id[i]};
}

// Lists longer than kDensCap: every sweep streams the list through one
// 64-record LDS tile.
template <bool MIXED>
__device__ void density_streaming(int lane, const Map &M, int i, int jb, int je,
                                  const int32_t *__restrict__ j_idx, const sphk_dens_src &pi,
                                  const sphk_dens_src &pj, const sphk_dens_out &out, double eta_ksrh,
                                  double hmax, double4 *sp, R4<real_t<MIXED>> *sv, double *red,
                                  double *shs)
{
    using R = real_t<MIXED>;
    const int ntile = (je - jb + kTile - 1) / kTile;
    const DensI I = dens_load_i(pi, i);
    double hs = I.ksr;
    double dens = 0.;
    long long np = 0;

    for (int rep = 0; rep < 3; rep++) {
        const double hi = 1. / hs;
        const double hi3 = hi * hi * hi;
        Pass1 a{0., 0};
        for (int t = 0; t < ntile; t++) {
            __syncthreads();
            dens_stage<MIXED>(lane, jb + t * kTile, je, j_idx, pj, sp, sv, false);
            __syncthreads();
            const int nt = min(kTile, je - jb - t * kTile);   // records of this tile
            int k = M.slice;
            for (; k + 3 * M.S < nt; k += 4 * M.S) {
                const double4 j0 = sp[k], j1 = sp[k + M.S], j2 = sp[k + 2 * M.S], j3 = sp[k + 3 * M.S];
                dens_pass1_pair(j0, I.x, I.y, I.z, hi, hi3, a);
                dens_pass1_pair(j1, I.x, I.y, I.z, hi, hi3, a);
                dens_pass1_pair(j2, I.x, I.y, I.z, hi, hi3, a);
                dens_pass1_pair(j3, I.x, I.y, I.z, hi, hi3, a);
            }
            for (; k < nt; k += M.S) dens_pass1_pair(sp[k], I.x, I.y, I.z, hi, hi3, a);
        }
        double v[2] = {a.rh, (double)a.nj};
        slice_reduce<2>(v, red, M, lane);
        if (M.lead) {
            dens = v[0];
            np = (long long)v[1];
            shs[M.il] = dens_next_h(eta_ksrh, I.m, dens, hmax);
        }
        __syncthreads();
        hs = shs[M.il];
    }

    const double hi_d = 1. / hs;
    const R hi = (R)hi_d;
    const R hi4 = (R)(hi_d * (hi_d * hi_d * hi_d));
    const R vx = (R)I.vx, vy = (R)I.vy, vz = (R)I.vz;
    double acc[5] = {0., 0., 0., 0., 0.};
    for (int t = 0; t < ntile; t++) {
        __syncthreads();
        dens_stage<MIXED>(lane, jb + t * kTile, je, j_idx, pj, sp, sv, true);
        __syncthreads();
        Pass2<R> a{R(0), R(0), R(0), R(0), R(0)};
        const int nt = min(kTile, je - jb - t * kTile);
        int k = M.slice;
        for (; k + M.S < nt; k += 2 * M.S) {
            dens_pass2_pair<R>(sp[k], sv[k], I.x, I.y, I.z, vx, vy, vz, I.id, hi, hi4, a);
            dens_pass2_pair<R>(sp[k + M.S], sv[k + M.S], I.x, I.y, I.z, vx, vy, vz, I.id, hi, hi4, a);
        }
        for (; k < nt; k += M.S)
            dens_pass2_pair<R>(sp[k], sv[k], I.x, I.y, I.z, vx, vy, vz, I.id, hi, hi4, a);
        acc[0] += (double)a.g;
        acc[1] += (double)a.dv;
        acc[2] += (double)a.rx;
        acc[3] += (double)a.ry;
        acc[4] += (double)a.rz;
    }
    dens_reduce_grad<MIXED>(acc, red, M, lane);
    if (M.lead) dens_finish(out, i, acc, dens, np, hs, I.rs);
}

// Launch bound: 3 waves per SIMD (at most 168 VGPRs; a few spill), the
// occupancy the 192-record LDS layout allows.
template <bool MIXED, int CAP, bool PERSIST>
__global__ __launch_bounds__(kWave, 3) void density_kernel(
    int n_walk, const int32_t *__restrict__ i_off, const int32_t *__restrict__ j_off,
    const int32_t *__restrict__ j_idx, sphk_dens_src pi, sphk_dens_src pj, sphk_dens_out out,
    double eta_ksrh, double hmax, int *queue)
{
    using R = real_t<MIXED>;
    static_assert(CAP % kTile == 0 && CAP <= kDensCap, "list capacity");
    constexpr int kWords = CAP / kTile;
    constexpr int kDensPad = CAP;   // inert record after the list, target of idle steps
    constexpr int kRed = MIXED ? 2 * kWave : 5 * kWave;   // doubles (mixed reduces two at a time)
    __shared__ double4 sp[CAP + 1];     // x, y, z, m
    __shared__ R4<R> sv[CAP + 1];       // vx, vy, vz, id
    __shared__ uint8_t cidx[kDensComp];
    __shared__ double red[kRed];
    __shared__ double shs[kWave];
    __shared__ int s_own[kWave];
    __shared__ uint16_t s_ci[kWave], s_coff[kWave];
    int *s_cnt = reinterpret_cast<int *>(red);   // membership counts, before any reduction

    auto walk = [&](const int w, const int lane) {
        const int ib = i_off[w], ie = i_off[w + 1];
        const int jb = j_off[w], je = j_off[w + 1];
        const int nj = je - jb;
        const int ntile = (nj + kTile - 1) / kTile;

        if (nj > CAP) {
            for (int c0 = ib; c0 < ie; c0 += kWave) {
                const Map M(min(kWave, ie - c0), lane);
                density_streaming<MIXED>(lane, M, c0 + M.il, jb, je, j_idx, pi, pj, out, eta_ksrh, hmax,
                                         sp, sv, red, shs);
            }
            return;
        }

        // one burst: every record of the list, padded to whole tiles, plus the inert record
        for (int e = lane; e < ntile * kTile; e += kWave) dens_stage<MIXED>(e, jb, je, j_idx, pj, sp, sv, true);
        if (lane == 0) dens_stage<MIXED>(kDensPad, 0, 0, j_idx, pj, sp, sv, true);

        for (int c0 = ib; c0 < ie;) {
            const Map M(min(kWave, ie - c0), lane);
            __syncthreads();   // staging done; LDS of the previous chunk no longer read

            // -- membership sweep: lane (i, slice) marks its records within R_i
            uint64_t wd[kWords];
            int cnt = 0;
            {
                const int i = c0 + M.il;
                const double px = pi.x[i], py = pi.y[i], pz = pi.z[i];
                const double h0 = pi.ksr[i], rs = pi.rs[i];
                const double rme = (h0 > rs ? h0 : rs) * (1. + 1e-12);
                const double rm2 = rme * rme;
    #pragma unroll
                for (int t = 0; t < kWords; t++) {
                    uint64_t b = 0;
                    if (t < ntile) {
                        const double4 *tp = sp + t * kTile;
                        const int nt = min(kTile, nj - t * kTile);
                        int k = M.slice;
                        for (; k + 3 * M.S < nt; k += 4 * M.S) {
                            const double4 j0 = tp[k], j1 = tp[k + M.S], j2 = tp[k + 2 * M.S], j3 = tp[k + 3 * M.S];
                            const double x0 = px - j0.x, y0 = py - j0.y, z0 = pz - j0.z;
                            const double x1 = px - j1.x, y1 = py - j1.y, z1 = pz - j1.z;
                            const double x2 = px - j2.x, y2 = py - j2.y, z2 = pz - j2.z;
                            const double x3 = px - j3.x, y3 = py - j3.y, z3 = pz - j3.z;
                            b |= (uint64_t)(x0 * x0 + y0 * y0 + z0 * z0 < rm2) << k;
                            b |= (uint64_t)(x1 * x1 + y1 * y1 + z1 * z1 < rm2) << (k + M.S);
                            b |= (uint64_t)(x2 * x2 + y2 * y2 + z2 * z2 < rm2) << (k + 2 * M.S);
                            b |= (uint64_t)(x3 * x3 + y3 * y3 + z3 * z3 < rm2) << (k + 3 * M.S);
                        }
                        for (; k < nt; k += M.S) {
                            const double dx = px - tp[k].x, dy = py - tp[k].y, dz = pz - tp[k].z;
                            b |= (uint64_t)(dx * dx + dy * dy + dz * dz < rm2) << k;
                        }
                    }
                    wd[t] = b;
                    cnt += __popcll(b);
                }
                s_cnt[lane] = cnt;
            }
            __syncthreads();
            int start = 0, ci = 0;
            if (M.slice < M.S)
                for (int s2 = 0; s2 < M.slice; s2++) start += s_cnt[M.il + M.ni * s2];
            if (M.lead)
                for (int s2 = 0; s2 < M.S; s2++) ci += s_cnt[M.il + M.ni * s2];
            // take the leading i-particles whose lists fit in kDensComp entries
            // (every single list fits: c_i <= kDensCap)
            const int incl = wave_scan_incl(ci, lane);   // leads are lanes 0 .. ni-1
            const bool fits = M.lead && incl <= kDensComp;
            const int n_take = __popcll(__ballot(fits));
            const int total = __shfl(incl, n_take - 1, kWave);
            const bool take = M.il < n_take;             // lane works on a taken i

            if (M.lead && take) {
                s_coff[M.il] = (uint16_t)(incl - ci);
                s_ci[M.il] = (uint16_t)ci;
            }
            s_own[lane] = -1;
            __syncthreads();
            if (take && M.slice < M.S) {
                int o = s_coff[M.il] + start;
    #pragma unroll
                for (int t = 0; t < kWords; t++) {
                    uint64_t b = wd[t];
                    while (b) {
                        cidx[o++] = (uint8_t)(t * kTile + __builtin_ctzll(b));
                        b &= b - 1;
                    }
                }
            }
            // lanes per taken i: max(1, ceil(c_i / T)), T = ceil(C / (64 - n)); they sum to <= 64
            const int tsteps = n_take < kWave ? max(1, (total + (kWave - n_take) - 1) / (kWave - n_take)) : 1;
            const int si = (M.lead && take) ? (n_take < kWave ? max(1, (ci + tsteps - 1) / tsteps) : 1) : 0;
            const int lincl = wave_scan_incl(si, lane);
            if (si > 0)
                for (int s2 = 0; s2 < si; s2++) s_own[lincl - si + s2] = M.il | (s2 << 8) | (si << 16);
            __syncthreads();
            const int own = s_own[lane];

            // lane's particle and compacted range; every lane runs the wave's
            // largest step count, idle steps read the inert record
            int my = own >= 0 ? (own & 0xff) : 0;
            int e0 = 0, n_e = 0, stride = 1;
            const uint8_t *myidx = cidx;
            if (own >= 0) {
                e0 = (own >> 8) & 0xff;
                stride = own >> 16;
                n_e = s_ci[my];
                myidx = cidx + s_coff[my];
            }
            const int nsteps = wave_max(n_e > e0 ? (n_e - e0 + stride - 1) / stride : 0);
            bool compact = true;
            bool owner = own >= 0 && e0 == 0;
            DensI I = dens_load_i(pi, c0 + my);
            const Map MF(n_take, lane);   // full sweeps over the taken i-particles
            double hs = I.ksr;
            double dens = 0.;
            long long np = 0;

            // -- three fixed-point passes: rho(H_k), n(H_k), H_{k+1} (hdr_density.hpp:264-302).
            // When every lane has at most kDensR members, their exact distances
            // and masses are held in registers for the passes (one sqrt per
            // member instead of one rsqrt refinement per member and pass); the
            // passes that remain after an H beyond R_i run below.
            int rep0 = 0;
            if (nsteps <= kDensR) {
                double rr[kDensR], mm[kDensR];
    #pragma unroll
                for (int t = 0; t < kDensR; t++) {
                    if (t < nsteps) {
                        const int e = e0 + t * stride;
                        const double4 j = sp[e < n_e ? myidx[e] : kDensPad];
                        rr[t] = exact_r(I.x - j.x, I.y - j.y, I.z - j.z);
                        mm[t] = j.w;
                    }
                }
                for (; rep0 < 3; rep0++) {
                    const double hi = 1. / hs;
                    const double hi3 = hi * hi * hi;
                    double rh = 0.;
                    int nq = 0;
    #pragma unroll
                    for (int t = 0; t < kDensR; t++) {
                        if (t < nsteps) {
                            const double q = rr[t] * hi;
                            rh += mm[t] * hi3 * kernel0th(q);
                            nq += (q < 1.) ? 1 : 0;
                        }
                    }
                    __syncthreads();
                    red[lane] = rh;
                    red[kWave + lane] = (double)nq;
                    __syncthreads();
                    double v0 = red[lane], v1 = red[kWave + lane];
                    if (owner)
                        for (int t = 1; t < stride; t++) {
                            v0 += red[lane + t];
                            v1 += red[kWave + lane + t];
                        }
                    int viol = 0;
                    if (owner) {
                        dens = v0;
                        np = (long long)v1;
                        const double hnew = dens_next_h(eta_ksrh, I.m, dens, hmax);
                        shs[my] = hnew;
                        viol = hnew > (I.ksr > I.rs ? I.ksr : I.rs);
                    }
                    if (__syncthreads_or(viol)) {   // full sweeps from the next pass on
                        if (owner) {
                            red[my] = dens;
                            red[kWave + my] = (double)np;
                        }
                        __syncthreads();
                        compact = false;
                        my = MF.il;
                        owner = MF.lead;
                        I = dens_load_i(pi, c0 + my);
                        dens = red[my];
                        np = (long long)red[kWave + my];
                        hs = shs[my];
                        rep0++;
                        break;
                    }
                    hs = shs[my];
                }
            }
            for (int rep = rep0; rep < 3; rep++) {
                const double hi = 1. / hs;
                const double hi3 = hi * hi * hi;
                Pass1 a{0., 0};
                double v[2];
                if (compact) {
                    int st = 0, e = e0;
                    for (; st + 4 <= nsteps; st += 4, e += 4 * stride) {
                        const int k0 = e < n_e ? myidx[e] : kDensPad;
                        const int k1 = e + stride < n_e ? myidx[e + stride] : kDensPad;
                        const int k2 = e + 2 * stride < n_e ? myidx[e + 2 * stride] : kDensPad;
                        const int k3 = e + 3 * stride < n_e ? myidx[e + 3 * stride] : kDensPad;
                        const double4 j0 = sp[k0], j1 = sp[k1], j2 = sp[k2], j3 = sp[k3];
                        dens_pass1_pair(j0, I.x, I.y, I.z, hi, hi3, a);
                        dens_pass1_pair(j1, I.x, I.y, I.z, hi, hi3, a);
                        dens_pass1_pair(j2, I.x, I.y, I.z, hi, hi3, a);
                        dens_pass1_pair(j3, I.x, I.y, I.z, hi, hi3, a);
                    }
                    for (; st < nsteps; st++, e += stride)
                        dens_pass1_pair(sp[e < n_e ? myidx[e] : kDensPad], I.x, I.y, I.z, hi, hi3, a);
                    __syncthreads();
                    red[lane] = a.rh;
                    red[kWave + lane] = (double)a.nj;
                    __syncthreads();
                    v[0] = red[lane];
                    v[1] = red[kWave + lane];
                    if (owner)
                        for (int t = 1; t < stride; t++) {
                            v[0] += red[lane + t];
                            v[1] += red[kWave + lane + t];
                        }
                } else {
                    for (int t = 0; t < ntile; t++) {
                        const double4 *tp = sp + t * kTile;
                        const int nt = min(kTile, nj - t * kTile);
                        int k = MF.slice;
                        for (; k + 3 * MF.S < nt; k += 4 * MF.S) {
                            const double4 j0 = tp[k], j1 = tp[k + MF.S], j2 = tp[k + 2 * MF.S], j3 = tp[k + 3 * MF.S];
                            dens_pass1_pair(j0, I.x, I.y, I.z, hi, hi3, a);
                            dens_pass1_pair(j1, I.x, I.y, I.z, hi, hi3, a);
                            dens_pass1_pair(j2, I.x, I.y, I.z, hi, hi3, a);
                            dens_pass1_pair(j3, I.x, I.y, I.z, hi, hi3, a);
                        }
                        for (; k < nt; k += MF.S) dens_pass1_pair(tp[k], I.x, I.y, I.z, hi, hi3, a);
                    }
                    v[0] = a.rh;
                    v[1] = (double)a.nj;
                    slice_reduce<2>(v, red, MF, lane);
                }
                int viol = 0;
                if (owner) {
                    dens = v[0];
                    np = (long long)v[1];
                    const double hnew = dens_next_h(eta_ksrh, I.m, dens, hmax);
                    shs[my] = hnew;
                    viol = compact && hnew > (I.ksr > I.rs ? I.ksr : I.rs);
                }
                // an H beyond R_i leaves the compacted lists incomplete: full
                // sweeps from here on, the lead lanes of MF take over ownership
                if (__syncthreads_or(viol)) {
                    if (owner) {
                        red[my] = dens;
                        red[kWave + my] = (double)np;
                    }
                    __syncthreads();
                    compact = false;
                    my = MF.il;
                    owner = MF.lead;
                    I = dens_load_i(pi, c0 + my);
                    dens = red[my];
                    np = (long long)red[kWave + my];
                }
                hs = shs[my];
            }

            // -- grad-h, divergence, curl with H_3 (hdr_density.hpp:304-349)
            const double hi_d = 1. / hs;
            const R hi = (R)hi_d;
            const R hi4 = (R)(hi_d * (hi_d * hi_d * hi_d));
            const R vx = (R)I.vx, vy = (R)I.vy, vz = (R)I.vz;
            double acc[5] = {0., 0., 0., 0., 0.};
            if (compact) {
                // pair terms summed in R over at most 16 entries, then folded
                // into FP64 (as the full sweep folds each 64-record tile)
                Pass2<R> a{R(0), R(0), R(0), R(0), R(0)};
                double part[5] = {0., 0., 0., 0., 0.};
                auto fold = [&]() {
                    part[0] += (double)a.g;
                    part[1] += (double)a.dv;
                    part[2] += (double)a.rx;
                    part[3] += (double)a.ry;
                    part[4] += (double)a.rz;
                    a = Pass2<R>{R(0), R(0), R(0), R(0), R(0)};
                };
                int st = 0, e = e0;
                for (; st + 2 <= nsteps; st += 2, e += 2 * stride) {
                    const int k0 = e < n_e ? myidx[e] : kDensPad;
                    const int k1 = e + stride < n_e ? myidx[e + stride] : kDensPad;
                    dens_pass2_pair<R>(sp[k0], sv[k0], I.x, I.y, I.z, vx, vy, vz, I.id, hi, hi4, a);
                    dens_pass2_pair<R>(sp[k1], sv[k1], I.x, I.y, I.z, vx, vy, vz, I.id, hi, hi4, a);
                    if ((st & 14) == 14) fold();
                }
                if (st < nsteps) {
                    const int k0 = e < n_e ? myidx[e] : kDensPad;
                    dens_pass2_pair<R>(sp[k0], sv[k0], I.x, I.y, I.z, vx, vy, vz, I.id, hi, hi4, a);
                }
                fold();
                // per-i sums over the i's lanes in FP64, two quantities per round
                for (int q0 = 0; q0 < 5; q0 += 2) {
                    const int nq = q0 + 1 < 5 ? 2 : 1;
                    __syncthreads();
                    for (int q = 0; q < nq; q++) red[q * kWave + lane] = part[q0 + q];
                    __syncthreads();
                    if (owner)
                        for (int q = 0; q < nq; q++) {
                            double s = red[q * kWave + lane];
                            for (int t = 1; t < stride; t++) s += red[q * kWave + lane + t];
                            acc[q0 + q] = s;
                        }
                }
            } else {
                for (int t = 0; t < ntile; t++) {
                    const double4 *tp = sp + t * kTile;
                    const R4<R> *tv = sv + t * kTile;
                    Pass2<R> a{R(0), R(0), R(0), R(0), R(0)};
                    const int nt = min(kTile, nj - t * kTile);   // records of this tile
                    int k = MF.slice;
                    for (; k + MF.S < nt; k += 2 * MF.S) {
                        dens_pass2_pair<R>(tp[k], tv[k], I.x, I.y, I.z, vx, vy, vz, I.id, hi, hi4, a);
                        dens_pass2_pair<R>(tp[k + MF.S], tv[k + MF.S], I.x, I.y, I.z, vx, vy, vz, I.id, hi,
                                           hi4, a);
                    }
                    for (; k < nt; k += MF.S)
                        dens_pass2_pair<R>(tp[k], tv[k], I.x, I.y, I.z, vx, vy, vz, I.id, hi, hi4, a);
                    acc[0] += (double)a.g;
                    acc[1] += (double)a.dv;
                    acc[2] += (double)a.rx;
                    acc[3] += (double)a.ry;
                    acc[4] += (double)a.rz;
                }
                dens_reduce_grad<MIXED>(acc, red, MF, lane);
            }
            if (owner) dens_finish(out, c0 + my, acc, dens, np, hs, I.rs);
            c0 += n_take;
        }
    };
    run_walks<PERSIST>(n_walk, queue, walk);
}

// ---------------------------------------------------------------------------
// Hydro: scalar calcHydro, sph_yamauchi/run/hdr_hydro.hpp:320-421.
// ---------------------------------------------------------------------------
// Position record of a hydro j-entry: FP64 absolute in the FP64 path; in the
// mixed path FP32 offsets from the walk's first particle, formed in FP64 (as
// for gravity: 16-byte records, no FP64 arithmetic per pair).
template <typename R>
struct HPos;
template <>
struct HPos<double> {
    using rec = double4;
};
template <>
struct HPos<float> {
    using rec = float4;
};

template <typename R>
struct HydroI {
    double px, py, pz;   // FP64 path: absolute; mixed path: offsets (in the FP32 range)
    R vx, vy, vz, hi, h5, pp, pc, bs, dn, cs, al, au, eg, et;
    int32_t id;
};

template <typename R>
struct HydroAcc {
    R ax, ay, az, ud, df, gx, gy, gz, vs;
};

template <typename R>
__device__ __forceinline__ void hydro_pair(const typename HPos<R>::rec &p, const R4<R> &A, const R4<R> &B,
                                           const R4<R> &Cc, R eh_j, R h5_j, int32_t jid,
                                           const HydroI<R> &I, HydroAcc<R> &a)
{
    // p = {x, y, z, -5 C1 m}; A = {vx, vy, vz, hinv}; B = {pres, presc, bswt, dens};
    // Cc = {vsnd, alph, alphu, thrm}; eh_j = eta_j h_j^5, h5_j = h_j^5 (formed
    // per record when it is staged)
    const R dpx = (R)I.px - (R)p.x;
    const R dpy = (R)I.py - (R)p.y;
    const R dpz = (R)I.pz - (R)p.z;
    const R dvx = I.vx - A.a;
    const R dvy = I.vy - A.b;
    const R dvz = I.vz - A.c;
    const R r2 = dpx * dpx + dpy * dpy + dpz * dpz;
    const bool other = I.id != jid;
    const R ri = other ? rsqrt_r(r2) : R(0);
    const R r1 = r2 * ri;
    const R hj = A.d;
    const R q_i = r1 * I.hi;
    const R q_j = r1 * hj;
    // W'(q) = -5 C1 q (1-q)_+^3, so dW_k / r = h_k^4 W'(r h_k) / r = -5 C1 F_k
    // with F_k = h_k^5 (1 - q_k)_+^3: ka/r and the eta term need no kernel
    // evaluation beyond F (the self entry contributes nothing, as ri = 0
    // makes it in the reference's form)
    R ti = R(1) - q_i, tj = R(1) - q_j;
    ti = ti > R(0) ? ti : R(0);
    tj = tj > R(0) ? tj : R(0);
    const R tj3 = tj * tj * tj;
    const R Fi = I.h5 * (ti * ti * ti);
    const R Fj = h5_j * tj3;
    const R mk = other ? (R)p.w : R(0);
    const R kri = mk * (Fi + Fj);   // ka * ri, ka = (dW_i + dW_j) m_j
    const R ka = kri * r1;

    const R rv = dpx * dvx + dpy * dvy + dpz * dvz;
    const R w_ij = rv * ri;
    const R w0 = w_ij < R(0) ? w_ij : R(0);
    const R vs = I.cs + Cc.a - R(3) * w0;
    const R rhi = rcp_r(I.dn + B.d);
    const R av0 = (I.bs + B.c) * (I.al + Cc.b) * vs * w0;
    a.vs = a.vs > vs ? a.vs : vs;

    const R ta = (I.pc + B.b - R(0.5) * av0 * rhi) * kri;
    a.ax -= ta * dpx;
    a.ay -= ta * dpy;
    a.az -= ta * dpz;

    const R vsu2 = abs_r(I.pp - B.a) * rhi * R(2);
    const R vsu = vsu2 != R(0) ? vsu2 * rsqrt_r(vsu2) : R(0);
    const R du = I.eg - Cc.d;
    a.ud += ka * (I.pc * w_ij - rhi * (R(0.25) * av0 * w0 - (I.au + Cc.c) * vsu * du));
    a.df += rhi * du * kri;

    const R dg = mk * (I.et * Fi + eh_j * tj3);   // m_j ri (eta_i dW_i + eta_j dW_j)
    a.gx += dg * dpx;
    a.gy += dg * dpy;
    a.gz += dg * dpz;
}

// One j-record of the hydro functor as staged in LDS (R: the pair-term
// type; positions as offsets from the walk's first particle in mixed mode).
template <typename R>
struct HRec {
    typename HPos<R>::rec p;
    R4<R> a, b, c;
    R e, h5;
    int32_t id;
};

// Record k of the list, or an inert record for k >= end (beyond every
// support, zero mass, signal speed -1e20 so it never sets vsmx).  The
// per-record factors of the pair term are formed here in FP64: -5 C1 m_j,
// h_j^5 = hinv_j^5 and eta_j h_j^5.
template <typename R>
__device__ __forceinline__ HRec<R> hydro_load(const sphk_hydro_src &pj, const int32_t *__restrict__ j_idx, int k,
                                              int end, double cx, double cy, double cz)
{
    using PR = typename HPos<R>::rec;
    if (k < end) {
        const int j = j_idx ? j_idx[k] : k;
        const double hid = pj.hinv[j];
        const double h5 = hid * (hid * hid * hid) * hid;
        return HRec<R>{PR{(R)(pj.x[j] - cx), (R)(pj.y[j] - cy), (R)(pj.z[j] - cz), (R)(-5. * kCeff1 * pj.m[j])},
                       R4<R>{(R)pj.vx[j], (R)pj.vy[j], (R)pj.vz[j], (R)hid},
                       R4<R>{(R)pj.pres[j], (R)pj.presc[j], (R)pj.bswt[j], (R)pj.dens[j]},
                       R4<R>{(R)pj.vsnd[j], (R)pj.alph[j], (R)pj.alphu[j], (R)pj.thrm[j]}, (R)(pj.eta[j] * h5),
                       (R)h5, pj.id[j]};
    }
    return HRec<R>{PR{(R)kFar, (R)kFar, (R)kFar, R(0)}, R4<R>{R(0), R(0), R(0), R(0)},
                   R4<R>{R(0), R(0), R(0), R(1)}, R4<R>{R(-1e20), R(0), R(0), R(0)}, R(0), R(0), -1};
}

// FP64: at most 168 VGPRs (3 waves per SIMD, a few spills): 11 % faster than
// 2 waves without spills.  Mixed needs 115 VGPRs (4 waves) as written.
template <bool MIXED, bool PERSIST>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(MIXED ? (PERSIST ? 4 : 1) : 3))) void hydro_kernel(
    int n_walk, const int32_t *__restrict__ i_off, const int32_t *__restrict__ j_off,
    const int32_t *__restrict__ j_idx, sphk_hydro_src pi, sphk_hydro_src pj, sphk_hydro_out out, int *queue)
{
    using R = real_t<MIXED>;
    // tile records and the slice-reduction buffer share one LDS block (the
    // reduction runs after the last tile)
    using PR = typename HPos<R>::rec;
    constexpr int kTileBytes = kTile * (sizeof(PR) + 3 * sizeof(R4<R>) + 2 * sizeof(R) + 4);
    constexpr int kRedBytes = 9 * kWave * sizeof(double);
    __shared__ __attribute__((aligned(32))) char smem[kTileBytes > kRedBytes ? kTileBytes : kRedBytes];
    PR *sp = reinterpret_cast<PR *>(smem);
    R4<R> *sa = reinterpret_cast<R4<R> *>(sp + kTile);
    R4<R> *sb = sa + kTile;
    R4<R> *sc = sb + kTile;
    R *se = reinterpret_cast<R *>(sc + kTile);
    R *sh = se + kTile;
    int32_t *sid = reinterpret_cast<int32_t *>(sh + kTile);
    double *red = reinterpret_cast<double *>(smem);

    auto walk = [&](const int w, const int lane) {
        const int ib = i_off[w], ie = i_off[w + 1];
        const int jb = j_off[w], je = j_off[w + 1];

        // mixed: positions as offsets from the walk's first particle
        const double cx = MIXED ? pi.x[ib] : 0., cy = MIXED ? pi.y[ib] : 0., cz = MIXED ? pi.z[ib] : 0.;
        for (int c0 = ib; c0 < ie; c0 += kWave) {
            const Map M(min(kWave, ie - c0), lane);
            const int i = c0 + M.il;
            HydroI<R> I;
            I.px = pi.x[i] - cx;
            I.py = pi.y[i] - cy;
            I.pz = pi.z[i] - cz;
            I.vx = (R)pi.vx[i];
            I.vy = (R)pi.vy[i];
            I.vz = (R)pi.vz[i];
            const double hid = pi.hinv[i];
            I.hi = (R)hid;
            I.h5 = (R)(hid * (hid * hid * hid) * hid);
            I.pp = (R)pi.pres[i];
            I.pc = (R)pi.presc[i];
            I.bs = (R)pi.bswt[i];
            I.dn = (R)pi.dens[i];
            I.cs = (R)pi.vsnd[i];
            I.al = (R)pi.alph[i];
            I.au = (R)pi.alphu[i];
            I.eg = (R)pi.thrm[i];
            I.et = (R)pi.eta[i];
            I.id = pi.id[i];

            double acc[9] = {0., 0., 0., 0., 0., 0., 0., 0., 0.};
            R vsmx = R(0);
            // each lane loads its record of the tile into registers before the
            // barrier (loading the next tile's record during the sweep measured
            // 3 % slower in mixed mode, and its registers spill in FP64)
            for (int t0 = jb; t0 < je; t0 += kTile) {
                const int nt = min(kTile, je - t0);
                const HRec<R> nx = hydro_load<R>(pj, j_idx, t0 + lane, je, cx, cy, cz);
                __syncthreads();
                sp[lane] = nx.p;
                sa[lane] = nx.a;
                sb[lane] = nx.b;
                sc[lane] = nx.c;
                se[lane] = nx.e;
                sh[lane] = nx.h5;
                sid[lane] = nx.id;
                __syncthreads();
                HydroAcc<R> a{R(0), R(0), R(0), R(0), R(0), R(0), R(0), R(0), vsmx};
                int k = M.slice;
                for (; k + M.S < nt; k += 2 * M.S) {
                    const int k1 = k + M.S;
                    const PR p0 = sp[k], p1 = sp[k1];
                    const R4<R> a0 = sa[k], a1 = sa[k1], b0 = sb[k], b1 = sb[k1], c0_ = sc[k], c1 = sc[k1];
                    const R e0 = se[k], e1 = se[k1], f0 = sh[k], f1 = sh[k1];
                    const int32_t d0 = sid[k], d1 = sid[k1];
                    hydro_pair<R>(p0, a0, b0, c0_, e0, f0, d0, I, a);
                    hydro_pair<R>(p1, a1, b1, c1, e1, f1, d1, I, a);
                }
                for (; k < nt; k += M.S) hydro_pair<R>(sp[k], sa[k], sb[k], sc[k], se[k], sh[k], sid[k], I, a);
                acc[0] += (double)a.ax;
                acc[1] += (double)a.ay;
                acc[2] += (double)a.az;
                acc[3] += (double)a.ud;
                acc[4] += (double)a.df;
                acc[5] += (double)a.gx;
                acc[6] += (double)a.gy;
                acc[7] += (double)a.gz;
                vsmx = a.vs;
            }
            acc[8] = (double)vsmx;
            slice_reduce<9>(acc, red, M, lane, 1 << 8);
            // hdr_hydro.hpp:402-419
            if (M.lead) {
                out.ax[i] = acc[0] * 0.5;
                out.ay[i] = acc[1] * 0.5;
                out.az[i] = acc[2] * 0.5;
                out.udot[i] = acc[3] * 0.5;
                out.diffu[i] = acc[4] * 2.;
                out.gx[i] = acc[5] * 0.5;
                out.gy[i] = acc[6] * 0.5;
                out.gz[i] = acc[7] * 0.5;
                out.vsmx[i] = acc[8];
            }
        }
    };
    run_walks<PERSIST>(n_walk, queue, walk);
}

// ---------------------------------------------------------------------------
// Gravity: scalar calcGravity<TJ>, sph_yamauchi/run/hdr_gravity.hpp:295-338,
// over the EPJ list and then the SPJ list of each walk.
// ---------------------------------------------------------------------------
// j-record staged in LDS: position FP64; mass and softening in the pair
// precision (one 32-byte record in the mixed path).
template <bool MIXED>
struct GRec;
template <>
struct alignas(16) GRec<true> {
    double x, y, z;
    float m, e;
};
template <>
struct alignas(8) GRec<false> {
    double x, y, z, m, e;
};

template <typename R>
struct GravAcc {
    R ax, ay, az, pt, et;
};

// General pair term (softening of i and of j may differ):
// acc -= m x (rei3_i + rei3_j), pot -= m (rei_i + rei_j), eta += m rei3_i;
// the per-particle totals are halved at the end (hdr_gravity.hpp:310-331).
template <bool MIXED, bool ETA>
__device__ __forceinline__ void gravity_pair(const GRec<MIXED> &p, double px, double py, double pz,
                                             real_t<MIXED> e2i, GravAcc<real_t<MIXED>> &a)
{
    using R = real_t<MIXED>;
    const R dpx = (R)(px - p.x);
    const R dpy = (R)(py - p.y);
    const R dpz = (R)(pz - p.z);
    const R r2 = dpx * dpx + dpy * dpy + dpz * dpz;
    const R rei_i = rsqrt_r(r2 + e2i);
    const R rei_j = rsqrt_r(r2 + (R)p.e);
    const R rei3_i = rei_i * rei_i * rei_i;
    const R rei3_j = rei_j * rei_j * rei_j;
    const R m_j = (R)p.m;
    const R dg2 = m_j * (rei3_i + rei3_j);
    a.pt -= m_j * (rei_i + rei_j);
    a.ax -= dpx * dg2;
    a.ay -= dpy * dg2;
    a.az -= dpz * dg2;
    if (ETA) a.et += m_j * rei3_i;
}

// Uniform softening (every eps2 equal, e.g. a constant gravitational
// softening): the two symmetrised terms coincide, one rsqrt per pair.
// Accumulates +m x rei3, +m rei, +m rei3; the caller finishes with the sign.
template <bool MIXED, bool ETA>
__device__ __forceinline__ void gravity_pair_uniform(const GRec<MIXED> &p, double px, double py,
                                                     double pz, real_t<MIXED> e2,
                                                     GravAcc<real_t<MIXED>> &a)
{
    using R = real_t<MIXED>;
    const R dpx = (R)(px - p.x);
    const R dpy = (R)(py - p.y);
    const R dpz = (R)(pz - p.z);
    // r^2 + eps^2 with the softening in the first FMA
    const R r2e = fma_r(dpz, dpz, fma_r(dpy, dpy, fma_r(dpx, dpx, e2)));
    const R rei = rsqrt_r(r2e);
    const R m_j = (R)p.m;
    const R mr = m_j * rei;
    const R w = mr * rei * rei;
    a.pt += mr;
    a.ax += dpx * w;
    a.ay += dpy * w;
    a.az += dpz * w;
    if (ETA) a.et += w;
}

template <bool MIXED, bool UNIFORM, bool ETA>
__device__ __forceinline__ void gravity_list(int lane, const Map &M, int kb, int ke,
                                             const int32_t *__restrict__ idx, sphk_grav_src src,
                                             double px, double py, double pz, double e2_i,
                                             GRec<MIXED> *sp, double (&acc)[5])
{
    using R = real_t<MIXED>;
    const R e2 = (R)e2_i;
    for (int t0 = kb; t0 < ke; t0 += kTile) {
        const int nt = min(kTile, ke - t0);
        __syncthreads();
        if (lane < nt) {
            const int j = idx ? idx[t0 + lane] : t0 + lane;
            sp[lane] = GRec<MIXED>{src.x[j], src.y[j], src.z[j], (R)src.m[j], (R)src.eps2[j]};
        } else {
            sp[lane] = GRec<MIXED>{kFar, kFar, kFar, R(0), R(1)};
        }
        __syncthreads();
        GravAcc<R> a{R(0), R(0), R(0), R(0), R(0)};
        int k = M.slice;
        for (; k + 3 * M.S < nt; k += 4 * M.S) {
            const GRec<MIXED> p0 = sp[k], p1 = sp[k + M.S], p2 = sp[k + 2 * M.S], p3 = sp[k + 3 * M.S];
            if (UNIFORM) {
                gravity_pair_uniform<MIXED, ETA>(p0, px, py, pz, e2, a);
                gravity_pair_uniform<MIXED, ETA>(p1, px, py, pz, e2, a);
                gravity_pair_uniform<MIXED, ETA>(p2, px, py, pz, e2, a);
                gravity_pair_uniform<MIXED, ETA>(p3, px, py, pz, e2, a);
            } else {
                gravity_pair<MIXED, ETA>(p0, px, py, pz, e2, a);
                gravity_pair<MIXED, ETA>(p1, px, py, pz, e2, a);
                gravity_pair<MIXED, ETA>(p2, px, py, pz, e2, a);
                gravity_pair<MIXED, ETA>(p3, px, py, pz, e2, a);
            }
        }
        for (; k < nt; k += M.S) {
            if (UNIFORM)
                gravity_pair_uniform<MIXED, ETA>(sp[k], px, py, pz, e2, a);
            else
                gravity_pair<MIXED, ETA>(sp[k], px, py, pz, e2, a);
        }
        acc[0] += (double)a.ax;
        acc[1] += (double)a.ay;
        acc[2] += (double)a.az;
        acc[3] += (double)a.pt;
        acc[4] += (double)a.et;
    }
}

// Mixed precision with uniform softening: the j-records of a walk are staged
// as FP32 offsets from a reference point of the walk (its first i-particle),
// formed in FP64, so the pair term needs no FP64 arithmetic.  The offset of
// a far record carries a relative rounding error of 2^-24 of its distance;
// for a near pair the error is 2^-24 of its distance from the reference point.
template <bool ETA>
__device__ __forceinline__ void gravity_pair_rel(const float4 &p, float xi, float yi, float zi, float e2,
                                                 GravAcc<float> &a)
{
    const float dx = xi - p.x;
    const float dy = yi - p.y;
    const float dz = zi - p.z;
    // r^2 + eps^2 with the softening in the first FMA
    const float rei = rsqrt_r(fma_r(dz, dz, fma_r(dy, dy, fma_r(dx, dx, e2))));
    const float mr = p.w * rei;
    const float w = mr * rei * rei;
    a.pt += mr;
    a.ax += dx * w;
    a.ay += dy * w;
    a.az += dz * w;
    if (ETA) a.et += w;
}

// Records are staged kGravBlock at a time (one barrier pair per block, four
// records in flight per lane); each lane's FP32 partial sums over a block
// (at most kGravBlock / S terms) are folded into FP64 once per block.
constexpr int kGravBlock = 4 * kTile;

template <bool ETA>
__device__ __forceinline__ void gravity_list_rel(const Map &M, int lane, int kb, int ke,
                                                 const int32_t *__restrict__ idx, const sphk_grav_src &src,
                                                 double cx, double cy, double cz, float xi, float yi,
                                                 float zi, float e2, float4 *sp, double (&acc)[5])
{
    const int k0 = M.slice < M.S ? M.slice : kGravBlock;   // idle lanes: empty sweeps
    for (int b0 = kb; b0 < ke; b0 += kGravBlock) {
        const int nb = min(kGravBlock, ke - b0);
        __syncthreads();
        for (int e = lane; e < nb; e += kWave) {
            const int j = idx ? idx[b0 + e] : b0 + e;
            sp[e] = make_float4((float)(src.x[j] - cx), (float)(src.y[j] - cy), (float)(src.z[j] - cz),
                                (float)src.m[j]);
        }
        __syncthreads();
        GravAcc<float> a{0.f, 0.f, 0.f, 0.f, 0.f};
        int k = k0;
        for (; k + 3 * M.S < nb; k += 4 * M.S) {
            const float4 p0 = sp[k], p1 = sp[k + M.S], p2 = sp[k + 2 * M.S], p3 = sp[k + 3 * M.S];
            gravity_pair_rel<ETA>(p0, xi, yi, zi, e2, a);
            gravity_pair_rel<ETA>(p1, xi, yi, zi, e2, a);
            gravity_pair_rel<ETA>(p2, xi, yi, zi, e2, a);
            gravity_pair_rel<ETA>(p3, xi, yi, zi, e2, a);
        }
        for (; k < nb; k += M.S) gravity_pair_rel<ETA>(sp[k], xi, yi, zi, e2, a);
        acc[0] += (double)a.ax;
        acc[1] += (double)a.ay;
        acc[2] += (double)a.az;
        acc[3] += (double)a.pt;
        acc[4] += (double)a.et;
    }
}

// ETA false: the caller does not need the softening-gradient sum (out.eta
// NULL, e.g. under a constant softening, where it does not enter).
template <bool MIXED, bool UNIFORM, bool ETA, bool PERSIST>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(PERSIST && MIXED ? 8 : 1))) void gravity_kernel(
    int n_walk, const int32_t *__restrict__ i_off, const int32_t *__restrict__ j_off,
    const int32_t *__restrict__ j_idx, sphk_grav_src pj, const int32_t *__restrict__ s_off,
    const int32_t *__restrict__ s_idx, sphk_grav_src ps, sphk_grav_src pi, sphk_grav_out out, int *queue)
{
    constexpr int kTileBytes = kTile * sizeof(GRec<MIXED>);
    constexpr int kRelBytes = (MIXED && UNIFORM) ? kGravBlock * (int)sizeof(float4) : 0;
    constexpr int kRedBytes = 5 * kWave * sizeof(double);
    constexpr int kMax1 = kTileBytes > kRedBytes ? kTileBytes : kRedBytes;
    __shared__ __attribute__((aligned(32))) char smem[kMax1 > kRelBytes ? kMax1 : kRelBytes];
    GRec<MIXED> *sp = reinterpret_cast<GRec<MIXED> *>(smem);
    double *red = reinterpret_cast<double *>(smem);

    auto walk = [&](const int w, const int lane) {
        const int ib = i_off[w], ie = i_off[w + 1];

        for (int c0 = ib; c0 < ie; c0 += kWave) {
            const Map M(min(kWave, ie - c0), lane);
            const int i = c0 + M.il;
            const double px = pi.x[i], py = pi.y[i], pz = pi.z[i], e2 = pi.eps2[i];
            double acc[5] = {0., 0., 0., 0., 0.};
            if (MIXED && UNIFORM) {
                const double cx = pi.x[ib], cy = pi.y[ib], cz = pi.z[ib];
                const float xi = (float)(px - cx), yi = (float)(py - cy), zi = (float)(pz - cz);
                float4 *sr = reinterpret_cast<float4 *>(smem);
                gravity_list_rel<ETA>(M, lane, j_off[w], j_off[w + 1], j_idx, pj, cx, cy, cz, xi, yi, zi,
                                      (float)e2, sr, acc);
                if (s_off)
                    gravity_list_rel<ETA>(M, lane, s_off[w], s_off[w + 1], s_idx, ps, cx, cy, cz, xi, yi, zi,
                                          (float)e2, sr, acc);
            } else {
                gravity_list<MIXED, UNIFORM, ETA>(lane, M, j_off[w], j_off[w + 1], j_idx, pj, px, py, pz, e2, sp,
                                                  acc);
                if (s_off)
                    gravity_list<MIXED, UNIFORM, ETA>(lane, M, s_off[w], s_off[w + 1], s_idx, ps, px, py, pz, e2,
                                                      sp, acc);
            }
            slice_reduce<5>(acc, red, M, lane);
            if (M.lead) {
                if (UNIFORM) {
                    // 0.5 * (-2 sum m rei3 x) etc.
                    out.ax[i] = -acc[0];
                    out.ay[i] = -acc[1];
                    out.az[i] = -acc[2];
                    out.pot[i] = -acc[3];
                } else {
                    out.ax[i] = acc[0] * 0.5;
                    out.ay[i] = acc[1] * 0.5;
                    out.az[i] = acc[2] * 0.5;
                    out.pot[i] = acc[3] * 0.5;
                }
                if (ETA) out.eta[i] = acc[4];
            }
        }
    };
    run_walks<PERSIST>(n_walk, queue, walk);
}

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------
// Grid of a persistent launch: the workgroups of this kernel the chip holds
// at once (a multiple of kXcd), at most one per walk of each XCD queue.
template <typename K>
static int persistent_grid(K kernel, int n_walk)
{
    int dev = 0, n_cu = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kWave, 0) != hipSuccess)
        return -1;
    const int full = (per_cu > 0 ? per_cu : 1) * (n_cu > 0 ? n_cu : 1);
    const int g = (full + kXcd - 1) / kXcd * kXcd;
    const int need = (n_walk + kXcd - 1) / kXcd * kXcd;
    return g < need ? g : need;
}

// One launch: ordinary (grid = n_walk) or, with a queue, persistent.
#define SPHK_LAUNCH(KERNEL, KERNEL_P, ...)                                                     \
    do {                                                                                        \
        if (queue) {                                                                            \
            const int g_ = persistent_grid(KERNEL_P, n_walk);                                   \
            if (g_ <= 0) return hipErrorInvalidValue;                                           \
            hipLaunchKernelGGL(KERNEL_P, dim3(g_), dim3(kWave), 0, st, __VA_ARGS__, queue);     \
        } else {                                                                                \
            hipLaunchKernelGGL(KERNEL, dim3(n_walk), dim3(kWave), 0, st, __VA_ARGS__, nullptr); \
        }                                                                                       \
    } while (0)

hipError_t launch_density(hipStream_t st, int n_walk, const int32_t *i_off, const int32_t *j_off,
                          const int32_t *j_idx, const sphk_dens_src &pi, const sphk_dens_src &pj,
                          const sphk_dens_out &out, double eta_ksrh, double hmax, bool mixed, int list_cap,
                          int *queue)
{
    if (n_walk <= 0) return hipSuccess;
#define SPHK_LAUNCH_D(M, CAPV)                                                                           \
    SPHK_LAUNCH((density_kernel<M, CAPV, false>), (density_kernel<M, CAPV, true>), n_walk, i_off, j_off, \
                j_idx, pi, pj, out, eta_ksrh, hmax)
    if (list_cap == 192) {
        if (mixed)
            SPHK_LAUNCH_D(true, 192);
        else
            SPHK_LAUNCH_D(false, 192);
    } else {
        if (mixed)
            SPHK_LAUNCH_D(true, 256);
        else
            SPHK_LAUNCH_D(false, 256);
    }
#undef SPHK_LAUNCH_D
    return hipGetLastError();
}

hipError_t launch_hydro(hipStream_t st, int n_walk, const int32_t *i_off, const int32_t *j_off,
                        const int32_t *j_idx, const sphk_hydro_src &pi, const sphk_hydro_src &pj,
                        const sphk_hydro_out &out, bool mixed, int *queue)
{
    if (n_walk <= 0) return hipSuccess;
    if (mixed)
        SPHK_LAUNCH((hydro_kernel<true, false>), (hydro_kernel<true, true>), n_walk, i_off, j_off, j_idx, pi, pj,
                    out);
    else
        SPHK_LAUNCH((hydro_kernel<false, false>), (hydro_kernel<false, true>), n_walk, i_off, j_off, j_idx, pi,
                    pj, out);
    return hipGetLastError();
}

hipError_t launch_gravity(hipStream_t st, int n_walk, const int32_t *i_off, const int32_t *j_off,
                          const int32_t *j_idx, const sphk_grav_src &pj, const int32_t *s_off,
                          const int32_t *s_idx, const sphk_grav_src &ps, const sphk_grav_src &pi,
                          const sphk_grav_out &out, bool mixed, bool uniform, int *queue)
{
    if (n_walk <= 0) return hipSuccess;
#define SPHK_LAUNCH_G(M, U, E)                                                                   \
    SPHK_LAUNCH((gravity_kernel<M, U, E, false>), (gravity_kernel<M, U, E, true>), n_walk, i_off, \
                j_off, j_idx, pj, s_off, s_idx, ps, pi, out)
    const bool eta = out.eta != nullptr;
    if (mixed && uniform) {
        if (eta) SPHK_LAUNCH_G(true, true, true);
        else SPHK_LAUNCH_G(true, true, false);
    } else if (mixed) {
        if (eta) SPHK_LAUNCH_G(true, false, true);
        else SPHK_LAUNCH_G(true, false, false);
    } else if (uniform) {
        if (eta) SPHK_LAUNCH_G(false, true, true);
        else SPHK_LAUNCH_G(false, true, false);
    } else {
        if (eta) SPHK_LAUNCH_G(false, false, true);
        else SPHK_LAUNCH_G(false, false, false);
    }
#undef SPHK_LAUNCH_G
    return hipGetLastError();
}
#undef SPHK_LAUNCH

}  // namespace sphk
