// fdps_driver.cpp -- test harness that drives libsphk.so through the
// reference's FDPS multiwalk API (include/sphk_fdps.hpp) the way FDPS does:
// for every batch of at most 200 walks it builds per-walk AoS EPI/EPJ/SPJ
// copies, calls DispatchKernel_* with tag 0 and then RetrieveKernel_* into
// per-walk Force arrays (hdr_density.hpp:387-394, main.cpp:337-344,
// hdr_gravity.hpp:390-397).
//
// usage: fdps_driver <dir> [reps]  -- reads <dir>/manifest.txt and <name>.bin
// raw arrays (written by tests/test_gpu_fdps_boundary.py or bench.py), writes
// out_<name>.bin.  With reps > 0 every functor's multiwalk pass runs once
// untimed and then reps times with its Dispatch + Retrieve calls timed (the
// per-walk AoS copies, FDPS's own work, are outside the timing), and one
// JSON line reports ms per pass and the bytes each pass moves across PCIe.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>
#include <vector>

#include "sphk.h"
#include "sphk_fdps.hpp"
#include "sphk_fdps_types.hpp"

namespace RunParameter {
double KernelSupportRadiusMaximum = 0.0;  // the reference's global (hdr_run.hpp:24)
}

static std::string g_dir;
static std::map<std::string, std::vector<char>> g_in;

template <typename T>
static const T *arr(const char *name, size_t *n = nullptr)
{
    auto it = g_in.find(name);
    if (it == g_in.end()) {
        std::fprintf(stderr, "fdps_driver: missing array %s\n", name);
        std::exit(2);
    }
    if (n) *n = it->second.size() / sizeof(T);
    return reinterpret_cast<const T *>(it->second.data());
}

template <typename T>
static void out(const char *name, const std::vector<T> &v)
{
    std::string p = g_dir + "/out_" + name + ".bin";
    FILE *f = std::fopen(p.c_str(), "wb");
    if (!f || std::fwrite(v.data(), sizeof(T), v.size(), f) != v.size()) {
        std::fprintf(stderr, "fdps_driver: cannot write %s\n", p.c_str());
        std::exit(2);
    }
    std::fclose(f);
}

static void load(const std::string &dir)
{
    FILE *m = std::fopen((dir + "/manifest.txt").c_str(), "r");
    if (!m) {
        std::fprintf(stderr, "fdps_driver: no manifest in %s\n", dir.c_str());
        std::exit(2);
    }
    char name[128];
    long long bytes;
    while (std::fscanf(m, "%127s %lld", name, &bytes) == 2) {
        std::vector<char> buf(bytes);
        FILE *f = std::fopen((dir + "/" + name + ".bin").c_str(), "rb");
        if (!f || (bytes && std::fread(buf.data(), 1, bytes, f) != (size_t)bytes)) {
            std::fprintf(stderr, "fdps_driver: cannot read %s\n", name);
            std::exit(2);
        }
        std::fclose(f);
        g_in[name] = std::move(buf);
    }
    std::fclose(m);
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: fdps_driver <dir>\n");
        return 2;
    }
    g_dir = argv[1];
    const int reps = argc > 2 ? std::atoi(argv[2]) : 0;
    load(g_dir);
    using clk = std::chrono::steady_clock;
    double t_ms[3] = {0, 0, 0};
    double h2d[3] = {0, 0, 0}, d2h[3] = {0, 0, 0};   // bytes per pass
    auto since = [](clk::time_point a) { return std::chrono::duration<double, std::milli>(clk::now() - a).count(); };
    const double *par = arr<double>("params");  // eta, ksrh, hmax
    size_t n;
    const int32_t *id = arr<int32_t>("id", &n);
    const double *m = arr<double>("m"), *x = arr<double>("x"), *y = arr<double>("y"),
                 *z = arr<double>("z"), *vx = arr<double>("vx"), *vy = arr<double>("vy"),
                 *vz = arr<double>("vz"), *ksr = arr<double>("ksr"), *rs = arr<double>("rs");
    const double *hinv = arr<double>("hinv"), *pres = arr<double>("pres"),
                 *presc = arr<double>("presc"), *bswt = arr<double>("bswt"),
                 *dens = arr<double>("dens"), *vsnd = arr<double>("vsnd"),
                 *alph = arr<double>("alph"), *alphu = arr<double>("alphu"),
                 *thrm = arr<double>("thrm"), *eta = arr<double>("eta"), *eps2 = arr<double>("eps2");
    size_t nw_d, nw_h, nw_g, nspj;
    const int32_t *d_ioff = arr<int32_t>("dens_i_off", &nw_d), *d_joff = arr<int32_t>("dens_j_off"),
                  *d_jidx = arr<int32_t>("dens_j_idx");
    const int32_t *h_ioff = arr<int32_t>("hydro_i_off", &nw_h), *h_joff = arr<int32_t>("hydro_j_off"),
                  *h_jidx = arr<int32_t>("hydro_j_idx");
    const int32_t *g_ioff = arr<int32_t>("grav_i_off", &nw_g), *g_joff = arr<int32_t>("grav_j_off"),
                  *g_jidx = arr<int32_t>("grav_j_idx"), *g_soff = arr<int32_t>("grav_s_off"),
                  *g_sidx = arr<int32_t>("grav_s_idx");
    const double *sm = arr<double>("spj_m", &nspj), *sx = arr<double>("spj_x"),
                 *sy = arr<double>("spj_y"), *sz = arr<double>("spj_z"), *se = arr<double>("spj_eps2");
    nw_d--, nw_h--, nw_g--;
    const int kWalkLimit = 200;  // n_walk_limit of the reference's calls

    setup_gpu(0);
    sphk_params prm{par[0], par[1], par[2]};
    if (sphk_set_params(sphk_fdps_context(), &prm)) return 3;
    RunParameter::KernelSupportRadiusMaximum = par[2];

    // ---- density (calcForceAllMultiWalk)
    std::vector<double> o_dens(n), o_divv(n), o_rotv(n), o_grdh(n), o_ksr(n);
    std::vector<long long> o_np(n);
    std::vector<unsigned char> o_itr(n);
    for (int rep = 0; rep <= reps; rep++)
    for (size_t w0 = 0; w0 < nw_d; w0 += kWalkLimit) {
        const int nb = (int)std::min<size_t>(kWalkLimit, nw_d - w0);
        std::vector<std::vector<DensityEPI>> epi(nb);
        std::vector<std::vector<DensityEPJ>> epj(nb);
        std::vector<std::vector<Density>> force(nb);
        std::vector<const DensityEPI *> pe(nb);
        std::vector<const DensityEPJ *> pj(nb);
        std::vector<Density *> pf(nb);
        std::vector<int> nip(nb), njp(nb);
        for (int b = 0; b < nb; b++) {
            const size_t w = w0 + b;
            for (int i = d_ioff[w]; i < d_ioff[w + 1]; i++)
                epi[b].push_back(DensityEPI{id[i], m[i], {x[i], y[i], z[i]}, {vx[i], vy[i], vz[i]}, ksr[i], rs[i]});
            for (int k = d_joff[w]; k < d_joff[w + 1]; k++) {
                const int j = d_jidx[k];
                epj[b].push_back(DensityEPJ{id[j], m[j], {x[j], y[j], z[j]}, {vx[j], vy[j], vz[j]}});
            }
            force[b].resize(epi[b].size());
            pe[b] = epi[b].data();
            pj[b] = epj[b].data();
            pf[b] = force[b].data();
            nip[b] = (int)epi[b].size();
            njp[b] = (int)epj[b].size();
        }
        const auto t0 = clk::now();
        DispatchKernel_Dens(0, nb, pe.data(), nip.data(), pj.data(), njp.data());
        RetrieveKernel_Dens(0, nb, nip.data(), pf.data());
        if (rep > 0) t_ms[0] += since(t0);
        if (rep == 1)
            for (int b = 0; b < nb; b++) {
                // SoA staged by sphk_density_dispatch: id + 9 reals per i, id + 7 per j; 5 reals, np, itr back
                h2d[0] += nip[b] * 76.0 + njp[b] * 60.0 + 8.0;
                d2h[0] += nip[b] * 49.0;
            }
        for (int b = 0; b < nb; b++)
            for (int i = 0; i < nip[b]; i++) {
                const size_t g = d_ioff[w0 + b] + i;
                const Density &d = force[b][i];
                o_dens[g] = d.dens, o_divv[g] = d.divv, o_rotv[g] = d.rotv, o_grdh[g] = d.grdh;
                o_ksr[g] = d.ksr, o_np[g] = d.np, o_itr[g] = d.itr;
            }
    }
    out("dens", o_dens), out("divv", o_divv), out("rotv", o_rotv), out("grdh", o_grdh);
    out("ksr", o_ksr), out("np", o_np), out("itr", o_itr);

    // ---- hydro (calcForceAllAndWriteBackMultiWalk)
    std::vector<double> h[9];
    for (auto &v : h) v.resize(n);
    for (int rep = 0; rep <= reps; rep++)
    for (size_t w0 = 0; w0 < nw_h; w0 += kWalkLimit) {
        const int nb = (int)std::min<size_t>(kWalkLimit, nw_h - w0);
        std::vector<std::vector<HydroEPI>> epi(nb);
        std::vector<std::vector<HydroEPJ>> epj(nb);
        std::vector<std::vector<Hydro>> force(nb);
        std::vector<const HydroEPI *> pe(nb);
        std::vector<const HydroEPJ *> pj(nb);
        std::vector<Hydro *> pf(nb);
        std::vector<int> nip(nb), njp(nb);
        for (int b = 0; b < nb; b++) {
            const size_t w = w0 + b;
            for (int i = h_ioff[w]; i < h_ioff[w + 1]; i++)
                epi[b].push_back(HydroEPI{id[i], {x[i], y[i], z[i]}, {vx[i], vy[i], vz[i]}, 1.0 / hinv[i],
                                          hinv[i], pres[i], presc[i], bswt[i], dens[i], vsnd[i], alph[i],
                                          alphu[i], thrm[i], eta[i]});
            for (int k = h_joff[w]; k < h_joff[w + 1]; k++) {
                const int j = h_jidx[k];
                epj[b].push_back(HydroEPJ{id[j], m[j], {x[j], y[j], z[j]}, {vx[j], vy[j], vz[j]},
                                          1.0 / hinv[j], hinv[j], pres[j], presc[j], bswt[j], dens[j],
                                          vsnd[j], alph[j], alphu[j], thrm[j], eta[j]});
            }
            force[b].resize(epi[b].size());
            pe[b] = epi[b].data();
            pj[b] = epj[b].data();
            pf[b] = force[b].data();
            nip[b] = (int)epi[b].size();
            njp[b] = (int)epj[b].size();
        }
        const auto t0 = clk::now();
        DispatchKernel_HydroForce(0, nb, pe.data(), nip.data(), pj.data(), njp.data());
        RetrieveKernel_HydroForce(0, nb, nip.data(), pf.data());
        if (rep > 0) t_ms[1] += since(t0);
        if (rep == 1)
            for (int b = 0; b < nb; b++) {
                h2d[1] += (nip[b] + njp[b]) * 140.0 + 8.0;   // id + 17 reals per record
                d2h[1] += nip[b] * 72.0;
            }
        for (int b = 0; b < nb; b++)
            for (int i = 0; i < nip[b]; i++) {
                const size_t g = h_ioff[w0 + b] + i;
                const Hydro &f = force[b][i];
                const double v[9] = {f.acch.x, f.acch.y, f.acch.z, f.accg.x, f.accg.y, f.accg.z,
                                     f.udot, f.vsmx, f.diffu};
                for (int k = 0; k < 9; k++) h[k][g] = v[k];
            }
    }
    const char *hn[9] = {"h_ax", "h_ay", "h_az", "h_gx", "h_gy", "h_gz", "h_udot", "h_vsmx", "h_diffu"};
    for (int k = 0; k < 9; k++) out(hn[k], h[k]);

    // ---- gravity (calcForceMakingTreeMultiWalk)
    std::vector<double> gr[5];
    for (auto &v : gr) v.resize(n);
    for (int rep = 0; rep <= reps; rep++)
    for (size_t w0 = 0; w0 < nw_g; w0 += kWalkLimit) {
        const int nb = (int)std::min<size_t>(kWalkLimit, nw_g - w0);
        std::vector<std::vector<GravityEPI>> epi(nb);
        std::vector<std::vector<GravityEPJ>> epj(nb);
        std::vector<std::vector<ParticleSimulator::GravitySPJ>> spj(nb);
        std::vector<std::vector<Gravity>> force(nb);
        std::vector<const GravityEPI *> pe(nb);
        std::vector<const GravityEPJ *> pj(nb);
        std::vector<const ParticleSimulator::GravitySPJ *> ps(nb);
        std::vector<Gravity *> pf(nb);
        std::vector<int> nip(nb), njp(nb), nsp(nb);
        for (int b = 0; b < nb; b++) {
            const size_t w = w0 + b;
            for (int i = g_ioff[w]; i < g_ioff[w + 1]; i++)
                epi[b].push_back(GravityEPI{{x[i], y[i], z[i]}, eps2[i]});
            for (int k = g_joff[w]; k < g_joff[w + 1]; k++) {
                const int j = g_jidx[k];
                epj[b].push_back(GravityEPJ{m[j], {x[j], y[j], z[j]}, eps2[j]});
            }
            for (int k = g_soff[w]; k < g_soff[w + 1]; k++) {
                const int s = g_sidx[k];
                spj[b].push_back(ParticleSimulator::GravitySPJ{sm[s], {sx[s], sy[s], sz[s]}, se[s]});
            }
            force[b].resize(epi[b].size());
            pe[b] = epi[b].data();
            pj[b] = epj[b].data();
            ps[b] = spj[b].data();
            pf[b] = force[b].data();
            nip[b] = (int)epi[b].size();
            njp[b] = (int)epj[b].size();
            nsp[b] = (int)spj[b].size();
        }
        const auto t0 = clk::now();
        DispatchKernel_Gravity(0, nb, pe.data(), nip.data(), pj.data(), njp.data(), ps.data(), nsp.data());
        RetrieveKernel_Gravity(0, nb, nip.data(), pf.data());
        if (rep > 0) t_ms[2] += since(t0);
        if (rep == 1)
            for (int b = 0; b < nb; b++) {
                h2d[2] += (nip[b] + njp[b] + nsp[b]) * 40.0 + 12.0;   // 5 reals per record
                d2h[2] += nip[b] * 40.0;
            }
        for (int b = 0; b < nb; b++)
            for (int i = 0; i < nip[b]; i++) {
                const size_t g = g_ioff[w0 + b] + i;
                const Gravity &f = force[b][i];
                gr[0][g] = f.acc.x, gr[1][g] = f.acc.y, gr[2][g] = f.acc.z, gr[3][g] = f.pot, gr[4][g] = f.eta;
            }
    }
    const char *gn[5] = {"g_ax", "g_ay", "g_az", "g_pot", "g_eta"};
    for (int k = 0; k < 5; k++) out(gn[k], gr[k]);
    reset_gpu(0);
    (void)nspj;
    if (reps > 0)
        std::printf("{\"density_ms\": %.3f, \"hydro_ms\": %.3f, \"gravity_ms\": %.3f, \"h2d_bytes\": [%.0f, %.0f, %.0f], "
                    "\"d2h_bytes\": [%.0f, %.0f, %.0f], \"walk_limit\": %d, \"reps\": %d}\n",
                    t_ms[0] / reps, t_ms[1] / reps, t_ms[2] / reps, h2d[0], h2d[1], h2d[2], d2h[0], d2h[1], d2h[2],
                    kWalkLimit, reps);
    std::printf("fdps_driver: %zu particles, %zu/%zu/%zu walks ok\n", n, nw_d, nw_h, nw_g);
    return 0;
}
