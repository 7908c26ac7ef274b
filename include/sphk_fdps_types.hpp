// sphk_fdps_types.hpp -- data layouts of the reference's EPI/EPJ/SPJ/Force
// classes, for code that has no FDPS installation (the adapter inside
// libsphk.so and its test driver).
//
// Member order and types follow sph_yamauchi/run/hdr_density.hpp:4-78,
// hdr_hydro.hpp:3-126 and hdr_gravity.hpp:63-192 (FDPS: S32 = int,
// S64 = long long, F64 = double, F64vec = three doubles).  Only the data
// members are reproduced -- the member functions of the reference classes
// touch the caller's particle class and do not change the layout.  A
// program that already has the reference headers must not include this
// file in the same translation unit.
#pragma once

#include <cstdint>

namespace sphk_fdps {
struct F64vec {
    double x, y, z;
};
}  // namespace sphk_fdps

class Density {
public:
    double dens, divv, rotv, grdh, ksr;
    long long np;
    bool itr;
};

class DensityEPI {
public:
    int id;
    double mass;
    sphk_fdps::F64vec pos, vel;
    double ksr, rs;
};

class DensityEPJ {
public:
    int id;
    double mass;
    sphk_fdps::F64vec pos, vel;
};

class Hydro {
public:
    sphk_fdps::F64vec acch, accg;
    double udot, vsmx, diffu;
};

class HydroEPI {
public:
    long long id;
    sphk_fdps::F64vec pos, vel;
    double ksr, hinv, pres, presc, bswt, dens, vsnd, alph, alphu, thrm, eta;
};

class HydroEPJ {
public:
    long long id;
    double mass;
    sphk_fdps::F64vec pos, vel;
    double ksr, hinv, pres, presc, bswt, dens, vsnd, alph, alphu, thrm, eta;
};

class Gravity {
public:
    sphk_fdps::F64vec acc;
    double pot, eta;
};

class GravityEPI {
public:
    sphk_fdps::F64vec pos;
    double eps2;
};

class GravityEPJ {
public:
    double mass;
    sphk_fdps::F64vec pos;
    double eps2;
};

namespace ParticleSimulator {
class GravitySPJ {
public:
    double mass;
    sphk_fdps::F64vec pos;
    double eps2;
};
}  // namespace ParticleSimulator

static_assert(sizeof(DensityEPI) == 80 && sizeof(DensityEPJ) == 64 && sizeof(Density) == 56,
              "density layouts");
static_assert(sizeof(HydroEPI) == 144 && sizeof(HydroEPJ) == 152 && sizeof(Hydro) == 72,
              "hydro layouts");
static_assert(sizeof(GravityEPI) == 32 && sizeof(GravityEPJ) == 40 &&
                  sizeof(ParticleSimulator::GravitySPJ) == 40 && sizeof(Gravity) == 40,
              "gravity layouts");
