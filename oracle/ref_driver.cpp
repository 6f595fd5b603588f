// ref_driver.cpp -- extern "C" entry points around the reference's own scheme
// classes, compiled together with the unmodified reference sources
// (src/weno5js.cpp, src/weno5z.cpp, src/weno7js.cpp, src/crweno5.cpp,
// src/weno3js.cpp, src/upwind1st.cpp, src/hccs78.cpp, src/hccs77.cpp,
// src/centraldifference2nd.cpp, src/centraldifference6th.cpp,
// src/centraldifference8th.cpp, src/TVD-RK3.cpp)
// into oracle/_ref/libnucref.so by oracle/Makefile.
//
// TEST INFRASTRUCTURE ONLY: used by tests/ to pin the C restatement in
// oracle/nuc_oracle.c against the reference implementation.  hccs.cpp does
// not compile as shipped (src/hccs.cpp:61-67 uses ibeg/iend before their
// declaration) and physicsModel/NaiverStokes3d need MPI and FFTW, which this
// image lacks; those parts are pinned by closed forms instead.
#include "weno5js.h"
#include "weno5z.hpp"
#include "weno7js.hpp"
#include "crweno5.hpp"
#include "weno3js.hpp"
#include "upwind1st.hpp"
#include "hccs78.hpp"
#include "hccs77.hpp"
#include "centraldifference2nd.hpp"
#include "centraldifference6th.h"
#include "centraldifference8th.hpp"
#include "TVD-RK3.h"
#include <cstring>

using nuc3d::Field;

static Field make_field(const double *src, int n1, int n2, int n3)
{
    Field f(n1, n2, n3);
    std::memcpy(f.getDataPtr(), src, sizeof(double) * (size_t)n1 * n2 * n3);
    return f;
}

extern "C" {

// scheme ids match oracle/nuc_oracle.h: 0 weno5js, 1 weno5z, 2 weno7js, 3 crweno5,
// 5 weno3js, 6 upwind1st, 7 hccs78, 8 hccs77
int ref_reconstruct(int scheme, int upwind, int tilesize, const double *in, int n1, int n2, int n3,
                    const double *gL, const double *gR, int bf, double *out)
{
    Field fin = make_field(in, n1, n2, n3);
    Field fl = make_field(gL, bf, n2, n3);
    Field fr = make_field(gR, bf, n2, n3);
    Field fo(n1 + 1, n2, n3);
    nuc3d::interoplation *op = nullptr;
    nuc3d::weno5js s0;
    nuc3d::weno5z s1;
    nuc3d::weno7js s2;
    nuc3d::crweno5 s3;
    nuc3d::weno3js s5;
    nuc3d::upwind1st s6;
    nuc3d::hccs78 s7;
    nuc3d::hccs77 s8;
    switch (scheme) {
    case 0: s0.interpolationInner(fin, upwind, fo, tilesize); s0.interpolationBoundaryL(fin, fl, upwind, fo, tilesize);
            s0.interpolationBoundaryR(fin, fr, upwind, fo, tilesize); break;
    case 1: s1.interpolationInner(fin, upwind, fo, tilesize); s1.interpolationBoundaryL(fin, fl, upwind, fo, tilesize);
            s1.interpolationBoundaryR(fin, fr, upwind, fo, tilesize); break;
    case 2: s2.interpolationInner(fin, upwind, fo, tilesize); s2.interpolationBoundaryL(fin, fl, upwind, fo, tilesize);
            s2.interpolationBoundaryR(fin, fr, upwind, fo, tilesize); break;
    case 3: s3.interpolationInner(fin, upwind, fo, tilesize); s3.interpolationBoundaryL(fin, fl, upwind, fo, tilesize);
            s3.interpolationBoundaryR(fin, fr, upwind, fo, tilesize); break;
    case 5: s5.interpolationInner(fin, upwind, fo, tilesize); s5.interpolationBoundaryL(fin, fl, upwind, fo, tilesize);
            s5.interpolationBoundaryR(fin, fr, upwind, fo, tilesize); break;
    case 6: s6.interpolationInner(fin, upwind, fo, tilesize); s6.interpolationBoundaryL(fin, fl, upwind, fo, tilesize);
            s6.interpolationBoundaryR(fin, fr, upwind, fo, tilesize); break;
    case 7: s7.interpolationInner(fin, upwind, fo, tilesize); s7.interpolationBoundaryL(fin, fl, upwind, fo, tilesize);
            s7.interpolationBoundaryR(fin, fr, upwind, fo, tilesize); break;
    case 8: s8.interpolationInner(fin, upwind, fo, tilesize); s8.interpolationBoundaryL(fin, fl, upwind, fo, tilesize);
            s8.interpolationBoundaryR(fin, fr, upwind, fo, tilesize); break;
    default: return -1;
    }
    (void)op;
    std::memcpy(out, fo.getDataPtr(), sizeof(double) * (size_t)(n1 + 1) * n2 * n3);
    return 0;
}

// scheme: 0 cd6, 1 cd8, 2 cd2
int ref_diff(int scheme, int tilesize, const double *in, double h, int n1, int n2, int n3,
             const double *gL, const double *gR, int bf, double *out)
{
    Field fin = make_field(in, n1, n2, n3);
    Field fl = make_field(gL, bf, n2, n3);
    Field fr = make_field(gR, bf, n2, n3);
    Field fo(n1, n2, n3);
    if (scheme == 0) {
        nuc3d::centraldifference6th d;
        d.differentialInner(fin, h, fo, tilesize);
        d.differentialBoundaryL(fin, h, fl, fo, tilesize);
        d.differentialBoundaryR(fin, h, fr, fo, tilesize);
    } else if (scheme == 1) {
        nuc3d::centraldifference8th d;
        d.differentialInner(fin, h, fo, tilesize);
        d.differentialBoundaryL(fin, h, fl, fo, tilesize);
        d.differentialBoundaryR(fin, h, fr, fo, tilesize);
    } else if (scheme == 2) {
        nuc3d::centraldifference2nd d;
        d.differentialInner(fin, h, fo, tilesize);
        d.differentialBoundaryL(fin, h, fl, fo, tilesize);
        d.differentialBoundaryR(fin, h, fr, fo, tilesize);
    } else {
        return -1;
    }
    std::memcpy(out, fo.getDataPtr(), sizeof(double) * (size_t)n1 * n2 * n3);
    return 0;
}

// one TVD-RK3 stage on nvar fields of n cells each (VectorField of Field(n,1,1))
int ref_rk3(int stage, const double *rhs, const double *un, double *ui, double dt, int nvar, int n)
{
    nuc3d::VectorField R, U, UI;
    for (int v = 0; v < nvar; v++) {
        R.push_back(make_field(rhs + (size_t)v * n, n, 1, 1));
        U.push_back(make_field(un + (size_t)v * n, n, 1, 1));
        UI.push_back(make_field(ui + (size_t)v * n, n, 1, 1));
    }
    nuc3d::tvdrk3rd rk;
    rk.integrationAll(R, U, UI, dt, stage);
    for (int v = 0; v < nvar; v++) std::memcpy(ui + (size_t)v * n, UI[v].getDataPtr(), sizeof(double) * n);
    return 0;
}

}  // extern "C"
