// nuc3d_host.hpp -- host C++ classes with the reference's operator/plug-in
// interface (src/field.h, src/fieldOperator.h, src/TVD-RK3.h,
// src/euler3d.h, src/NaiverStokes3d.h, src/PDEData3d.hpp), forwarding to the
// gfx950 C ABI in include/nuc3d_gpu.h.  Code written against the reference's
// classes keeps its call sites; the arithmetic runs on the GPU.
#pragma once
#include "nuc3d_gpu.h"

#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace nuc3d {

// field3d<double>: x-fastest host array (src/field.h:9-79, src/field.cpp:37-108)
class Field {
    std::vector<double> d_;
    int nx_, ny_, nz_;

public:
    Field(int nx, int ny, int nz) : d_((size_t)nx * ny * nz, 1.0), nx_(nx), ny_(ny), nz_(nz) {}  // fill 1.0 (src/field.cpp:52)
    Field(int nx, int ny, int nz, double v) : d_((size_t)nx * ny * nz, v), nx_(nx), ny_(ny), nz_(nz) {}
    int getSizeX() const { return nx_; }
    int getSizeY() const { return ny_; }
    int getSizeZ() const { return nz_; }
    int getSize() const { return nx_ * ny_ * nz_; }
    int getIndex(int i, int j, int k) const
    {
        if (i < 0 || i >= nx_ || j < 0 || j >= ny_ || k < 0 || k >= nz_) throw std::out_of_range("Field index");
        return nx_ * ny_ * k + nx_ * j + i;
    }
    double getValue(int i, int j, int k) const { return d_[getIndex(i, j, k)]; }
    void setValue(int i, int j, int k, double v) { d_[getIndex(i, j, k)] = v; }
    double *getDataPtr() const { return const_cast<double *>(d_.data()); }
};
typedef std::vector<Field> VectorField;

// error from the C ABI; the reference exits, the host classes throw instead
struct nuc_error : std::runtime_error {
    int code;
    nuc_error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};
void check(int rc, const char *what);

// ---- plug-in interfaces (src/fieldOperator.h:13-81) ----------------------
class interoplation {
public:
    virtual ~interoplation() = default;
    virtual void interpolationInner(const Field &in, const int upwind, Field &out, const int tilesize) = 0;
    virtual void interpolationBoundaryL(const Field &in, const Field &ghostL, const int upwind, Field &out,
                                        const int tilesize) = 0;
    virtual void interpolationBoundaryR(const Field &in, const Field &ghostR, const int upwind, Field &out,
                                        const int tilesize) = 0;
};

class differential {
public:
    virtual ~differential() = default;
    virtual void differentialInner(const Field &in, const double &h, Field &out, const int tilesize) = 0;
    virtual void differentialBoundaryL(const Field &in, const double &h, const Field &ghostL, Field &out,
                                       const int tilesize) = 0;
    virtual void differentialBoundaryR(const Field &in, const double &h, const Field &ghostR, Field &out,
                                       const int tilesize) = 0;
};

class integration {
public:
    int nstep;
    explicit integration(int n) : nstep(n) {}
    virtual ~integration() = default;
    virtual void integrationAll(const VectorField &rhs, const VectorField &u_n, VectorField &u_i, double dt,
                                int stage) = 0;
};

// one GPU-backed reconstruction class per scheme id
class gpuInterpolation : public interoplation {
    int scheme_;

public:
    explicit gpuInterpolation(int scheme) : scheme_(scheme) {}
    int scheme() const { return scheme_; }
    void interpolationInner(const Field &in, const int upwind, Field &out, const int tilesize) override;
    void interpolationBoundaryL(const Field &in, const Field &ghostL, const int upwind, Field &out,
                                const int tilesize) override;
    void interpolationBoundaryR(const Field &in, const Field &ghostR, const int upwind, Field &out,
                                const int tilesize) override;
};
struct weno5js : gpuInterpolation { weno5js() : gpuInterpolation(NUC_WENO5JS) {} };
struct weno5z : gpuInterpolation { weno5z() : gpuInterpolation(NUC_WENO5Z) {} };
struct weno7js : gpuInterpolation { weno7js() : gpuInterpolation(NUC_WENO7JS) {} };
struct crweno5 : gpuInterpolation { crweno5() : gpuInterpolation(NUC_CRWENO5) {} };
struct hccs : gpuInterpolation { hccs() : gpuInterpolation(NUC_HCCS) {} };

class gpuDifferential : public differential {
    int scheme_;

public:
    explicit gpuDifferential(int s) : scheme_(s) {}
    void differentialInner(const Field &in, const double &h, Field &out, const int tilesize) override;
    void differentialBoundaryL(const Field &in, const double &h, const Field &ghostL, Field &out,
                               const int tilesize) override;
    void differentialBoundaryR(const Field &in, const double &h, const Field &ghostR, Field &out,
                               const int tilesize) override;
};
struct centraldifference6th : gpuDifferential { centraldifference6th() : gpuDifferential(NUC_CD6) {} };
struct centraldifference8th : gpuDifferential { centraldifference8th() : gpuDifferential(NUC_CD8) {} };

// TVD-RK3 (src/TVD-RK3.h:7-37)
class tvdrk3rd : public integration {
public:
    tvdrk3rd() : integration(3) {}
    void integrationAll(const VectorField &rhs, const VectorField &u_n, VectorField &u_i, double dt,
                        int stage) override;
};

// run control read from inp/*.in
struct RunControl {
    int start_step = 0, end_step = 100000, post_step = 1, save_step = 1000, benchmark = 0;
    bool binary = true, tecplot = true, postproc = true;   // IOController.in file types
    int post_save_step = 100;
    // PostProc.in, six positional values (src/postproc.cpp:36-82): TimeAverageType,
    // TimeAverageStartStep, SimutaneousData, AveragedSaveType, SimutaneousSaveType,
    // fft post type; absent file: no post-processing
    int average_type = 0, average_start = 0, simultaneous = 0, averaged_save_type = 0, simultaneous_save = 0;
    double tl = 0.0;
    double origin[3] = {0, 0, 0};
    double mach_t = 0.3, k_peak = 8.0;
};
int read_inp(const std::string &dir, nuc_config &c, RunControl &rc, std::string &err);
int scheme_from_name(const std::string &s, int *halo);

// fieldOperator3d: scheme selection from inp/Method.in and dispatch per
// direction (src/fieldOperator.h:83-200, src/fieldOperator.cpp)
class fieldOperator3d {
    std::vector<std::shared_ptr<interoplation>> interp_;
    std::vector<std::shared_ptr<differential>> diff_;
    std::shared_ptr<integration> integ_;

public:
    int bufferSize_xi = 3, bufferSize_eta = 3, bufferSize_zeta = 3;
    explicit fieldOperator3d(const nuc_config &c);
    int getBufferSize() const;
    int getSteps() const { return integ_->nstep; }
    void reconstructionInner(const Field &in, int dir, int upwind, Field &out);
    void reconstructionBoundary(const Field &in, const Field &gL, const Field &gR, int dir, int upwind, Field &out);
    void differenceInner(const Field &in, const double &h, int dir, Field &out);
    void differenceBoundary(const Field &in, const double &h, const Field &gL, const Field &gR, int dir, Field &out);
    void timeIntegral(const VectorField &rhs, const VectorField &u_n, VectorField &u_i, double dt, int stage);

private:
    int ts(int dir) const { return dir == 0 ? bufferSize_xi : (dir == 1 ? bufferSize_eta : bufferSize_zeta); }
    int tsv(int dir) const;
    int vis_[3];
};

// PDEData3d + NaiverStokesData3d/EulerData3D facade over one device context:
// the state lives in HBM, the host copies only on request.
class PDEData3d {
    nuc_ctx *ctx_ = nullptr;
    nuc_config cfg_;

public:
    explicit PDEData3d(const nuc_config &c);
    ~PDEData3d();
    PDEData3d(const PDEData3d &) = delete;
    PDEData3d &operator=(const PDEData3d &) = delete;
    nuc_ctx *ctx() { return ctx_; }
    const nuc_config &config() const { return cfg_; }
    void setQ(const VectorField &Q);
    void getQ(VectorField &Q);
    // one RHS evaluation (NaiverStokesData3d::solve); returns dt_local
    double solveRHS(VectorField *rhs);
    // one RK3 step (block::solve)
    nuc_step_info step(int n = 1);
    // one RK stage: RHS of the current stage state and the stage update
    // (NaiverStokesData3d::solve + PDEData3d::solve, src/block.cpp:198-202)
    void solve(int stage);
};

// Fluid at rest of src/block.cpp:285-314
void initial_default(const nuc_config &c, std::vector<double> &q);
// Isentropic vortex of src/block.cpp:316-357 on the rank grid
void initial_ivc(const nuc_config &c, const double origin[3], std::vector<double> &q);
// Taylor-Green initial condition of src/block.cpp:359-395 on the rank grid
void initial_taylorgreen(const nuc_config &c, const double origin[3], std::vector<double> &q);
// Seeded random-phase solenoidal field with E(k) ~ k^4 exp(-2 (k/k0)^2)
// (src/block.cpp:397-551 with a fixed seed); rho = 1, p = 1/(gamma M^2)
void initial_hit(const nuc_config &c, double k0, double u_rms, unsigned long long seed, std::vector<double> &q);
// Shu-Osher profile along x with a seeded solenoidal velocity perturbation
void initial_shu_osher(const nuc_config &c, double pert, unsigned long long seed, std::vector<double> &q);

}  // namespace nuc3d
