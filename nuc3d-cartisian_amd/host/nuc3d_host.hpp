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
struct hccs77 : gpuInterpolation { hccs77() : gpuInterpolation(NUC_HCCS77) {} };
struct hccs78 : gpuInterpolation { hccs78() : gpuInterpolation(NUC_HCCS78) {} };
struct weno3js : gpuInterpolation { weno3js() : gpuInterpolation(NUC_WENO3JS) {} };
struct upwind1st : gpuInterpolation { upwind1st() : gpuInterpolation(NUC_UPWIND1ST) {} };

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
struct centraldifference2nd : gpuDifferential { centraldifference2nd() : gpuDifferential(NUC_CD2) {} };

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
// <dir>/Method.in: inviscid and viscous scheme ids per direction
int read_method(const std::string &dir, int scheme[3], int vis_scheme[3], std::string &err);
int scheme_from_name(const std::string &s, int *halo);

// fieldOperator3d: scheme selection from inp/Method.in and dispatch per
// direction (src/fieldOperator.h:83-200, src/fieldOperator.cpp)
class fieldOperator3d {
    std::vector<std::shared_ptr<interoplation>> interp_;
    std::vector<std::shared_ptr<differential>> diff_;
    std::shared_ptr<integration> integ_;

public:
    // bufferSize per scheme: weno7js/hccs/hccs77/hccs78 4, weno5js/weno5z/crweno5 3,
    // weno3js 2, upwind1st 1 (src/fieldOperator.cpp:128-220)
    int bufferSize_xi = 3, bufferSize_eta = 3, bufferSize_zeta = 3;
    // reads inp/Method.in of the working directory, as the reference's
    // constructor does (src/fieldOperator.cpp:21-78); throws if it is absent
    fieldOperator3d();
    // the same from <inp_dir>/Method.in
    explicit fieldOperator3d(const std::string &inp_dir);
    explicit fieldOperator3d(const nuc_config &c);
    int getBufferSize() const;
    int getSteps() const { return integ_->nstep; }
    // scheme ids in use (NUC_WENO5JS..., NUC_CD6...)
    int schemeId(int dir) const { return scheme_[dir]; }
    int visSchemeId(int dir) const { return vis_[dir]; }
    void reconstructionInner(const Field &in, int dir, int upwind, Field &out);
    void reconstructionBoundary(const Field &in, const Field &gL, const Field &gR, int dir, int upwind, Field &out);
    void differenceInner(const Field &in, const double &h, int dir, Field &out);
    void differenceBoundary(const Field &in, const double &h, const Field &gL, const Field &gR, int dir, Field &out);
    void timeIntegral(const VectorField &rhs, const VectorField &u_n, VectorField &u_i, double dt, int stage);

private:
    void build(const int scheme[3], const int vis[3]);
    int ts(int dir) const { return dir == 0 ? bufferSize_xi : (dir == 1 ? bufferSize_eta : bufferSize_zeta); }
    int tsv(int dir) const;
    int scheme_[3], vis_[3];
};

// Stand-ins for the reference's communicator, halo buffers, physics model and
// FFT object in the PDE-level signatures.  On the device the periodic halo is
// an index wrap and the forcing transform lives in the context, so these carry
// only what the call sites read: the rank grid, the model constants.
class MPIComunicator3d_nonblocking {
    int dims_[3];

public:
    MPIComunicator3d_nonblocking(int npx = 1, int npy = 1, int npz = 1) : dims_{npx, npy, npz} {}
    explicit MPIComunicator3d_nonblocking(const nuc_config &c) : dims_{c.npx, c.npy, c.npz} {}
    int getMyDim(int d) const { return dims_[d]; }
    int getMyId() const { return 0; }
    void barrier() const {}
};
struct bufferData {};
typedef std::vector<bufferData> VectorBuffer;
class fft {};
// physicsModel (src/physicsModel.h): the constants of inp/PhysModel.in
class physicsModel {
    nuc_config c_;

public:
    explicit physicsModel(const nuc_config &c) : c_(c) {}
    const nuc_config &config() const { return c_; }
    double getGamma() const { return c_.gamma; }
    double getMach() const { return c_.mach; }
    double getReynolds() const { return c_.reynolds; }
    double getPrandtl() const { return c_.prandtl; }
};

// PDEData3d (src/PDEData3d.hpp:21-81) over one device context: Q, the stage
// values and RHS live in HBM; the VectorField accessors copy to the host.
class PDEData3d {
    nuc_ctx *ctx_ = nullptr;
    nuc_config cfg_;
    VectorField Q_, RHS_;
    bool rhs_pending_ = false;   // EulerData3D::solve ran for the current stage
    friend class EulerData3D;

public:
    int nEquations = 5;
    double dt_local = 0.0, dt_global = 0.0, res_global = 0.0;
    explicit PDEData3d(const nuc_config &c);
    ~PDEData3d();
    PDEData3d(const PDEData3d &) = delete;
    PDEData3d &operator=(const PDEData3d &) = delete;
    nuc_ctx *ctx() { return ctx_; }
    const nuc_config &config() const { return cfg_; }
    void setQ(const VectorField &Q);
    // the current stage state (Q_work, src/PDEData3d.cpp:41-58)
    VectorField &getQ();
    // RHS of the current stage state (dQ/dt = -RHS)
    VectorField &getRHS();
    double getDt() const { return dt_global; }
    double getRes() const { return res_global; }
    // PDEData3d::solve (src/PDEData3d.cpp:70-91): the stage update of stage
    // `step` from the RHS that EulerData3D::solve requested; dt_global = cfl *
    // dt_local at stage 0, the residual at stage 2.  The device fuses the RHS
    // and the update of a stage (nuc_rk3_stage); cfl must be the configured one.
    void solve(fieldOperator3d &op, MPIComunicator3d_nonblocking &mpi, double cfl, int step);
    // whole RK3 steps (block::solve, src/block.cpp:191-215)
    nuc_step_info step(int n = 1);
};

// EulerData3D / NaiverStokesData3d (src/euler3d.h:84-90, src/NaiverStokes3d.h:51-57):
// solve() requests the RHS of the PDE's current stage state; the model
// (Euler or Navier-Stokes) is the context's.  A following PDEData3d::solve
// runs it fused with the stage update; getRHS() runs it alone.
class EulerData3D {
protected:
    int model_;

public:
    explicit EulerData3D(int model = NUC_EULER3D) : model_(model) {}
    virtual ~EulerData3D() = default;
    virtual void solve(PDEData3d &pde, fieldOperator3d &op, VectorBuffer &buf, physicsModel &phys,
                       MPIComunicator3d_nonblocking &mpi, fft &myFFT, const VectorField &xyz_center);
};
class NaiverStokesData3d : public EulerData3D {
public:
    NaiverStokesData3d() : EulerData3D(NUC_NAVIERSTOKES3D) {}
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
