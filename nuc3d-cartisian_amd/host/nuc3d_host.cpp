// nuc3d_host.cpp -- host classes of nuc3d_host.hpp and host utilities
// (initial conditions) exported through the C ABI.
#include "nuc3d_host.hpp"
#include "hit_spectrum.hpp"

#include <cmath>
#include <complex>
#include <cstdio>
#include <fstream>
#include <random>

namespace nuc3d {

void check(int rc, const char *what)
{
    if (rc != NUC_OK) throw nuc_error(rc, std::string(what) + " failed with status " + std::to_string(rc));
}

// ---- operators --------------------------------------------------------------
static void recon(int scheme, const Field &in, const Field *gL, const Field *gR, int upwind, Field &out, int ts,
                  int zones)
{
    int bf = gL ? gL->getSizeX() : (gR ? gR->getSizeX() : 0);
    check(nuc_op_reconstruct(scheme, upwind, ts, in.getDataPtr(), in.getSizeX(), in.getSizeY(), in.getSizeZ(),
                             gL ? gL->getDataPtr() : nullptr, gR ? gR->getDataPtr() : nullptr, bf,
                             out.getDataPtr(), zones),
          "nuc_op_reconstruct");
}

void gpuInterpolation::interpolationInner(const Field &in, const int upwind, Field &out, const int tilesize)
{
    recon(scheme_, in, nullptr, nullptr, upwind, out, tilesize, 1);
}
void gpuInterpolation::interpolationBoundaryL(const Field &in, const Field &gL, const int upwind, Field &out,
                                              const int tilesize)
{
    recon(scheme_, in, &gL, nullptr, upwind, out, tilesize, 2);
}
void gpuInterpolation::interpolationBoundaryR(const Field &in, const Field &gR, const int upwind, Field &out,
                                              const int tilesize)
{
    recon(scheme_, in, nullptr, &gR, upwind, out, tilesize, 4);
}

static void cdiff(int scheme, const Field &in, double h, const Field *gL, const Field *gR, Field &out, int ts,
                  int zones)
{
    int bf = gL ? gL->getSizeX() : (gR ? gR->getSizeX() : 0);
    check(nuc_op_diff(scheme, ts, in.getDataPtr(), h, in.getSizeX(), in.getSizeY(), in.getSizeZ(),
                      gL ? gL->getDataPtr() : nullptr, gR ? gR->getDataPtr() : nullptr, bf, out.getDataPtr(), zones),
          "nuc_op_diff");
}
void gpuDifferential::differentialInner(const Field &in, const double &h, Field &out, const int tilesize)
{
    cdiff(scheme_, in, h, nullptr, nullptr, out, tilesize, 1);
}
void gpuDifferential::differentialBoundaryL(const Field &in, const double &h, const Field &gL, Field &out,
                                            const int tilesize)
{
    cdiff(scheme_, in, h, &gL, nullptr, out, tilesize, 2);
}
void gpuDifferential::differentialBoundaryR(const Field &in, const double &h, const Field &gR, Field &out,
                                            const int tilesize)
{
    cdiff(scheme_, in, h, nullptr, &gR, out, tilesize, 4);
}

void tvdrk3rd::integrationAll(const VectorField &rhs, const VectorField &u_n, VectorField &u_i, double dt, int stage)
{
    for (size_t v = 0; v < u_i.size(); v++)
        check(nuc_op_rk3(stage, rhs[v].getDataPtr(), u_n[v].getDataPtr(), u_i[v].getDataPtr(), dt,
                         u_i[v].getSize()),
              "nuc_op_rk3");
}

// bufferSize of an inviscid scheme id (src/fieldOperator.cpp:128-220)
static int scheme_buffer(int scheme)
{
    if (scheme == NUC_WENO7JS || scheme == NUC_HCCS || scheme == NUC_HCCS77 || scheme == NUC_HCCS78) return 4;
    if (scheme == NUC_WENO3JS) return 2;
    if (scheme == NUC_UPWIND1ST) return 1;
    return 3;
}

void fieldOperator3d::build(const int scheme[3], const int vis[3])
{
    int *bs[3] = {&bufferSize_xi, &bufferSize_eta, &bufferSize_zeta};
    interp_.clear();
    diff_.clear();
    for (int d = 0; d < 3; d++) {
        *bs[d] = scheme_buffer(scheme[d]);
        interp_.push_back(std::make_shared<gpuInterpolation>(scheme[d]));
        diff_.push_back(std::make_shared<gpuDifferential>(vis[d]));
        scheme_[d] = scheme[d];
        vis_[d] = vis[d];
    }
    integ_ = std::make_shared<tvdrk3rd>();   // scheme_time RK3, the only choice (src/fieldOperator.cpp:366-385)
}

fieldOperator3d::fieldOperator3d(const nuc_config &c) { build(c.scheme, c.vis_scheme); }

fieldOperator3d::fieldOperator3d(const std::string &inp_dir)
{
    int sch[3], vis[3];
    std::string err;
    if (read_method(inp_dir, sch, vis, err) != NUC_OK) throw nuc_error(NUC_EINVAL, err);
    build(sch, vis);
}

fieldOperator3d::fieldOperator3d() : fieldOperator3d(std::string("inp")) {}

int fieldOperator3d::getBufferSize() const
{
    int b = std::max(bufferSize_xi, std::max(bufferSize_eta, bufferSize_zeta));
    for (int d = 0; d < 3; d++) b = std::max(b, tsv(d));
    return b;
}

int fieldOperator3d::tsv(int dir) const
{
    // the viscous stencil's reach (cd2 1, cd6 3, cd8 4) when the inviscid halo
    // is narrower (the reference passes the inviscid halo, src/fieldOperator.cpp:520-526)
    const int reach = vis_[dir] == NUC_CD8 ? 4 : (vis_[dir] == NUC_CD2 ? 1 : 3);
    return std::max(ts(dir), reach);
}

void fieldOperator3d::reconstructionInner(const Field &in, int dir, int upwind, Field &out)
{
    interp_[dir]->interpolationInner(in, upwind, out, ts(dir));
}
void fieldOperator3d::reconstructionBoundary(const Field &in, const Field &gL, const Field &gR, int dir, int upwind,
                                             Field &out)
{
    interp_[dir]->interpolationBoundaryL(in, gL, upwind, out, ts(dir));
    interp_[dir]->interpolationBoundaryR(in, gR, upwind, out, ts(dir));
}
void fieldOperator3d::differenceInner(const Field &in, const double &h, int dir, Field &out)
{
    diff_[dir]->differentialInner(in, h, out, tsv(dir));
}
void fieldOperator3d::differenceBoundary(const Field &in, const double &h, const Field &gL, const Field &gR, int dir,
                                         Field &out)
{
    diff_[dir]->differentialBoundaryL(in, h, gL, out, tsv(dir));
    diff_[dir]->differentialBoundaryR(in, h, gR, out, tsv(dir));
}
void fieldOperator3d::timeIntegral(const VectorField &rhs, const VectorField &u_n, VectorField &u_i, double dt,
                                   int stage)
{
    integ_->integrationAll(rhs, u_n, u_i, dt, stage);
}

// ---- PDE level ------------------------------------------------------------------
PDEData3d::PDEData3d(const nuc_config &c) : cfg_(c) { check(nuc_create(&c, &ctx_), "nuc_create"); }
PDEData3d::~PDEData3d() { nuc_destroy(ctx_); }

static void to_fields(const nuc_config &c, const std::vector<double> &q, VectorField &F)
{
    const size_t N = (size_t)c.nx * c.ny * c.nz;
    F.clear();
    for (int v = 0; v < 5; v++) {
        F.emplace_back(c.nx, c.ny, c.nz);
        std::memcpy(F[v].getDataPtr(), q.data() + v * N, N * sizeof(double));
    }
}

void PDEData3d::setQ(const VectorField &Q)
{
    size_t N = (size_t)cfg_.nx * cfg_.ny * cfg_.nz;
    if (Q.size() != 5) throw nuc_error(NUC_EINVAL, "setQ: five conserved variables expected");
    std::vector<double> q(5 * N);
    for (int v = 0; v < 5; v++) {
        if ((size_t)Q[v].getSize() != N) throw nuc_error(NUC_EINVAL, "setQ: field size differs from the grid");
        std::memcpy(q.data() + v * N, Q[v].getDataPtr(), N * sizeof(double));
    }
    check(nuc_upload_state(ctx_, q.data()), "nuc_upload_state");
    rhs_pending_ = false;
}

VectorField &PDEData3d::getQ()
{
    std::vector<double> q(5 * (size_t)cfg_.nx * cfg_.ny * cfg_.nz);
    check(nuc_download_state(ctx_, q.data()), "nuc_download_state");
    to_fields(cfg_, q, Q_);
    return Q_;
}

VectorField &PDEData3d::getRHS()
{
    std::vector<double> r(5 * (size_t)cfg_.nx * cfg_.ny * cfg_.nz);
    check(nuc_rhs(ctx_, r.data(), &dt_local), "nuc_rhs");
    to_fields(cfg_, r, RHS_);
    return RHS_;
}

void PDEData3d::solve(fieldOperator3d &, MPIComunicator3d_nonblocking &, double cfl, int step)
{
    if (!rhs_pending_)
        throw nuc_error(NUC_EINVAL, "PDEData3d::solve: call the fluid model's solve for this stage first "
                                    "(src/block.cpp:198-202)");
    if (cfl != cfg_.cfl)
        throw nuc_error(NUC_EINVAL, "PDEData3d::solve: cfl differs from the context's configuration");
    rhs_pending_ = false;
    check(nuc_rk3_stage(ctx_, step), "nuc_rk3_stage");
    // dt_global set at stage 0, the residual at stage 2 (PDEData3d::setRES,
    // src/PDEData3d.cpp:93-128); both live on the device
    nuc_step_info info{};
    check(nuc_sync(ctx_, &info), "nuc_sync");
    dt_global = info.dt;
    if (step == 2) res_global = info.res;
}

nuc_step_info PDEData3d::step(int n)
{
    nuc_step_info info{};
    check(nuc_step(ctx_, n, &info), "nuc_step");
    dt_global = info.dt;
    res_global = info.res;
    rhs_pending_ = false;
    return info;
}

void EulerData3D::solve(PDEData3d &pde, fieldOperator3d &op, VectorBuffer &, physicsModel &phys,
                        MPIComunicator3d_nonblocking &mpi, fft &, const VectorField &)
{
    const nuc_config &c = pde.config();
    if (c.model != model_)
        throw nuc_error(NUC_EINVAL, "EulerData3D::solve: the PDE's context runs the other model");
    for (int d = 0; d < 3; d++)
        if (op.schemeId(d) != c.scheme[d] || op.visSchemeId(d) != c.vis_scheme[d])
            throw nuc_error(NUC_EINVAL, "EulerData3D::solve: fieldOperator3d schemes differ from the context's");
    if (phys.getGamma() != c.gamma || phys.getMach() != c.mach || phys.getReynolds() != c.reynolds)
        throw nuc_error(NUC_EINVAL, "EulerData3D::solve: physicsModel constants differ from the context's");
    if (mpi.getMyDim(0) != c.npx || mpi.getMyDim(1) != c.npy || mpi.getMyDim(2) != c.npz)
        throw nuc_error(NUC_EINVAL, "EulerData3D::solve: rank grid differs from the context's");
    pde.rhs_pending_ = true;
}

// ---- initial conditions -------------------------------------------------------
void initial_taylorgreen(const nuc_config &c, const double origin[3], std::vector<double> &q)
{
    const long N = (long)c.nx * c.ny * c.nz;
    q.assign(5 * N, 0.0);
    int nl[3] = {c.nx / c.npx, c.ny / c.npy, c.nz / c.npz};
    double dx = c.lx / c.nx, dy = c.ly / c.ny, dz = c.lz / c.nz;
    for (int k = 0; k < c.nz; k++)
        for (int j = 0; j < c.ny; j++)
            for (int i = 0; i < c.nx; i++) {
                // cell centres per rank, src/block.cpp:116-118
                double x0 = c.lx / c.npx * (i / nl[0]) + dx * (0.5 + i % nl[0]) - origin[0];
                double y0 = c.ly / c.npy * (j / nl[1]) + dy * (0.5 + j % nl[1]) - origin[1];
                double z0 = c.lz / c.npz * (k / nl[2]) + dz * (0.5 + k % nl[2]) - origin[2];
                double g = c.gamma, m = c.mach;
                double u0 = std::sin(x0) * std::cos(y0) * std::cos(z0);
                double v0 = -std::cos(x0) * std::sin(y0) * std::cos(z0);
                double p0 = 1.0 / (g * m * m)
                          + ((std::cos(2.0 * z0) + 2.0) * (std::cos(2.0 * x0) + std::cos(2.0 * y0)) - 2.0) / 16.0;
                long id = i + (long)c.nx * (j + (long)c.ny * k);
                q[id] = 1.0;
                q[N + id] = u0;
                q[2 * N + id] = v0;
                q[3 * N + id] = 0.0;
                q[4 * N + id] = p0 / (g - 1.0) + 0.5 * (u0 * u0 + v0 * v0);
            }
}

// Fluid at rest, p = 1/(gamma M^2) (src/block.cpp:285-314).  The reference's
// u_init, v_init, w_init come from getValue on a map that does not hold them
// and that PhysModel.in cannot extend, so they are always 0.
void initial_default(const nuc_config &c, std::vector<double> &q)
{
    const long N = (long)c.nx * c.ny * c.nz;
    q.assign(5 * N, 0.0);
    const double e = 1.0 / (c.mach * c.mach * c.gamma) / (c.gamma - 1.0);
    for (long id = 0; id < N; id++) {
        q[id] = 1.0;
        q[4 * N + id] = e;
    }
}

// Isentropic vortex of initial_ivc (src/block.cpp:316-357): centre (5, 5) in
// the x-y plane, strength 0.5, carried by u = 0.5; p = rho^gamma with the
// density exponent 2.5 as the reference writes it.
void initial_ivc(const nuc_config &c, const double origin[3], std::vector<double> &q)
{
    const long N = (long)c.nx * c.ny * c.nz;
    q.assign(5 * N, 0.0);
    int nl[2] = {c.nx / c.npx, c.ny / c.npy};
    const double dx = c.lx / c.nx, dy = c.ly / c.ny, g = c.gamma;
    const double pi = 4.0 * std::atan(1.0), b = 0.5, xc = 5.0, yc = 5.0;
    for (int k = 0; k < c.nz; k++)
        for (int j = 0; j < c.ny; j++)
            for (int i = 0; i < c.nx; i++) {
                const double x0 = c.lx / c.npx * (i / nl[0]) + dx * (0.5 + i % nl[0]) - origin[0] - xc;
                const double y0 = c.ly / c.npy * (j / nl[1]) + dy * (0.5 + j % nl[1]) - origin[1] - yc;
                const double r2 = x0 * x0 + y0 * y0;
                const double rho = std::pow(1.0 - (g - 1.0) * b * b * std::exp(1.0 - r2) / (8.0 * g * pi * pi), 2.5);
                const double a = b / (2.0 * pi) * std::exp(0.5 * (1.0 - r2));
                const double u = 0.5 - a * y0, v = a * x0;
                const long id = i + (long)c.nx * (j + (long)c.ny * k);
                q[id] = rho;
                q[N + id] = rho * u;
                q[2 * N + id] = rho * v;
                q[4 * N + id] = std::pow(rho, g) / (g - 1.0) + 0.5 * rho * (u * u + v * v);
            }
}

typedef std::complex<double> cplx;

// Seeded random-phase solenoidal field.  The spectrum
// E(k) = 16 sqrt(2/pi) (k/k0)^4 / k0 exp(-2 (k/k0)^2) (src/block.cpp:458), k the
// physical wave number, is populated for mode numbers
// |k_x|,|k_y|,|k_z| <= K = min(n/2-1, 4 k0) -- beyond 4 k0 it is
// below 1e-13 of its peak -- and synthesised by a separable inverse DFT with
// FFTW's backward sign (e^{+i}) and no normalisation, keeping the real part
// (src/block.cpp:504-510).  Phases come from mt19937_64(seed), three draws per
// mode in (kz, ky, kx) order, instead of std::random_device (src/block.cpp:416-418).
void hit_spectrum(const nuc_config &c, double k0, unsigned long long seed, HitSpectrum &S)
{
    S.n[0] = c.nx; S.n[1] = c.ny; S.n[2] = c.nz;
    const int *n = S.n;
    const double pie = 4.0 * std::atan(1.0);
    int *K = S.K, *W = S.W;
    for (int d = 0; d < 3; d++) {
        K[d] = std::max(1, std::min(n[d] / 2 - 1, (int)std::ceil(4.0 * k0)));
        W[d] = 2 * K[d] + 1;
    }
    std::mt19937_64 gen(seed);
    std::uniform_real_distribution<double> dis(0.0, 2.0 * pie);
    // spectral amplitudes, comp c at [(kz*W1 + ky)*W0 + kx]
    std::vector<cplx> *spec = S.spec;
    for (int cc = 0; cc < 3; cc++) spec[cc].assign((size_t)W[0] * W[1] * W[2], 0.0);
    for (int kz = -K[2]; kz <= K[2]; kz++)
        for (int ky = -K[1]; ky <= K[1]; ky++)
            for (int kx = -K[0]; kx <= K[0]; kx++) {
                // physical wave numbers 2 pi k / L, so that the field is
                // solenoidal on any box (the factor is exactly 1 on [0, 2 pi)^3)
                double kx0 = kx * (2.0 * pie / c.lx), ky0 = ky * (2.0 * pie / c.ly), kz0 = kz * (2.0 * pie / c.lz);
                double kmag = std::sqrt(kx0 * kx0 + ky0 * ky0 + kz0 * kz0);
                double k12 = std::sqrt(kx0 * kx0 + ky0 * ky0);
                if (kmag < 1e-6) kmag = 1.0;
                double Ek0 = 16.0 * std::sqrt(2.0 / pie) * std::pow(kmag / k0, 4) / k0 *
                             std::exp(-2.0 * kmag * kmag / k0 / k0);
                double phi1 = dis(gen), phi2 = dis(gen), phi3 = dis(gen);
                double amp = std::sqrt(2.0 * Ek0 / kmag / kmag / 4.0 / pie);
                cplx a(amp * std::cos(phi1) * std::cos(phi3), amp * std::sin(phi1) * std::cos(phi3));
                cplx b(amp * std::cos(phi2) * std::sin(phi3), amp * std::sin(phi2) * std::sin(phi3));
                if (kx == 0 && ky == 0 && kz == 0) { a = 0; b = 0; }  // no mean flow
                double k212 = (k12 < 1e-6) ? 1.0 : ky0 / k12;
                double k112 = (k12 < 1e-6) ? 0.0 : kx0 / k12;
                size_t id = ((size_t)(kz + K[2]) * W[1] + (ky + K[1])) * W[0] + (kx + K[0]);
                spec[0][id] = k212 * a + k112 * kz0 / kmag * b;   // src/block.cpp:488-493
                spec[1][id] = k212 * kz0 / kmag * b - k112 * a;
                spec[2][id] = -k12 / kmag * b;
            }
    // twiddles e^{+i 2 pi k x / n}
    std::vector<cplx> *tw = S.tw;
    for (int d = 0; d < 3; d++) {
        tw[d].resize((size_t)W[d] * n[d]);
        for (int k = -K[d]; k <= K[d]; k++)
            for (int x = 0; x < n[d]; x++) {
                long ph = ((long)k * x) % n[d];
                if (ph < 0) ph += n[d];
                double ang = 2.0 * pie * (double)ph / n[d];
                tw[d][(size_t)(k + K[d]) * n[d] + x] = cplx(std::cos(ang), std::sin(ang));
            }
    }
}

void initial_hit(const nuc_config &c, double k0, double u_rms, unsigned long long seed, std::vector<double> &q)
{
    HitSpectrum S;
    hit_spectrum(c, k0, seed, S);
    const int *n = S.n, *W = S.W;
    const std::vector<cplx> *spec = S.spec, *tw = S.tw;
    const long N = (long)n[0] * n[1] * n[2];
    std::vector<double> vel[3];
    for (int cc = 0; cc < 3; cc++) {
        // x pass: g[kz][ky][x]
        std::vector<cplx> g((size_t)W[2] * W[1] * n[0]);
        for (int kz = 0; kz < W[2]; kz++)
            for (int ky = 0; ky < W[1]; ky++)
                for (int x = 0; x < n[0]; x++) {
                    cplx acc = 0;
                    const cplx *sp = &spec[cc][((size_t)kz * W[1] + ky) * W[0]];
                    for (int kx = 0; kx < W[0]; kx++) acc += sp[kx] * tw[0][(size_t)kx * n[0] + x];
                    g[((size_t)kz * W[1] + ky) * n[0] + x] = acc;
                }
        // y pass: h[kz][y][x]
        std::vector<cplx> h((size_t)W[2] * n[1] * n[0]);
#pragma omp parallel for schedule(static)
        for (int kz = 0; kz < W[2]; kz++)
            for (int y = 0; y < n[1]; y++)
                for (int x = 0; x < n[0]; x++) {
                    cplx acc = 0;
                    for (int ky = 0; ky < W[1]; ky++)
                        acc += g[((size_t)kz * W[1] + ky) * n[0] + x] * tw[1][(size_t)ky * n[1] + y];
                    h[((size_t)kz * n[1] + y) * n[0] + x] = acc;
                }
        // z pass, real part
        vel[cc].assign(N, 0.0);
#pragma omp parallel for schedule(static)
        for (int z = 0; z < n[2]; z++)
            for (int y = 0; y < n[1]; y++)
                for (int x = 0; x < n[0]; x++) {
                    double acc = 0;
                    for (int kz = 0; kz < W[2]; kz++) {
                        const cplx hv = h[((size_t)kz * n[1] + y) * n[0] + x];
                        const cplx t = tw[2][(size_t)kz * n[2] + z];
                        acc += hv.real() * t.real() - hv.imag() * t.imag();
                    }
                    vel[cc][(size_t)x + (size_t)n[0] * (y + (size_t)n[1] * z)] = acc;
                }
    }
    double s2 = 0;
    for (long id = 0; id < N; id++) s2 += vel[0][id] * vel[0][id] + vel[1][id] * vel[1][id] + vel[2][id] * vel[2][id];
    double v = std::sqrt(s2 / N / 3.0);
    double scale = (v > 0) ? u_rms / v : 0.0;
    q.assign(5 * N, 0.0);
    double p0 = 1.0 / (c.gamma * c.mach * c.mach);
    for (long id = 0; id < N; id++) {
        double u0 = vel[0][id] * scale, v0 = vel[1][id] * scale, w0 = vel[2][id] * scale;
        q[id] = 1.0;
        q[N + id] = u0;
        q[2 * N + id] = v0;
        q[3 * N + id] = w0;
        q[4 * N + id] = p0 / (c.gamma - 1.0) + 0.5 * (u0 * u0 + v0 * v0 + w0 * w0);
    }
}

// Shu-Osher shock / entropy-wave profile along x (the reference has no such
// initial condition; its boundaries are periodic only, so the box also holds
// the seam at x = +-lx/2 as a second discontinuity).  x is measured from the
// box centre: (rho, u, p) = (3.857143, 2.629369, 10.33333) for x < -4, else
// (1 + 0.2 sin 5x, 0, 1).  A solenoidal velocity perturbation from initial_hit
// (k0 = 4, seeded) with rms pert * 2.629369 is added to all three components.
// With Mach = gamma^-1/2 the solver's temperature T = gamma M^2 p / rho equals
// p / rho.
void initial_shu_osher(const nuc_config &c, double pert, unsigned long long seed, std::vector<double> &q)
{
    const long N = (long)c.nx * c.ny * c.nz;
    const double uL = 2.629369, rhoL = 3.857143, pL = 10.33333;
    std::vector<double> v;
    initial_hit(c, 4.0, pert * uL, seed, v);
    q.assign(5 * N, 0.0);
    const int nl = c.nx / c.npx;
    const double dx = c.lx / c.nx, g = c.gamma;
    for (int k = 0; k < c.nz; k++)
        for (int j = 0; j < c.ny; j++)
            for (int i = 0; i < c.nx; i++) {
                // cell centres per rank as initial_taylorgreen, shifted to the box centre
                const double x = c.lx / c.npx * (i / nl) + dx * (0.5 + i % nl) - 0.5 * c.lx;
                const long id = i + (long)c.nx * (j + (long)c.ny * k);
                double rho, u, p;
                if (x < -4.0) { rho = rhoL; u = uL; p = pL; }
                else { rho = 1.0 + 0.2 * std::sin(5.0 * x); u = 0.0; p = 1.0; }
                const double uu = u + v[N + id], vv = v[2 * N + id], ww = v[3 * N + id];
                q[id] = rho;
                q[N + id] = rho * uu;
                q[2 * N + id] = rho * vv;
                q[3 * N + id] = rho * ww;
                q[4 * N + id] = p / (g - 1.0) + 0.5 * rho * (uu * uu + vv * vv + ww * ww);
            }
}

}  // namespace nuc3d

// ---- C ABI host utilities -------------------------------------------------------
extern "C" int nuc_init_tgv(const nuc_config *cfg, double *q)
{
    if (!cfg || !q) return NUC_EINVAL;
    std::vector<double> v;
    nuc3d::initial_taylorgreen(*cfg, cfg->origin, v);
    std::memcpy(q, v.data(), v.size() * sizeof(double));
    return NUC_OK;
}

extern "C" int nuc_init_default(const nuc_config *cfg, double *q)
{
    if (!cfg || !q) return NUC_EINVAL;
    std::vector<double> v;
    nuc3d::initial_default(*cfg, v);
    std::memcpy(q, v.data(), v.size() * sizeof(double));
    return NUC_OK;
}

extern "C" int nuc_init_ivc(const nuc_config *cfg, double *q)
{
    if (!cfg || !q) return NUC_EINVAL;
    if (cfg->npx < 1 || cfg->npy < 1 || cfg->nx % cfg->npx != 0 || cfg->ny % cfg->npy != 0) return NUC_EINVAL;
    std::vector<double> v;
    nuc3d::initial_ivc(*cfg, cfg->origin, v);
    std::memcpy(q, v.data(), v.size() * sizeof(double));
    return NUC_OK;
}

extern "C" int nuc_init_hit(const nuc_config *cfg, double k0, double u_rms, unsigned long long seed, double *q)
{
    if (!cfg || !q || !(k0 > 0) || !(u_rms >= 0)) return NUC_EINVAL;
    std::vector<double> v;
    nuc3d::initial_hit(*cfg, k0, u_rms, seed, v);
    std::memcpy(q, v.data(), v.size() * sizeof(double));
    return NUC_OK;
}

extern "C" int nuc_init_shu_osher(const nuc_config *cfg, double pert, unsigned long long seed, double *q)
{
    if (!cfg || !q || !(pert >= 0)) return NUC_EINVAL;
    if (cfg->nx < 1 || cfg->ny < 1 || cfg->nz < 1 || cfg->npx < 1 || cfg->nx % cfg->npx != 0) return NUC_EINVAL;
    std::vector<double> v;
    nuc3d::initial_shu_osher(*cfg, pert, seed, v);
    std::memcpy(q, v.data(), v.size() * sizeof(double));
    return NUC_OK;
}
