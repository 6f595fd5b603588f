// nuc3d_host.cpp -- host classes of nuc3d_host.hpp and host utilities
// (initial conditions, restart files) exported through the C ABI.
#include "nuc3d_host.hpp"

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

fieldOperator3d::fieldOperator3d(const nuc_config &c)
{
    int *bs[3] = {&bufferSize_xi, &bufferSize_eta, &bufferSize_zeta};
    for (int d = 0; d < 3; d++) {
        *bs[d] = (c.scheme[d] == NUC_WENO7JS || c.scheme[d] == NUC_HCCS) ? 4 : 3;
        interp_.push_back(std::make_shared<gpuInterpolation>(c.scheme[d]));
        diff_.push_back(std::make_shared<gpuDifferential>(c.vis_scheme[d]));
        vis_[d] = c.vis_scheme[d];
    }
    integ_ = std::make_shared<tvdrk3rd>();
}

int fieldOperator3d::getBufferSize() const
{
    int b = std::max(bufferSize_xi, std::max(bufferSize_eta, bufferSize_zeta));
    for (int d = 0; d < 3; d++)
        if (vis_[d] == NUC_CD8) b = std::max(b, 4);
    return b;
}

int fieldOperator3d::tsv(int dir) const
{
    // cd8 runs with four ghost cells (the reference passes the inviscid halo, src/fieldOperator.cpp:520-526)
    return (vis_[dir] == NUC_CD8 && ts(dir) < 4) ? 4 : ts(dir);
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

// ---- PDE facade ---------------------------------------------------------------
PDEData3d::PDEData3d(const nuc_config &c) : cfg_(c) { check(nuc_create(&c, &ctx_), "nuc_create"); }
PDEData3d::~PDEData3d() { nuc_destroy(ctx_); }

void PDEData3d::setQ(const VectorField &Q)
{
    size_t N = (size_t)cfg_.nx * cfg_.ny * cfg_.nz;
    std::vector<double> q(5 * N);
    for (int v = 0; v < 5; v++) std::memcpy(q.data() + v * N, Q[v].getDataPtr(), N * sizeof(double));
    check(nuc_upload_state(ctx_, q.data()), "nuc_upload_state");
}

void PDEData3d::getQ(VectorField &Q)
{
    size_t N = (size_t)cfg_.nx * cfg_.ny * cfg_.nz;
    std::vector<double> q(5 * N);
    check(nuc_download_state(ctx_, q.data()), "nuc_download_state");
    Q.clear();
    for (int v = 0; v < 5; v++) {
        Q.emplace_back(cfg_.nx, cfg_.ny, cfg_.nz);
        std::memcpy(Q[v].getDataPtr(), q.data() + v * N, N * sizeof(double));
    }
}

double PDEData3d::solveRHS(VectorField *rhs)
{
    size_t N = (size_t)cfg_.nx * cfg_.ny * cfg_.nz;
    std::vector<double> r(rhs ? 5 * N : 0);
    double dt = 0.0;
    check(nuc_rhs(ctx_, rhs ? r.data() : nullptr, &dt), "nuc_rhs");
    if (rhs) {
        rhs->clear();
        for (int v = 0; v < 5; v++) {
            rhs->emplace_back(cfg_.nx, cfg_.ny, cfg_.nz);
            std::memcpy((*rhs)[v].getDataPtr(), r.data() + v * N, N * sizeof(double));
        }
    }
    return dt;
}

nuc_step_info PDEData3d::step(int n)
{
    nuc_step_info info{};
    check(nuc_step(ctx_, n, &info), "nuc_step");
    return info;
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

typedef std::complex<double> cplx;

// in-place unnormalised DFT of a strided line, sign +1 (FFTW_BACKWARD);
// mixed radix 2/3/5 recursion, direct sums for other factors
static void dft_rec(cplx *x, int n, int sign, std::vector<cplx> &tmp)
{
    if (n == 1) return;
    int p = (n % 2 == 0) ? 2 : (n % 3 == 0) ? 3 : (n % 5 == 0) ? 5 : n;
    if (p == n) {
        std::vector<cplx> y(n);
        for (int k = 0; k < n; k++) {
            cplx s = 0;
            for (int t = 0; t < n; t++) {
                double a = sign * 2.0 * M_PI * (double)((long)k * t % n) / n;
                s += x[t] * cplx(std::cos(a), std::sin(a));
            }
            y[k] = s;
        }
        for (int k = 0; k < n; k++) x[k] = y[k];
        return;
    }
    int m = n / p;
    std::vector<cplx> sub((size_t)n);
    for (int r = 0; r < p; r++)
        for (int t = 0; t < m; t++) sub[(size_t)r * m + t] = x[t * p + r];
    for (int r = 0; r < p; r++) dft_rec(sub.data() + (size_t)r * m, m, sign, tmp);
    for (int k = 0; k < n; k++) {
        cplx s = 0;
        for (int r = 0; r < p; r++) {
            double a = sign * 2.0 * M_PI * (double)((long)r * k % n) / n;
            s += sub[(size_t)r * m + (k % m)] * cplx(std::cos(a), std::sin(a));
        }
        x[k] = s;
    }
}

static void ifft3(std::vector<cplx> &a, int nx, int ny, int nz)
{
    std::vector<cplx> line, tmp;
    line.resize(std::max(nx, std::max(ny, nz)));
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++) {
            cplx *p = a.data() + (size_t)nx * (j + (size_t)ny * k);
            dft_rec(p, nx, +1, tmp);
        }
    for (int k = 0; k < nz; k++)
        for (int i = 0; i < nx; i++) {
            for (int j = 0; j < ny; j++) line[j] = a[i + (size_t)nx * (j + (size_t)ny * k)];
            dft_rec(line.data(), ny, +1, tmp);
            for (int j = 0; j < ny; j++) a[i + (size_t)nx * (j + (size_t)ny * k)] = line[j];
        }
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            for (int k = 0; k < nz; k++) line[k] = a[i + (size_t)nx * (j + (size_t)ny * k)];
            dft_rec(line.data(), nz, +1, tmp);
            for (int k = 0; k < nz; k++) a[i + (size_t)nx * (j + (size_t)ny * k)] = line[k];
        }
}

void initial_hit(const nuc_config &c, double k0, double u_rms, unsigned long long seed, std::vector<double> &q)
{
    const int nx = c.nx, ny = c.ny, nz = c.nz;
    const long N = (long)nx * ny * nz;
    const double pie = 4.0 * std::atan(1.0);
    std::mt19937_64 gen(seed);
    std::uniform_real_distribution<double> dis(0.0, 2.0 * pie);
    std::vector<cplx> U(N), V(N), W(N);
    // spectrum and solenoidal random-phase amplitudes, src/block.cpp:448-499
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                long id = i + (long)nx * (j + (long)ny * k);
                double kx0 = (i >= nx / 2.0) ? i - nx : i;
                double ky0 = (j >= ny / 2.0) ? j - ny : j;
                double kz0 = (k >= nz / 2.0) ? k - nz : k;
                double kmag = std::sqrt(kx0 * kx0 + ky0 * ky0 + kz0 * kz0);
                double k12 = std::sqrt(kx0 * kx0 + ky0 * ky0);
                if (kmag < 1e-6) kmag = 1.0;
                double Ek0 = 16.0 * std::sqrt(2.0 / pie) * std::pow(kmag / k0, 4) / k0 * std::exp(-2.0 * kmag * kmag / k0 / k0);
                double phi1 = dis(gen), phi2 = dis(gen), phi3 = dis(gen);
                double amp = std::sqrt(2.0 * Ek0 / kmag / kmag / 4.0 / pie);
                cplx a(amp * std::cos(phi1) * std::cos(phi3), amp * std::sin(phi1) * std::cos(phi3));
                cplx b(amp * std::cos(phi2) * std::sin(phi3), amp * std::sin(phi2) * std::sin(phi3));
                double k212 = (k12 < 1e-6) ? 1.0 : ky0 / k12;
                double k112 = (k12 < 1e-6) ? 0.0 : kx0 / k12;
                if (i == 0 && j == 0 && k == 0) { a = 0; b = 0; }  // no mean flow
                U[id] = k212 * a + k112 * kz0 / kmag * b;
                V[id] = k212 * kz0 / kmag * b - k112 * a;
                W[id] = -k12 / kmag * b;
            }
    ifft3(U, nx, ny, nz);
    ifft3(V, nx, ny, nz);
    ifft3(W, nx, ny, nz);
    double s2 = 0;
    for (long id = 0; id < N; id++) s2 += U[id].real() * U[id].real() + V[id].real() * V[id].real() + W[id].real() * W[id].real();
    double vel = std::sqrt(s2 / N / 3.0);
    double scale = (vel > 0) ? u_rms / vel : 0.0;
    q.assign(5 * N, 0.0);
    double p0 = 1.0 / (c.gamma * c.mach * c.mach);
    for (long id = 0; id < N; id++) {
        double u0 = U[id].real() * scale, v0 = V[id].real() * scale, w0 = W[id].real() * scale;
        q[id] = 1.0;
        q[N + id] = u0;
        q[2 * N + id] = v0;
        q[3 * N + id] = w0;
        q[4 * N + id] = p0 / (c.gamma - 1.0) + 0.5 * (u0 * u0 + v0 * v0 + w0 * w0);
    }
}

bool write_restart(const std::string &path, double time, const std::vector<double> &q)
{
    std::ofstream f(path, std::ios::binary);
    if (!f) return false;
    f.write(reinterpret_cast<const char *>(&time), sizeof(double));
    f.write(reinterpret_cast<const char *>(q.data()), q.size() * sizeof(double));
    return (bool)f;
}

bool read_restart(const std::string &path, double &time, std::vector<double> &q)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    f.read(reinterpret_cast<char *>(&time), sizeof(double));
    f.read(reinterpret_cast<char *>(q.data()), q.size() * sizeof(double));
    return (bool)f;
}

}  // namespace nuc3d

// ---- C ABI host utilities -------------------------------------------------------
extern "C" int nuc_init_tgv(const nuc_config *cfg, double *q)
{
    if (!cfg || !q) return NUC_EINVAL;
    std::vector<double> v;
    double o[3] = {0, 0, 0};
    nuc3d::initial_taylorgreen(*cfg, o, v);
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
