// test_host_classes.cpp -- drives the reference-shaped host classes of
// nuc3d_host.hpp (fieldOperator3d and its scheme classes, tvdrk3rd,
// PDEData3d, NaiverStokesData3d / EulerData3D) the way the reference's call
// sites do, and checks them against the oracle (oracle/nuc_oracle.h, the CPU
// restatement; test infrastructure).  Needs a GPU.  Exit status 0 = pass.
//
// usage: test_host_classes <inp dir with Method.in>
#include "nuc3d_host.hpp"
#include "nuc_oracle.h"

#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

using namespace nuc3d;

static int g_fail = 0;

static void expect(bool ok, const char *what, double v, double bar)
{
    printf("%-60s %10.3e  (bar %.1e)  %s\n", what, v, bar, ok ? "ok" : "FAIL");
    if (!ok) g_fail++;
}

static double rel_err(const double *a, const double *b, size_t n)
{
    double e = 0.0, m = 0.0;
    for (size_t i = 0; i < n; i++) {
        e = std::max(e, std::fabs(a[i] - b[i]));
        m = std::max(m, std::fabs(b[i]));
    }
    return e / std::max(m, 1e-300);
}

static void fill(Field &f, std::mt19937_64 &rng, double scale, double off)
{
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    const int n1 = f.getSizeX(), n2 = f.getSizeY(), n3 = f.getSizeZ();
    for (int k = 0; k < n3; k++)
        for (int j = 0; j < n2; j++)
            for (int i = 0; i < n1; i++)
                f.setValue(i, j, k, off + scale * (std::sin(0.37 * i + 0.11 * j - 0.23 * k) + 0.05 * U(rng)));
}

// operator level: every inviscid scheme through the fieldOperator3d dispatch
// (reconstructionInner + reconstructionBoundary, src/fieldOperator.cpp:387-465)
static void operators()
{
    std::mt19937_64 rng(20261203);
    const int n1 = 37, n2 = 3, n3 = 2;
    const int names[9] = {NUC_WENO5JS, NUC_WENO5Z, NUC_WENO7JS, NUC_CRWENO5, NUC_HCCS,
                          NUC_WENO3JS, NUC_UPWIND1ST, NUC_HCCS78, NUC_HCCS77};
    for (int s : names) {
        nuc_config c{};
        c.scheme[0] = c.scheme[1] = c.scheme[2] = s;
        c.vis_scheme[0] = c.vis_scheme[1] = c.vis_scheme[2] = NUC_CD8;
        fieldOperator3d op(c);
        const int bf = op.getBufferSize();
        Field in(n1, n2, n3), gL(bf, n2, n3), gR(bf, n2, n3);
        fill(in, rng, 1.0, 2.0);
        fill(gL, rng, 1.0, 2.0);
        fill(gR, rng, 1.0, 2.0);
        for (int up : {1, -1}) {
            Field out(n1 + 1, n2, n3, 0.0);
            op.reconstructionInner(in, 0, up, out);
            op.reconstructionBoundary(in, gL, gR, 0, up, out);
            std::vector<double> ref((size_t)(n1 + 1) * n2 * n3);
            orc_field_reconstruct(s, up, op.bufferSize_xi, in.getDataPtr(), n1, n2, n3, gL.getDataPtr(),
                                  gR.getDataPtr(), bf, ref.data());
            char what[96];
            snprintf(what, sizeof what, "reconstruction scheme %d upwind %+d (bufferSize %d)", s, up, op.bufferSize_xi);
            // the 7-point schemes' indicators cancel terms up to 17246 f^2
            const double bar = (s == NUC_WENO7JS || s == NUC_HCCS || s == NUC_HCCS77 || s == NUC_HCCS78) ? 1e-11 : 1e-12;
            expect(rel_err(out.getDataPtr(), ref.data(), ref.size()) <= bar, what,
                   rel_err(out.getDataPtr(), ref.data(), ref.size()), bar);
        }
    }
    // viscous differences (differenceInner + differenceBoundary)
    for (int v : {NUC_CD2, NUC_CD6, NUC_CD8}) {
        nuc_config c{};
        c.scheme[0] = c.scheme[1] = c.scheme[2] = NUC_WENO5JS;
        c.vis_scheme[0] = c.vis_scheme[1] = c.vis_scheme[2] = v;
        fieldOperator3d op(c);
        const int bf = op.getBufferSize();
        Field in(n1, n2, n3), gL(bf, n2, n3), gR(bf, n2, n3), out(n1, n2, n3, 0.0);
        fill(in, rng, 1.0, 0.0);
        fill(gL, rng, 1.0, 0.0);
        fill(gR, rng, 1.0, 0.0);
        const double h = 0.173;
        op.differenceInner(in, h, 0, out);
        op.differenceBoundary(in, h, gL, gR, 0, out);
        std::vector<double> ref((size_t)n1 * n2 * n3);
        const int tsv = v == NUC_CD8 ? 4 : 3;   // cd2 sits behind weno5js's 3-cell halo
        orc_field_diff(v, tsv, in.getDataPtr(), h, n1, n2, n3, gL.getDataPtr(), gR.getDataPtr(), bf, ref.data());
        char what[96];
        snprintf(what, sizeof what, "difference scheme %d", v);
        expect(rel_err(out.getDataPtr(), ref.data(), ref.size()) <= 1e-13, what,
               rel_err(out.getDataPtr(), ref.data(), ref.size()), 1e-13);
    }
    // TVD-RK3 through fieldOperator3d::timeIntegral (src/TVD-RK3.cpp:16-157)
    {
        nuc_config c{};
        fieldOperator3d op(c);
        VectorField rhs, un, ui;
        for (int v = 0; v < 5; v++) {
            rhs.emplace_back(9, 4, 3);
            un.emplace_back(9, 4, 3);
            ui.emplace_back(9, 4, 3);
            fill(rhs[v], rng, 1.0, 0.0);
            fill(un[v], rng, 1.0, 1.0);
            fill(ui[v], rng, 1.0, 1.0);
        }
        double worst = 0.0;
        for (int stage = 0; stage < op.getSteps(); stage++) {
            VectorField ref = ui;
            op.timeIntegral(rhs, un, ui, 0.013, stage);
            for (int v = 0; v < 5; v++) {
                orc_rk3(stage, rhs[v].getDataPtr(), un[v].getDataPtr(), ref[v].getDataPtr(), 0.013, ref[v].getSize());
                worst = std::max(worst, rel_err(ui[v].getDataPtr(), ref[v].getDataPtr(), ref[v].getSize()));
            }
        }
        expect(worst <= 1e-15, "tvdrk3rd stages 0, 1, 2", worst, 1e-15);
    }
}

// fieldOperator3d from an inp directory: the scheme names and buffer sizes
static void from_inp(const char *dir)
{
    fieldOperator3d op{std::string(dir)};
    int sch[3], vis[3];
    std::string err;
    read_method(dir, sch, vis, err);
    bool ok = true;
    for (int d = 0; d < 3; d++) ok = ok && op.schemeId(d) == sch[d] && op.visSchemeId(d) == vis[d];
    // hccs78 / weno3js / upwind1st (tests/golden/inp_operators): 4, 2, 1
    printf("fieldOperator3d(%s): bufferSize %d %d %d, getBufferSize %d\n", dir, op.bufferSize_xi,
           op.bufferSize_eta, op.bufferSize_zeta, op.getBufferSize());
    expect(ok, "fieldOperator3d schemes equal Method.in's", 0.0, 0.0);
}

// PDE level: block::solve's stage loop (src/block.cpp:191-215) through
// NaiverStokesData3d::solve + PDEData3d::solve, against the oracle's steps
static void pde_level()
{
    nuc_config c{};
    c.nx = 24; c.ny = 20; c.nz = 16;
    c.npx = c.npy = c.npz = 1;
    c.lx = c.ly = c.lz = 2.0 * M_PI;
    for (int d = 0; d < 3; d++) { c.scheme[d] = NUC_CRWENO5; c.vis_scheme[d] = NUC_CD8; }
    c.riemann = NUC_LF; c.model = NUC_NAVIERSTOKES3D; c.viscosity = NUC_SUTHERLAND; c.precision = NUC_FP64;
    c.gamma = 1.4; c.mach = 0.2; c.reynolds = 800.0; c.prandtl = 0.72; c.t_ref = 110.4; c.t_inf = 288.0;
    c.cfl = 0.5;
    const size_t N = (size_t)c.nx * c.ny * c.nz;
    std::vector<double> q;
    const double origin[3] = {0, 0, 0};
    initial_taylorgreen(c, origin, q);
    for (size_t i = 0; i < N; i++) {
        q[3 * N + i] += 0.1 * std::sin(0.013 * (double)i);
        q[i] *= 1.0 + 0.01 * std::cos(0.007 * (double)i);
    }
    orc_config oc{};
    oc.nx = c.nx; oc.ny = c.ny; oc.nz = c.nz; oc.npx = oc.npy = oc.npz = 1;
    oc.lx = c.lx; oc.ly = c.ly; oc.lz = c.lz;
    for (int d = 0; d < 3; d++) { oc.scheme[d] = ORC_CRWENO5; oc.vis_scheme[d] = ORC_CD8; }
    oc.riemann = ORC_LF; oc.model = ORC_NS; oc.viscosity = ORC_SUTHERLAND; oc.forcing = ORC_FORCE_NONE;
    oc.gamma = c.gamma; oc.mach = c.mach; oc.reynolds = c.reynolds; oc.prandtl = c.prandtl;
    oc.t_ref = c.t_ref; oc.t_inf = c.t_inf; oc.cfl = c.cfl; oc.nthreads = 1;
    orc_solver *s = orc_create(&oc);
    orc_set_state(s, q.data());

    PDEData3d pde(c);
    VectorField Q;
    for (int v = 0; v < 5; v++) {
        Q.emplace_back(c.nx, c.ny, c.nz);
        std::memcpy(Q[v].getDataPtr(), q.data() + v * N, N * sizeof(double));
    }
    pde.setQ(Q);
    fieldOperator3d op(c);
    physicsModel phys(c);
    MPIComunicator3d_nonblocking mpi(c);
    VectorBuffer buf;
    fft myFFT;
    VectorField xyz;
    NaiverStokesData3d fluid;

    // one RHS evaluation (getRHS) against orc_rhs
    {
        std::vector<double> ref(5 * N);
        double dt_ref = 0.0;
        orc_rhs(s, ref.data(), &dt_ref);
        VectorField &R = pde.getRHS();
        double worst = 0.0;
        for (int v = 0; v < 5; v++) worst = std::max(worst, rel_err(R[v].getDataPtr(), ref.data() + v * N, N));
        expect(worst <= 1e-11, "PDEData3d::getRHS vs oracle RHS", worst, 1e-11);
        expect(std::fabs(pde.dt_local - dt_ref) <= 1e-14 * dt_ref, "dt_local", std::fabs(pde.dt_local - dt_ref) / dt_ref, 1e-14);
    }
    // the stage loop for 5 steps
    const int nsteps = 5;
    for (int it = 0; it < nsteps; it++)
        for (int step = 0; step < op.getSteps(); step++) {
            fluid.solve(pde, op, buf, phys, mpi, myFFT, xyz);
            pde.solve(op, mpi, c.cfl, step);
        }
    double dt_ref, res_ref, t_ref;
    orc_step(s, nsteps, &dt_ref, &res_ref, &t_ref);
    std::vector<double> qref(5 * N);
    orc_get_state(s, qref.data());
    VectorField &Qg = pde.getQ();
    double worst = 0.0;
    for (int v = 0; v < 5; v++) worst = std::max(worst, rel_err(Qg[v].getDataPtr(), qref.data() + v * N, N));
    expect(worst <= 1e-12, "5 steps of fluid.solve + PDEData3d::solve vs oracle", worst, 1e-12);
    expect(std::fabs(pde.getDt() - dt_ref) <= 1e-13 * dt_ref, "dt_global", std::fabs(pde.getDt() - dt_ref) / dt_ref, 1e-13);
    expect(std::fabs(pde.getRes() - res_ref) <= 1e-9 * std::fabs(res_ref), "residual",
           std::fabs(pde.getRes() - res_ref) / std::fabs(res_ref), 1e-9);

    // misuse is refused, not silently run: a stage without the fluid's solve,
    // a model or cfl that differ from the context's
    bool threw = false;
    try { pde.solve(op, mpi, c.cfl, 0); } catch (const nuc_error &) { threw = true; }
    expect(threw, "PDEData3d::solve without the fluid's solve throws", 0.0, 0.0);
    threw = false;
    EulerData3D euler;
    try { euler.solve(pde, op, buf, phys, mpi, myFFT, xyz); } catch (const nuc_error &) { threw = true; }
    expect(threw, "EulerData3D::solve on a Navier-Stokes context throws", 0.0, 0.0);
    threw = false;
    fluid.solve(pde, op, buf, phys, mpi, myFFT, xyz);
    try { pde.solve(op, mpi, 0.3, 0); } catch (const nuc_error &) { threw = true; }
    expect(threw, "PDEData3d::solve with another cfl throws", 0.0, 0.0);
    orc_destroy(s);
}

int main(int argc, char **argv)
{
    try {
        if (argc > 1) from_inp(argv[1]);
        operators();
        pde_level();
    } catch (const std::exception &e) {
        printf("exception: %s\n", e.what());
        return 2;
    }
    printf("%s: %d failure(s)\n", g_fail ? "FAIL" : "PASS", g_fail);
    return g_fail ? 1 : 0;
}
