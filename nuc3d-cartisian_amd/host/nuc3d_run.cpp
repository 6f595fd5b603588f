// nuc3d_run -- the reference's executable flow (src/NUC3d.cpp:3-14,
// src/singleBlock.cpp:31-71) on one GPU: read inp/*.in, set the initial
// condition (Benchmark 0 = at rest, 1 = isentropic vortex, 2 = Taylor-Green, 3 = seeded HIT) or read the restart
// file when startStep != 0, then step while ifsolve() holds
// (src/IOcontroller.cpp:182-192), printing the status line of
// block::printStatus (src/block.cpp:217-220) and writing restart files every
// saveStep (src/block.cpp:609-638).
//
// usage: nuc3d_run [inp_dir] [--seed S]
#include "nuc3d_host.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

int main(int argc, char **argv)
{
    std::string dir = "inp";
    unsigned long long seed = 20261202ull;
    for (int a = 1; a < argc; a++) {
        std::string s = argv[a];
        if (s == "--seed" && a + 1 < argc) seed = std::strtoull(argv[++a], nullptr, 10);
        else dir = s;
    }
    nuc_config cfg;
    nuc3d::RunControl rc;
    std::string err;
    if (nuc3d::read_inp(dir, cfg, rc, err) != NUC_OK) {
        fprintf(stderr, "%s\n", err.c_str());
        return 1;
    }
    const long N = (long)cfg.nx * cfg.ny * cfg.nz;
    std::vector<double> q(5 * N);
    double time = 0.0;
    int step = rc.start_step;
    if (rc.start_step != 0) {
        char path[512];
        snprintf(path, sizeof(path), "flowData/NUC3d_step_%d_ID_0.bin", rc.start_step);
        if (!nuc3d::read_restart(path, time, q)) {
            fprintf(stderr, "cannot read restart file %s\n", path);
            return 1;
        }
    } else if (rc.benchmark == 0) {
        nuc3d::initial_default(cfg, q);
    } else if (rc.benchmark == 1) {
        nuc3d::initial_ivc(cfg, rc.origin, q);
    } else if (rc.benchmark == 2) {
        nuc3d::initial_taylorgreen(cfg, rc.origin, q);
    } else if (rc.benchmark == 3) {
        // u_rms from the turbulent Mach number with c = 1/Mach
        double u_rms = rc.mach_t / cfg.mach / std::sqrt(3.0);
        nuc3d::initial_hit(cfg, rc.k_peak, u_rms, seed, q);
    } else {
        fprintf(stderr, "Benchmark %d is not available (0 = at rest, 1 = isentropic vortex, 2 = Taylor-Green, 3 = HIT, or a restart)\n", rc.benchmark);
        return 1;
    }
    try {
        nuc3d::PDEData3d pde(cfg);
        nuc3d::check(nuc_upload_state(pde.ctx(), q.data()), "nuc_upload_state");
        while (time <= rc.tl && step < rc.end_step) {
            auto t0 = std::chrono::steady_clock::now();
            nuc_step_info info = pde.step(1);
            double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            step++;
            time += info.dt;
            printf("========step = %d\n time = %g, dt = %g, residual = %g, GPU time = %g(s)\n", step, time, info.dt,
                   info.res, wall);
            if (rc.save_step > 0 && step % rc.save_step == 0) {
                nuc3d::check(nuc_download_state(pde.ctx(), q.data()), "nuc_download_state");
                char path[512];
                snprintf(path, sizeof(path), "flowData/NUC3d_step_%d_ID_0.bin", step);
                if (!nuc3d::write_restart(path, time, q)) fprintf(stderr, "cannot write %s\n", path);
            }
        }
    } catch (const std::exception &e) {
        fprintf(stderr, "%s\n", e.what());
        return 2;
    }
    return 0;
}
