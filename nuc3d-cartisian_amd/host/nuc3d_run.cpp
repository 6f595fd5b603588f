// nuc3d_run -- the reference's executable flow (src/NUC3d.cpp:3-14,
// src/singleBlock.cpp:31-71) on one GPU: read inp/*.in, set the initial
// condition (Benchmark 0 = at rest, 1 = isentropic vortex, 2 = Taylor-Green, 3 = seeded HIT) or read the restart
// file when startStep != 0, then step while ifsolve() holds
// (src/IOcontroller.cpp:182-192), printing the status line of
// block::printStatus (src/block.cpp:217-220) and writing the restart and
// Tecplot files IOController.in selects every saveStep and at the end
// (src/singleBlock.cpp:50-68, src/block.cpp:238-266).
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
        if (nuc_read_restart(&cfg, "flowData", rc.start_step, &time, q.data()) != NUC_OK) {
            fprintf(stderr, "cannot read the restart files flowData/NUC3d_step_%d_ID_*.bin\n", rc.start_step);
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
    // block::Output (src/block.cpp:238-266): the files IOController.in asks for
    auto output = [&](nuc3d::PDEData3d &pde) {
        nuc3d::check(nuc_download_state(pde.ctx(), q.data()), "nuc_download_state");
        if (rc.binary && nuc_write_restart(&cfg, "flowData", step, time, q.data()) != NUC_OK)
            fprintf(stderr, "cannot write flowData/NUC3d_step_%d_ID_*.bin\n", step);
        if (rc.tecplot && nuc_write_tecplot(&cfg, "flowData", step, rc.origin, q.data()) != NUC_OK)
            fprintf(stderr, "cannot write flowData/NUC3d_step_%d_id_*.dat\n", step);
    };
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
            if (rc.save_step > 0 && step % rc.save_step == 0) output(pde);
        }
        output(pde);   // and once more after the loop (src/singleBlock.cpp:67)
    } catch (const std::exception &e) {
        fprintf(stderr, "%s\n", e.what());
        return 2;
    }
    return 0;
}
