// nuc3d_run -- the reference's executable flow (src/NUC3d.cpp:3-14,
// src/singleBlock.cpp:31-71) on one GPU: read inp/*.in, set the initial
// condition (Benchmark 0 = at rest, 1 = isentropic vortex, 2 = Taylor-Green, 3 = seeded HIT) or read the restart
// file when startStep != 0, then step while ifsolve() holds
// (src/IOcontroller.cpp:182-192), printing the status line of
// block::printStatus (src/block.cpp:217-220), post-processing every postStep
// as PostProc.in selects (running averages; enstrophy.dat, turb_para.dat and
// the shell spectra files),
// and writing the restart, Tecplot and averaged-statistics files
// IOController.in selects every saveStep and at the end (src/singleBlock.cpp:50-68,
// src/block.cpp:238-266), with the averaged Tecplot files when AveragedSaveType
// is 1 and of solveQ's per-point fields when SimutaneousSaveType is set.
//
// usage: nuc3d_run [inp_dir] [--seed S] [--precision fp64|mixed|fp32]
// (the inp files select no precision: the reference computes in FP64, the default)
#include "nuc3d_host.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

int main(int argc, char **argv)
{
    std::string dir = "inp", precision = "fp64";
    unsigned long long seed = 20261202ull;
    for (int a = 1; a < argc; a++) {
        std::string s = argv[a];
        if (s == "--seed" && a + 1 < argc) seed = std::strtoull(argv[++a], nullptr, 10);
        else if (s == "--precision" && a + 1 < argc) precision = argv[++a];
        else dir = s;
    }
    nuc_config cfg;
    nuc3d::RunControl rc;
    std::string err;
    if (nuc3d::read_inp(dir, cfg, rc, err) != NUC_OK) {
        fprintf(stderr, "%s\n", err.c_str());
        return 1;
    }
    if (precision == "fp64") cfg.precision = NUC_FP64;
    else if (precision == "mixed") cfg.precision = NUC_MIXED;
    else if (precision == "fp32") cfg.precision = NUC_FP32;
    else {
        fprintf(stderr, "--precision takes fp64, mixed or fp32, not %s\n", precision.c_str());
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
    // block::Output (src/block.cpp:238-266): the files IOController.in asks for,
    // and postproc::OutputPost's averaged-statistics files (src/postproc.cpp:310-343)
    auto output = [&](nuc3d::PDEData3d &pde) {
        nuc3d::check(nuc_download_state(pde.ctx(), q.data()), "nuc_download_state");
        if (rc.binary && nuc_write_restart(&cfg, "flowData", step, time, q.data()) != NUC_OK)
            fprintf(stderr, "cannot write flowData/NUC3d_step_%d_ID_*.bin\n", step);
        if (rc.tecplot && nuc_write_tecplot(&cfg, "flowData", step, rc.origin, q.data()) != NUC_OK)
            fprintf(stderr, "cannot write flowData/NUC3d_step_%d_id_*.dat\n", step);
        if (rc.postproc && rc.average_type != 0) {
            if (nuc_average_write(pde.ctx(), &cfg, rc.average_type, "AveragedFlowData") != NUC_OK)
                fprintf(stderr, "cannot write AveragedFlowData/Averaged_step_*_id_*.bin\n");
            // AveragedSaveType 1: binary and Tecplot (src/postproc.cpp:322-334)
            if (rc.averaged_save_type == 1 &&
                nuc_average_write_tecplot(pde.ctx(), &cfg, rc.average_type, "AveragedFlowData", time, rc.origin) != NUC_OK)
                fprintf(stderr, "cannot write AveragedFlowData/Averaged_tec_step_*_id_*.dat\n");
        }
        // SimutaneousSaveType: the Q fields (src/postproc.cpp:336-339), here of the
        // current state (the reference keeps those of its last post step, the
        // same state when postStep divides saveStep)
        if (rc.postproc && rc.simultaneous_save != 0 &&
            nuc_write_q_tecplot(pde.ctx(), &cfg, "flowDataTemp", step, time, rc.origin) != NUC_OK)
            fprintf(stderr, "cannot write flowDataTemp/Q_tec_step_%d_id_*.dat\n", step);
    };
    // postproc::solvePost (src/postproc.cpp:281-308): the running averages, and
    // solveQ's global statistics appended to enstrophy.dat and turb_para.dat
    // (src/postproc.cpp:880-920)
    auto post = [&](nuc3d::PDEData3d &pde) {
        if (rc.average_type != 0) nuc3d::check(nuc_average_update(pde.ctx()), "nuc_average_update");
        if (rc.simultaneous != 0) {
            double st[25];
            nuc3d::check(nuc_stats(pde.ctx(), st), "nuc_stats");
            FILE *e = fopen("enstrophy.dat", "a"), *t = fopen("turb_para.dat", "a");
            if (e) {
                fprintf(e, "%d %.12g", step, time);
                for (int i = 0; i < 16; i++) fprintf(e, " %.12g", st[i]);
                fprintf(e, "\n");
                fclose(e);
            }
            if (t) {
                fprintf(t, "%d %.12g", step, time);
                for (int i = 16; i < 25; i++) fprintf(t, " %.12g", st[i]);
                fprintf(t, "\n");
                fclose(t);
            }
            // postproc::solveFFT's shell spectra (src/postproc.cpp:931-1184),
            // written as fft::solveShell does (src/fft.cpp:449-470):
            // flowDataTemp/power_shell_<step>_<name>.dat, "k E E/sum(E)" per shell
            const int kmax = std::min(cfg.nx, std::min(cfg.ny, cfg.nz)) / 2;
            std::vector<double> kin(5 * (size_t)kmax), sc((size_t)kmax);
            nuc3d::check(nuc_spectra_kinetic(pde.ctx(), kin.data(), kmax), "nuc_spectra_kinetic");
            auto shell_file = [&](const char *name, const double *e) {
                char path[256];
                snprintf(path, sizeof(path), "flowDataTemp/power_shell_%d_%s.dat", step, name);
                FILE *f = fopen(path, "w");
                if (!f) { fprintf(stderr, "cannot write %s\n", path); return; }
                double sum = 0.0;
                for (int k = 0; k < kmax; k++) sum += e[k];
                for (int k = 0; k < kmax; k++) fprintf(f, "%d %.12g %.12g\n", k, e[k], e[k] / sum);
                fprintf(f, "\n");
                fclose(f);
            };
            const char *kn[5] = {"kin", "kinc", "kins", "kinl", "kinr"};
            for (int s5 = 0; s5 < 5; s5++) shell_file(kn[s5], kin.data() + (size_t)s5 * kmax);
            const int sf[4] = {NUC_SPEC_HELICITY, NUC_SPEC_ENSTROPHY, NUC_SPEC_DENSITY, NUC_SPEC_PRESSURE};
            const char *sn[4] = {"hel", "ens", "den", "pre"};
            for (int i = 0; i < 4; i++) {
                nuc3d::check(nuc_spectrum_scalar(pde.ctx(), sf[i], sc.data(), kmax), "nuc_spectrum_scalar");
                shell_file(sn[i], sc.data());
            }
        }
    };
    try {
        nuc3d::PDEData3d pde(cfg);
        nuc3d::check(nuc_upload_state(pde.ctx(), q.data()), "nuc_upload_state");
        if (rc.average_type != 0) {
            // a new average, or the one saved at TimeAverageStartStep (src/postproc.cpp:157-256)
            if (rc.average_start != 0)
                nuc3d::check(nuc_average_read(pde.ctx(), &cfg, rc.average_type, "AveragedFlowData", rc.average_start),
                             "nuc_average_read");
            else
                nuc3d::check(nuc_average_begin(pde.ctx(), rc.average_type), "nuc_average_begin");
        }
        while (time <= rc.tl && step < rc.end_step) {
            auto t0 = std::chrono::steady_clock::now();
            nuc_step_info info = pde.step(1);
            double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            step++;
            time += info.dt;
            printf("========step = %d\n time = %g, dt = %g, residual = %g, GPU time = %g(s)\n", step, time, info.dt,
                   info.res, wall);
            // singleBlock::solvePDE's cadence (src/singleBlock.cpp:52-66, src/IOcontroller.cpp:194-213)
            if (rc.post_step > 0 && step % rc.post_step == 0) post(pde);
            if (rc.save_step > 0 && step % rc.save_step == 0) output(pde);
        }
        output(pde);   // and once more after the loop (src/singleBlock.cpp:67)
    } catch (const std::exception &e) {
        fprintf(stderr, "%s\n", e.what());
        return 2;
    }
    return 0;
}
