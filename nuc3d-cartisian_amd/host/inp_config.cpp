// inp_config.cpp -- readers for the reference's inp/*.in run configuration,
// with the reference's parsing rules:
//   inp/block.in        npx npy npz / nx ny nz / xl yl zl / x0 y0 z0
//                       (src/block.cpp:49-61; a missing origin line reads as 0)
//   inp/Method.in       key value (src/fieldOperator.cpp:21-125)
//   inp/PhysModel.in    six positional pairs Neqs EoS Riemann PDEModel Viscous
//                       Forcing, then key value (src/physicsModel.cpp:72-156)
//   inp/IOController.in key value (src/IOcontroller.cpp:12-110)
#include "nuc3d_gpu.h"
#include "nuc3d_host.hpp"

#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>
#include <string>

namespace nuc3d {

static bool read_pairs(const std::string &path, std::vector<std::pair<std::string, std::string>> &out)
{
    std::ifstream f(path);
    if (!f) return false;
    std::string a, b;
    while (f >> a >> b) out.emplace_back(a, b);
    return true;
}

int scheme_from_name(const std::string &s, int *halo)
{
    // src/fieldOperator.cpp:130-176; an unknown name would select weno5js there but
    // leave its halo width unset (:177-180), so it is rejected here (-1)
    if (s == "weno7js") { *halo = 4; return NUC_WENO7JS; }
    if (s == "hccs") { *halo = 4; return NUC_HCCS; }
    if (s == "weno5js") { *halo = 3; return NUC_WENO5JS; }
    if (s == "weno5z") { *halo = 3; return NUC_WENO5Z; }
    if (s == "crweno5") { *halo = 3; return NUC_CRWENO5; }
    if (s == "hccs78") { *halo = 4; return NUC_HCCS78; }
    if (s == "hccs77") { *halo = 4; return NUC_HCCS77; }
    if (s == "weno3js") { *halo = 2; return NUC_WENO3JS; }
    if (s == "upwind1st") { *halo = 1; return NUC_UPWIND1ST; }
    *halo = 3;
    return -1;
}

int read_method(const std::string &dir, int scheme[3], int vis_scheme[3], std::string &err)
{
    // defaults src/fieldOperator.cpp:22-33
    std::map<std::string, std::string> method = {
        {"scheme_time", "RK3"}, {"scheme_x_ivs", "weno5js"}, {"scheme_y_ivs", "weno5js"},
        {"scheme_z_ivs", "weno5js"}, {"scheme_x_vis", "cd6"}, {"scheme_y_vis", "cd6"}, {"scheme_z_vis", "cd6"}};
    std::vector<std::pair<std::string, std::string>> kv;
    if (!read_pairs(dir + "/Method.in", kv)) { err = "inp/Method.in does not exist"; return NUC_EINVAL; }
    for (auto &p : kv) {
        if (method.count(p.first)) method[p.first] = p.second;
        else fprintf(stderr, "Method.in: parameter '%s' is not supported\n", p.first.c_str());
    }
    const char *ivs[3] = {"scheme_x_ivs", "scheme_y_ivs", "scheme_z_ivs"};
    const char *vis[3] = {"scheme_x_vis", "scheme_y_vis", "scheme_z_vis"};
    for (int d = 0; d < 3; d++) {
        int halo;
        int s = scheme_from_name(method[ivs[d]], &halo);
        if (s < 0) {
            err = "Method.in: inviscid scheme '" + method[ivs[d]] + "' is not available";
            return NUC_EINVAL;
        }
        scheme[d] = s;
        const std::string &v = method[vis[d]];
        if (v == "cd8") vis_scheme[d] = NUC_CD8;
        else if (v == "cd6") vis_scheme[d] = NUC_CD6;
        else if (v == "cd2") vis_scheme[d] = NUC_CD2;
        else vis_scheme[d] = NUC_CD6;  // src/fieldOperator.cpp:309-312 default
    }
    return NUC_OK;
}

int read_inp(const std::string &dir, nuc_config &c, RunControl &rc, std::string &err)
{
    std::memset(&c, 0, sizeof(c));
    // ---- block.in
    {
        std::ifstream f(dir + "/block.in");
        if (!f) { err = "inp/block.in does not exist"; return NUC_EINVAL; }
        double x0 = 0, y0 = 0, z0 = 0;
        f >> c.npx >> c.npy >> c.npz >> c.nx >> c.ny >> c.nz >> c.lx >> c.ly >> c.lz;
        if (!f) { err = "inp/block.in: expected ranks, grid and lengths"; return NUC_EINVAL; }
        f >> x0 >> y0 >> z0;  // optional origin line
        rc.origin[0] = f ? x0 : 0.0;
        rc.origin[1] = f ? y0 : 0.0;
        rc.origin[2] = f ? z0 : 0.0;
        for (int d = 0; d < 3; d++) c.origin[d] = rc.origin[d];
    }
    // ---- Method.in
    if (read_method(dir, c.scheme, c.vis_scheme, err) != NUC_OK) return NUC_EINVAL;
    // ---- PhysModel.in (defaults src/physicsModel.cpp:22-43)
    std::map<std::string, double> par = {
        {"Reynolds", 1000}, {"Mach", 0.5}, {"Pt", 0.72}, {"Rossby", 0.0}, {"x_center", 3.1415926},
        {"y_center", 3.1415926}, {"z_center", 0.0}, {"well_rad", 1.0}, {"Gamma", 1.4}, {"T_ref", 110.4},
        {"T_inf", 288.0}, {"T_wall", 1.0}, {"Forcing_amplitude", 0.01}, {"Forcing_amplitude_hel", 0.01},
        {"Forcing_wave_k", 1.0}, {"Mach_t", 0.3}, {"k_peak", 8.0}};
    std::map<std::string, std::string> sw = {{"Coriolis", "no"}, {"Centrifugal", "no"}};
    {
        std::ifstream f(dir + "/PhysModel.in");
        if (!f) { err = "inp/PhysModel.in does not exist"; return NUC_EINVAL; }
        std::string k, neqs, eos, riemann, model, visc, forcing;
        f >> k >> neqs >> k >> eos >> k >> riemann >> k >> model >> k >> visc >> k >> forcing;
        if (riemann == "LF") c.riemann = NUC_LF;
        else if (riemann == "AUSM") c.riemann = NUC_AUSM;
        else if (riemann == "AUSMp") c.riemann = NUC_AUSMP;
        else { err = "PhysModel.in: Riemann solver '" + riemann + "' is not available"; return NUC_EINVAL; }
        if (model == "NavierStokes3d") c.model = NUC_NAVIERSTOKES3D;
        else if (model == "Euler3d") c.model = NUC_EULER3D;
        else { err = "PhysModel.in: PDE model '" + model + "' does not exist"; return NUC_EINVAL; }
        c.viscosity = (visc == "Constant") ? NUC_CONSTANT : NUC_SUTHERLAND;  // unknown -> Sutherland
        if (forcing == "Const") c.forcing = NUC_FORCING_CONST;
        else if (forcing == "ABC") c.forcing = NUC_FORCING_ABC;
        else if (forcing == "TG") c.forcing = NUC_FORCING_TG;
        else c.forcing = NUC_FORCING_NONE;  // "None", or an unknown name -> None (src/physicsModel.cpp:104-110)
        std::string a, b;
        while (f >> a >> b) {
            if (par.count(a)) {
                std::istringstream vs(b);
                double v;
                vs >> v;
                par[a] = v;
            } else if (sw.count(a)) {
                sw[a] = b;
            } else {
                err = "PhysModel.in: word " + a + " does not exist";
                return NUC_EINVAL;
            }
        }
    }
    // any value other than "no" switches a source on (src/physicsModel.cpp:1295,1307)
    c.coriolis = (sw["Coriolis"] != "no") ? 1 : 0;
    c.centrifugal = (sw["Centrifugal"] != "no") ? 1 : 0;
    c.rossby = par["Rossby"]; c.x_center = par["x_center"]; c.y_center = par["y_center"];
    c.well_rad = par["well_rad"];
    c.gamma = par["Gamma"]; c.mach = par["Mach"]; c.reynolds = par["Reynolds"]; c.prandtl = par["Pt"];
    c.t_ref = par["T_ref"]; c.t_inf = par["T_inf"];
    c.forcing_amp = par["Forcing_amplitude"]; c.forcing_amp_hel = par["Forcing_amplitude_hel"];
    c.forcing_k = par["Forcing_wave_k"];
    rc.mach_t = par["Mach_t"]; rc.k_peak = par["k_peak"];
    // ---- IOController.in (defaults src/IOcontroller.cpp:13-38)
    std::map<std::string, int> io = {{"startStep", 0}, {"endStep", 100000}, {"postStep", 1},
                                     {"saveStep", 1000}, {"postSaveStep", 100}, {"Benchmark", 0}};
    std::map<std::string, double> tc = {{"cfl", 0.5}, {"tl", 0.0}, {"dt", 0.01}};
    std::map<std::string, std::string> ft = {{"Binary", "yes"}, {"Tecplot", "yes"}, {"PostProc", "yes"}};
    {
        std::vector<std::pair<std::string, std::string>> kv;
        if (!read_pairs(dir + "/IOController.in", kv)) { err = "inp/IOController.in does not exist"; return NUC_EINVAL; }
        for (auto &p : kv) {
            std::istringstream vs(p.second);
            if (io.count(p.first)) { int v = 0; vs >> v; io[p.first] = v; }
            else if (tc.count(p.first)) { double v = 0; vs >> v; tc[p.first] = v; }
            else if (ft.count(p.first)) ft[p.first] = p.second;
        }
    }
    c.cfl = tc["cfl"];
    rc.start_step = io["startStep"]; rc.end_step = io["endStep"]; rc.post_step = io["postStep"];
    rc.save_step = io["saveStep"]; rc.benchmark = io["Benchmark"]; rc.tl = tc["tl"];
    // a type is written when its value is "yes" (src/block.cpp:245-263)
    rc.binary = ft["Binary"] == "yes"; rc.tecplot = ft["Tecplot"] == "yes"; rc.postproc = ft["PostProc"] == "yes";
    rc.post_save_step = io["postSaveStep"];
    // ---- PostProc.in: values in order, the names ignored (src/postproc.cpp:62-77)
    {
        std::vector<std::pair<std::string, std::string>> kv;
        if (read_pairs(dir + "/PostProc.in", kv)) {
            std::vector<int> vals;
            for (auto &p : kv) { std::istringstream vs(p.second); int v = 0; vs >> v; vals.push_back(v); }
            if (vals.size() > 0) rc.average_type = vals[0];
            if (vals.size() > 1) rc.average_start = vals[1];
            if (vals.size() > 2) rc.simultaneous = vals[2];
            if (vals.size() > 3) rc.averaged_save_type = vals[3];
            if (vals.size() > 4) rc.simultaneous_save = vals[4];
            if (rc.average_type < 0 || rc.average_type > 3) {
                err = "PostProc.in: no such postproc for " + std::to_string(rc.average_type);
                return NUC_EINVAL;
            }
        }
    }
    c.precision = NUC_FP64;
    return NUC_OK;
}

}  // namespace nuc3d

extern "C" int nuc_config_from_inp(const char *dir, nuc_config *cfg, int *start_step, int *end_step)
{
    if (!dir || !cfg) return NUC_EINVAL;
    nuc3d::RunControl rc;
    std::string err;
    int r = nuc3d::read_inp(dir, *cfg, rc, err);
    if (r != NUC_OK) {
        fprintf(stderr, "nuc_config_from_inp: %s\n", err.c_str());
        return r;
    }
    if (start_step) *start_step = rc.start_step;
    if (end_step) *end_step = rc.end_step;
    return NUC_OK;
}
