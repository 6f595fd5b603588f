// nuc3d_io.cpp -- the reference's file formats, per rank of its rank grid:
//   * binary restart files, block::outputQ_binary / inputQ_binary
//     (src/block.cpp:609-671): [double time][5 fields of the rank's block,
//     x fastest], named NUC3d_step_<step>_ID_<rank>.bin;
//   * Tecplot output, block::outputQ_tecplot (src/block.cpp:553-607): node
//     coordinates and 8 cell-centred primitive / acoustic fields in BLOCK
//     packing, 12 significant digits, named NUC3d_step_<step>_id_<rank>.dat;
//   * averaged-statistics files, postproc::OutputAveraged_bin and the reader in
//     postproc's constructor (src/postproc.cpp:157-256,412-522): [int averStep]
//     [the averaged groups of the averaging type], written as
//     Averaged_step_<averStep>_id_<rank>.bin and read as
//     Averaged_step_<step>_ID_<rank>.bin (the reference's writer and reader
//     differ in the case of "id"; the reader here accepts both);
//   * Tecplot files of the averaged fields, postproc::OutputAveraged_tec
//     (src/postproc.cpp:523-720), Averaged_tec_step_<averStep>_id_<rank>.dat;
//   * Tecplot files of solveQ's per-point fields, postproc::OutputTemporal
//     (src/postproc.cpp:346-405), Q_tec_step_<step>_id_<rank>.dat.
//
// Rank numbering: the reference's Cartesian communicator has dims
// (npz, npy, npx) (src/MPICommunicator.cpp:26-33) and MPI numbers Cartesian
// ranks row-major, so rank = (cz*npy + cy)*npx + cx; rank (cx, cy, cz) holds
// cells [cx*nlx, (cx+1)*nlx) x [cy*nly, ...) x [cz*nlz, ...) of the global grid
// (src/block.cpp:56-58,103-118).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "nuc3d_gpu.h"

namespace {

struct RankGrid {
    int n[3], np[3], nl[3];
    long N, Nl;
    int ranks;
};

bool rank_grid(const nuc_config *c, RankGrid &g)
{
    g.n[0] = c->nx; g.n[1] = c->ny; g.n[2] = c->nz;
    g.np[0] = c->npx; g.np[1] = c->npy; g.np[2] = c->npz;
    for (int d = 0; d < 3; d++) {
        if (g.n[d] < 1 || g.np[d] < 1 || g.n[d] % g.np[d] != 0) return false;
        g.nl[d] = g.n[d] / g.np[d];
    }
    g.N = (long)g.n[0] * g.n[1] * g.n[2];
    g.Nl = (long)g.nl[0] * g.nl[1] * g.nl[2];
    g.ranks = g.np[0] * g.np[1] * g.np[2];
    return true;
}

void rank_coords(const RankGrid &g, int rank, int c[3])
{
    c[0] = rank % g.np[0];
    c[1] = (rank / g.np[0]) % g.np[1];
    c[2] = rank / (g.np[0] * g.np[1]);
}

// copy field f of the rank's block between the global layout and a local one
void block_copy(const RankGrid &g, int rank, const double *glob, double *loc, bool to_local)
{
    int c[3];
    rank_coords(g, rank, c);
    for (int k = 0; k < g.nl[2]; k++)
        for (int j = 0; j < g.nl[1]; j++) {
            const long gi = (long)c[0] * g.nl[0]
                          + (long)g.n[0] * (((long)c[1] * g.nl[1] + j) + (long)g.n[1] * ((long)c[2] * g.nl[2] + k));
            const long li = (long)g.nl[0] * (j + (long)g.nl[1] * k);
            if (to_local) std::memcpy(loc + li, glob + gi, g.nl[0] * sizeof(double));
            else std::memcpy(const_cast<double *>(glob) + gi, loc + li, g.nl[0] * sizeof(double));
        }
}

std::string file_name(const char *dir, const char *stem, long step, const char *mid, int rank, const char *tail)
{
    std::ostringstream s;
    s << dir << "/" << stem << step << mid << rank << tail;
    return s.str();
}

const int kGroupCount[5] = {10, 21, 28, 49, 45};

// Tecplot variable names of the averaged groups, in field order
// (src/postproc.hpp:49-121): part of the file format
const char *const kAvgNames[153] = {
    // primary
    "rho", "u", "v", "w", "p", "T", "e", "alpha", "miu", "coeff",
    // gradients, vorticity and its squares
    "dudx", "dudy", "dudz", "dvdx", "dvdy", "dvdz", "dwdx", "dwdy", "dwdz", "dpdx", "dpdy", "dpdz",
    "dTdx", "dTdy", "dTdz", "omega0", "omega1", "omega2", "omega0^2", "omega1^2", "omega2^2",
    // derived
    "rhou", "rhov", "rhow", "rhoe", "uu", "uv", "uw", "vv", "vw", "ww",
    "rhouu", "rhouv", "rhouw", "rhovv", "rhovw", "rhoww",
    "tau_xx", "tau_xy", "tau_xz", "tau_yy", "tau_yz", "tau_zz",
    "rhotau_xx", "rhotau_xy", "rhotau_xz", "rhotau_yy", "rhotau_yz", "rhotau_zz",
    // TKE fields
    "up", "vp", "wp", "pdudx", "pdvdy", "pdwdz",
    "dtau_xxdx", "dtau_yxdx", "dtau_zxdx", "dtau_xydy", "dtau_yydy", "dtau_zydy",
    "dtau_xzdz", "dtau_yzdz", "dtau_zzdz",
    "utau_xx", "vtau_yx", "wtau_zx", "uktau_kx", "duktau_kxdx",
    "utau_xy", "vtau_yy", "wtau_zy", "uktau_ky", "duktau_kydy",
    "utau_xz", "vtau_yz", "wtau_zz", "uktau_kz", "duktau_kzdz",
    "tau_xxdudx", "tau_yxdvdx", "tau_zxdwdx", "tau_xydudy", "tau_yydvdy", "tau_zydwdy",
    "tau_xzdudz", "tau_yzdvdz", "tau_zzdwdz",
    "rhouuu", "rhouuv", "rhouuw", "rhovuv", "rhouvw", "rhouww", "rhovvv", "rhovvw", "rhovww", "rhowww",
    // TKE budget
    "tke", "production", "turbulen_transportation", "pressure_diffusion", "p1", "p2", "p3",
    "viscous_diffusion", "v1", "v2", "v3", "convection",
    "drhoudx", "drhoudy", "drhoudz", "drhovdx", "drhovdy", "drhovdz", "drhowdx", "drhowdy", "drhowdz",
    "up_t", "vp_t", "wp_t", "dupdx_t", "dvpdy_t", "dwpdz_t",
    "uktau_kx_t", "uktau_ky_t", "uktau_kz_t", "duktau_kxdx_t", "duktau_kydy_t", "duktau_kzdz_t",
    "rhouk", "rhovk", "rhowk", "drhoukdx", "drhovkdy", "drhowkdz",
    "u_tke", "v_tke", "w_tke", "du_tkedx", "dv_tkedy", "dw_tkedz"};

// node coordinates of a rank's block, x then y then z (src/block.cpp:124-138)
void write_nodes(std::ofstream &f, const RankGrid &g, const double L[3], const double o[3], int rank)
{
    int c[3];
    rank_coords(g, rank, c);
    for (int d = 0; d < 3; d++) {
        const double h = L[d] / g.n[d], base = L[d] / g.np[d] * c[d];
        for (int k = 0; k <= g.nl[2]; k++)
            for (int j = 0; j <= g.nl[1]; j++)
                for (int i = 0; i <= g.nl[0]; i++) {
                    const int ix = d == 0 ? i : (d == 1 ? j : k);
                    f << base + h * ix - o[d] << "\n";
                }
    }
}

int groups_of(int type) { return type == 1 ? 2 : (type == 2 ? 3 : 5); }

}  // namespace

extern "C" int nuc_write_restart(const nuc_config *cfg, const char *dir, int step, double time, const double *q)
{
    RankGrid g;
    if (!cfg || !dir || !q || !rank_grid(cfg, g)) return NUC_EINVAL;
    std::vector<double> loc(g.Nl);
    for (int r = 0; r < g.ranks; r++) {
        std::ofstream f(file_name(dir, "NUC3d_step_", step, "_ID_", r, ".bin"), std::ios::binary);
        if (!f) return NUC_EINVAL;
        f.write(reinterpret_cast<const char *>(&time), sizeof(double));
        for (int v = 0; v < 5; v++) {
            block_copy(g, r, q + (size_t)v * g.N, loc.data(), true);
            f.write(reinterpret_cast<const char *>(loc.data()), g.Nl * sizeof(double));
        }
        if (!f) return NUC_EINVAL;
    }
    return NUC_OK;
}

extern "C" int nuc_read_restart(const nuc_config *cfg, const char *dir, int step, double *time, double *q)
{
    RankGrid g;
    if (!cfg || !dir || !q || !rank_grid(cfg, g)) return NUC_EINVAL;
    std::vector<double> loc(g.Nl);
    for (int r = 0; r < g.ranks; r++) {
        std::ifstream f(file_name(dir, "NUC3d_step_", step, "_ID_", r, ".bin"), std::ios::binary);
        if (!f) return NUC_EINVAL;
        double t = 0.0;
        f.read(reinterpret_cast<char *>(&t), sizeof(double));
        if (r == 0 && time) *time = t;
        for (int v = 0; v < 5; v++) {
            f.read(reinterpret_cast<char *>(loc.data()), g.Nl * sizeof(double));
            if (!f) return NUC_EINVAL;   // short file: the wrong grid or rank grid
            block_copy(g, r, q + (size_t)v * g.N, loc.data(), false);
        }
    }
    return NUC_OK;
}

extern "C" int nuc_write_tecplot(const nuc_config *cfg, const char *dir, int step, const double *origin,
                                 const double *q)
{
    RankGrid g;
    if (!cfg || !dir || !q || !rank_grid(cfg, g)) return NUC_EINVAL;
    const double o[3] = {origin ? origin[0] : 0.0, origin ? origin[1] : 0.0, origin ? origin[2] : 0.0};
    const double L[3] = {cfg->lx, cfg->ly, cfg->lz};
    const double gamma = cfg->gamma, mach = cfg->mach;
    std::vector<double> loc(g.Nl * 5);
    for (int r = 0; r < g.ranks; r++) {
        int c[3];
        rank_coords(g, r, c);
        for (int v = 0; v < 5; v++) block_copy(g, r, q + (size_t)v * g.N, loc.data() + (size_t)v * g.Nl, true);
        std::ofstream f(file_name(dir, "NUC3d_step_", step, "_id_", r, ".dat"));
        if (!f) return NUC_EINVAL;
        f << "title=NUC3d\n" << "variables=x,y,z";
        for (const char *name : {"rho", "u", "v", "w", "pressure", "temperature", "e", "sound_speed"}) f << "," << name;
        f << "\n Zone I = " << g.nl[0] + 1 << ", J= " << g.nl[1] + 1 << ", K=" << g.nl[2] + 1
          << "\n DATAPACKING=BLOCK, VARLOCATION=([4-11]=CELLCENTERED)\n";
        f << std::setprecision(12);
        write_nodes(f, g, L, o, r);
        // con2prim's primitive and acoustic fields (src/physicsModel.cpp:331-344)
        for (int v = 0; v < 8; v++)
            for (long id = 0; id < g.Nl; id++) {
                const double rho = loc[id], u = loc[g.Nl + id] / rho, vv = loc[2 * g.Nl + id] / rho;
                const double w = loc[3 * g.Nl + id] / rho, E = loc[4 * g.Nl + id];
                const double p = std::fabs((E - 0.5 * rho * (u * u + vv * vv + w * w)) * (gamma - 1.0));
                double x = 0.0;
                switch (v) {
                case 0: x = rho; break;
                case 1: x = u; break;
                case 2: x = vv; break;
                case 3: x = w; break;
                case 4: x = p; break;
                case 5: x = gamma * mach * mach * p / rho; break;
                case 6: x = p / rho / (gamma - 1.0); break;
                default: x = std::sqrt(gamma * p / rho); break;
                }
                f << x << "\n";
            }
        if (!f) return NUC_EINVAL;
    }
    return NUC_OK;
}

extern "C" int nuc_average_write(nuc_ctx *ctx, const nuc_config *cfg, int type, const char *dir)
{
    RankGrid g;
    if (!ctx || !cfg || !dir || type < 1 || type > 3 || !rank_grid(cfg, g)) return NUC_EINVAL;
    const long long steps = nuc_average_steps(ctx);
    if (steps < 0) return NUC_EINVAL;
    const int averStep = (int)steps;
    const int ng = groups_of(type);
    std::vector<std::ofstream> files(g.ranks);
    for (int r = 0; r < g.ranks; r++) {
        files[r].open(file_name(dir, "Averaged_step_", averStep, "_id_", r, ".bin"), std::ios::binary);
        if (!files[r]) return NUC_EINVAL;
        files[r].write(reinterpret_cast<const char *>(&averStep), sizeof(int));
    }
    std::vector<double> field(g.N), loc(g.Nl);
    for (int gr = 0; gr < ng; gr++)
        for (int i = 0; i < kGroupCount[gr]; i++) {
            int rc = nuc_average_field(ctx, gr, i, field.data());
            if (rc != NUC_OK) return rc;
            for (int r = 0; r < g.ranks; r++) {
                block_copy(g, r, field.data(), loc.data(), true);
                files[r].write(reinterpret_cast<const char *>(loc.data()), g.Nl * sizeof(double));
            }
        }
    for (auto &f : files)
        if (!f) return NUC_EINVAL;
    return NUC_OK;
}

extern "C" int nuc_average_read(nuc_ctx *ctx, const nuc_config *cfg, int type, const char *dir, int step)
{
    RankGrid g;
    if (!ctx || !cfg || !dir || type < 1 || type > 3 || !rank_grid(cfg, g)) return NUC_EINVAL;
    const int ng = groups_of(type);
    long nf = 0;
    for (int gr = 0; gr < ng; gr++) nf += kGroupCount[gr];
    std::vector<double> all((size_t)nf * g.N), loc(g.Nl);
    int averStep = 0;
    for (int r = 0; r < g.ranks; r++) {
        std::ifstream f(file_name(dir, "Averaged_step_", step, "_ID_", r, ".bin"), std::ios::binary);
        if (!f) f.open(file_name(dir, "Averaged_step_", step, "_id_", r, ".bin"), std::ios::binary);
        if (!f) return NUC_EINVAL;
        int s = 0;
        f.read(reinterpret_cast<char *>(&s), sizeof(int));
        if (r == 0) averStep = s;
        for (long i = 0; i < nf; i++) {
            f.read(reinterpret_cast<char *>(loc.data()), g.Nl * sizeof(double));
            if (!f) return NUC_EINVAL;
            block_copy(g, r, all.data() + (size_t)i * g.N, loc.data(), false);
        }
    }
    return nuc_average_restore(ctx, type, averStep, all.data());
}

extern "C" int nuc_average_write_tecplot(nuc_ctx *ctx, const nuc_config *cfg, int type, const char *dir, double time,
                                         const double *origin)
{
    RankGrid g;
    if (!ctx || !cfg || !dir || type < 1 || type > 3 || !rank_grid(cfg, g)) return NUC_EINVAL;
    const long long steps = nuc_average_steps(ctx);
    if (steps < 0) return NUC_EINVAL;
    const int ng = groups_of(type);
    int nf = 0;
    for (int gr = 0; gr < ng; gr++) nf += kGroupCount[gr];
    const double o[3] = {origin ? origin[0] : 0.0, origin ? origin[1] : 0.0, origin ? origin[2] : 0.0};
    const double L[3] = {cfg->lx, cfg->ly, cfg->lz};
    std::vector<std::ofstream> files(g.ranks);
    for (int r = 0; r < g.ranks; r++) {
        files[r].open(file_name(dir, "Averaged_tec_step_", steps, "_id_", r, ".dat"));
        if (!files[r]) return NUC_EINVAL;
        std::ofstream &f = files[r];
        // header of postproc::OutputAveraged_tec (src/postproc.cpp:554-641); the
        // time at the stream's default 6 digits, the values at 12
        f << "title=time_" << time << "\n" << "variables=x,y,z";
        for (int i = 0; i < nf; i++) f << "," << kAvgNames[i];
        f << "\n Zone I = " << g.nl[0] + 1 << ", J= " << g.nl[1] + 1 << ", K=" << g.nl[2] + 1
          << "\n DATAPACKING=BLOCK, VARLOCATION=([4-" << 3 + nf << "]=CELLCENTERED)\n";
        f << std::setprecision(12);
        write_nodes(f, g, L, o, r);
    }
    std::vector<double> field(g.N), loc(g.Nl);
    for (int gr = 0; gr < ng; gr++)
        for (int i = 0; i < kGroupCount[gr]; i++) {
            int rc = nuc_average_field(ctx, gr, i, field.data());
            if (rc != NUC_OK) return rc;
            for (int r = 0; r < g.ranks; r++) {
                block_copy(g, r, field.data(), loc.data(), true);
                for (long id = 0; id < g.Nl; id++) files[r] << loc[id] << "\n";
            }
        }
    for (auto &f : files)
        if (!f) return NUC_EINVAL;
    return NUC_OK;
}

extern "C" int nuc_write_q_tecplot(nuc_ctx *ctx, const nuc_config *cfg, const char *dir, int step, double time,
                                   const double *origin)
{
    RankGrid g;
    if (!ctx || !cfg || !dir || !rank_grid(cfg, g)) return NUC_EINVAL;
    std::vector<double> f((size_t)16 * g.N), loc(g.Nl);
    int rc = nuc_q_fields(ctx, f.data());
    if (rc != NUC_OK) return rc;
    static const char *const names[16] = {"dudx", "dudy", "dudz", "dvdx", "dvdy", "dvdz", "dwdx", "dwdy", "dwdz",
                                          "omega0", "omega1", "omega2", "enstrophy", "Q", "helicity", "kinetic"};
    const double o[3] = {origin ? origin[0] : 0.0, origin ? origin[1] : 0.0, origin ? origin[2] : 0.0};
    const double L[3] = {cfg->lx, cfg->ly, cfg->lz};
    for (int r = 0; r < g.ranks; r++) {
        std::ofstream out(file_name(dir, "Q_tec_step_", step, "_id_", r, ".dat"));
        if (!out) return NUC_EINVAL;
        out << "title=time_" << time << "\n" << "variables=x,y,z";
        for (const char *nm : names) out << "," << nm;
        out << "\n Zone I = " << g.nl[0] + 1 << ", J= " << g.nl[1] + 1 << ", K=" << g.nl[2] + 1
            << "\n DATAPACKING=BLOCK, VARLOCATION=([4-19]=CELLCENTERED)\n";
        out << std::setprecision(12);
        write_nodes(out, g, L, o, r);
        for (int v = 0; v < 16; v++) {
            block_copy(g, r, f.data() + (size_t)v * g.N, loc.data(), true);
            for (long id = 0; id < g.Nl; id++) out << loc[id] << "\n";
        }
        if (!out) return NUC_EINVAL;
    }
    return NUC_OK;
}
