// nuc_force.h -- band-limited spectral forcing on the device (the Const
// forcing of src/physicsModel.cpp:1368-1677).
//
// The reference transforms u, v, w with full 3-D FFTs (src/fft.cpp:182-382),
// touches only the modes of one shell |k - k_f| < 0.5, and transforms back.
// Here the forward transform is evaluated only at the wave numbers the shell
// uses, one direction at a time (x, then y, then z: a pruned DFT), and the
// synthesis runs the same separable route backwards.  Work per stage is
// O(N * n_kx) instead of O(N log N) with three all-to-all transposes.
#pragma once
#include "nuc3d_gpu.h"

#include <hip/hip_runtime.h>
#include <string>

namespace nucforce {

constexpr int MAXMODES = 1024;  // shell modes
constexpr int MAXK = 32;        // distinct wave numbers per direction

struct Plan {
    int nx = 0, ny = 0, nz = 0;
    long N = 0;
    int nm = 0, na = 0, nb = 0, nc = 0;
    double amp = 0.0, amp_hel = 0.0;
    // device buffers
    int *modes = nullptr;   // per mode: ia, ib, ic, forced (sorted by ia)
    int *astart = nullptr;  // na+1 offsets of the ia groups in modes
    int *kv = nullptr;      // [3][MAXK] distinct wave numbers kx, ky, kz
    double *tw = nullptr;   // e^{-2 pi i p/n}: x [2nx], y [2ny], z [2nz]
    double *X = nullptr;    // 3 x na x (ny nz) complex; reused for the synthesis slabs
    double *Y = nullptr;    // 3 x na x nb x nz complex
    double *fh = nullptr;   // forcing spectrum per mode: 3 complex
    double *part = nullptr; // per x line: sum of f.u
    double *diag = nullptr; // E_s, E_c, theta of the last evaluation
};

// builds the mode list of the shell; returns NUC_OK or NUC_EINVAL / NUC_ENOMEM with err set
// allocations are stream-ordered on s
int plan_create(Plan &p, const nuc_config &c, hipStream_t s, std::string &err);
void plan_destroy(Plan &p, hipStream_t s);
long long plan_bytes(const nuc_config &c);

// rhs[0] = 0, rhs[1..3] = -f, rhs[4] = -(f.u); mean_out = <f.u> (added back to
// rhs[4] by the stage's last sweep, src/physicsModel.cpp:1605-1667)
void enqueue(const Plan &p, const double *u, const double *v, const double *w, double *const rhs[5],
             double *mean_out, hipStream_t s);

}  // namespace nucforce
