// nuc3d.hip -- gfx950 kernels and the C ABI (include/nuc3d_gpu.h) for NUC3d's
// per-stage right-hand side and TVD-RK3 advance.
//
// Per RK stage (FP64 path):
//   k_prim   : at stage 0 the per-rank maximum eigenvalues (src/physicsModel.cpp:546-654,
//              LF z-into-eta quirk kept); with forcing also u,v,w,T of the stage input
//   k_dt     : dt_local = min over ranks of getDt (src/euler3d.cpp:459-466), dt_global = cfl*dt_local
//   k_gradq  : (Navier-Stokes) con2prim, gradients of u,v,w,T with cd6/cd8,
//              Sutherland viscosity, stress and heat flux (src/NaiverStokes3d.cpp:87-303)
//              straight from Q; writes the 9 distinct viscous flux components
//   k_sweep  : one kernel per direction.  Each workgroup holds LPB whole lines
//              in LDS; per variable it splits the flux, reconstructs f^+ and
//              f^- (explicit WENO or compact tridiagonal solved per wave),
//              differences the interface flux and the viscous flux and
//              accumulates the RHS.  The periodic halo of MPIComunicator is an
//              index wrap inside the line buffer.  The sweeps run z, y, x;
//              the x sweep ends with the TVD-RK3 update
//              (src/TVD-RK3.cpp:37-157) and the residual.
#include "nuc3d_gpu.h"
#include "nuc_device.h"
#include "nuc_lines.h"
#include "nuc_sweep.h"
#include "nuc_force.h"
#include "../host/hit_spectrum.hpp"

#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

using namespace nucdev;
using namespace nucsweep;

// the sweep launchers (and their kernels) are instantiated per scheme in
// sweep_inst.hip, one translation unit each
namespace nucsweep {
#define NUC_SWEEP_EXTERN(SCH)                                                       \
    extern template void launch_sweep_s<SCH>(int, const SweepArgs &, hipStream_t); \
    extern template void launch_sweep_pack_s<SCH>(int, int, const PackLaunch &, hipStream_t);
NUC_SWEEP_EXTERN(WENO5JS)
NUC_SWEEP_EXTERN(WENO5Z)
NUC_SWEEP_EXTERN(WENO7JS)
NUC_SWEEP_EXTERN(CRWENO5)
NUC_SWEEP_EXTERN(HCCS)
NUC_SWEEP_EXTERN(WENO3JS)
NUC_SWEEP_EXTERN(UPWIND1ST)
NUC_SWEEP_EXTERN(HCCS78)
NUC_SWEEP_EXTERN(HCCS77)
#undef NUC_SWEEP_EXTERN
}  // namespace nucsweep

namespace {

constexpr int MAXRANK = 64;
constexpr int NUC_PROF_SLOTS = 8;


constexpr int kSweepOrder[3] = {2, 1, 0};

struct PrimArgs {
    const void GMEM *q[5];    // state, element type S of the kernel
    double GMEM *prim[4];
    long N;
    int nx, ny, nlx, nly, nlz, npx, npy, P;
    int riemann, do_lam, write_prim;
    double gamma, mach;
    unsigned long long GMEM *lam;
};

struct GradArgs {
    const void GMEM *q[5];    // conserved state, element type S of the kernel
    const double GMEM *prim[4];
    void GMEM *out[9];        // 9 viscous flux components, T of the kernel
    int nx, ny, nz, gkz;
    int nloc[3], tsv[3], vis[3];
    double inv_h[3];
    int law;
    double re, mach, pr, gamma, tref, tinf;
};

__device__ __forceinline__ double bits2d(unsigned long long b) { return __longlong_as_double((long long)b); }
__device__ __forceinline__ unsigned long long d2bits(double d) { return (unsigned long long)__double_as_longlong(d); }

// ---------------------------------------------------------------------------
// con2prim (src/physicsModel.cpp:331-344) and eigenvalue maxima
// ---------------------------------------------------------------------------
// bid/nbk: this workgroup's index and the workgroup count of the job's
// grid-stride loop
template <typename S>
__device__ __forceinline__ void prim_body(const PrimArgs &A, long bid, long nbk)
{
    const S GMEM *const q0 = static_cast<const S GMEM *>(A.q[0]), *const q1 = static_cast<const S GMEM *>(A.q[1]);
    const S GMEM *const q2 = static_cast<const S GMEM *>(A.q[2]), *const q3 = static_cast<const S GMEM *>(A.q[3]);
    const S GMEM *const q4 = static_cast<const S GMEM *>(A.q[4]);
    __shared__ unsigned long long sl[3 * MAXRANK];
    const int t = threadIdx.x;
    if (A.do_lam) {
        for (int i = t; i < 3 * A.P; i += 256) sl[i] = 0ull;
        __syncthreads();
    }
    int cur = -1;
    double m0 = 0.0, m1 = 0.0, m2 = 0.0;
    const double gm1 = A.gamma - 1.0;
    for (long id = bid * 256 + t; id < A.N; id += nbk * 256) {
        double rho = q0[id];
        double ir = 1.0 / rho;
        double u = double(q1[id]) * ir, v = double(q2[id]) * ir, w = double(q3[id]) * ir;
        double E = q4[id];
        double p = fabs((E - 0.5 * rho * (u * u + v * v + w * w)) * gm1);
        if (A.write_prim) {
            A.prim[0][id] = u;
            A.prim[1][id] = v;
            A.prim[2][id] = w;
            A.prim[3][id] = A.gamma * A.mach * A.mach * p * ir;
        }
        if (A.do_lam) {
            double c = sqrt(A.gamma * p * ir);
            // owning rank of the cell; a box holds at most 2^29 cells, so
            // 32-bit index arithmetic suffices (and one rank needs none)
            int rank = 0;
            if (A.P > 1) {
                const unsigned ui = (unsigned)id, nx = (unsigned)A.nx, ny = (unsigned)A.ny;
                const unsigned i = ui % nx, r2 = ui / nx;
                const unsigned j = r2 % ny, k = r2 / ny;
                rank = (int)(i / (unsigned)A.nlx) + A.npx * ((int)(j / (unsigned)A.nly) + A.npy * (int)(k / (unsigned)A.nlz));
            }
            if (rank != cur) {
                if (cur >= 0) {
                    atomicMax(&sl[3 * cur + 0], d2bits(m0));
                    atomicMax(&sl[3 * cur + 1], d2bits(m1));
                    atomicMax(&sl[3 * cur + 2], d2bits(m2));
                }
                cur = rank; m0 = m1 = m2 = 0.0;
            }
            double lx = fabs(u) + c, ly = fabs(v) + c, lz = fabs(w) + c;
            m0 = fmax(lx, m0);
            m1 = fmax(ly, m1);
            if (A.riemann == RIE_LF) m1 = fmax(lz, m1);  // src/physicsModel.cpp:653-654
            else m2 = fmax(lz, m2);
        }
    }
    if (A.do_lam) {
        if (cur >= 0) {
            atomicMax(&sl[3 * cur + 0], d2bits(m0));
            atomicMax(&sl[3 * cur + 1], d2bits(m1));
            atomicMax(&sl[3 * cur + 2], d2bits(m2));
        }
        __syncthreads();
        for (int i = t; i < 3 * A.P; i += 256)
            if (sl[i]) atomicMax((unsigned long long *)&A.lam[i], sl[i]);
    }
}

template <typename S>
__global__ __launch_bounds__(256) void k_prim(PrimArgs A) { prim_body<S>(A, blockIdx.x, gridDim.x); }

// dt per rank and min over ranks; resets the eigenvalue slots for the next step
struct DtArgs {
    unsigned long long GMEM *lam;
    int P, set_global;
    double hx, hy, hz, cfl;
    double GMEM *scal;
};

__device__ __forceinline__ void dt_body(const DtArgs &A)
{
    if (threadIdx.x == 0) {
        double dtmin = 0.0;
        for (int r = 0; r < A.P; r++) {
            double dt_xi = A.hx / bits2d(A.lam[3 * r + 0]);
            double dt_eta = A.hy / bits2d(A.lam[3 * r + 1]);
            double dt_zeta = A.hz / bits2d(A.lam[3 * r + 2]);
            double dt = 1.0 / (1.0 / dt_xi + 1.0 / dt_eta + 1.0 / dt_zeta);
            if (r == 0 || dt < dtmin) dtmin = dt;
        }
        A.scal[3] = dtmin;
        if (A.set_global) A.scal[0] = dtmin * A.cfl;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 3 * A.P; i += blockDim.x) A.lam[i] = 0ull;
}

__global__ void k_dt(DtArgs A) { dt_body(A); }

// ---------------------------------------------------------------------------
// gradients + viscous fluxes
// ---------------------------------------------------------------------------
__device__ __forceinline__ int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// Gradients and viscous fluxes straight from the conserved state, as a
// plane-marching stencil: a workgroup owns a GTX x GTY column of cells and
// walks A.gkz planes along z.  x/y derivatives read a (GTX+8) x (GTY+8) tile of
// the current plane's primitives in LDS (halos wrapped periodically); the z
// derivative reads a 9-plane register queue of the thread's own column.  Each
// state value is fetched from HBM about once (plus the z overlap of
// neighbouring workgroups), instead of the 27 stencil reads per field of a
// gather kernel, and no primitive field round-trips through HBM.
constexpr int GTX = 32, GTY = 16;
// planes per workgroup: long marches amortise the 8-plane z halo on large
// boxes; short ones give small boxes enough workgroups
inline int grad_planes(int nz) { return nz >= 256 ? 32 : (nz >= 128 ? 16 : (nz >= 64 ? 8 : 4)); }

// primitives u, v, w, T of one cell from its five conserved values
// (src/physicsModel.cpp:331-344, the same expression as k_prim)
__device__ __forceinline__ void con2prim_uvwT(const double qc[5], double gamma, double mach, double &u, double &v,
                                              double &w, double &T)
{
    const double gm1 = gamma - 1.0;
    double rho = qc[0];
    double ir = 1.0 / rho;
    u = qc[1] * ir; v = qc[2] * ir; w = qc[3] * ir;
    double E = qc[4];
    double p = fabs((E - 0.5 * rho * (u * u + v * v + w * w)) * gm1);
    T = gamma * mach * mach * p * ir;
}

template <typename S>
__device__ __forceinline__ void load_cell(const void GMEM *const *q, long id, double qc[5])
{
#pragma unroll
    for (int f = 0; f < 5; f++) qc[f] = static_cast<const S GMEM *>(q[f])[id];
}

template <typename T, typename S>
__device__ __forceinline__ void gradq_body(const GradArgs &A, int bx, int by, int bz)
{
    constexpr int PX = GTX + 8, PY = GTY + 8;
    __shared__ T tile[4][PY][PX];
    const int tx = threadIdx.x, ty = threadIdx.y, t = tx + GTX * ty;
    const int i0 = bx * GTX, j0 = by * GTY, k0 = bz * A.gkz;
    const int i = i0 + tx, j = j0 + ty;
    const int nx = A.nx, ny = A.ny, nz = A.nz;
    const long sy = nx, sz = (long)nx * ny;
    const int iw = wrapi(i % nx, nx), jw = wrapi(j % ny, ny);
    const bool own = i < nx && j < ny;
    const double gamma = A.gamma, mach = A.mach;
    const int kend = min(k0 + A.gkz, nz);
    // register queue: planes k-4 .. k+4 of u, v, w, T at (iw, jw)
    T qz[4][9];
#pragma unroll
    for (int o = 0; o < 8; o++) {
        int kk = wrapi((k0 - 4 + o) % nz, nz);
        double qc[5], u, v, w, Tt;
        load_cell<S>(A.q, iw + sy * jw + sz * kk, qc);
        con2prim_uvwT(qc, gamma, mach, u, v, w, Tt);
        qz[0][o + 1] = T(u); qz[1][o + 1] = T(v); qz[2][o + 1] = T(w); qz[3][o + 1] = T(Tt);
    }
    // The tile's 4-cell ring around the current plane: (RING points, at most
    // one per thread, the same (x, y) on every plane).  Its state and the
    // column's plane k+4 are loaded one plane ahead (ring_q, col_q), so that
    // a plane step does not wait for its own loads: a 512-thread workgroup
    // is alone on its CU at this register count.
    constexpr int RING_ROWS = 8 * PX;                 // 4 rows above and below the tile
    constexpr int RING = RING_ROWS + GTY * 8;          // + 4 columns left and right
    static_assert(RING <= GTX * GTY, "one ring point per thread");
    const bool has_ring = t < RING;
    int rcx = 0, rcy = 0;
    if (t < RING_ROWS) {
        rcx = t % PX;
        rcy = t / PX;
        rcy = rcy < 4 ? rcy : rcy + GTY;
    } else {
        const int u8 = t - RING_ROWS;
        rcy = 4 + u8 / 8;
        rcx = u8 % 8;
        rcx = rcx < 4 ? rcx : rcx + GTX;
    }
    const long ring_xy = wrapi((i0 + rcx - 4 + 4 * nx) % nx, nx) + sy * wrapi((j0 + rcy - 4 + 4 * ny) % ny, ny);
    double ring_q[5], col_q[5];
    if (has_ring) load_cell<S>(A.q, ring_xy + sz * k0, ring_q);
    load_cell<S>(A.q, iw + sy * jw + sz * wrapi((k0 + 4) % nz, nz), col_q);
    const T Ssu = T(A.tref / A.tinf);   // Sutherland constant over T_inf
    const T coeff_den = T((A.gamma - 1.0) * A.mach * A.mach * A.pr);
    const int cli = iw % A.nloc[0], clj = jw % A.nloc[1];
    const bool inx = cli >= A.tsv[0] && cli < A.nloc[0] - A.tsv[0];
    const bool iny = clj >= A.tsv[1] && clj < A.nloc[1] - A.tsv[1];
    T GMEM *const *out = reinterpret_cast<T GMEM *const *>(A.out);
    for (int k = k0; k < kend; k++) {
        // advance the queue by one plane
#pragma unroll
        for (int f = 0; f < 4; f++)
#pragma unroll
            for (int o = 0; o < 8; o++) qz[f][o] = qz[f][o + 1];
        {
            double u, v, w, Tt;
            con2prim_uvwT(col_q, gamma, mach, u, v, w, Tt);   // plane k+4, loaded one step ago
            qz[0][8] = T(u); qz[1][8] = T(v); qz[2][8] = T(w); qz[3][8] = T(Tt);
        }
        if (k + 1 < kend) load_cell<S>(A.q, iw + sy * jw + sz * wrapi((k + 5) % nz, nz), col_q);
        // tile of plane k with wrapped x/y halos: the thread's own (wrapped)
        // cell is the centre of its register queue; only the 4-cell ring
        // around the tile is read from the state
#pragma unroll
        for (int f = 0; f < 4; f++) tile[f][ty + 4][tx + 4] = qz[f][4];
        if (has_ring) {
            double u, v, w, Tt;
            con2prim_uvwT(ring_q, gamma, mach, u, v, w, Tt);   // plane k, loaded one step ago
            tile[0][rcy][rcx] = T(u); tile[1][rcy][rcx] = T(v); tile[2][rcy][rcx] = T(w); tile[3][rcy][rcx] = T(Tt);
            if (k + 1 < kend) load_cell<S>(A.q, ring_xy + sz * (k + 1), ring_q);
        }
        __syncthreads();
        if (own) {
            const int clk = k % A.nloc[2];
            const bool inz = clk >= A.tsv[2] && clk < A.nloc[2] - A.tsv[2];
            T g[4][3];
#pragma unroll
            for (int f = 0; f < 4; f++) {
                T Lx[9], Ly[9];
#pragma unroll
                for (int o = 0; o < 9; o++) {
                    Lx[o] = tile[f][ty + 4][tx + o];
                    Ly[o] = tile[f][ty + o][tx + 4];
                }
                g[f][0] = central_diff<T>(Lx + 4, 0, inx, A.vis[0], T(A.inv_h[0]));
                g[f][1] = central_diff<T>(Ly + 4, 0, iny, A.vis[1], T(A.inv_h[1]));
                g[f][2] = central_diff<T>(qz[f] + 4, 0, inz, A.vis[2], T(A.inv_h[2]));
            }
            T ux = g[0][0], uy = g[0][1], uz = g[0][2];
            T vx = g[1][0], vy = g[1][1], vz = g[1][2];
            T wx = g[2][0], wy = g[2][1], wz = g[2][2];
            T tx_ = g[3][0], ty_ = g[3][1], tz_ = g[3][2];
            T Tt = qz[3][4];
            T miu0;
            if (A.law == 0) miu0 = (T(1) + Ssu) / (Tt + Ssu) * (Tt * sqrt(Tt)) / T(A.re);   // src/physicsModel.cpp:220
            else miu0 = T(1) / T(A.re);
            T coeff0 = miu0 / coeff_den;
            T div = ux + vy + wz;
            // src/NaiverStokes3d.cpp:265-283 (off-diagonals carry -2/3 div u as coded)
            T txx = miu0 * (T(2) * ux - T(2.0 / 3.0) * div);
            T txy = miu0 * (uy + vx - T(2.0 / 3.0) * div);
            T txz = miu0 * (uz + wx - T(2.0 / 3.0) * div);
            T tyy = miu0 * (T(2) * vy - T(2.0 / 3.0) * div);
            T tyz = miu0 * (vz + wy - T(2.0 / 3.0) * div);
            T tzz = miu0 * (T(2) * wz - T(2.0 / 3.0) * div);
            T u = qz[0][4], v = qz[1][4], w = qz[2][4];
            const long id = i + sy * j + sz * k;
            out[0][id] = txx;
            out[1][id] = tyy;
            out[2][id] = tzz;
            out[3][id] = txy;
            out[4][id] = txz;
            out[5][id] = tyz;
            out[6][id] = u * txx + v * txy + w * txz + coeff0 * tx_;
            out[7][id] = u * txy + v * tyy + w * tyz + coeff0 * ty_;
            out[8][id] = u * txz + v * tyz + w * tzz + coeff0 * tz_;
        }
        __syncthreads();
    }
}

template <typename T, typename S>
__global__ __launch_bounds__(GTX * GTY) void k_gradq(GradArgs A) { gradq_body<T, S>(A, blockIdx.x, blockIdx.y, blockIdx.z); }

struct ResArgs {
    const double GMEM *part;
    int nb;
    double inv_count;
    double GMEM *scal;
};

__device__ __forceinline__ void res_body(const ResArgs &A)
{
    const double GMEM *part = A.part;
    const int nb = A.nb;
    double GMEM *scal = A.scal;
    const double inv_count = A.inv_count;
    __shared__ double sm[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < nb; i += 256) acc += part[i];
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        scal[1] = sqrt(sm[0] * inv_count) / scal[0];
        scal[2] += scal[0];  // time += dt (src/block.cpp:208-209)
    }
}

__global__ void k_res(ResArgs A) { res_body(A); }

// ---------------------------------------------------------------------------
// statistics (postproc::solveQ set), per-block partial sums then a fixed-order reduce
// ---------------------------------------------------------------------------
constexpr int NSTAT = 24;
struct StatArgs {
    const double *prim[4];   // u v w T
    const double *rho;
    int nx, ny, nz;
    int nloc[3], tsv[3], vis[3];
    double inv_h[3];
    double gamma, mach, tref, tinf;
    double *part;
};

__global__ __launch_bounds__(256) void k_stats(StatArgs A)
{
    __shared__ double sm[NSTAT][4];
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y;
    const int k = blockIdx.z;
    double acc[NSTAT];
    for (int s = 0; s < NSTAT; s++) acc[s] = 0.0;
    if (i < A.nx && j < A.ny) {
        const long sy = A.nx, sz = (long)A.nx * A.ny;
        const long id = i + sy * j + sz * k;
        const int pos[3] = {i, j, k};
        const int ext[3] = {A.nx, A.ny, A.nz};
        const long str[3] = {1, sy, sz};
        double g[3][3];
        for (int d = 0; d < 3; d++) {
            int cl = pos[d] % A.nloc[d];
            for (int f = 0; f < 3; f++) {
                double L[9];
                for (int o = -4; o <= 4; o++) {
                    int p = wrapi(pos[d] + o, ext[d]);
                    L[o + 4] = A.prim[f][id + (long)(p - pos[d]) * str[d]];
                }
                g[f][d] = central_diff<double>(L + 4, 0, cl >= A.tsv[d] && cl < A.nloc[d] - A.tsv[d], A.vis[d], A.inv_h[d]);
            }
        }
        double ux = g[0][0], uy = g[0][1], uz = g[0][2];
        double vx = g[1][0], vy = g[1][1], vz = g[1][2];
        double wx = g[2][0], wy = g[2][1], wz = g[2][2];
        double u0 = A.prim[0][id], u1 = A.prim[1][id], u2 = A.prim[2][id], T = A.prim[3][id];
        double rho = A.rho[id];
        double w0 = wy - vz, w1 = uz - wx, w2 = vx - uy;
        double q0 = ux * vy + ux * wz + vy * wz - uy * vx - uz * wx - vz * wy;
        // sound speed from T: c^2 = gamma p/rho = T/M^2
        double alpha = sqrt(T) / A.mach;
        double S = A.tref / A.tinf;
        double mu = (1.0 + S) / (T + S) * (T * sqrt(T));
        double div = ux + vy + wz;
        double vals[NSTAT] = {ux, uy, uz, vx, vy, vz, wx, wy, wz, w0, w1, w2,
                              (w0 * w0 + w1 * w1 + w2 * w2) * 0.5, q0, w0 * u0 + w1 * u1 + w2 * u2,
                              0.5 * rho * (u0 * u0 + u1 * u1 + u2 * u2),
                              (u0 * u0 + u1 * u1 + u2 * u2) / 3.0,
                              (ux * ux + vy * vy + wz * wz) / 3.0,
                              (ux * ux * ux + vy * vy * vy + wz * wz * wz) / 3.0,
                              alpha,
                              mu, rho,
                              // solenoidal and dilatational dissipation integrands
                              mu * (w0 * w0 + w1 * w1 + w2 * w2), mu * div * div};
        for (int s = 0; s < NSTAT; s++) acc[s] = vals[s];
    }
    const int t = threadIdx.x + 64 * threadIdx.y, lane = t & 63, wave = t >> 6;
    for (int s = 0; s < NSTAT; s++) {
        double x = acc[s];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o);
        if (lane == 0) sm[s][wave] = x;
    }
    __syncthreads();
    if (t < NSTAT) {
        long b = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
        A.part[b * NSTAT + t] = ((sm[t][0] + sm[t][1]) + sm[t][2]) + sm[t][3];
    }
}

// one workgroup per statistic: strided partial sums in a fixed order, then a
// fixed-shape tree (deterministic run to run)
__global__ __launch_bounds__(256) void k_stats_reduce(const double *part, long nb, double *out)
{
    __shared__ double sm[256];
    const int s = blockIdx.x, t = threadIdx.x;
    double acc = 0.0;
    for (long b = t; b < nb; b += 256) acc += part[b * NSTAT + s];
    sm[t] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) sm[t] += sm[t + w];
        __syncthreads();
    }
    if (t == 0) out[s] = sm[0];
}

// ---------------------------------------------------------------------------
// operator hooks: one wave per line of a Field with explicit ghost strips
// ---------------------------------------------------------------------------
struct OpRecArgs {
    const double *in, *gL, *gR;
    double *out;
    int n1, nlines, bf, ts, upwind, zones;
};

template <int SCH, int R>
__global__ __launch_bounds__(64) void k_op_recon(OpRecArgs A)
{
    constexpr int NP = Skew<R>::pitch(32 * R);
    constexpr int RAD = Radius<SCH>::value;
    __shared__ double sL[NP];
    __shared__ double sLm[NP];
    const int lane = threadIdx.x;
    const long line = blockIdx.x;
    const int n = A.n1;
    double *L = sL;
    for (int c = lane - HALO; c < 32 * R + HALO; c += 64) {
        double x = 0.0;
        if (c >= 0 && c < n) x = A.in[line * n + c];
        else if (c < 0 && A.gL && -c <= A.bf) x = A.gL[line * A.bf + (A.bf + c)];
        else if (c >= n && A.gR && c - n < A.bf && c < n + HALO) x = A.gR[line * A.bf + (c - n)];
        L[Skew<R>::at(c)] = x;
        sLm[Skew<R>::mir(c)] = x;
    }
    __syncthreads();
    const int h = lane & 31;
    const bool neg = lane >= 32;
    const int chunk = neg ? 31 - h : h;
    const int m0 = chunk * R;
    const int m0r = neg ? m0 + R - 1 : m0;
    const int mstep = neg ? -1 : 1;
    const Zone z{n + 1, n, A.ts};
    double out[R];
    line_half<SCH, R, double, double, double>((neg ? sLm : L) + h * (R + 1), m0r, mstep, z, h, neg, out);
    if (neg != (A.upwind < 0)) return;
    for (int r = 0; r < R; r++) {
        int m = m0r + mstep * r;
        if (m > n) continue;
        int zbit = (m < A.ts) ? 2 : (m > n - A.ts ? 4 : 1);
        if (A.zones & zbit) A.out[line * (n + 1) + m] = out[r];
    }
}

struct OpDiffArgs {
    const double *in, *gL, *gR;
    double *out;
    int n1, bf, ts, vis, zones;
    long nlines;
    double inv_h;
};

__global__ void k_op_diff(OpDiffArgs A)
{
    long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= A.nlines * A.n1) return;
    long line = gid / A.n1;
    int c = (int)(gid % A.n1);
    int n = A.n1;
    int zbit = (c < A.ts) ? 2 : (c >= n - A.ts ? 4 : 1);
    if (!(A.zones & zbit)) return;
    double Lb[9];
    for (int o = -4; o <= 4; o++) {
        int cc = c + o;
        double x = 0.0;
        if (cc >= 0 && cc < n) x = A.in[line * n + cc];
        else if (cc < 0 && A.gL && -cc <= A.bf) x = A.gL[line * A.bf + (A.bf + cc)];
        else if (cc >= n && A.gR && cc - n < A.bf) x = A.gR[line * A.bf + (cc - n)];
        Lb[o + 4] = x;
    }
    A.out[gid] = central_diff<double>(Lb + 4, 0, c >= A.ts && c < n - A.ts, A.vis, A.inv_h);
}

__global__ void k_op_rk3(int stage, const double *rhs, const double *un, double *ui, double dt, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double rr = -rhs[i];
    if (stage == 0) ui[i] = ui[i] + rr * dt;
    else if (stage == 1) ui[i] = 0.75 * un[i] + 0.25 * (rr * dt + ui[i]);
    else ui[i] = (1.0 / 3.0) * un[i] + (2.0 / 3.0) * (rr * dt + ui[i]);
}

// ---------------------------------------------------------------------------
// dispatch helpers
// ---------------------------------------------------------------------------
// rows per lane of a half wave for a line of nrows interfaces (32 lanes per sign)
int rows_for(int nrows)
{
    int r = (nrows + 31) / 32;
    if (r <= 2) return 2;
    if (r <= 4) return 4;
    if (r <= 6) return 6;
    if (r <= 8) return 8;
    if (r <= 16) return 16;
    return -1;
}

// rows per lane of a sweep over a periodic line of n cells: also 3 and 5, so
// that lines of 65-96 and 129-160 cells do not leave up to a quarter of the
// lanes idle, and 12 for 289-384 cells (on 16 rows a 384-cell line leaves a
// quarter idle, and the compact schemes' workgroups hold 103 KB of LDS, one
// per CU, against 77 KB).  7 rows for 193-224 cells measured no faster than
// 8 (odd R has no affine LDS slots and no tail-free form).
int sweep_rows(int n)
{
    int r = (n + 31) / 32;
    if (r == 3 || r == 5) return r;
    if (r >= 10 && r <= 12) return 12;
    return rows_for(n);
}

// lines per workgroup: 4 (one wave per line).  With the sweep's register
// budget two such workgroups share a CU, so one workgroup's barrier waits
// overlap the other's work; the y/z load mapping still reads 4 consecutive
// x cells per line cell, and xcd_block keeps the neighbouring groups that
// share the cache line on one XCD.
int sweep_lpb(int) { return SWEEP_LPB; }

void launch_sweep(int scheme, int R, const SweepArgs &A, hipStream_t s)
{
    switch (scheme) {
    case WENO5JS: launch_sweep_s<WENO5JS>(R, A, s); break;
    case WENO5Z: launch_sweep_s<WENO5Z>(R, A, s); break;
    case WENO7JS: launch_sweep_s<WENO7JS>(R, A, s); break;
    case CRWENO5: launch_sweep_s<CRWENO5>(R, A, s); break;
    case WENO3JS: launch_sweep_s<WENO3JS>(R, A, s); break;
    case UPWIND1ST: launch_sweep_s<UPWIND1ST>(R, A, s); break;
    case HCCS78: launch_sweep_s<HCCS78>(R, A, s); break;
    case HCCS77: launch_sweep_s<HCCS77>(R, A, s); break;
    default: launch_sweep_s<HCCS>(R, A, s); break;
    }
}

// ---------------------------------------------------------------------------
// packed launches: one grid covers the workgroups of several jobs (contexts
// with the same kernel configuration); workgroup b belongs to job j with
// off[j] <= b < off[j+1] and runs that job's body with that job's arguments
// at job-relative index b - off[j], i.e. exactly the work of the same
// workgroup in the job's own launch.  Only short lines (R <= PACK_RMAX) are
// packed: longer ones fill the device alone.  The tables are const
// __restrict__, so their entries are read with scalar loads.
// ---------------------------------------------------------------------------
constexpr int MAXPACK = 64;
constexpr int PACK_RMAX = 4;

__global__ __launch_bounds__(256) void k_prim_pack(const PrimArgs *__restrict__ tab, const int *__restrict__ off, int nj)
{
    const int j = pack_job(off, nj, blockIdx.x);
    prim_body<double>(tab[j], (long)blockIdx.x - off[j], (long)off[j + 1] - off[j]);
}

__global__ void k_dt_pack(const DtArgs *__restrict__ tab) { dt_body(tab[blockIdx.x]); }

__global__ void k_res_pack(const ResArgs *__restrict__ tab) { res_body(tab[blockIdx.x]); }

template <typename T>
__global__ __launch_bounds__(GTX * GTY) void k_gradq_pack(const GradArgs *__restrict__ tab, const int *__restrict__ off, int nj)
{
    const int j = pack_job(off, nj, blockIdx.x);
    const GradArgs &A = tab[j];
    const int rel = (int)blockIdx.x - off[j];
    const int gx = (A.nx + GTX - 1) / GTX, gy = (A.ny + GTY - 1) / GTY;
    gradq_body<T, double>(A, rel % gx, (rel / gx) % gy, rel / (gx * gy));
}

void launch_sweep_pack(int scheme, int riemann, int R, const PackLaunch &P, hipStream_t s)
{
    switch (scheme) {
    case WENO5JS: launch_sweep_pack_s<WENO5JS>(riemann, R, P, s); break;
    case WENO5Z: launch_sweep_pack_s<WENO5Z>(riemann, R, P, s); break;
    case WENO7JS: launch_sweep_pack_s<WENO7JS>(riemann, R, P, s); break;
    case CRWENO5: launch_sweep_pack_s<CRWENO5>(riemann, R, P, s); break;
    case WENO3JS: launch_sweep_pack_s<WENO3JS>(riemann, R, P, s); break;
    case UPWIND1ST: launch_sweep_pack_s<UPWIND1ST>(riemann, R, P, s); break;
    case HCCS78: launch_sweep_pack_s<HCCS78>(riemann, R, P, s); break;
    case HCCS77: launch_sweep_pack_s<HCCS77>(riemann, R, P, s); break;
    default: launch_sweep_pack_s<HCCS>(riemann, R, P, s); break;
    }
}

template <int SCH, int R>
void launch_op_r(const OpRecArgs &A, hipStream_t s)
{
    hipLaunchKernelGGL((k_op_recon<SCH, R>), dim3(A.nlines), dim3(64), 0, s, A);
}

template <int SCH>
void launch_op_sch(int R, const OpRecArgs &A, hipStream_t s)
{
    switch (R) {
    case 2: launch_op_r<SCH, 2>(A, s); break;
    case 4: launch_op_r<SCH, 4>(A, s); break;
    case 6: launch_op_r<SCH, 6>(A, s); break;
    case 8: launch_op_r<SCH, 8>(A, s); break;
    default: launch_op_r<SCH, 16>(A, s); break;
    }
}

}  // namespace

// ===========================================================================
// context
// ===========================================================================
struct nuc_ctx {
    nuc_config c;
    hipStream_t stream = nullptr;
    long N = 0;
    int n[3], np[3], nloc[3], ts[3], tsv[3], cpt[3];
    int seg[3], nseg[3];   // sweep segment length and count per line (nseg = 1 up to 512 cells)
    double h[3];
    int P = 1;
    double *q[3] = {nullptr, nullptr, nullptr};
    double *rhs = nullptr, *prim = nullptr, *fv = nullptr;
    unsigned long long *lam = nullptr;
    double *scal = nullptr;
    double *res_part = nullptr;
    long res_nb = 0;
    double *stat_part = nullptr;
    long stat_nb = 0;
    double *stat_out = nullptr;
    // time-averaged statistics (nuc_average_*): temporal and averaged fields of
    // each group, budget fields, reduction partials; type 0 = not running
    int post_type = 0;
    long post_step = 0;
    double *post_buf = nullptr;
    bool forcing = false;
    bool sources = false;   // ABC/TG forcing or rotating-frame sources (k_source)
    bool mixed = false;     // NUC_MIXED (and NUC_FP32): FP32 stencils, FP64 weights/solves/RK3
    // viscous fluxes stored FP32 (k_gradq<float>): the mixed paths when some
    // direction runs FP32 stencils.  hccs and hccs77 keep FP64 stencils, and
    // their FP64 sweeps read FP32 fluxes more slowly than they save in the
    // gradient kernel (forced 512^3 hccs: 277 against 257 ms per step)
    bool fv32 = false;
    bool fp32 = false;      // NUC_FP32: state and RHS stored in FP32 (the first half of each buffer)
    double *q64 = nullptr;  // NUC_FP32: the state widened to FP64 for I/O, statistics, post-processing
    // planes per workgroup of the gradient kernel (nuc_autotune; 0: grad_planes)
    int gkz = 0;
    // next RK stage of nuc_rk3_stage; the current state is q[next_stage]
    int next_stage = 0;
    // one captured RK3 step, replayed by nuc_step/nuc_step_async (nuc_set_graph)
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    nucforce::Plan fplan;
    std::string err;
    long step = 0;
    double time = 0.0;
    int pending = 0;  // steps enqueued but not yet accounted in time
    // per-kernel timing with HIP events on the launch stream (nuc_profile_*)
    bool prof = false;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev;
    double prof_ms[NUC_PROF_SLOTS] = {0};
    long prof_n[NUC_PROF_SLOTS] = {0};
};

static const char *kProfNames[NUC_PROF_SLOTS] = {"k_prim", "k_dt", "k_grad", "k_sweep_x", "k_sweep_y",
                                                 "k_sweep_z", "k_res", "k_force"};

// brackets one launch with events when profiling is on
struct ProfScope {
    nuc_ctx *x;
    int slot;
    hipEvent_t a = nullptr, b = nullptr;
    ProfScope(nuc_ctx *x_, int s_) : x(x_), slot(s_)
    {
        if (x->prof) {
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
            (void)hipEventRecord(a, x->stream);
        }
    }
    ~ProfScope()
    {
        if (x->prof) {
            (void)hipEventRecord(b, x->stream);
            x->ev.push_back({slot, {a, b}});
        }
    }
};

#define NUC_CHECK(ctx, call)                                                                  \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) {                                                               \
            if (ctx) (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);          \
            return NUC_EHIP;                                                                  \
        }                                                                                     \
    } while (0)

// bufferSize per scheme (src/fieldOperator.cpp:128-181)
static int scheme_halo(int scheme)
{
    if (scheme == NUC_WENO7JS || scheme == NUC_HCCS || scheme == NUC_HCCS78 || scheme == NUC_HCCS77) return 4;
    if (scheme == NUC_WENO3JS) return 2;
    if (scheme == NUC_UPWIND1ST) return 1;
    return 3;
}

// cells on each side a viscous scheme reads (cd2 1, cd6 3, cd8 4)
static int vis_reach(int vis) { return vis == NUC_CD8 ? 4 : (vis == NUC_CD2 ? 1 : 3); }

static int vis_id(int vis) { return vis == NUC_CD8 ? VIS_CD8 : (vis == NUC_CD2 ? VIS_CD2 : VIS_CD6); }

// The reference hands the viscous operators the inviscid halo as their
// tilesize (src/fieldOperator.cpp:520-526) and sizes the ghost strips by the
// inviscid halos (src/block.cpp:170-171): cd8 behind a 3-wide scheme and cd6
// behind weno3js / upwind1st read outside the strip.  Widened to the reach.
static int vis_tilesize(int scheme, int vis) { return std::max(scheme_halo(scheme), vis_reach(vis)); }

// Lines longer than 512 cells are swept in segments of ls cells, one wave
// each, with the neighbouring segments' cells as halo.  Explicit schemes: the
// largest divisor of n in [64, 255].  Compact schemes, whose tridiagonal
// systems end at rank-block edges (src/crweno5.cpp:83-206): whole rank blocks,
// k blocks of together at most 255 cells, or one block of at most 511.
// 0: no such segment.  Up to 512 cells a line is one segment.
static int segment_length(int n, int nloc, int np, int scheme)
{
    if (n <= 512) return n;
    if (compact_scheme(scheme)) {
        if (nloc > 511) return 0;
        if (nloc > 255) return nloc;
        int best = nloc;
        for (int k = 1; k <= np; k++)
            if (np % k == 0 && (long)k * nloc <= 255) best = k * nloc;
        return best;
    }
    for (int ls = 255; ls >= 64; ls--)
        if (n % ls == 0) return ls;
    return 0;
}

static int validate(const nuc_config *c, std::string &err)
{
    int n[3] = {c->nx, c->ny, c->nz}, np[3] = {c->npx, c->npy, c->npz};
    for (int d = 0; d < 3; d++) {
        if (np[d] < 1 || n[d] % np[d] != 0) { err = "grid size must divide evenly by the rank grid (src/block.cpp:56-58)"; return NUC_EINVAL; }
        if (c->scheme[d] < 0 || c->scheme[d] > NUC_HCCS77) { err = "unknown inviscid scheme id"; return NUC_EINVAL; }
        if (c->vis_scheme[d] < 0 || c->vis_scheme[d] > NUC_CD2) { err = "unknown viscous scheme id"; return NUC_EINVAL; }
        int ts = scheme_halo(c->scheme[d]), tsv = vis_tilesize(c->scheme[d], c->vis_scheme[d]);
        if (n[d] / np[d] < 2 * ts + 2 || n[d] / np[d] < 2 * tsv) {
            err = "rank-local line too short for the scheme's halo";
            return NUC_EINVAL;
        }
        if (segment_length(n[d], n[d] / np[d], np[d], c->scheme[d]) == 0) {
            err = compact_scheme(c->scheme[d])
                      ? "a compact scheme on a line above 512 cells needs rank blocks of at most 511 cells"
                      : "a line above 512 cells needs a divisor in [64, 255] to be swept in segments";
            return NUC_EINVAL;
        }
    }
    // the sweeps address a field with 32-bit byte offsets
    if ((long)c->nx * c->ny * c->nz > (1l << 29)) { err = "box above 2^29 cells (4 GiB per FP64 field)"; return NUC_EINVAL; }
    if ((long)c->npx * c->npy * c->npz > MAXRANK) { err = "at most 64 ranks"; return NUC_EINVAL; }
    if (c->riemann < 0 || c->riemann > NUC_AUSMP || c->model < 0 || c->model > 1) { err = "bad riemann/model id"; return NUC_EINVAL; }
    if (c->forcing < NUC_FORCING_NONE || c->forcing > NUC_FORCING_TG) { err = "unknown forcing id"; return NUC_EINVAL; }
    if (c->centrifugal && !(c->well_rad > 0)) { err = "centrifugal well radius must be positive"; return NUC_EINVAL; }
    if (c->precision != NUC_FP64 && c->precision != NUC_MIXED && c->precision != NUC_FP32) {
        err = "unknown precision id";
        return NUC_EINVAL;
    }
    if (c->precision == NUC_FP32) {
        for (int d = 0; d < 3; d++)
            if (c->scheme[d] == NUC_HCCS || c->scheme[d] == NUC_HCCS77) {
                err = "NUC_FP32 does not take hccs or hccs77 (their theta blend needs FP64 stencils)";
                return NUC_EINVAL;
            }
        if (c->forcing != NUC_FORCING_NONE || c->coriolis || c->centrifugal) {
            err = "NUC_FP32 does not take forcing or rotating-frame sources";
            return NUC_EINVAL;
        }
    }
    if (!(c->lx > 0 && c->ly > 0 && c->lz > 0)) { err = "box lengths must be positive"; return NUC_EINVAL; }
    return NUC_OK;
}

// whether the viscous fluxes are stored FP32: a mixed precision with FP32
// stencils in some direction (hccs and hccs77 keep FP64 stencils)
static bool fv_fp32(const nuc_config *c)
{
    if (c->precision == NUC_FP64) return false;
    for (int d = 0; d < 3; d++)
        if (c->scheme[d] != NUC_HCCS && c->scheme[d] != NUC_HCCS77) return true;
    return false;
}

extern "C" long long nuc_footprint(const nuc_config *c)
{
    long long N = (long long)c->nx * c->ny * c->nz;
    // 3 states + rhs (FP32 with NUC_FP32, and its FP64 view); 9 viscous flux
    // fields (Navier-Stokes; FP32 when some direction's stencils are); u,v,w,T (forcing)
    const long long es = c->precision == NUC_FP32 ? 4 : 8, fes = fv_fp32(c) ? 4 : 8;
    long long bytes = 20 * N * es + (c->model == NUC_NAVIERSTOKES3D ? 9 * N * fes : 0);
    if (c->precision == NUC_FP32) bytes += 5 * N * 8;
    if (c->forcing == NUC_FORCING_CONST) bytes += 4 * N * 8;
    long long fb = (c->forcing == NUC_FORCING_CONST) ? nucforce::plan_bytes(*c) : 0;
    return bytes + (N / 32 + 4096) * 8 + 4096 + fb;
}

// message of the last failed call without a context on this thread
// (nuc_create, nuc_rhs_noalloc), returned by nuc_last_error(NULL)
static thread_local std::string g_create_err;

// configuration, grid geometry and precision flags of a context (no device work)
static void ctx_geometry(nuc_ctx *x, const nuc_config *cfg)
{
    x->c = *cfg;
    x->n[0] = cfg->nx; x->n[1] = cfg->ny; x->n[2] = cfg->nz;
    x->np[0] = cfg->npx; x->np[1] = cfg->npy; x->np[2] = cfg->npz;
    double Lb[3] = {cfg->lx, cfg->ly, cfg->lz};
    x->N = (long)cfg->nx * cfg->ny * cfg->nz;
    x->P = cfg->npx * cfg->npy * cfg->npz;
    for (int d = 0; d < 3; d++) {
        x->nloc[d] = x->n[d] / x->np[d];
        x->h[d] = Lb[d] / x->n[d];
        x->ts[d] = scheme_halo(cfg->scheme[d]);
        x->tsv[d] = vis_tilesize(cfg->scheme[d], cfg->vis_scheme[d]);
        x->seg[d] = segment_length(x->n[d], x->nloc[d], x->np[d], cfg->scheme[d]);
        x->nseg[d] = x->n[d] / x->seg[d];
        x->cpt[d] = x->nseg[d] > 1 ? (x->seg[d] <= 255 ? 8 : 16) : sweep_rows(x->n[d]);
    }
}

extern "C" int nuc_create(const nuc_config *cfg, nuc_ctx **out)
{
    g_create_err.clear();
    if (!cfg || !out) { g_create_err = "nuc_create: null config or output pointer"; return NUC_EINVAL; }
    *out = nullptr;
    std::string err;
    int rc = validate(cfg, err);
    if (rc != NUC_OK) {
        g_create_err = "nuc_create: " + err;
        return rc;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        g_create_err = "nuc_create: no usable gfx950 device";
        return NUC_ENODEV;
    }
    nuc_ctx *x = new nuc_ctx();
    ctx_geometry(x, cfg);
    const size_t N = (size_t)x->N;
    auto fail = [&](const char *what) {
        g_create_err = std::string("nuc_create: ") + what;
        nuc_destroy(x);
        return NUC_ENOMEM;
    };
    // a non-blocking stream and stream-ordered allocations: contexts created
    // and destroyed while others run (the serving loop) neither wait for nor
    // stall the other contexts' streams
    if (hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking) != hipSuccess) {
        g_create_err = "nuc_create: hipStreamCreateWithFlags failed";
        delete x;
        return NUC_EHIP;
    }
    auto dalloc = [&](auto **ptr, size_t bytes) {
        return hipMallocAsync(reinterpret_cast<void **>(ptr), bytes, x->stream) == hipSuccess;
    };
    x->forcing = (cfg->forcing == NUC_FORCING_CONST);
    x->sources = cfg->forcing == NUC_FORCING_ABC || cfg->forcing == NUC_FORCING_TG || cfg->coriolis || cfg->centrifugal;
    x->mixed = (cfg->precision != NUC_FP64);
    x->fp32 = (cfg->precision == NUC_FP32);
    x->fv32 = fv_fp32(cfg);
    // element bytes of the state/rhs fields and of the viscous-flux fields
    const size_t es = x->fp32 ? sizeof(float) : sizeof(double), fes = x->fv32 ? sizeof(float) : sizeof(double);
    for (int b = 0; b < 3; b++)
        if (!dalloc(&x->q[b], 5 * N * es)) return fail("device allocation (state)");
    if (!dalloc(&x->rhs, 5 * N * es)) return fail("device allocation (rhs)");
    if (x->fp32)
        if (!dalloc(&x->q64, 5 * N * sizeof(double))) return fail("device allocation (FP64 view)");
    if (x->forcing)
        if (!dalloc(&x->prim, 4 * N * sizeof(double))) return fail("device allocation (prim)");
    if (cfg->model == NUC_NAVIERSTOKES3D)
        if (!dalloc(&x->fv, 9 * N * fes)) return fail("device allocation (fv)");
    if (x->forcing) {
        std::string ferr;
        int frc = nucforce::plan_create(x->fplan, *cfg, x->stream, ferr);
        if (frc != NUC_OK) {
            g_create_err = "nuc_create: " + ferr;
            x->forcing = false;
            nuc_destroy(x);
            return frc;
        }
    }
    if (!dalloc(&x->lam, 3 * MAXRANK * sizeof(unsigned long long))) return fail("device allocation (lam)");
    if (!dalloc(&x->scal, 8 * sizeof(double))) return fail("device allocation (scal)");
    // residual partials: one per workgroup of the last sweep
    const int dl = kSweepOrder[2];
    long lines_last = x->N / x->n[dl] * x->nseg[dl];
    x->res_nb = (lines_last + sweep_lpb(x->cpt[dl]) - 1) / sweep_lpb(x->cpt[dl]);
    if (!dalloc(&x->res_part, x->res_nb * sizeof(double))) return fail("device allocation (residual)");
    (void)hipMemsetAsync(x->lam, 0, 3 * MAXRANK * sizeof(unsigned long long), x->stream);
    (void)hipMemsetAsync(x->scal, 0, 8 * sizeof(double), x->stream);
    if (hipStreamSynchronize(x->stream) != hipSuccess) return fail("device init");
    *out = x;
    return NUC_OK;
}

extern "C" int nuc_destroy(nuc_ctx *x)
{
    if (!x) return NUC_OK;
    if (!x->stream) { delete x; return NUC_OK; }
    (void)hipStreamSynchronize(x->stream);
    if (x->gexec) (void)hipGraphExecDestroy(x->gexec);
    if (x->graph) (void)hipGraphDestroy(x->graph);
    void *bufs[] = {x->q[0], x->q[1], x->q[2], x->rhs, x->prim, x->fv, x->lam, x->scal, x->res_part,
                    x->stat_part, x->stat_out, x->post_buf, x->q64};
    for (void *b : bufs)
        if (b) (void)hipFreeAsync(b, x->stream);
    if (x->forcing) nucforce::plan_destroy(x->fplan, x->stream);
    (void)hipStreamSynchronize(x->stream);
    (void)hipStreamDestroy(x->stream);
    delete x;
    return NUC_OK;
}

extern "C" const char *nuc_last_error(const nuc_ctx *x) { return x ? x->err.c_str() : g_create_err.c_str(); }

// NUC_FP32 conversions between an FP64 field set and the FP32 state
__global__ __launch_bounds__(256) void k_narrow(const double *a, float *b, long n)
{
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) b[i] = float(a[i]);
}
__global__ __launch_bounds__(256) void k_widen(const float *a, double *b, long n)
{
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) b[i] = double(a[i]);
}
static dim3 conv_grid(long n) { return dim3((unsigned)std::min<long>((n + 255) / 256, 65536)); }

// field v of RK slot b and of the RHS in the kernels' state type (FP32 with
// NUC_FP32: the buffers then hold 5 FP32 fields)
static const void *slot_field(const nuc_ctx *x, int b, int v)
{
    return x->fp32 ? (const void *)(reinterpret_cast<const float *>(x->q[b]) + (size_t)v * x->N)
                   : (const void *)(x->q[b] + (size_t)v * x->N);
}
static void *slot_field_w(nuc_ctx *x, int b, int v) { return const_cast<void *>(slot_field(x, b, v)); }
static void *rhs_field(nuc_ctx *x, int v)
{
    return x->fp32 ? (void *)(reinterpret_cast<float *>(x->rhs) + (size_t)v * x->N) : (void *)(x->rhs + (size_t)v * x->N);
}
// viscous-flux component o (FP32 fields, k_gradq<float>, when the stencils are FP32)
static void *fv_field(const nuc_ctx *x, int o)
{
    return x->fv32 ? (void *)(reinterpret_cast<float *>(x->fv) + (size_t)o * x->N) : (void *)(x->fv + (size_t)o * x->N);
}

// The current state in FP64 for the statistics, post-processing and
// downloads: the state itself, or with NUC_FP32 its widened copy (enqueued)
static const double *state64(nuc_ctx *x)
{
    if (!x->fp32) return x->q[x->next_stage];
    hipLaunchKernelGGL(k_widen, conv_grid(5 * x->N), dim3(256), 0, x->stream,
                       reinterpret_cast<const float *>(x->q[x->next_stage]), x->q64, 5 * x->N);
    return x->q64;
}
// Where the initial-condition generators write a new FP64 state (RK slot 0,
// or with NUC_FP32 the FP64 view), and its narrowing into slot 0 afterwards
static double *state64_w(nuc_ctx *x) { return x->fp32 ? x->q64 : x->q[0]; }
static void state64_commit(nuc_ctx *x)
{
    if (x->fp32)
        hipLaunchKernelGGL(k_narrow, conv_grid(5 * x->N), dim3(256), 0, x->stream, x->q64,
                           reinterpret_cast<float *>(x->q[0]), 5 * x->N);
}

extern "C" int nuc_upload_state(nuc_ctx *x, const double *q)
{
    if (!x || !q) return NUC_EINVAL;
    x->next_stage = 0;   // an upload starts a new step
    NUC_CHECK(x, hipMemcpyAsync(state64_w(x), q, 5 * x->N * sizeof(double), hipMemcpyHostToDevice, x->stream));
    state64_commit(x);
    NUC_CHECK(x, hipStreamSynchronize(x->stream));
    return NUC_OK;
}

extern "C" int nuc_download_state(nuc_ctx *x, double *q)
{
    if (!x || !q) return NUC_EINVAL;
    NUC_CHECK(x, hipMemcpyAsync(q, state64(x), 5 * x->N * sizeof(double), hipMemcpyDeviceToHost, x->stream));
    NUC_CHECK(x, hipStreamSynchronize(x->stream));
    return NUC_OK;
}

// Arguments of every launch of one RHS evaluation of state buffer `in`; with
// epilogue the last sweep performs RK stage `stage` from `in` (and q[0] = u^n)
// into `outb`.  The single-context launches and the packed launches are built
// from the same arguments.
struct StageArgs {
    PrimArgs prim{};
    long prim_nb = 0;
    bool do_prim = false;
    DtArgs dt{};
    bool do_dt = false;
    GradArgs grad{};
    dim3 ggrid;
    bool do_grad = false;
    SweepArgs sw[3]{};
    long sw_nb[3] = {0, 0, 0};
    ResArgs res{};
    bool do_res = false;
};

static void stage_args(nuc_ctx *x, int in, int stage, int outb, bool epilogue, StageArgs &sa)
{
    const nuc_config &c = x->c;
    const long N = x->N;
    const bool ns = (c.model == NUC_NAVIERSTOKES3D);
    const bool lam_needed = (stage == 0);
    {
        PrimArgs &P = sa.prim;
        for (int v = 0; v < 5; v++) P.q[v] = gp(slot_field(x, in, v));
        const bool wp = x->forcing;   // the gradient kernel reads Q itself
        if (wp) for (int f = 0; f < 4; f++) P.prim[f] = gp(x->prim + f * N);
        P.N = N; P.nx = x->n[0]; P.ny = x->n[1];
        P.nlx = x->nloc[0]; P.nly = x->nloc[1]; P.nlz = x->nloc[2];
        P.npx = x->np[0]; P.npy = x->np[1]; P.P = x->P;
        P.riemann = c.riemann; P.do_lam = lam_needed ? 1 : 0; P.write_prim = wp ? 1 : 0;
        P.gamma = c.gamma; P.mach = c.mach; P.lam = gp(x->lam);
        sa.do_prim = P.do_lam || P.write_prim;
        sa.prim_nb = std::min<long>((N + 4095) / 4096, 65536);
        sa.do_dt = lam_needed;
        sa.dt = DtArgs{gp(x->lam), x->P, epilogue ? 1 : 0, x->h[0], x->h[1], x->h[2], c.cfl, gp(x->scal)};
    }
    if (ns) {
        GradArgs &G = sa.grad;
        for (int v = 0; v < 5; v++) G.q[v] = gp(slot_field(x, in, v));
        for (int o = 0; o < 9; o++) G.out[o] = gp(fv_field(x, o));
        G.nx = x->n[0]; G.ny = x->n[1]; G.nz = x->n[2];
        for (int d = 0; d < 3; d++) {
            G.nloc[d] = x->nloc[d]; G.tsv[d] = x->tsv[d];
            G.vis[d] = vis_id(c.vis_scheme[d]);
            G.inv_h[d] = 1.0 / x->h[d];
        }
        G.law = c.viscosity; G.re = c.reynolds; G.mach = c.mach; G.pr = c.prandtl; G.gamma = c.gamma;
        G.tref = c.t_ref; G.tinf = c.t_inf;
        G.gkz = x->gkz > 0 ? x->gkz : grad_planes(x->n[2]);
        sa.ggrid = dim3((x->n[0] + GTX - 1) / GTX, (x->n[1] + GTY - 1) / GTY, (x->n[2] + G.gkz - 1) / G.gkz);
        sa.do_grad = true;
    }
    // fv component order per direction: x (txx,txy,txz,ex) y (txy,tyy,tyz,ey) z (txz,tyz,tzz,ez)
    static const int fvmap[3][4] = {{0, 3, 4, 6}, {3, 1, 5, 7}, {4, 5, 2, 8}};
    for (int d = 0; d < 3; d++) {
        SweepArgs &S = sa.sw[d];
        S.dir = d; S.n = x->seg[d]; S.nloc = x->nloc[d]; S.ts = x->ts[d]; S.tsv = x->tsv[d];
        S.nseg = x->nseg[d]; S.nfull = x->n[d];
        S.vis = ns ? vis_id(c.vis_scheme[d]) : VIS_NONE;
        S.riemann = c.riemann;
        S.nx = x->n[0]; S.ny = x->n[1]; S.nz = x->n[2];
        S.nlines = N / x->n[d];
        S.inv_h = 1.0 / x->h[d];
        S.gamma = c.gamma;
        for (int v = 0; v < 5; v++) { S.q[v] = gp(slot_field(x, in, v)); S.rhs[v] = gp(rhs_field(x, v)); }
        if (ns) for (int o = 0; o < 4; o++) S.fv[o] = gp(fv_field(x, fvmap[d][o]));
        S.stage = stage;
        const bool first = (d == kSweepOrder[0]), last = (d == kSweepOrder[2]);
        S.first = first ? 1 : 0;
        S.accum = (first && (x->forcing || x->sources)) ? 1 : 0;
        S.mixed = x->mixed ? 1 : 0;
        S.fp32 = x->fp32 ? 1 : 0;
        S.fv_f32 = x->fv32 ? 1 : 0;   // k_gradq<float>
        S.fmean = (last && x->forcing) ? gp(x->scal + 4) : nullptr;
        S.epilogue = (last && epilogue) ? 1 : 0;
        if (S.epilogue) {
            for (int v = 0; v < 5; v++) { S.qn[v] = gp(slot_field(x, 0, v)); S.qout[v] = gp(slot_field_w(x, outb, v)); }
            S.dtp = gp(x->scal);
            S.res_part = gp(x->res_part);
        }
        sa.sw_nb[d] = (S.nlines * S.nseg + SWEEP_LPB - 1) / SWEEP_LPB;
    }
    sa.do_res = epilogue && stage == 2;
    sa.res = ResArgs{gp(x->res_part), (int)x->res_nb, 1.0 / ((double)N * 5.0), gp(x->scal)};
}

// AddForcing_ABC / AddForcing_TG (src/physicsModel.cpp:1679-1799),
// AddCoriolisForce (:1801-1840) and AddCentrifugalForce (:1842-1913, whose
// energy term goes to the y-momentum RHS as written), in solveSource's order
// (:1280-1319), at the per-rank cell centres (src/block.cpp:116-118)
struct SourceArgs {
    const double GMEM *q;
    double GMEM *rhs;
    long N;
    int nx, ny, nl[3], forcing, coriolis, centrifugal, accumulate;
    double dx, dy, dz, lxr, lyr, lzr, A, kf, Ro, xc, yc, r0;
    double ox, oy, oz;   // block.in origin, subtracted from the centres
};

__global__ __launch_bounds__(256) void k_source(SourceArgs A)
{
    const long id = (long)blockIdx.x * 256 + threadIdx.x;
    if (id >= A.N) return;
    const int i = (int)(id % A.nx);
    const long rr = id / A.nx;
    const int j = (int)(rr % A.ny), k = (int)(rr / A.ny);
    const double x0 = A.lxr * (i / A.nl[0]) + A.dx * (0.5 + i % A.nl[0]) - A.ox;
    const double y0 = A.lyr * (j / A.nl[1]) + A.dy * (0.5 + j % A.nl[1]) - A.oy;
    const double z0 = A.lzr * (k / A.nl[2]) + A.dz * (0.5 + k % A.nl[2]) - A.oz;
    const double rho = A.q[id];
    const double u0 = A.q[A.N + id] / rho, v0 = A.q[2 * A.N + id] / rho, w0 = A.q[3 * A.N + id] / rho;
    double r1 = 0.0, r2 = 0.0, r3 = 0.0, r4 = 0.0;
    if (A.accumulate) { r1 = A.rhs[A.N + id]; r2 = A.rhs[2 * A.N + id]; r3 = A.rhs[3 * A.N + id]; r4 = A.rhs[4 * A.N + id]; }
    if (A.forcing == NUC_FORCING_ABC) {
        const double fx = A.A * 1.0 * (cos(A.kf * y0) + 1.1 * sin(A.kf * z0));
        const double fy = A.A * 1.0 * (1.1 * cos(A.kf * z0) + 0.9 * sin(A.kf * x0));
        const double fz = A.A * 1.0 * (0.9 * cos(A.kf * x0) + sin(A.kf * y0));
        r1 -= fx; r2 -= fy; r3 -= fz;
        r4 -= fx * u0 + fy * v0 + fz * w0;
    } else if (A.forcing == NUC_FORCING_TG) {
        const double fx = A.A * 2.0 * (sin(A.kf * x0) * cos(A.kf * y0) * sin(A.kf * z0));
        const double fy = -A.A * 2.0 * (cos(A.kf * x0) * sin(A.kf * y0) * sin(A.kf * z0));
        r1 -= fx; r2 -= fy;
        r4 -= (fx * u0 + fy * v0);
    }
    if (A.coriolis) {
        r1 -= rho * A.Ro * v0;
        r2 -= -rho * A.Ro * u0;
    }
    if (A.centrifugal) {
        const double r = sqrt((x0 - A.xc) * (x0 - A.xc) + (y0 - A.yc) * (y0 - A.yc));
        const double fx = rho * A.Ro * A.Ro * (x0 - A.xc) / (1.0 + pow(r / A.r0, 20.0)) / 4.0;
        const double fy = rho * A.Ro * A.Ro * (y0 - A.yc) / (1.0 + pow(r / A.r0, 20.0)) / 4.0;
        r1 -= fx;
        r2 -= fy;
        r2 -= fx * u0 + fy * v0;
    }
    if (!A.accumulate) A.rhs[id] = 0.0;
    A.rhs[A.N + id] = r1; A.rhs[2 * A.N + id] = r2; A.rhs[3 * A.N + id] = r3; A.rhs[4 * A.N + id] = r4;
}

static SourceArgs source_args(nuc_ctx *x, int in, bool accumulate)
{
    const nuc_config &c = x->c;
    SourceArgs A{};
    A.q = gp((const double *)x->q[in]);
    A.rhs = gp(x->rhs);
    A.N = x->N;
    A.nx = c.nx; A.ny = c.ny;
    for (int d = 0; d < 3; d++) A.nl[d] = x->nloc[d];
    A.forcing = c.forcing; A.coriolis = c.coriolis; A.centrifugal = c.centrifugal;
    A.accumulate = accumulate ? 1 : 0;
    A.dx = c.lx / c.nx; A.dy = c.ly / c.ny; A.dz = c.lz / c.nz;
    A.lxr = c.lx / c.npx; A.lyr = c.ly / c.npy; A.lzr = c.lz / c.npz;
    A.A = c.forcing_amp; A.kf = c.forcing_k;
    A.Ro = c.rossby; A.xc = c.x_center; A.yc = c.y_center; A.r0 = c.well_rad;
    A.ox = c.origin[0]; A.oy = c.origin[1]; A.oz = c.origin[2];
    return A;
}

// the gradient kernel for the context's precision (flux type, state type)
static void launch_gradq(nuc_ctx *x, dim3 grid, const GradArgs &G)
{
    if (x->fp32) hipLaunchKernelGGL((k_gradq<float, float>), grid, dim3(GTX, GTY, 1), 0, x->stream, G);
    else if (x->fv32) hipLaunchKernelGGL((k_gradq<float, double>), grid, dim3(GTX, GTY, 1), 0, x->stream, G);
    else hipLaunchKernelGGL((k_gradq<double, double>), grid, dim3(GTX, GTY, 1), 0, x->stream, G);
}

static int enqueue_stage(nuc_ctx *x, int in, int stage, int outb, bool epilogue)
{
    StageArgs sa;
    stage_args(x, in, stage, outb, epilogue, sa);
    const long N = x->N;
    if (sa.do_prim) {
        ProfScope ps(x, 0);
        if (x->fp32) hipLaunchKernelGGL(k_prim<float>, dim3((unsigned)sa.prim_nb), dim3(256), 0, x->stream, sa.prim);
        else hipLaunchKernelGGL(k_prim<double>, dim3((unsigned)sa.prim_nb), dim3(256), 0, x->stream, sa.prim);
    }
    if (sa.do_dt) {
        ProfScope ps(x, 1);
        hipLaunchKernelGGL(k_dt, dim3(1), dim3(64), 0, x->stream, sa.dt);
    }
    if (sa.do_grad) {
        ProfScope ps(x, 2);
        launch_gradq(x, sa.ggrid, sa.grad);
    }
    if (x->forcing) {
        // Const forcing (src/physicsModel.cpp:1368-1677) initialises rhs; the
        // first sweep accumulates onto it and the last adds <f.u> back
        double *r[5];
        for (int v = 0; v < 5; v++) r[v] = x->rhs + v * N;
        ProfScope ps(x, 7);
        nucforce::enqueue(x->fplan, x->prim, x->prim + N, x->prim + 2 * N, r, x->scal + 4, x->stream);
    }
    if (x->sources) {
        // ABC/TG forcing and the rotating-frame sources initialise rhs (or add
        // to the Const forcing's initialisation); the first sweep accumulates onto it
        SourceArgs SA = source_args(x, in, x->forcing);
        ProfScope ps(x, 7);
        hipLaunchKernelGGL(k_source, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, x->stream, SA);
    }
    for (int i = 0; i < 3; i++) {
        const int d = kSweepOrder[i];
        ProfScope ps(x, 3 + d);
        launch_sweep(x->c.scheme[d], x->cpt[d], sa.sw[d], x->stream);
    }
    if (sa.do_res) {
        ProfScope ps(x, 6);
        hipLaunchKernelGGL(k_res, dim3(1), dim3(256), 0, x->stream, sa.res);
    }
    return NUC_OK;
}

static void enqueue_one_step(nuc_ctx *x)
{
    enqueue_stage(x, 0, 0, 1, true);
    enqueue_stage(x, 1, 1, 2, true);
    enqueue_stage(x, 2, 2, 0, true);
}

static int enqueue_steps(nuc_ctx *x, int nsteps)
{
    if (x->use_graph && !x->prof) {
        // the step's launches depend only on the context's buffers and
        // configuration (dt, residual and time live on the device), so one
        // capture serves every later step; it holds kernel nodes only
        if (!x->gexec) {
            NUC_CHECK(x, hipStreamBeginCapture(x->stream, hipStreamCaptureModeRelaxed));
            enqueue_one_step(x);
            hipError_t e = hipStreamEndCapture(x->stream, &x->graph);
            if (e != hipSuccess) { x->err = std::string("graph capture: ") + hipGetErrorString(e); return NUC_EHIP; }
            NUC_CHECK(x, hipGraphInstantiate(&x->gexec, x->graph, nullptr, nullptr, 0));
        }
        for (int it = 0; it < nsteps; it++) NUC_CHECK(x, hipGraphLaunch(x->gexec, x->stream));
        return NUC_OK;
    }
    for (int it = 0; it < nsteps; it++) enqueue_one_step(x);
    NUC_CHECK(x, hipGetLastError());
    return NUC_OK;
}

// ---- autotuning of tile shapes ----------------------------------------------
// The gradient kernel's z march (planes per workgroup) trades the 8-plane halo
// re-read against the number of workgroups; its best value depends on the box
// shape.  nuc_autotune times every candidate on the context's own state and
// keeps the fastest; results are shared by all contexts of the same shape and
// precision in the process.  The tile shape does not change any result.
static std::mutex g_tune_mu;
static std::map<std::tuple<int, int, int, int>, int> g_tune_gkz;

extern "C" int nuc_autotune(nuc_ctx *x)
{
    if (!x) return NUC_EINVAL;
    if (x->c.model != NUC_NAVIERSTOKES3D) return NUC_OK;   // no gradient kernel
    const auto key = std::make_tuple(x->n[0], x->n[1], x->n[2], (int)x->c.precision * 2 + (x->fv32 ? 1 : 0));
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_gkz.find(key);
        if (it != g_tune_gkz.end()) {
            x->gkz = it->second;
            if (x->gexec) { (void)hipGraphExecDestroy(x->gexec); x->gexec = nullptr; }
            if (x->graph) { (void)hipGraphDestroy(x->graph); x->graph = nullptr; }
            return NUC_OK;
        }
    }
    StageArgs sa;
    stage_args(x, 0, 0, 0, false, sa);
    hipEvent_t e0, e1;
    NUC_CHECK(x, hipEventCreate(&e0));
    NUC_CHECK(x, hipEventCreate(&e1));
    const int cand[4] = {4, 8, 16, 32};
    int best = grad_planes(x->n[2]);
    float best_ms = 1e30f;
    for (int gkz : cand) {
        if (gkz > x->n[2]) continue;
        GradArgs G = sa.grad;
        G.gkz = gkz;
        const dim3 grid(sa.ggrid.x, sa.ggrid.y, (x->n[2] + gkz - 1) / gkz);
        float tot = 0.f;
        for (int rep = 0; rep < 4; rep++) {
            (void)hipEventRecord(e0, x->stream);
            launch_gradq(x, grid, G);
            (void)hipEventRecord(e1, x->stream);
            NUC_CHECK(x, hipEventSynchronize(e1));
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (rep > 0) tot += ms;   // the first launch warms caches
        }
        if (tot < best_ms) { best_ms = tot; best = gkz; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    NUC_CHECK(x, hipGetLastError());
    x->gkz = best;
    // a captured step holds the old launch shape
    if (x->gexec) { (void)hipGraphExecDestroy(x->gexec); x->gexec = nullptr; }
    if (x->graph) { (void)hipGraphDestroy(x->graph); x->graph = nullptr; }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_gkz[key] = best;
    return NUC_OK;
}

extern "C" int nuc_set_tile_shape(nuc_ctx *x, int planes)
{
    if (!x || planes < 1 || planes > x->n[2]) return NUC_EINVAL;
    x->gkz = planes;
    if (x->gexec) { (void)hipGraphExecDestroy(x->gexec); x->gexec = nullptr; }
    if (x->graph) { (void)hipGraphDestroy(x->graph); x->graph = nullptr; }
    return NUC_OK;
}

extern "C" int nuc_tile_shape(const nuc_ctx *x, int *grad_planes_out)
{
    if (!x || !grad_planes_out) return NUC_EINVAL;
    *grad_planes_out = x->gkz > 0 ? x->gkz : grad_planes(x->n[2]);
    return NUC_OK;
}

extern "C" int nuc_set_graph(nuc_ctx *x, int enable)
{
    if (!x) return NUC_EINVAL;
    x->use_graph = enable != 0;
    return NUC_OK;
}

extern "C" int nuc_query(nuc_ctx *x)
{
    if (!x) return NUC_EINVAL;
    hipError_t e = hipStreamQuery(x->stream);
    if (e == hipSuccess) return NUC_OK;
    if (e == hipErrorNotReady) return NUC_EBUSY;
    x->err = std::string("hipStreamQuery: ") + hipGetErrorString(e);
    return NUC_EHIP;
}

static bool mid_step(nuc_ctx *x)
{
    if (x->next_stage == 0) return false;
    x->err = "a step is in progress: run its remaining stages with nuc_rk3_stage";
    return true;
}

extern "C" int nuc_step_async(nuc_ctx *x, int nsteps)
{
    if (!x || nsteps < 0) return NUC_EINVAL;
    if (mid_step(x)) return NUC_EINVAL;
    int rc = enqueue_steps(x, nsteps);
    if (rc == NUC_OK) x->pending += nsteps;
    return rc;
}

extern "C" int nuc_sync(nuc_ctx *x, nuc_step_info *info)
{
    if (!x) return NUC_EINVAL;
    NUC_CHECK(x, hipStreamSynchronize(x->stream));
    for (auto &e : x->ev) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, e.second.first, e.second.second) == hipSuccess) {
            x->prof_ms[e.first] += ms;
            x->prof_n[e.first] += 1;
        }
        (void)hipEventDestroy(e.second.first);
        (void)hipEventDestroy(e.second.second);
    }
    x->ev.clear();
    double sc[8];
    NUC_CHECK(x, hipMemcpyAsync(sc, x->scal, sizeof(sc), hipMemcpyDeviceToHost, x->stream));
    NUC_CHECK(x, hipStreamSynchronize(x->stream));
    x->time = sc[2];
    x->step += x->pending;
    x->pending = 0;
    if (info) { info->dt = sc[0]; info->res = sc[1]; info->time = x->time; info->step = x->step; }
    if (!std::isfinite(sc[0]) || !std::isfinite(sc[1])) {
        x->err = "non-finite dt or residual";
        return NUC_ENONFINITE;
    }
    return NUC_OK;
}

extern "C" int nuc_step(nuc_ctx *x, int nsteps, nuc_step_info *info)
{
    if (!x || nsteps < 0) return NUC_EINVAL;
    if (nsteps > 0 && mid_step(x)) return NUC_EINVAL;
    for (int it = 0; it < nsteps; it++) {
        int rc = enqueue_steps(x, 1);
        if (rc != NUC_OK) return rc;
        x->pending = 1;
        rc = nuc_sync(x, info);
        if (rc != NUC_OK) return rc;
    }
    if (nsteps == 0 && info) return nuc_sync(x, info);
    return NUC_OK;
}

// one RK stage of block::solve (src/block.cpp:198-202): the RHS of the current
// stage state (NaiverStokesData3d::solve) and PDEData3d::solve's update
// (src/PDEData3d.cpp:70-91) into the next stage state
extern "C" int nuc_rk3_stage(nuc_ctx *x, int stage)
{
    if (!x) return NUC_EINVAL;
    if (stage != x->next_stage) {
        x->err = "RK stages run in order 0, 1, 2 (src/block.cpp:198-202); next is stage " +
                 std::to_string(x->next_stage);
        return NUC_EINVAL;
    }
    enqueue_stage(x, stage, stage, (stage + 1) % 3, true);
    NUC_CHECK(x, hipGetLastError());
    x->next_stage = (stage + 1) % 3;
    if (stage < 2) {
        NUC_CHECK(x, hipStreamSynchronize(x->stream));
        return NUC_OK;
    }
    x->pending += 1;   // the step is complete: time, step count, dt and residual
    return nuc_sync(x, nullptr);
}

extern "C" int nuc_rhs(nuc_ctx *x, double *rhs_host, double *dt_local)
{
    if (!x) return NUC_EINVAL;
    enqueue_stage(x, x->next_stage, 0, 0, false);
    NUC_CHECK(x, hipGetLastError());
    NUC_CHECK(x, hipStreamSynchronize(x->stream));
    if (rhs_host) {
        const double *r = x->rhs;
        if (x->fp32) {   // widened through the FP64 view (the state itself is not touched)
            hipLaunchKernelGGL(k_widen, conv_grid(5 * x->N), dim3(256), 0, x->stream,
                               reinterpret_cast<const float *>(x->rhs), x->q64, 5 * x->N);
            r = x->q64;
        }
        NUC_CHECK(x, hipMemcpyAsync(rhs_host, r, 5 * x->N * sizeof(double), hipMemcpyDeviceToHost, x->stream));
    }
    double sc[8];
    NUC_CHECK(x, hipMemcpyAsync(sc, x->scal, sizeof(sc), hipMemcpyDeviceToHost, x->stream));
    NUC_CHECK(x, hipStreamSynchronize(x->stream));
    if (dt_local) *dt_local = sc[3];
    return NUC_OK;
}

// ---- nuc_rhs_noalloc: one RHS evaluation without any device allocation ----
// The intermediates of an RHS evaluation that nuc_rhs keeps in its context
// (the 9 viscous flux fields, the per-rank eigenvalue maxima, the scalar
// slots) live in one fixed array of this code object instead.  Every GPU the
// library runs on holds its own copy until the process ends; its address is
// looked up once per device.  Calls that overlap on one GPU share the array
// and race.
constexpr long NOALLOC_MAX_CELLS = NUC_NOALLOC_MAX_CELLS;
constexpr long NOALLOC_FV = 9 * NOALLOC_MAX_CELLS;              // doubles (float in the mixed path)
constexpr long NOALLOC_LAM = 3 * MAXRANK;                       // uint64 slots
constexpr long NOALLOC_DOUBLES = NOALLOC_FV + NOALLOC_LAM + 8;  // + 8 scalar slots
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

static std::mutex g_scratch_mu;
static std::map<int, double *> g_scratch_ptr;   // device ordinal -> device address of scratch

__global__ void k_clear_u64(unsigned long long *p, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0ull;
}

__global__ void k_copy_dt(const double *scal, double *dt_out) { *dt_out = scal[3]; }

extern "C" int nuc_rhs_noalloc(const nuc_config *cfg, const double *q_dev, double *rhs_dev, double *dt_local_dev,
                               void *stream)
{
    g_create_err.clear();
    auto fail = [](int rc, const std::string &m) {
        g_create_err = "nuc_rhs_noalloc: " + m;
        return rc;
    };
    if (!cfg || !q_dev || !rhs_dev) return fail(NUC_EINVAL, "null config, state or rhs pointer");
    std::string err;
    int rc = validate(cfg, err);
    if (rc != NUC_OK) return fail(rc, err);
    const long N = (long)cfg->nx * cfg->ny * cfg->nz;
    if (N > NOALLOC_MAX_CELLS)
        return fail(NUC_EINVAL, std::to_string(N) + " cells exceed the scratch capacity of " +
                                    std::to_string(NOALLOC_MAX_CELLS) + " cells (NUC_NOALLOC_MAX_CELLS)");
    if (cfg->forcing == NUC_FORCING_CONST)
        return fail(NUC_EINVAL, "Const forcing needs its transform plan's buffers: use a context");
    if (cfg->precision == NUC_FP32)
        return fail(NUC_EINVAL, "the state and rhs buffers are FP64: NUC_FP32 needs a context");
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return fail(NUC_ENODEV, "no current device");
    double *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        auto it = g_scratch_ptr.find(dev);
        if (it == g_scratch_ptr.end()) {
            // set-up, once per device with that device current: the host-side
            // name of a __device__ array is not a device address
            void *p = nullptr;
            hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p)
                return fail(NUC_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            it = g_scratch_ptr.emplace(dev, static_cast<double *>(p)).first;
        }
        base = it->second;
    }
    // a context on the host stack that borrows the caller's buffers and the
    // scratch array: the same launches as nuc_rhs, no device allocation
    nuc_ctx x;
    ctx_geometry(&x, cfg);
    x.stream = static_cast<hipStream_t>(stream);
    x.sources = cfg->forcing == NUC_FORCING_ABC || cfg->forcing == NUC_FORCING_TG || cfg->coriolis || cfg->centrifugal;
    x.mixed = (cfg->precision == NUC_MIXED);
    x.fv32 = fv_fp32(cfg);
    x.q[0] = const_cast<double *>(q_dev);
    x.rhs = rhs_dev;
    x.fv = base;
    x.lam = reinterpret_cast<unsigned long long *>(base + NOALLOC_FV);
    x.scal = base + NOALLOC_FV + NOALLOC_LAM;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_gkz.find(std::make_tuple(x.n[0], x.n[1], x.n[2], (int)cfg->precision * 2 + (x.fv32 ? 1 : 0)));
        if (it != g_tune_gkz.end()) x.gkz = it->second;
    }
    // the array keeps the last call's contents: the eigenvalue maxima start
    // from zero (a fill kernel, so that the call can be captured in a graph)
    hipLaunchKernelGGL(k_clear_u64, dim3(1), dim3(256), 0, x.stream, x.lam, (int)NOALLOC_LAM);
    enqueue_stage(&x, 0, 0, 0, false);
    if (dt_local_dev) hipLaunchKernelGGL(k_copy_dt, dim3(1), dim3(1), 0, x.stream, x.scal, dt_local_dev);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(NUC_EHIP, std::string("launch: ") + hipGetErrorString(e));
    return NUC_OK;
}

extern "C" int nuc_stats(nuc_ctx *x, double *out)
{
    if (!out) return NUC_EINVAL;
    double ex[NUC_NSTATS_EX];
    int rc = nuc_stats_ex(x, ex, NUC_NSTATS_EX);
    if (rc == NUC_OK) memcpy(out, ex, 25 * sizeof(double));
    return rc;
}

extern "C" int nuc_stats_ex(nuc_ctx *x, double *out, int nout)
{
    if (!x || !out || nout < 1) return NUC_EINVAL;
    const long N = x->N;
    const nuc_config &c = x->c;
    double *prim = x->prim;
    bool own_prim = false;
    if (!prim) {
        NUC_CHECK(x, hipMallocAsync(reinterpret_cast<void **>(&prim), 4 * N * sizeof(double), x->stream));
        own_prim = true;
    }
    const double *q64 = state64(x);
    PrimArgs P{};
    for (int v = 0; v < 5; v++) P.q[v] = gp(q64 + v * N);
    for (int f = 0; f < 4; f++) P.prim[f] = gp(prim + f * N);
    P.N = N; P.nx = x->n[0]; P.ny = x->n[1];
    P.nlx = x->nloc[0]; P.nly = x->nloc[1]; P.nlz = x->nloc[2]; P.npx = x->np[0]; P.npy = x->np[1]; P.P = x->P;
    P.riemann = c.riemann; P.do_lam = 0; P.write_prim = 1; P.gamma = c.gamma; P.mach = c.mach; P.lam = gp(x->lam);
    long nbp = (N + 4095) / 4096;
    if (nbp > 65536) nbp = 65536;
    hipLaunchKernelGGL(k_prim<double>, dim3((unsigned)nbp), dim3(256), 0, x->stream, P);
    dim3 grid((x->n[0] + 63) / 64, (x->n[1] + 3) / 4, x->n[2]);
    long nb = (long)grid.x * grid.y * grid.z;
    if (nb > x->stat_nb) {
        if (x->stat_part) (void)hipFreeAsync(x->stat_part, x->stream);
        x->stat_part = nullptr;
        NUC_CHECK(x, hipMallocAsync(reinterpret_cast<void **>(&x->stat_part), nb * NSTAT * sizeof(double), x->stream));
        x->stat_nb = nb;
    }
    if (!x->stat_out)
        NUC_CHECK(x, hipMallocAsync(reinterpret_cast<void **>(&x->stat_out), NSTAT * sizeof(double), x->stream));
    StatArgs S{};
    for (int f = 0; f < 4; f++) S.prim[f] = prim + f * N;
    S.rho = q64;
    S.nx = x->n[0]; S.ny = x->n[1]; S.nz = x->n[2];
    for (int d = 0; d < 3; d++) {
        S.nloc[d] = x->nloc[d]; S.tsv[d] = x->tsv[d];
        S.vis[d] = vis_id(c.vis_scheme[d]);
        S.inv_h[d] = 1.0 / x->h[d];
    }
    S.gamma = c.gamma; S.mach = c.mach; S.tref = c.t_ref; S.tinf = c.t_inf; S.part = x->stat_part;
    hipLaunchKernelGGL(k_stats, grid, dim3(64, 4, 1), 0, x->stream, S);
    hipLaunchKernelGGL(k_stats_reduce, dim3(NSTAT), dim3(256), 0, x->stream, x->stat_part, nb, x->stat_out);
    NUC_CHECK(x, hipGetLastError());
    double s[NSTAT];
    NUC_CHECK(x, hipMemcpyAsync(s, x->stat_out, sizeof(s), hipMemcpyDeviceToHost, x->stream));
    if (own_prim) (void)hipFreeAsync(prim, x->stream);
    NUC_CHECK(x, hipStreamSynchronize(x->stream));
    const double Nd = (double)N;
    double o[NUC_NSTATS_EX];
    for (int t = 0; t < 16; t++) o[t] = s[t] / Nd;
    double u_rms_aver = s[16] / Nd, div_aver = s[17] / Nd, Skew3 = s[18] / Nd;
    double alpha_aver = s[19] / Nd, miu_aver = s[20] / Nd, rho_bar = s[21] / Nd, Skew2 = div_aver;
    // src/postproc.cpp:858-872
    double skew = Skew3 / std::pow(Skew2, 1.5);
    double lambda = std::sqrt(u_rms_aver / div_aver);
    double Re_lambda = std::sqrt(u_rms_aver) * lambda * rho_bar / miu_aver * c.reynolds;
    double tau_t = lambda / std::sqrt(u_rms_aver);
    o[16] = std::sqrt(u_rms_aver);
    o[17] = std::sqrt(u_rms_aver * 3.0) / alpha_aver;
    o[18] = std::sqrt(div_aver);
    o[19] = miu_aver;
    o[20] = Re_lambda;
    o[21] = lambda;
    o[22] = rho_bar;
    o[23] = tau_t;
    o[24] = skew;
    // dissipation eps = <mu |omega|^2>/Re + 4/3 <mu (div u)^2>/Re with the
    // Sutherland viscosity of getMiu (src/physicsModel.cpp:193-225)
    o[25] = s[22] / Nd / c.reynolds;
    o[26] = 4.0 / 3.0 * s[23] / Nd / c.reynolds;
    memcpy(out, o, (size_t)std::min(nout, NUC_NSTATS_EX) * sizeof(double));
    return NUC_OK;
}

// ===========================================================================
// operator hooks
// ===========================================================================
namespace {
struct DevBuf {
    double *p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
    bool alloc(size_t n) { return hipMalloc(&p, n * sizeof(double)) == hipSuccess; }
};
}  // namespace

extern "C" int nuc_op_reconstruct(int scheme, int upwind, int tilesize, const double *in, int n1, int n2, int n3,
                                  const double *gL, const double *gR, int bf, double *out, int zones)
{
    if (!in || !out || n1 < 2 * tilesize + 1 || n2 < 1 || n3 < 1 || (upwind != 1 && upwind != -1)) return NUC_EINVAL;
    if (((zones & 2) && !gL) || ((zones & 4) && !gR) || ((zones & 6) && bf < tilesize)) return NUC_EINVAL;
    if (scheme < 0 || scheme > NUC_HCCS77 || tilesize < 1 || tilesize > 4 || bf > 16) return NUC_EINVAL;
    int cpt = rows_for(n1 + 1);
    if (cpt < 0) return NUC_EINVAL;
    long nl = (long)n2 * n3;
    DevBuf di, dl, dr, dout;
    if (!di.alloc(nl * n1) || !dout.alloc(nl * (n1 + 1))) return NUC_ENOMEM;
    if (hipMemcpy(di.p, in, nl * n1 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return NUC_EHIP;
    if (hipMemcpy(dout.p, out, nl * (n1 + 1) * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return NUC_EHIP;
    if (gL && bf > 0) {
        if (!dl.alloc(nl * bf)) return NUC_ENOMEM;
        if (hipMemcpy(dl.p, gL, nl * bf * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return NUC_EHIP;
    }
    if (gR && bf > 0) {
        if (!dr.alloc(nl * bf)) return NUC_ENOMEM;
        if (hipMemcpy(dr.p, gR, nl * bf * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return NUC_EHIP;
    }
    OpRecArgs A{di.p, dl.p, dr.p, dout.p, n1, (int)nl, bf, tilesize, upwind, zones};
    switch (scheme) {
    case WENO5JS: launch_op_sch<WENO5JS>(cpt, A, 0); break;
    case WENO5Z: launch_op_sch<WENO5Z>(cpt, A, 0); break;
    case WENO7JS: launch_op_sch<WENO7JS>(cpt, A, 0); break;
    case CRWENO5: launch_op_sch<CRWENO5>(cpt, A, 0); break;
    case WENO3JS: launch_op_sch<WENO3JS>(cpt, A, 0); break;
    case UPWIND1ST: launch_op_sch<UPWIND1ST>(cpt, A, 0); break;
    case HCCS78: launch_op_sch<HCCS78>(cpt, A, 0); break;
    case HCCS77: launch_op_sch<HCCS77>(cpt, A, 0); break;
    default: launch_op_sch<HCCS>(cpt, A, 0); break;
    }
    if (hipGetLastError() != hipSuccess) return NUC_EHIP;
    if (hipMemcpy(out, dout.p, nl * (n1 + 1) * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return NUC_EHIP;
    return NUC_OK;
}

extern "C" int nuc_op_diff(int vis_scheme, int tilesize, const double *in, double h, int n1, int n2, int n3,
                           const double *gL, const double *gR, int bf, double *out, int zones)
{
    if (!in || !out || n1 < 2 * tilesize || n2 < 1 || n3 < 1 || !(h != 0.0)) return NUC_EINVAL;
    if (vis_scheme < 0 || vis_scheme > NUC_CD2) return NUC_EINVAL;
    int need = vis_reach(vis_scheme);
    if (((zones & 2) && !gL) || ((zones & 4) && !gR) || ((zones & 6) && bf < need)) return NUC_EINVAL;
    if (tilesize < need) return NUC_EINVAL;  // the reference's out-of-stencil read
    long nl = (long)n2 * n3;
    DevBuf di, dl, dr, dout;
    if (!di.alloc(nl * n1) || !dout.alloc(nl * n1)) return NUC_ENOMEM;
    if (hipMemcpy(di.p, in, nl * n1 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return NUC_EHIP;
    if (hipMemcpy(dout.p, out, nl * n1 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return NUC_EHIP;
    if (gL && bf > 0) {
        if (!dl.alloc(nl * bf)) return NUC_ENOMEM;
        if (hipMemcpy(dl.p, gL, nl * bf * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return NUC_EHIP;
    }
    if (gR && bf > 0) {
        if (!dr.alloc(nl * bf)) return NUC_ENOMEM;
        if (hipMemcpy(dr.p, gR, nl * bf * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return NUC_EHIP;
    }
    OpDiffArgs A{di.p, dl.p, dr.p, dout.p, n1, bf, tilesize, vis_id(vis_scheme), zones, nl,
                 1.0 / h};
    long tot = nl * n1;
    hipLaunchKernelGGL(k_op_diff, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, 0, A);
    if (hipGetLastError() != hipSuccess) return NUC_EHIP;
    if (hipMemcpy(out, dout.p, nl * n1 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return NUC_EHIP;
    return NUC_OK;
}

extern "C" int nuc_op_rk3(int stage, const double *rhs, const double *un, double *ui, double dt, long n)
{
    if (!rhs || !un || !ui || n < 1 || stage < 0 || stage > 2) return NUC_EINVAL;
    DevBuf dr, du, di;
    if (!dr.alloc(n) || !du.alloc(n) || !di.alloc(n)) return NUC_ENOMEM;
    if (hipMemcpy(dr.p, rhs, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return NUC_EHIP;
    if (hipMemcpy(du.p, un, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return NUC_EHIP;
    if (hipMemcpy(di.p, ui, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return NUC_EHIP;
    hipLaunchKernelGGL(k_op_rk3, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, stage, dr.p, du.p, di.p, dt, n);
    if (hipGetLastError() != hipSuccess) return NUC_EHIP;
    if (hipMemcpy(ui, di.p, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return NUC_EHIP;
    return NUC_OK;
}

extern "C" int nuc_profile(nuc_ctx *x, int enable)
{
    if (!x) return NUC_EINVAL;
    x->prof = enable != 0;
    for (int i = 0; i < NUC_PROF_SLOTS; i++) { x->prof_ms[i] = 0.0; x->prof_n[i] = 0; }
    return NUC_OK;
}

extern "C" int nuc_profile_read(const nuc_ctx *x, int slot, double *total_ms, long *launches, const char **name)
{
    if (!x || slot < 0 || slot >= NUC_PROF_SLOTS) return NUC_EINVAL;
    if (total_ms) *total_ms = x->prof_ms[slot];
    if (launches) *launches = x->prof_n[slot];
    if (name) *name = kProfNames[slot];
    return NUC_OK;
}


// ---------------------------------------------------------------------------
// time-averaged statistics and TKE budget: postproc::solveAveraged0/1/2
// (src/postproc.cpp:1186-2110), field lists of src/postproc.hpp:49-121.
// All in FP64 on the FP64 state of either precision path.  Each group is one
// contiguous block of fields (field f of a group at base + f*N); every
// derivative is solveGrad (src/postproc.cpp:2111-2124,2284-2298): the viscous
// central difference per direction with the periodic wrap, inner/boundary
// forms chosen by the rank-local position as in the sweeps.
// ---------------------------------------------------------------------------
constexpr int NPRIM = 10, NGRAD = 21, NDER = 28, NTKE = 49, NBUD = 45;
constexpr int POST_PART = 256;   // reduction partials per field

struct PostArgs {
    const double GMEM *q;
    double GMEM *tp, *ap, *tg, *ag, *td, *ad, *tt, *at, *bu;
    long N;
    double gamma, mach, re, pr, tref, tinf;
    int law;
};

struct PostDiffArgs {
    const double GMEM *in;
    double GMEM *out;
    long N;
    int nx, ny, d, n, nloc, tsv, vis;
    double inv_h;
};

__global__ __launch_bounds__(256) void k_post_diff(PostDiffArgs A)
{
    const long id = (long)blockIdx.x * 256 + threadIdx.x;
    if (id >= A.N) return;
    const long i = id % A.nx, r = id / A.nx, j = r % A.ny, k = r / A.ny;
    long pos, stride;
    if (A.d == 0) { pos = i; stride = 1; }
    else if (A.d == 1) { pos = j; stride = A.nx; }
    else { pos = k; stride = (long)A.nx * A.ny; }
    const long base = id - pos * stride;
    double L[9];
#pragma unroll
    for (int o = 0; o < 9; o++) {
        long m = pos + o - 4;
        m = m < 0 ? m + A.n : (m >= A.n ? m - A.n : m);
        L[o] = A.in[base + m * stride];
    }
    const int cl = (int)(pos % A.nloc);
    A.out[id] = central_diff<double>(L + 4, 0, cl >= A.tsv && cl < A.nloc - A.tsv, A.vis, A.inv_h);
}

// f_aver = (f_aver*averStep + f_temp)/(averStep + 1), src/postproc.cpp:1318
// (kernel parameters are plain pointers: an address-space qualifier would
// change the device symbol's mangled name but not the host stub's)
__global__ __launch_bounds__(256) void k_post_avg(double *a, const double *t, long n, double step)
{
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256)
        a[i] = (a[i] * step + t[i]) / (step + 1.0);
}

#define PF(base, f) (base)[(long)(f) * A.N + id]

// solveAveraged0: primaries (con2prim, src/physicsModel.cpp:331-344) and getMiu
__global__ __launch_bounds__(256) void k_post_prim(PostArgs A)
{
    const long id = (long)blockIdx.x * 256 + threadIdx.x;
    if (id >= A.N) return;
    const double rho = A.q[id];
    const double u = A.q[A.N + id] / rho, v = A.q[2 * A.N + id] / rho, w = A.q[3 * A.N + id] / rho;
    const double E = A.q[4 * A.N + id];
    const double p = fabs((E - 0.5 * rho * (u * u + v * v + w * w)) * (A.gamma - 1.0));
    const double T = A.gamma * A.mach * A.mach * p / rho;
    double miu0;
    const double S = A.tref / A.tinf;
    if (A.law == 0) miu0 = (1.0 + S) / (T + S) * pow(T, 1.5) / A.re;   // src/physicsModel.cpp:193-266
    else miu0 = 1.0 / A.re;
    PF(A.tp, 0) = rho; PF(A.tp, 1) = u; PF(A.tp, 2) = v; PF(A.tp, 3) = w; PF(A.tp, 4) = p; PF(A.tp, 5) = T;
    PF(A.tp, 6) = p / rho / (A.gamma - 1.0);
    PF(A.tp, 7) = sqrt(A.gamma * p / rho);
    PF(A.tp, 8) = miu0;
    PF(A.tp, 9) = miu0 / ((A.gamma - 1.0) * A.mach * A.mach * A.pr);
}

// vorticity and its squares, src/postproc.cpp:1296-1302
__global__ __launch_bounds__(256) void k_post_vort(PostArgs A)
{
    const long id = (long)blockIdx.x * 256 + threadIdx.x;
    if (id >= A.N) return;
    const double o0 = PF(A.tg, 7) - PF(A.tg, 5), o1 = PF(A.tg, 2) - PF(A.tg, 6), o2 = PF(A.tg, 3) - PF(A.tg, 1);
    PF(A.tg, 15) = o0; PF(A.tg, 16) = o1; PF(A.tg, 17) = o2;
    PF(A.tg, 18) = o0 * o0; PF(A.tg, 19) = o1 * o1; PF(A.tg, 20) = o2 * o2;
}

// solveAveraged1: derived fields, src/postproc.cpp:1414-1465
__global__ __launch_bounds__(256) void k_post_derived(PostArgs A)
{
    const long id = (long)blockIdx.x * 256 + threadIdx.x;
    if (id >= A.N) return;
    const double rho = PF(A.tp, 0), u = PF(A.tp, 1), v = PF(A.tp, 2), w = PF(A.tp, 3), e = PF(A.tp, 6);
    const double miu = PF(A.tp, 8);
    const double ux = PF(A.tg, 0), uy = PF(A.tg, 1), uz = PF(A.tg, 2), vx = PF(A.tg, 3), vy = PF(A.tg, 4);
    const double vz = PF(A.tg, 5), wx = PF(A.tg, 6), wy = PF(A.tg, 7), wz = PF(A.tg, 8);
    PF(A.td, 0) = rho * u; PF(A.td, 1) = rho * v; PF(A.td, 2) = rho * w; PF(A.td, 3) = rho * e;
    PF(A.td, 4) = u * u; PF(A.td, 5) = u * v; PF(A.td, 6) = u * w;
    PF(A.td, 7) = v * v; PF(A.td, 8) = v * w; PF(A.td, 9) = w * w;
    PF(A.td, 10) = rho * u * u; PF(A.td, 11) = rho * u * v; PF(A.td, 12) = rho * u * w;
    PF(A.td, 13) = rho * v * v; PF(A.td, 14) = rho * v * w; PF(A.td, 15) = rho * w * w;
    const double div = ux + vy + wz;
    double t[6];
    t[0] = miu * (2.0 * ux - 2.0 / 3.0 * div);
    t[1] = miu * (uy + vx - 2.0 / 3.0 * div);
    t[2] = miu * (uz + wx - 2.0 / 3.0 * div);
    t[3] = miu * (2.0 * vy - 2.0 / 3.0 * div);
    t[4] = miu * (vz + wy - 2.0 / 3.0 * div);
    t[5] = miu * (2.0 * wz - 2.0 / 3.0 * div);
#pragma unroll
    for (int c = 0; c < 6; c++) { PF(A.td, 16 + c) = t[c]; PF(A.td, 22 + c) = rho * t[c]; }
}

// solveAveraged2: pointwise TKE fields, src/postproc.cpp:1641-1700
__global__ __launch_bounds__(256) void k_post_tke(PostArgs A)
{
    const long id = (long)blockIdx.x * 256 + threadIdx.x;
    if (id >= A.N) return;
    const double rho = PF(A.tp, 0), u = PF(A.tp, 1), v = PF(A.tp, 2), w = PF(A.tp, 3), p = PF(A.tp, 4);
    const double ux = PF(A.tg, 0), uy = PF(A.tg, 1), uz = PF(A.tg, 2), vx = PF(A.tg, 3), vy = PF(A.tg, 4);
    const double vz = PF(A.tg, 5), wx = PF(A.tg, 6), wy = PF(A.tg, 7), wz = PF(A.tg, 8);
    const double txx = PF(A.td, 16), txy = PF(A.td, 17), txz = PF(A.td, 18);
    const double tyy = PF(A.td, 19), tyz = PF(A.td, 20), tzz = PF(A.td, 21);
    PF(A.tt, 0) = u * p; PF(A.tt, 1) = v * p; PF(A.tt, 2) = w * p;
    PF(A.tt, 3) = p * ux; PF(A.tt, 4) = p * vy; PF(A.tt, 5) = p * wz;
    double a = u * txx, b = v * txy, c = w * txz;
    PF(A.tt, 15) = a; PF(A.tt, 16) = b; PF(A.tt, 17) = c; PF(A.tt, 18) = a + b + c;
    a = u * txy; b = v * tyy; c = w * tyz;
    PF(A.tt, 20) = a; PF(A.tt, 21) = b; PF(A.tt, 22) = c; PF(A.tt, 23) = a + b + c;
    a = u * txz; b = v * tyz; c = w * tzz;
    PF(A.tt, 25) = a; PF(A.tt, 26) = b; PF(A.tt, 27) = c; PF(A.tt, 28) = a + b + c;
    PF(A.tt, 30) = txx * ux; PF(A.tt, 31) = txy * vx; PF(A.tt, 32) = txz * wx;
    PF(A.tt, 33) = txy * uy; PF(A.tt, 34) = tyy * vy; PF(A.tt, 35) = tyz * wy;
    PF(A.tt, 36) = txz * uz; PF(A.tt, 37) = tyz * vz; PF(A.tt, 38) = tzz * wz;
    PF(A.tt, 39) = rho * u * u * u; PF(A.tt, 40) = rho * u * u * v; PF(A.tt, 41) = rho * u * u * w;
    PF(A.tt, 42) = rho * u * v * v; PF(A.tt, 43) = rho * u * v * w; PF(A.tt, 44) = rho * u * w * w;
    PF(A.tt, 45) = rho * v * v * v; PF(A.tt, 46) = rho * v * v * w; PF(A.tt, 47) = rho * v * w * w;
    PF(A.tt, 48) = rho * w * w * w;
}

// budget, stage 1: <u'p'> = <up> - <u><p>, src/postproc.cpp:1829-1833
__global__ __launch_bounds__(256) void k_post_bud1(PostArgs A)
{
    const long id = (long)blockIdx.x * 256 + threadIdx.x;
    if (id >= A.N) return;
    const double pa = PF(A.ap, 4);
    PF(A.bu, 21) = PF(A.at, 0) - PF(A.ap, 1) * pa;
    PF(A.bu, 22) = PF(A.at, 1) - PF(A.ap, 2) * pa;
    PF(A.bu, 23) = PF(A.at, 2) - PF(A.ap, 3) * pa;
}

// budget, stage 2: <u_k><tau_kj>, src/postproc.cpp:1851-1861
__global__ __launch_bounds__(256) void k_post_bud2(PostArgs A)
{
    const long id = (long)blockIdx.x * 256 + threadIdx.x;
    if (id >= A.N) return;
    const double u = PF(A.ap, 1), v = PF(A.ap, 2), w = PF(A.ap, 3);
    const double txx = PF(A.ad, 16), txy = PF(A.ad, 17), txz = PF(A.ad, 18);
    const double tyy = PF(A.ad, 19), tyz = PF(A.ad, 20), tzz = PF(A.ad, 21);
    PF(A.bu, 27) = u * txx + v * txy + w * txz;
    PF(A.bu, 28) = u * txy + v * tyy + w * tyz;
    PF(A.bu, 29) = u * txz + v * tyz + w * tzz;
}

// budget, stage 3: tke, production, pressure and viscous diffusion, rho_k u k,
// src/postproc.cpp:1880-1960
__global__ __launch_bounds__(256) void k_post_bud3(PostArgs A)
{
    const long id = (long)blockIdx.x * 256 + threadIdx.x;
    if (id >= A.N) return;
    const double rho = PF(A.ap, 0), u = PF(A.ap, 1), v = PF(A.ap, 2), w = PF(A.ap, 3), pa = PF(A.ap, 4);
    const double ux = PF(A.ag, 0), uy = PF(A.ag, 1), uz = PF(A.ag, 2), vx = PF(A.ag, 3), vy = PF(A.ag, 4);
    const double vz = PF(A.ag, 5), wx = PF(A.ag, 6), wy = PF(A.ag, 7), wz = PF(A.ag, 8);
    const double px = PF(A.ag, 9), py = PF(A.ag, 10), pz = PF(A.ag, 11);
    const double ru = PF(A.ad, 0), rv = PF(A.ad, 1), rw = PF(A.ad, 2);
    const double ruu = PF(A.ad, 10), ruv = PF(A.ad, 11), ruw = PF(A.ad, 12);
    const double rvv = PF(A.ad, 13), rvw = PF(A.ad, 14), rww = PF(A.ad, 15);
    const double txx = PF(A.ad, 16), txy = PF(A.ad, 17), txz = PF(A.ad, 18);
    const double tyy = PF(A.ad, 19), tyz = PF(A.ad, 20), tzz = PF(A.ad, 21);
    const double fu = ru / rho, fv = rv / rho, fw = rw / rho;
    const double tke = 0.5 * (ruu / rho - fu * fu) + 0.5 * (rvv / rho - fv * fv) + 0.5 * (rww / rho - fw * fw);
    const double prod = (ruu - ru * ru / rho) * PF(A.bu, 12) + (ruv - ru * rv / rho) * PF(A.bu, 13) +
                        (ruw - ru * rw / rho) * PF(A.bu, 14) + (ruv - ru * rv / rho) * PF(A.bu, 15) +
                        (rvv - rv * rv / rho) * PF(A.bu, 16) + (rvw - rv * rw / rho) * PF(A.bu, 17) +
                        (ruw - ru * rw / rho) * PF(A.bu, 18) + (rvw - rv * rw / rho) * PF(A.bu, 19) +
                        (rww - rw * rw / rho) * PF(A.bu, 20);
    const double du = u - fu, dv = v - fv, dw = w - fw;
    const double p1 = -du * px - dv * py - dw * pz;
    const double p2 = -PF(A.bu, 24) - PF(A.bu, 25) - PF(A.bu, 26);
    const double p3 = (PF(A.at, 3) - ux * pa) + (PF(A.at, 4) - vy * pa) + (PF(A.at, 5) - wz * pa);
    const double v1 = du * PF(A.at, 6) + dv * PF(A.at, 7) + dw * PF(A.at, 8) + du * PF(A.at, 9) + dv * PF(A.at, 10) +
                      dw * PF(A.at, 11) + du * PF(A.at, 12) + dv * PF(A.at, 13) + dw * PF(A.at, 14);
    const double v2 = (PF(A.at, 19) - PF(A.bu, 30)) + (PF(A.at, 24) - PF(A.bu, 31)) + (PF(A.at, 29) - PF(A.bu, 32));
    const double v3 = -(PF(A.at, 30) - ux * txx) - (PF(A.at, 31) - vx * txy) - (PF(A.at, 32) - wx * txz) -
                      (PF(A.at, 33) - uy * txy) - (PF(A.at, 34) - vy * tyy) - (PF(A.at, 35) - wy * tyz) -
                      (PF(A.at, 36) - uz * txz) - (PF(A.at, 37) - vz * tyz) - (PF(A.at, 38) - wz * tzz);
    PF(A.bu, 0) = tke;
    PF(A.bu, 1) = prod;
    PF(A.bu, 4) = p1; PF(A.bu, 5) = p2; PF(A.bu, 6) = p3;
    PF(A.bu, 3) = p1 + p2 + p3;
    PF(A.bu, 8) = v1; PF(A.bu, 9) = v2; PF(A.bu, 10) = v3;
    PF(A.bu, 7) = v1 + v2 + v3;
    PF(A.bu, 33) = ru * tke; PF(A.bu, 34) = rv * tke; PF(A.bu, 35) = rw * tke;
}

// budget, stage 4: convection and the turbulent transport fluxes,
// src/postproc.cpp:1973-2004
__global__ __launch_bounds__(256) void k_post_bud4(PostArgs A)
{
    const long id = (long)blockIdx.x * 256 + threadIdx.x;
    if (id >= A.N) return;
    const double rho = PF(A.ap, 0);
    const double ru = PF(A.ad, 0), rv = PF(A.ad, 1), rw = PF(A.ad, 2);
    const double ruu = PF(A.ad, 10), ruv = PF(A.ad, 11), ruw = PF(A.ad, 12);
    const double rvv = PF(A.ad, 13), rvw = PF(A.ad, 14), rww = PF(A.ad, 15);
    PF(A.bu, 11) = -PF(A.bu, 36) - PF(A.bu, 37) - PF(A.bu, 38);
    const double kk = ruu - ru * ru / rho + rvv - rv * rv / rho + rww - rw * rw / rho;
    PF(A.bu, 39) = 0.5 * PF(A.at, 39) - ruu * ru / rho + 0.5 * ru * ru * ru / rho / rho +
                   0.5 * PF(A.at, 42) - ruv * rv / rho + 0.5 * ru * rv * rv / rho / rho +
                   0.5 * PF(A.at, 44) - ruw * rw / rho + 0.5 * ru * rw * rw / rho / rho - ru / rho * 0.5 * kk;
    PF(A.bu, 40) = 0.5 * PF(A.at, 40) - ruv * ru / rho + 0.5 * rv * ru * ru / rho / rho +
                   0.5 * PF(A.at, 45) - rvv * rv / rho + 0.5 * rv * rv * rv / rho / rho +
                   0.5 * PF(A.at, 47) - rvw * rw / rho + 0.5 * rv * rw * rw / rho / rho - rv / rho * 0.5 * kk;
    PF(A.bu, 41) = 0.5 * PF(A.at, 41) - ruw * ru / rho + 0.5 * rw * ru * ru / rho / rho +
                   0.5 * PF(A.at, 46) - rvw * rv / rho + 0.5 * rw * rv * rv / rho / rho +
                   0.5 * PF(A.at, 48) - rww * rw / rho + 0.5 * rw * rw * rw / rho / rho - rw / rho * 0.5 * kk;
}

// budget, stage 5: turbulent transport, src/postproc.cpp:2097-2105
__global__ __launch_bounds__(256) void k_post_bud5(PostArgs A)
{
    const long id = (long)blockIdx.x * 256 + threadIdx.x;
    if (id >= A.N) return;
    PF(A.bu, 2) = -PF(A.bu, 42) - PF(A.bu, 43) - PF(A.bu, 44);
}
#undef PF

// volume means of nf consecutive fields: per-field partial sums of a fixed
// grid, then a fixed-order reduction (deterministic run to run)
__global__ __launch_bounds__(256) void k_post_mean_part(const double *f, long N, double *part)
{
    __shared__ double sm[256];
    const double *g = f + (long)blockIdx.y * N;
    double acc = 0.0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < N; i += (long)POST_PART * 256) acc += g[i];
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(long)blockIdx.y * POST_PART + blockIdx.x] = sm[0];
}

__global__ __launch_bounds__(256) void k_post_mean_fin(const double *part, double inv_n, double *out)
{
    __shared__ double sm[256];
    sm[threadIdx.x] = part[(long)blockIdx.x * POST_PART + threadIdx.x];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sm[0] * inv_n;
}


// segment offsets (in doubles, multiples of N) of the averaging buffer
struct PostLayout {
    long tp, ap, tg, ag, td, ad, tt, at, bu, part, mean, total;
};

static PostLayout post_layout(int type, long N)
{
    PostLayout L{};
    long o = 0;
    auto seg = [&](long nf) { long r = o; o += nf * N; return r; };
    L.tp = seg(NPRIM); L.ap = seg(NPRIM); L.tg = seg(NGRAD); L.ag = seg(NGRAD);
    L.td = L.ad = L.tt = L.at = L.bu = -1;
    if (type >= 2) { L.td = seg(NDER); L.ad = seg(NDER); }
    if (type >= 3) { L.tt = seg(NTKE); L.at = seg(NTKE); L.bu = seg(NBUD); }
    L.part = o; o += (long)NTKE * POST_PART;
    L.mean = o; o += NTKE;
    L.total = o;
    return L;
}

extern "C" int nuc_average_end(nuc_ctx *x)
{
    if (!x) return NUC_EINVAL;
    if (x->post_buf) (void)hipFreeAsync(x->post_buf, x->stream);
    x->post_buf = nullptr;
    x->post_type = 0;
    x->post_step = 0;
    NUC_CHECK(x, hipStreamSynchronize(x->stream));
    return NUC_OK;
}

extern "C" int nuc_average_begin(nuc_ctx *x, int type)
{
    if (!x) return NUC_EINVAL;
    if (type < 1 || type > 3) { x->err = "averaging type must be 1, 2 or 3 (src/postproc.cpp:37-60)"; return NUC_EINVAL; }
    int rc = nuc_average_end(x);
    if (rc != NUC_OK) return rc;
    const PostLayout L = post_layout(type, x->N);
    if (hipMallocAsync(reinterpret_cast<void **>(&x->post_buf), (size_t)L.total * sizeof(double), x->stream) !=
        hipSuccess) {
        x->post_buf = nullptr;
        x->err = "device allocation (averaged statistics)";
        return NUC_ENOMEM;
    }
    // a new average starts from zero fields (TimeAverageStartStep 0)
    NUC_CHECK(x, hipMemsetAsync(x->post_buf, 0, (size_t)L.total * sizeof(double), x->stream));
    NUC_CHECK(x, hipStreamSynchronize(x->stream));
    x->post_type = type;
    x->post_step = 0;
    return NUC_OK;
}

extern "C" long long nuc_average_steps(const nuc_ctx *x) { return x ? x->post_step : -1; }

extern "C" int nuc_average_update(nuc_ctx *x)
{
    if (!x) return NUC_EINVAL;
    if (!x->post_type) { x->err = "no average running (nuc_average_begin)"; return NUC_EINVAL; }
    const long N = x->N;
    const PostLayout L = post_layout(x->post_type, N);
    double *b = x->post_buf;
    const nuc_config &c = x->c;
    PostArgs A{};
    A.q = gp(state64(x));
    A.tp = gp(b + L.tp); A.ap = gp(b + L.ap); A.tg = gp(b + L.tg); A.ag = gp(b + L.ag);
    if (x->post_type >= 2) { A.td = gp(b + L.td); A.ad = gp(b + L.ad); }
    if (x->post_type >= 3) { A.tt = gp(b + L.tt); A.at = gp(b + L.at); A.bu = gp(b + L.bu); }
    A.N = N;
    A.gamma = c.gamma; A.mach = c.mach; A.re = c.reynolds; A.pr = c.prandtl; A.tref = c.t_ref; A.tinf = c.t_inf;
    A.law = c.viscosity;
    const dim3 grid((unsigned)((N + 255) / 256)), blk(256);
    auto diff = [&](const double *in, double *out, int d) {
        PostDiffArgs D{gp(in), gp(out), N, x->n[0], x->n[1], d, x->n[d], x->nloc[d], x->tsv[d],
                       vis_id(c.vis_scheme[d]), 1.0 / x->h[d]};
        hipLaunchKernelGGL(k_post_diff, grid, blk, 0, x->stream, D);
    };
    auto F = [&](long seg, int f) { return b + seg + (long)f * N; };
    const double step = (double)x->post_step;
    auto avg = [&](long a_seg, long t_seg, int nf) {
        long n = (long)nf * N;
        long nb = std::min<long>((n + 255) / 256, 65536);
        hipLaunchKernelGGL(k_post_avg, dim3((unsigned)nb), blk, 0, x->stream, b + a_seg, (const double *)(b + t_seg), n,
                           step);
    };
    // solveAveraged0 (src/postproc.cpp:1207-1349)
    hipLaunchKernelGGL(k_post_prim, grid, blk, 0, x->stream, A);
    for (int f = 0; f < 5; f++)            // d(u, v, w, p, T)/d(x, y, z)
        for (int d = 0; d < 3; d++) diff(F(L.tp, 1 + f), F(L.tg, 3 * f + d), d);
    hipLaunchKernelGGL(k_post_vort, grid, blk, 0, x->stream, A);
    avg(L.ap, L.tp, NPRIM);
    avg(L.ag, L.tg, NGRAD);
    if (x->post_type >= 2) {              // solveAveraged1 (src/postproc.cpp:1351-1483)
        hipLaunchKernelGGL(k_post_derived, grid, blk, 0, x->stream, A);
        avg(L.ad, L.td, NDER);
    }
    if (x->post_type >= 3) {              // solveAveraged2 (src/postproc.cpp:1485-2110)
        diff(F(L.td, 16), F(L.tt, 6), 0);
        diff(F(L.td, 17), F(L.tt, 7), 0);
        diff(F(L.td, 18), F(L.tt, 8), 0);
        diff(F(L.td, 17), F(L.tt, 9), 1);
        diff(F(L.td, 19), F(L.tt, 10), 1);
        diff(F(L.td, 20), F(L.tt, 11), 1);
        diff(F(L.td, 18), F(L.tt, 12), 2);
        diff(F(L.td, 20), F(L.tt, 13), 2);
        diff(F(L.td, 21), F(L.tt, 14), 2);
        hipLaunchKernelGGL(k_post_tke, grid, blk, 0, x->stream, A);
        diff(F(L.tt, 18), F(L.tt, 19), 0);
        diff(F(L.tt, 23), F(L.tt, 24), 1);
        diff(F(L.tt, 28), F(L.tt, 29), 2);
        avg(L.at, L.tt, NTKE);
        for (int c3 = 0; c3 < 3; c3++)
            for (int d = 0; d < 3; d++) diff(F(L.ad, c3), F(L.bu, 12 + 3 * c3 + d), d);
        hipLaunchKernelGGL(k_post_bud1, grid, blk, 0, x->stream, A);
        for (int d = 0; d < 3; d++) diff(F(L.bu, 21 + d), F(L.bu, 24 + d), d);
        hipLaunchKernelGGL(k_post_bud2, grid, blk, 0, x->stream, A);
        for (int d = 0; d < 3; d++) diff(F(L.bu, 27 + d), F(L.bu, 30 + d), d);
        hipLaunchKernelGGL(k_post_bud3, grid, blk, 0, x->stream, A);
        for (int d = 0; d < 3; d++) diff(F(L.bu, 33 + d), F(L.bu, 36 + d), d);
        hipLaunchKernelGGL(k_post_bud4, grid, blk, 0, x->stream, A);
        for (int d = 0; d < 3; d++) diff(F(L.bu, 39 + d), F(L.bu, 42 + d), d);
        hipLaunchKernelGGL(k_post_bud5, grid, blk, 0, x->stream, A);
    }
    NUC_CHECK(x, hipGetLastError());
    NUC_CHECK(x, hipStreamSynchronize(x->stream));
    x->post_step++;
    return NUC_OK;
}

extern "C" int nuc_average_restore(nuc_ctx *x, int type, long long steps, const double *fields)
{
    if (!x || !fields || steps < 0) return NUC_EINVAL;
    int rc = nuc_average_begin(x, type);
    if (rc != NUC_OK) return rc;
    static const int cnt[5] = {NPRIM, NGRAD, NDER, NTKE, NBUD};
    const PostLayout L = post_layout(type, x->N);
    const long seg[5] = {L.ap, L.ag, L.ad, L.at, L.bu};
    const int ng = type == 1 ? 2 : (type == 2 ? 3 : 5);
    size_t off = 0;
    for (int g = 0; g < ng; g++) {
        const size_t n = (size_t)cnt[g] * x->N;
        NUC_CHECK(x, hipMemcpyAsync(x->post_buf + seg[g], fields + off, n * sizeof(double), hipMemcpyHostToDevice,
                                    x->stream));
        off += n;
    }
    NUC_CHECK(x, hipStreamSynchronize(x->stream));
    x->post_step = steps;
    return NUC_OK;
}

static long post_group(const nuc_ctx *x, int group, int *count)
{
    static const int cnt[5] = {NPRIM, NGRAD, NDER, NTKE, NBUD};
    if (!x->post_type || group < 0 || group > 4) return -1;
    const PostLayout L = post_layout(x->post_type, x->N);
    const long seg[5] = {L.ap, L.ag, L.ad, L.at, L.bu};
    if (seg[group] < 0) return -1;
    *count = cnt[group];
    return seg[group];
}

extern "C" int nuc_average_field(nuc_ctx *x, int group, int index, double *out)
{
    if (!x || !out) return NUC_EINVAL;
    int count = 0;
    const long seg = post_group(x, group, &count);
    if (seg < 0 || index < 0 || index >= count) {
        x->err = "no such averaged field at this averaging type";
        return NUC_EINVAL;
    }
    NUC_CHECK(x, hipMemcpyAsync(out, x->post_buf + seg + (long)index * x->N, x->N * sizeof(double),
                                hipMemcpyDeviceToHost, x->stream));
    NUC_CHECK(x, hipStreamSynchronize(x->stream));
    return NUC_OK;
}

extern "C" int nuc_average_means(nuc_ctx *x, int group, double *out, int nout)
{
    if (!x || !out) return NUC_EINVAL;
    int count = 0;
    const long seg = post_group(x, group, &count);
    if (seg < 0 || nout < count) {
        x->err = "no such averaged group at this averaging type, or out too short";
        return NUC_EINVAL;
    }
    const PostLayout L = post_layout(x->post_type, x->N);
    double *b = x->post_buf;
    hipLaunchKernelGGL(k_post_mean_part, dim3(POST_PART, count), dim3(256), 0, x->stream, (const double *)(b + seg),
                       x->N, b + L.part);
    hipLaunchKernelGGL(k_post_mean_fin, dim3(count), dim3(256), 0, x->stream, (const double *)(b + L.part),
                       1.0 / (double)x->N, b + L.mean);
    NUC_CHECK(x, hipGetLastError());
    NUC_CHECK(x, hipMemcpyAsync(out, b + L.mean, count * sizeof(double), hipMemcpyDeviceToHost, x->stream));
    NUC_CHECK(x, hipStreamSynchronize(x->stream));
    return NUC_OK;
}


// ---------------------------------------------------------------------------
// initial conditions on the device: the decaying-turbulence field of
// initial_hit (src/block.cpp:397-551, the host generator's seeded spectrum)
// synthesised by a separable inverse DFT in three passes (x, then y, then z
// keeping the real part), rescaled to u_rms with a fixed-order reduction; and
// the Taylor-Green field of initial_taylorgreen (src/block.cpp:359-395)
// ---------------------------------------------------------------------------
struct IcArgs {
    const double2 GMEM *spec[3];   // mode amplitudes per velocity component
    const double2 GMEM *tw[3];     // twiddles e^{+i 2 pi k x / n}, [(k+K)*n + x]
    double2 GMEM *g;               // x pass: [c][kz][ky][x]
    double2 GMEM *h;               // y pass: [c][kz][y][x]
    double GMEM *vel;              // z pass: [c][cell]
    double GMEM *q;                // state out
    const double GMEM *mean_sq;    // <|u|^2> of the synthesised field
    int n[3], W[3];
    long N;
    double gamma, mach, scale_num;
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__global__ __launch_bounds__(256) void k_ic_x(IcArgs A)
{
    const long tot = 3L * A.W[2] * A.W[1] * A.n[0];
    const long id = (long)blockIdx.x * 256 + threadIdx.x;
    if (id >= tot) return;
    const int x = (int)(id % A.n[0]);
    const long r = id / A.n[0];
    const long kzky = r % ((long)A.W[2] * A.W[1]);
    const int c = (int)(r / ((long)A.W[2] * A.W[1]));
    const double2 GMEM *sp = A.spec[c] + kzky * A.W[0];
    double2 acc = make_double2(0.0, 0.0);
    for (int kx = 0; kx < A.W[0]; kx++) {
        const double2 t = cmul(sp[kx], A.tw[0][(long)kx * A.n[0] + x]);
        acc.x += t.x; acc.y += t.y;
    }
    A.g[id] = acc;
}

__global__ __launch_bounds__(256) void k_ic_y(IcArgs A)
{
    const long tot = 3L * A.W[2] * A.n[1] * A.n[0];
    const long id = (long)blockIdx.x * 256 + threadIdx.x;
    if (id >= tot) return;
    const int x = (int)(id % A.n[0]);
    long r = id / A.n[0];
    const int y = (int)(r % A.n[1]);
    r /= A.n[1];                                  // c * W2 + kz
    double2 acc = make_double2(0.0, 0.0);
    for (int ky = 0; ky < A.W[1]; ky++) {
        const double2 t = cmul(A.g[(r * A.W[1] + ky) * A.n[0] + x], A.tw[1][(long)ky * A.n[1] + y]);
        acc.x += t.x; acc.y += t.y;
    }
    A.h[id] = acc;
}

__global__ __launch_bounds__(256) void k_ic_z(IcArgs A)
{
    const long id = (long)blockIdx.x * 256 + threadIdx.x;
    if (id >= 3 * A.N) return;
    const int c = (int)(id / A.N);
    const long cell = id % A.N;
    const long nxy = (long)A.n[0] * A.n[1];
    const int z = (int)(cell / nxy);
    const long xy = cell % nxy;
    double acc = 0.0;
    for (int kz = 0; kz < A.W[2]; kz++) {
        const double2 hv = A.h[((long)c * A.W[2] + kz) * nxy + xy];
        const double2 t = A.tw[2][(long)kz * A.n[2] + z];
        acc += hv.x * t.x - hv.y * t.y;
    }
    A.vel[id] = acc;
}

// |u|^2 per cell into the state's rho slot, reduced by the post mean kernels
__global__ __launch_bounds__(256) void k_ic_sq(IcArgs A)
{
    const long id = (long)blockIdx.x * 256 + threadIdx.x;
    if (id >= A.N) return;
    const double u = A.vel[id], v = A.vel[A.N + id], w = A.vel[2 * A.N + id];
    A.q[id] = u * u + v * v + w * w;
}

// rho = 1, p = 1/(gamma M^2), velocity rescaled to u_rms: scale = u_rms / sqrt(<|u|^2>/3)
__global__ __launch_bounds__(256) void k_ic_q(IcArgs A)
{
    const long id = (long)blockIdx.x * 256 + threadIdx.x;
    if (id >= A.N) return;
    const double v0 = sqrt(*A.mean_sq / 3.0);
    const double scale = (v0 > 0) ? A.scale_num / v0 : 0.0;
    const double u = A.vel[id] * scale, v = A.vel[A.N + id] * scale, w = A.vel[2 * A.N + id] * scale;
    const double p0 = 1.0 / (A.gamma * A.mach * A.mach);
    A.q[id] = 1.0;
    A.q[A.N + id] = u;
    A.q[2 * A.N + id] = v;
    A.q[3 * A.N + id] = w;
    A.q[4 * A.N + id] = p0 / (A.gamma - 1.0) + 0.5 * (u * u + v * v + w * w);
}

// analytic initial conditions of src/block.cpp:285-395, by IcKind
enum IcKind { IC_DEFAULT = 0, IC_IVC = 1, IC_TGV = 2 };

struct AnalyticIcArgs {
    double GMEM *q;
    int nx, ny, nz, nl[3], kind;
    long N;
    double dx, dy, dz, lxr, lyr, lzr, gamma, mach;   // cell sizes, rank box sizes
    double ox, oy, oz;                               // block.in origin
};

// initial_default (src/block.cpp:285-314): at rest, p = 1/(gamma M^2)
__device__ void ic_default(const AnalyticIcArgs &A, long id)
{
    A.q[id] = 1.0;
    A.q[A.N + id] = 0.0;
    A.q[2 * A.N + id] = 0.0;
    A.q[3 * A.N + id] = 0.0;
    A.q[4 * A.N + id] = 1.0 / (A.mach * A.mach * A.gamma) / (A.gamma - 1.0);
}

// initial_ivc (src/block.cpp:316-357): vortex at (5, 5), strength 0.5, on u = 0.5
__device__ void ic_ivc(const AnalyticIcArgs &A, long id, double x0, double y0)
{
    const double g = A.gamma, b = 0.5, pi = 3.141592653589793;
    const double dx = x0 - 5.0, dy = y0 - 5.0;
    const double r2 = dx * dx + dy * dy;
    const double rho = pow(1.0 - (g - 1.0) * b * b * exp(1.0 - r2) / (8.0 * g * pi * pi), 2.5);
    const double a = b / (2.0 * pi) * exp(0.5 * (1.0 - r2));
    const double u = 0.5 - a * dy, v = a * dx;
    A.q[id] = rho;
    A.q[A.N + id] = rho * u;
    A.q[2 * A.N + id] = rho * v;
    A.q[3 * A.N + id] = 0.0;
    A.q[4 * A.N + id] = pow(rho, g) / (g - 1.0) + 0.5 * rho * (u * u + v * v);
}

__global__ __launch_bounds__(256) void k_ic_analytic(AnalyticIcArgs A)
{
    const long id = (long)blockIdx.x * 256 + threadIdx.x;
    if (id >= A.N) return;
    const int i = (int)(id % A.nx);
    const long r = id / A.nx;
    const int j = (int)(r % A.ny), k = (int)(r / A.ny);
    // cell centres per rank, src/block.cpp:116-118
    const double x0 = A.lxr * (i / A.nl[0]) + A.dx * (0.5 + i % A.nl[0]) - A.ox;
    const double y0 = A.lyr * (j / A.nl[1]) + A.dy * (0.5 + j % A.nl[1]) - A.oy;
    const double z0 = A.lzr * (k / A.nl[2]) + A.dz * (0.5 + k % A.nl[2]) - A.oz;
    if (A.kind == IC_DEFAULT) { ic_default(A, id); return; }
    if (A.kind == IC_IVC) { ic_ivc(A, id, x0, y0); return; }
    const double g = A.gamma, m = A.mach;
    const double u0 = sin(x0) * cos(y0) * cos(z0);
    const double v0 = -cos(x0) * sin(y0) * cos(z0);
    const double p0 = 1.0 / (g * m * m) + ((cos(2.0 * z0) + 2.0) * (cos(2.0 * x0) + cos(2.0 * y0)) - 2.0) / 16.0;
    A.q[id] = 1.0;
    A.q[A.N + id] = u0;
    A.q[2 * A.N + id] = v0;
    A.q[3 * A.N + id] = 0.0;
    A.q[4 * A.N + id] = p0 / (g - 1.0) + 0.5 * (u0 * u0 + v0 * v0);
}

static int init_analytic(nuc_ctx *x, int kind)
{
    if (!x) return NUC_EINVAL;
    const nuc_config &c = x->c;
    AnalyticIcArgs A{};
    A.kind = kind;
    A.q = gp(state64_w(x));
    A.nx = c.nx; A.ny = c.ny; A.nz = c.nz;
    for (int d = 0; d < 3; d++) A.nl[d] = x->nloc[d];
    A.N = x->N;
    A.dx = c.lx / c.nx; A.dy = c.ly / c.ny; A.dz = c.lz / c.nz;
    A.lxr = c.lx / c.npx; A.lyr = c.ly / c.npy; A.lzr = c.lz / c.npz;
    A.gamma = c.gamma; A.mach = c.mach;
    A.ox = c.origin[0]; A.oy = c.origin[1]; A.oz = c.origin[2];
    x->next_stage = 0;
    hipLaunchKernelGGL(k_ic_analytic, dim3((unsigned)((x->N + 255) / 256)), dim3(256), 0, x->stream, A);
    state64_commit(x);
    NUC_CHECK(x, hipGetLastError());
    NUC_CHECK(x, hipStreamSynchronize(x->stream));
    return NUC_OK;
}

extern "C" int nuc_init_tgv_device(nuc_ctx *x) { return init_analytic(x, IC_TGV); }
extern "C" int nuc_init_default_device(nuc_ctx *x) { return init_analytic(x, IC_DEFAULT); }
extern "C" int nuc_init_ivc_device(nuc_ctx *x) { return init_analytic(x, IC_IVC); }

extern "C" int nuc_init_hit_device(nuc_ctx *x, double k0, double u_rms, unsigned long long seed)
{
    if (!x || !(k0 > 0) || !(u_rms >= 0)) return NUC_EINVAL;
    nuc3d::HitSpectrum S;
    nuc3d::hit_spectrum(x->c, k0, seed, S);
    const long N = x->N;
    const size_t nspec = (size_t)S.W[0] * S.W[1] * S.W[2];
    const size_t ng = 3ull * S.W[2] * S.W[1] * S.n[0], nh = 3ull * S.W[2] * S.n[1] * S.n[0];
    size_t ntw = 0;
    for (int d = 0; d < 3; d++) ntw += S.tw[d].size();
    // one scratch allocation: spectra, twiddles, both passes, the velocity, partials
    const size_t n2 = 3 * nspec + ntw + ng + nh;            // double2 entries
    const size_t bytes = n2 * sizeof(double2) + (3 * (size_t)N + POST_PART + 1) * sizeof(double);
    char *buf = nullptr;
    if (hipMallocAsync(reinterpret_cast<void **>(&buf), bytes, x->stream) != hipSuccess) {
        x->err = "device allocation (initial condition)";
        return NUC_ENOMEM;
    }
    double2 *spec = reinterpret_cast<double2 *>(buf);
    double2 *tw = spec + 3 * nspec;
    double2 *g = tw + ntw;
    double2 *h = g + ng;
    double *vel = reinterpret_cast<double *>(h + nh);
    double *part = vel + 3 * N;
    double *mean = part + POST_PART;
    IcArgs A{};
    for (int cc = 0; cc < 3; cc++) {
        NUC_CHECK(x, hipMemcpyAsync(spec + cc * nspec, S.spec[cc].data(), nspec * sizeof(double2),
                                    hipMemcpyHostToDevice, x->stream));
        A.spec[cc] = gp((const double2 *)(spec + cc * nspec));
    }
    size_t o = 0;
    for (int d = 0; d < 3; d++) {
        NUC_CHECK(x, hipMemcpyAsync(tw + o, S.tw[d].data(), S.tw[d].size() * sizeof(double2), hipMemcpyHostToDevice,
                                    x->stream));
        A.tw[d] = gp((const double2 *)(tw + o));
        o += S.tw[d].size();
    }
    double *q0 = state64_w(x);
    A.g = gp(g); A.h = gp(h); A.vel = gp(vel); A.q = gp(q0);
    for (int d = 0; d < 3; d++) { A.n[d] = S.n[d]; A.W[d] = S.W[d]; }
    A.N = N;
    A.gamma = x->c.gamma; A.mach = x->c.mach; A.scale_num = u_rms;
    x->next_stage = 0;
    auto blocks = [](size_t n) { return dim3((unsigned)((n + 255) / 256)); };
    hipLaunchKernelGGL(k_ic_x, blocks(ng), dim3(256), 0, x->stream, A);
    hipLaunchKernelGGL(k_ic_y, blocks(nh), dim3(256), 0, x->stream, A);
    hipLaunchKernelGGL(k_ic_z, blocks(3 * (size_t)N), dim3(256), 0, x->stream, A);
    hipLaunchKernelGGL(k_ic_sq, blocks((size_t)N), dim3(256), 0, x->stream, A);
    hipLaunchKernelGGL(k_post_mean_part, dim3(POST_PART, 1), dim3(256), 0, x->stream, (const double *)q0, N, part);
    hipLaunchKernelGGL(k_post_mean_fin, dim3(1), dim3(256), 0, x->stream, (const double *)part, 1.0 / (double)N, mean);
    A.mean_sq = gp((const double *)mean);
    hipLaunchKernelGGL(k_ic_q, blocks((size_t)N), dim3(256), 0, x->stream, A);
    state64_commit(x);
    NUC_CHECK(x, hipGetLastError());
    (void)hipFreeAsync(buf, x->stream);
    NUC_CHECK(x, hipStreamSynchronize(x->stream));
    return NUC_OK;
}


// ---------------------------------------------------------------------------
// shell spectra: postproc::solveFFT_kinetic/density/pressure/enstrophy/helicity
// (src/postproc.cpp:931-1184) with fft::solveShell (src/fft.cpp:384-572).
// The 3-D transforms are hipFFT Z2Z, forward (e^{-i}), divided by N as
// src/fft.cpp:626-641; the split, the energies and the shell sums are kernels
// of their own: per-workgroup shell histograms in LDS, then a fixed-order sum
// over workgroups (the LDS accumulation order within a workgroup is not fixed,
// so results may differ by rounding from run to run).
// ---------------------------------------------------------------------------
constexpr int SPEC_BLOCKS = 512;

struct SpecArgs {
    const double2 GMEM *f[3];      // transformed fields (not yet divided by N)
    double GMEM *part;             // [block][spectrum][kmax+1]
    int nx, ny, nz, kmax, nspec;
    long N;
};

__device__ __forceinline__ double sgnk(int i, int n) { return (i >= n / 2.0) ? (double)(i - n) : (double)i; }

// shell of solveShell's rounding: floor(k) or floor(k)+1 (src/fft.cpp:430-435)
__device__ __forceinline__ int shell_of(double kmag)
{
    const double fl = floor(kmag);
    return ((kmag - fl) - 0.5 <= 0) ? (int)fl : (int)fl + 1;
}

__global__ __launch_bounds__(256) void k_spec_kinetic(SpecArgs A)
{
    extern __shared__ double hist[];                 // 5 * (kmax + 1)
    const int nh = 5 * (A.kmax + 1);
    for (int i = threadIdx.x; i < nh; i += 256) hist[i] = 0.0;
    __syncthreads();
    const double invN = (double)A.N;
    for (long id = (long)blockIdx.x * 256 + threadIdx.x; id < A.N; id += (long)gridDim.x * 256) {
        const int i = (int)(id % A.nx);
        const long r = id / A.nx;
        const int j = (int)(r % A.ny), k = (int)(r / A.ny);
        const double kx = sgnk(i, A.nx), ky = sgnk(j, A.ny), kz = sgnk(k, A.nz);
        double kmag = sqrt(kx * kx + ky * ky + kz * kz);
        if (kmag < 1e-6) kmag = 1.0;
        if (!(kmag < A.kmax)) continue;
        const double2 u = A.f[0][id], v = A.f[1][id], w = A.f[2][id];
        const double ur = u.x / invN, ui = u.y / invN, vr = v.x / invN, vi = v.y / invN;
        const double wr = w.x / invN, wi = w.y / invN;
        // src/postproc.cpp:1019-1075
        const double ku_r = (kx * ur + ky * vr + kz * wr) / kmag, ku_i = (kx * ui + ky * vi + kz * wi) / kmag;
        const double s0r = ur - ku_r * kx / kmag, s1r = vr - ku_r * ky / kmag, s2r = wr - ku_r * kz / kmag;
        const double s0i = ui - ku_i * kx / kmag, s1i = vi - ku_i * ky / kmag, s2i = wi - ku_i * kz / kmag;
        const double c0r = ur - s0r, c1r = vr - s1r, c2r = wr - s2r, c0i = ui - s0i, c1i = vi - s1i, c2i = wi - s2i;
        const double x0r = -(ky * s2i - kz * s1i) / kmag, x1r = (kx * s2i - kz * s0i) / kmag;
        const double x2r = -(kx * s1i - ky * s0i) / kmag;
        const double x0i = (ky * s2r - kz * s1r) / kmag, x1i = -(kx * s2r - kz * s0r) / kmag;
        const double x2i = (kx * s1r - ky * s0r) / kmag;
        auto e6 = [](double a, double b, double c, double d, double e, double f) {
            return 0.5 * (a * a + b * b + c * c + d * d + e * e + f * f);
        };
        const double E = e6(ur, vr, wr, ui, vi, wi);
        const double Ec = e6(c0r, c1r, c2r, c0i, c1i, c2i);
        const double Es = e6(s0r, s1r, s2r, s0i, s1i, s2i);
        const double El = e6(0.5 * (s0r - x0r), 0.5 * (s1r - x1r), 0.5 * (s2r - x2r), 0.5 * (s0i - x0i),
                             0.5 * (s1i - x1i), 0.5 * (s2i - x2i));
        const double Er = e6(0.5 * (s0r + x0r), 0.5 * (s1r + x1r), 0.5 * (s2r + x2r), 0.5 * (s0i + x0i),
                             0.5 * (s1i + x1i), 0.5 * (s2i + x2i));
        const int ki = shell_of(kmag);
        const int K1 = A.kmax + 1;
        atomicAdd(&hist[0 * K1 + ki], E);
        atomicAdd(&hist[1 * K1 + ki], Ec);
        atomicAdd(&hist[2 * K1 + ki], Es);
        atomicAdd(&hist[3 * K1 + ki], El);
        atomicAdd(&hist[4 * K1 + ki], Er);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nh; i += 256) A.part[(long)blockIdx.x * nh + i] = hist[i];
}

// |f^|^2 over non-negative indices with i^2+j^2+k^2 < kmax^2 (src/fft.cpp:513-528)
__global__ __launch_bounds__(256) void k_spec_scalar(SpecArgs A)
{
    extern __shared__ double hist[];
    const int nh = A.kmax + 1;
    for (int i = threadIdx.x; i < nh; i += 256) hist[i] = 0.0;
    __syncthreads();
    const double invN = (double)A.N;
    for (long id = (long)blockIdx.x * 256 + threadIdx.x; id < A.N; id += (long)gridDim.x * 256) {
        const int i = (int)(id % A.nx);
        const long r = id / A.nx;
        const int j = (int)(r % A.ny), k = (int)(r / A.ny);
        const long k0 = (long)i * i + (long)j * j + (long)k * k;
        if (!(k0 < (long)A.kmax * A.kmax)) continue;
        const double2 f = A.f[0][id];
        const double fr = f.x / invN, fi = f.y / invN;
        atomicAdd(&hist[shell_of(sqrt((double)k0))], fr * fr + fi * fi);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nh; i += 256) A.part[(long)blockIdx.x * nh + i] = hist[i];
}

__global__ __launch_bounds__(256) void k_spec_sum(const double *part, int nh, int nblocks, double *out)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= nh) return;
    double acc = 0.0;
    for (int b = 0; b < nblocks; b++) acc += part[(long)b * nh + t];
    out[t] = acc;
}

// real field(s) -> complex, and the velocity / pressure / vorticity inputs
struct SpecPrep {
    const double GMEM *q;
    double2 GMEM *c[3];
    double GMEM *uvw;              // 3N: u, v, w (vorticity inputs)
    const double GMEM *grad;       // 9N: d(u,v,w)/d(x,y,z)
    int what;                      // 0 velocity, 1 density, 2 pressure, 3 enstrophy, 4 helicity
    long N;
    double gamma;
};

__global__ __launch_bounds__(256) void k_spec_prep(SpecPrep A)
{
    const long id = (long)blockIdx.x * 256 + threadIdx.x;
    if (id >= A.N) return;
    const double rho = A.q[id];
    const double u = A.q[A.N + id] / rho, v = A.q[2 * A.N + id] / rho, w = A.q[3 * A.N + id] / rho;
    if (A.what == 0) {
        A.c[0][id] = make_double2(u, 0.0);
        A.c[1][id] = make_double2(v, 0.0);
        A.c[2][id] = make_double2(w, 0.0);
    } else if (A.what == 1) {
        A.c[0][id] = make_double2(rho, 0.0);
    } else if (A.what == 2) {
        const double p = fabs((A.q[4 * A.N + id] - 0.5 * rho * (u * u + v * v + w * w)) * (A.gamma - 1.0));
        A.c[0][id] = make_double2(p, 0.0);
    } else if (A.grad == nullptr) {
        A.uvw[id] = u; A.uvw[A.N + id] = v; A.uvw[2 * A.N + id] = w;
    } else {
        // vorticity from the gradients, src/postproc.cpp:815-835
        const double GMEM *g = A.grad;
        const double uy = g[1 * A.N + id], uz = g[2 * A.N + id], vx = g[3 * A.N + id], vz = g[5 * A.N + id];
        const double wx = g[6 * A.N + id], wy = g[7 * A.N + id];
        const double w0 = wy - vz, w1 = uz - wx, w2 = vx - uy;
        const double val = (A.what == 3) ? (w0 * w0 + w1 * w1 + w2 * w2) * 0.5 : w0 * u + w1 * v + w2 * w;
        A.c[0][id] = make_double2(val, 0.0);
    }
}

// the per-point fields of postproc::solveQ after the gradients
// (src/postproc.cpp:815-835): vorticity, enstrophy, Q, helicity, kinetic energy
struct QArgs {
    const double GMEM *q;
    const double GMEM *uvw;
    double GMEM *f;                // 16N: 9 gradients (input), then the 7 outputs
    long N;
};

__global__ __launch_bounds__(256) void k_qfields(QArgs A)
{
    const long id = (long)blockIdx.x * 256 + threadIdx.x;
    if (id >= A.N) return;
    const long N = A.N;
    const double ux = A.f[id], uy = A.f[N + id], uz = A.f[2 * N + id];
    const double vx = A.f[3 * N + id], vy = A.f[4 * N + id], vz = A.f[5 * N + id];
    const double wx = A.f[6 * N + id], wy = A.f[7 * N + id], wz = A.f[8 * N + id];
    const double u0 = A.uvw[id], u1 = A.uvw[N + id], u2 = A.uvw[2 * N + id];
    const double w0 = wy - vz, w1 = uz - wx, w2 = vx - uy;
    const double q0 = ux * vy + ux * wz + vy * wz - uy * vx - uz * wx - vz * wy;
    A.f[9 * N + id] = w0;
    A.f[10 * N + id] = w1;
    A.f[11 * N + id] = w2;
    A.f[12 * N + id] = (w0 * w0 + w1 * w1 + w2 * w2) * 0.5;
    A.f[13 * N + id] = q0;
    A.f[14 * N + id] = w0 * u0 + w1 * u1 + w2 * u2;
    A.f[15 * N + id] = 0.5 * A.q[id] * (u0 * u0 + u1 * u1 + u2 * u2);
}

extern "C" int nuc_q_fields(nuc_ctx *x, double *out)
{
    if (!x || !out) return NUC_EINVAL;
    const long N = x->N;
    double *buf = nullptr;
    if (hipMallocAsync(reinterpret_cast<void **>(&buf), 19 * (size_t)N * sizeof(double), x->stream) != hipSuccess) {
        x->err = "device allocation (Q fields)";
        return NUC_ENOMEM;
    }
    double *uvw = buf, *f = buf + 3 * N;
    const dim3 grid((unsigned)((N + 255) / 256)), blk(256);
    const double *q64 = state64(x);
    SpecPrep P{};
    P.q = gp(q64);
    P.uvw = gp(uvw);
    P.what = 3;
    P.N = N;
    P.gamma = x->c.gamma;
    hipLaunchKernelGGL(k_spec_prep, grid, blk, 0, x->stream, P);
    // the gradients with the viscous scheme (solveGrad_xyz, src/postproc.cpp:759-768)
    for (int c = 0; c < 3; c++)
        for (int d = 0; d < 3; d++) {
            PostDiffArgs D{gp((const double *)(uvw + (size_t)c * N)), gp(f + (size_t)(3 * c + d) * N), N, x->n[0],
                           x->n[1], d, x->n[d], x->nloc[d], x->tsv[d], vis_id(x->c.vis_scheme[d]), 1.0 / x->h[d]};
            hipLaunchKernelGGL(k_post_diff, grid, blk, 0, x->stream, D);
        }
    QArgs Q{gp(q64), gp((const double *)uvw), gp(f), N};
    hipLaunchKernelGGL(k_qfields, grid, blk, 0, x->stream, Q);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, f, 16 * (size_t)N * sizeof(double), hipMemcpyDeviceToHost, x->stream);
    (void)hipFreeAsync(buf, x->stream);
    hipError_t e2 = hipStreamSynchronize(x->stream);
    if (e != hipSuccess || e2 != hipSuccess) {
        x->err = "Q fields";
        return NUC_EHIP;
    }
    return NUC_OK;
}

static int spectra_run(nuc_ctx *x, int what, double *out, int kmax_out)
{
    const long N = x->N;
    const int kmax = std::min(std::min(x->n[0], x->n[1]), x->n[2]) / 2;
    if (kmax_out < kmax) { x->err = "output holds fewer than kmax = min(n)/2 shells"; return NUC_EINVAL; }
    const int ns = (what == 0) ? 5 : 1;
    const int nf = (what == 0) ? 3 : 1;
    const int nh = ns * (kmax + 1);
    const bool vort = (what >= 3);
    size_t bytes = (size_t)nf * N * sizeof(double2) + (size_t)SPEC_BLOCKS * nh * sizeof(double) + nh * sizeof(double);
    if (vort) bytes += 12 * (size_t)N * sizeof(double);
    char *buf = nullptr;
    if (hipMallocAsync(reinterpret_cast<void **>(&buf), bytes, x->stream) != hipSuccess) {
        x->err = "device allocation (spectra)";
        return NUC_ENOMEM;
    }
    double2 *cf = reinterpret_cast<double2 *>(buf);
    double *part = reinterpret_cast<double *>(cf + (size_t)nf * N);
    double *sum = part + (size_t)SPEC_BLOCKS * nh;
    double *uvw = sum + nh, *grad = uvw + 3 * N;
    SpecPrep P{};
    P.q = gp(state64(x));
    for (int c = 0; c < nf; c++) P.c[c] = gp(cf + (size_t)c * N);
    P.what = what;
    P.N = N;
    P.gamma = x->c.gamma;
    const dim3 grid((unsigned)((N + 255) / 256)), blk(256);
    if (vort) {
        P.uvw = gp(uvw);
        P.grad = nullptr;
        hipLaunchKernelGGL(k_spec_prep, grid, blk, 0, x->stream, P);
        for (int f = 0; f < 3; f++)
            for (int d = 0; d < 3; d++) {
                PostDiffArgs D{gp((const double *)(uvw + (size_t)f * N)), gp(grad + (size_t)(3 * f + d) * N), N,
                               x->n[0], x->n[1], d, x->n[d], x->nloc[d], x->tsv[d],
                               vis_id(x->c.vis_scheme[d]), 1.0 / x->h[d]};
                hipLaunchKernelGGL(k_post_diff, grid, blk, 0, x->stream, D);
            }
        P.grad = gp((const double *)grad);
    }
    hipLaunchKernelGGL(k_spec_prep, grid, blk, 0, x->stream, P);
    hipfftHandle plan;
    if (hipfftPlan3d(&plan, x->n[2], x->n[1], x->n[0], HIPFFT_Z2Z) != HIPFFT_SUCCESS) {
        (void)hipFreeAsync(buf, x->stream);
        x->err = "hipfftPlan3d";
        return NUC_EHIP;
    }
    int frc = hipfftSetStream(plan, x->stream);
    for (int c = 0; c < nf && frc == HIPFFT_SUCCESS; c++) {
        hipfftDoubleComplex *d = reinterpret_cast<hipfftDoubleComplex *>(cf + (size_t)c * N);
        frc = hipfftExecZ2Z(plan, d, d, HIPFFT_FORWARD);
    }
    SpecArgs S{};
    for (int c = 0; c < nf; c++) S.f[c] = gp((const double2 *)(cf + (size_t)c * N));
    S.part = gp(part);
    S.nx = x->n[0]; S.ny = x->n[1]; S.nz = x->n[2];
    S.kmax = kmax; S.nspec = ns; S.N = N;
    const size_t lds = (size_t)nh * sizeof(double);
    if (frc == HIPFFT_SUCCESS) {
        if (what == 0) hipLaunchKernelGGL(k_spec_kinetic, dim3(SPEC_BLOCKS), blk, lds, x->stream, S);
        else hipLaunchKernelGGL(k_spec_scalar, dim3(SPEC_BLOCKS), blk, lds, x->stream, S);
        hipLaunchKernelGGL(k_spec_sum, dim3((nh + 255) / 256), blk, 0, x->stream, (const double *)part, nh,
                           SPEC_BLOCKS, sum);
    }
    std::vector<double> h(nh);
    hipError_t e = hipMemcpyAsync(h.data(), sum, nh * sizeof(double), hipMemcpyDeviceToHost, x->stream);
    (void)hipStreamSynchronize(x->stream);
    (void)hipfftDestroy(plan);
    (void)hipFreeAsync(buf, x->stream);
    (void)hipStreamSynchronize(x->stream);
    if (frc != HIPFFT_SUCCESS) { x->err = "hipfftExecZ2Z"; return NUC_EHIP; }
    NUC_CHECK(x, e);
    NUC_CHECK(x, hipGetLastError());
    for (int sidx = 0; sidx < ns; sidx++)
        for (int k = 0; k < kmax; k++) out[sidx * kmax_out + k] = h[sidx * (kmax + 1) + k];
    return NUC_OK;
}

extern "C" int nuc_spectra_kinetic(nuc_ctx *x, double *out, int kmax)
{
    if (!x || !out) return NUC_EINVAL;
    return spectra_run(x, 0, out, kmax);
}

extern "C" int nuc_spectrum_scalar(nuc_ctx *x, int field, double *out, int kmax)
{
    if (!x || !out || field < NUC_SPEC_DENSITY || field > NUC_SPEC_HELICITY) return NUC_EINVAL;
    return spectra_run(x, 1 + field, out, kmax);
}

// ---------------------------------------------------------------------------
// packs: several contexts advanced by one launch per kernel (include/nuc3d_gpu.h)
// ---------------------------------------------------------------------------
static bool packable(const nuc_ctx *x)
{
    // the packed kernels take FP64 state (the FP32-state path runs alone)
    if (x->forcing || x->sources || x->fp32) return false;
    for (int d = 0; d < 3; d++)
        if (x->cpt[d] > PACK_RMAX || x->nseg[d] > 1) return false;
    return true;
}

static bool sweep_full(const nuc_ctx *x, int d)
{
    return x->cpt[d] % 2 == 0 && x->n[d] == 32 * x->cpt[d] && (x->N / x->n[d]) % SWEEP_LPB == 0;
}

// same kernels for every launch of a step: scheme, splitting, model,
// viscous schemes, precision, chunk length and tail-free variant per direction
static bool pack_key_equal(const nuc_ctx *a, const nuc_ctx *b)
{
    if (a->c.riemann != b->c.riemann || a->c.model != b->c.model || a->mixed != b->mixed) return false;
    for (int d = 0; d < 3; d++) {
        if (a->c.scheme[d] != b->c.scheme[d] || a->c.vis_scheme[d] != b->c.vis_scheme[d]) return false;
        if (a->cpt[d] != b->cpt[d] || sweep_full(a, d) != sweep_full(b, d)) return false;
    }
    return true;
}

struct nuc_pack {
    std::vector<nuc_ctx *> m;
    hipStream_t stream = nullptr;
    std::vector<hipEvent_t> ev_in;
    hipEvent_t ev_out = nullptr;
    char *dbuf = nullptr;
    // device tables (pointers into dbuf)
    const PrimArgs *prim = nullptr;
    const DtArgs *dt = nullptr;
    const GradArgs *grad[3] = {nullptr, nullptr, nullptr};
    const SweepArgs *sw[3][3] = {};
    const ResArgs *res = nullptr;
    const int *prim_off = nullptr, *grad_off = nullptr, *sw_off[3] = {nullptr, nullptr, nullptr};
    long prim_total = 0, grad_total = 0, sw_total[3] = {0, 0, 0};
    bool ns = false, mixed = false, fv32 = false;
    int scheme[3] = {0, 0, 0}, riemann = 0, cpt[3] = {0, 0, 0};
    bool full[3] = {false, false, false};
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::string err;
};

extern "C" int nuc_pack_compatible(const nuc_ctx *a, const nuc_ctx *b)
{
    if (!a || !b) return 0;
    return (packable(a) && packable(b) && pack_key_equal(a, b)) ? 1 : 0;
}

static thread_local std::string g_pack_err;

extern "C" const char *nuc_pack_last_error(const nuc_pack *p) { return p ? p->err.c_str() : g_pack_err.c_str(); }

extern "C" int nuc_pack_destroy(nuc_pack *p)
{
    if (!p) return NUC_OK;
    // stream-ordered free: a device-wide hipFree would stall every other
    // context's work (the serving loop drops packs while jobs run)
    if (p->dbuf && p->stream) (void)hipFreeAsync(p->dbuf, p->stream);
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    for (hipEvent_t e : p->ev_in) (void)hipEventDestroy(e);
    if (p->ev_out) (void)hipEventDestroy(p->ev_out);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
    return NUC_OK;
}

extern "C" int nuc_pack_create(nuc_ctx *const *members, int n, int use_graph, nuc_pack **out)
{
    if (!members || !out || n < 1 || n > MAXPACK) { g_pack_err = "pack needs 1..64 contexts"; return NUC_EINVAL; }
    *out = nullptr;
    for (int i = 0; i < n; i++) {
        if (!members[i]) { g_pack_err = "null context in pack"; return NUC_EINVAL; }
        for (int k = 0; k < i; k++)
            if (members[k] == members[i]) { g_pack_err = "context listed twice in pack"; return NUC_EINVAL; }
        if (!packable(members[i])) {
            g_pack_err = "context not packable (forcing or sources on, NUC_FP32, or a line longer than 128 cells)";
            return NUC_EINVAL;
        }
        if (members[i]->next_stage != 0) {
            g_pack_err = "context is in the middle of a step (nuc_rk3_stage)";
            return NUC_EINVAL;
        }
        if (!pack_key_equal(members[0], members[i])) {
            g_pack_err = "contexts in a pack must share scheme, splitting, model, viscous schemes, precision "
                         "and line-length class per direction";
            return NUC_EINVAL;
        }
    }
    nuc_pack *p = new nuc_pack();
    p->m.assign(members, members + n);
    const nuc_ctx *x0 = members[0];
    p->ns = (x0->c.model == NUC_NAVIERSTOKES3D);
    p->mixed = x0->mixed;
    p->fv32 = x0->fv32;
    p->riemann = x0->c.riemann;
    for (int d = 0; d < 3; d++) { p->scheme[d] = x0->c.scheme[d]; p->cpt[d] = x0->cpt[d]; p->full[d] = sweep_full(x0, d); }
    p->use_graph = use_graph != 0;
    auto fail = [&](const char *what, int rc) {
        g_pack_err = what;
        nuc_pack_destroy(p);
        return rc;
    };
    if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) return fail("pack stream", NUC_EHIP);
    p->ev_in.resize(n, nullptr);
    for (int i = 0; i < n; i++)
        if (hipEventCreateWithFlags(&p->ev_in[i], hipEventDisableTiming) != hipSuccess) return fail("pack event", NUC_EHIP);
    if (hipEventCreateWithFlags(&p->ev_out, hipEventDisableTiming) != hipSuccess) return fail("pack event", NUC_EHIP);

    // the three stages of one RK3 step, per member: stage s reads q[s] and
    // writes q[(s+1)%3] (as enqueue_one_step)
    std::vector<StageArgs> sa((size_t)3 * n);
    for (int i = 0; i < n; i++)
        for (int st = 0; st < 3; st++) stage_args(members[i], st, st, (st + 1) % 3, true, sa[(size_t)st * n + i]);
    // host image of the tables: every block 256-byte aligned
    std::vector<char> img;
    auto put = [&](const void *src, size_t bytes) {
        size_t o = (img.size() + 255) & ~(size_t)255;
        img.resize(o + bytes);
        memcpy(img.data() + o, src, bytes);
        return o;
    };
    std::vector<PrimArgs> vp(n);
    std::vector<DtArgs> vd(n);
    std::vector<ResArgs> vr(n);
    std::vector<int> poff(n + 1), goff(n + 1), soff[3];
    for (int d = 0; d < 3; d++) soff[d].resize(n + 1);
    poff[0] = goff[0] = soff[0][0] = soff[1][0] = soff[2][0] = 0;
    for (int i = 0; i < n; i++) {
        const StageArgs &a0 = sa[i];
        vp[i] = a0.prim;
        vd[i] = a0.dt;
        vr[i] = sa[(size_t)2 * n + i].res;
        poff[i + 1] = poff[i] + (int)a0.prim_nb;
        goff[i + 1] = goff[i] + (int)(a0.ggrid.x * a0.ggrid.y * a0.ggrid.z);
        for (int d = 0; d < 3; d++) soff[d][i + 1] = soff[d][i] + (int)a0.sw_nb[d];
    }
    size_t o_prim = put(vp.data(), sizeof(PrimArgs) * n);
    size_t o_dt = put(vd.data(), sizeof(DtArgs) * n);
    size_t o_res = put(vr.data(), sizeof(ResArgs) * n);
    size_t o_poff = put(poff.data(), sizeof(int) * (n + 1));
    size_t o_goff = put(goff.data(), sizeof(int) * (n + 1));
    size_t o_soff[3], o_grad[3], o_sw[3][3];
    for (int d = 0; d < 3; d++) o_soff[d] = put(soff[d].data(), sizeof(int) * (n + 1));
    for (int st = 0; st < 3; st++) {
        std::vector<GradArgs> vg(n);
        for (int i = 0; i < n; i++) vg[i] = sa[(size_t)st * n + i].grad;
        o_grad[st] = put(vg.data(), sizeof(GradArgs) * n);
        for (int d = 0; d < 3; d++) {
            std::vector<SweepArgs> vs(n);
            for (int i = 0; i < n; i++) vs[i] = sa[(size_t)st * n + i].sw[d];
            o_sw[st][d] = put(vs.data(), sizeof(SweepArgs) * n);
        }
    }
    if (hipMallocAsync(reinterpret_cast<void **>(&p->dbuf), img.size(), p->stream) != hipSuccess)
        return fail("pack table allocation", NUC_ENOMEM);
    if (hipMemcpyAsync(p->dbuf, img.data(), img.size(), hipMemcpyHostToDevice, p->stream) != hipSuccess ||
        hipStreamSynchronize(p->stream) != hipSuccess)
        return fail("pack table upload", NUC_EHIP);
    p->prim = reinterpret_cast<const PrimArgs *>(p->dbuf + o_prim);
    p->dt = reinterpret_cast<const DtArgs *>(p->dbuf + o_dt);
    p->res = reinterpret_cast<const ResArgs *>(p->dbuf + o_res);
    p->prim_off = reinterpret_cast<const int *>(p->dbuf + o_poff);
    p->grad_off = reinterpret_cast<const int *>(p->dbuf + o_goff);
    for (int d = 0; d < 3; d++) p->sw_off[d] = reinterpret_cast<const int *>(p->dbuf + o_soff[d]);
    for (int st = 0; st < 3; st++) {
        p->grad[st] = reinterpret_cast<const GradArgs *>(p->dbuf + o_grad[st]);
        for (int d = 0; d < 3; d++) p->sw[st][d] = reinterpret_cast<const SweepArgs *>(p->dbuf + o_sw[st][d]);
    }
    p->prim_total = poff[n];
    p->grad_total = goff[n];
    for (int d = 0; d < 3; d++) p->sw_total[d] = soff[d][n];
    *out = p;
    return NUC_OK;
}

// one RK3 step of every member, the launch order of enqueue_one_step
static void pack_enqueue_step(nuc_pack *p)
{
    const int nj = (int)p->m.size();
    for (int st = 0; st < 3; st++) {
        if (st == 0) {
            hipLaunchKernelGGL(k_prim_pack, dim3((unsigned)p->prim_total), dim3(256), 0, p->stream, p->prim,
                               p->prim_off, nj);
            hipLaunchKernelGGL(k_dt_pack, dim3(nj), dim3(64), 0, p->stream, p->dt);
        }
        if (p->ns) {
            if (p->fv32)
                hipLaunchKernelGGL(k_gradq_pack<float>, dim3((unsigned)p->grad_total), dim3(GTX, GTY, 1), 0, p->stream,
                                   p->grad[st], p->grad_off, nj);
            else
                hipLaunchKernelGGL(k_gradq_pack<double>, dim3((unsigned)p->grad_total), dim3(GTX, GTY, 1), 0, p->stream,
                                   p->grad[st], p->grad_off, nj);
        }
        for (int i = 0; i < 3; i++) {
            const int d = kSweepOrder[i];
            PackLaunch L{p->sw[st][d], p->sw_off[d], nj, p->sw_total[d], p->mixed, p->full[d]};
            launch_sweep_pack(p->scheme[d], p->riemann, p->cpt[d], L, p->stream);
        }
        if (st == 2) hipLaunchKernelGGL(k_res_pack, dim3(nj), dim3(256), 0, p->stream, p->res);
    }
}

extern "C" int nuc_pack_step_async(nuc_pack *p, int nsteps)
{
    if (!p || nsteps < 0) return NUC_EINVAL;
    if (nsteps == 0) return NUC_OK;
    for (nuc_ctx *x : p->m)
        if (x->next_stage != 0) { p->err = "a member is in the middle of a step (nuc_rk3_stage)"; return NUC_EINVAL; }
    // the pack's stream waits for each member's earlier work ...
    for (size_t i = 0; i < p->m.size(); i++) {
        NUC_CHECK(p, hipEventRecord(p->ev_in[i], p->m[i]->stream));
        NUC_CHECK(p, hipStreamWaitEvent(p->stream, p->ev_in[i], 0));
    }
    if (p->use_graph) {
        if (!p->gexec) {
            NUC_CHECK(p, hipStreamBeginCapture(p->stream, hipStreamCaptureModeRelaxed));
            pack_enqueue_step(p);
            hipError_t e = hipStreamEndCapture(p->stream, &p->graph);
            if (e != hipSuccess) { p->err = std::string("pack graph capture: ") + hipGetErrorString(e); return NUC_EHIP; }
            NUC_CHECK(p, hipGraphInstantiate(&p->gexec, p->graph, nullptr, nullptr, 0));
        }
        for (int it = 0; it < nsteps; it++) NUC_CHECK(p, hipGraphLaunch(p->gexec, p->stream));
    } else {
        for (int it = 0; it < nsteps; it++) pack_enqueue_step(p);
        NUC_CHECK(p, hipGetLastError());
    }
    // ... and each member's later work waits for the pack's steps
    NUC_CHECK(p, hipEventRecord(p->ev_out, p->stream));
    for (nuc_ctx *x : p->m) {
        NUC_CHECK(p, hipStreamWaitEvent(x->stream, p->ev_out, 0));
        x->pending += nsteps;
    }
    return NUC_OK;
}
