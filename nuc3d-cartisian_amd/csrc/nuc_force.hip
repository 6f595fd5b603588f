// nuc_force.hip -- Const forcing (src/physicsModel.cpp:1368-1677) as a pruned
// separable DFT on the device.  See nuc_force.h for the data flow.
//
// Reference steps and where they live here:
//   forward FFT of u,v,w, normalised by 1/N (src/fft.cpp:182-288)   k_fx, k_fy, k_fband
//   solenoidal/compressive split, helical mix sigma (:1476-1531)     k_fband
//   band energies E_s, E_c over |kmag-k_f|<0.5 with the kmag<1e-6
//   -> 1 substitution (:1483-1484,1534-1557), theta (:1559)          k_fband
//   forcing spectrum (theta-1) v_s on the shell, unnormalised
//   backward FFT, real part (:1560-1603)                              k_fg, k_fapply
//   <f.u> correction and RHS update, rho taken as 1 (:1605-1667)     k_fapply, k_fmean, last sweep
#include "nuc_force.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace nucforce {
namespace {

struct Dev {
    int nx, ny, nz, na, nb, nc, nm;
    long NL;  // ny*nz x lines
    double invN, amp, amp_hel;
    const int *modes, *astart, *order, *kv;
    const double *tw;  // x at 0, y at 2nx, z at 2(nx+ny)
    double *X, *Y, *fh, *part, *diag;
};

__device__ __forceinline__ int wrapmod(int a, int n)
{
    int r = a % n;
    return r < 0 ? r + n : r;
}

// sums K values over a 256-thread block; thread k < K returns sum k
template <int K>
__device__ __forceinline__ void block_sum(double v[K], double *sred)
{
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
#pragma unroll
    for (int k = 0; k < K; k++)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; k++) sred[wave * K + k] = v[k];
    __syncthreads();
    if (t < K) v[0] = ((sred[t] + sred[K + t]) + sred[2 * K + t]) + sred[3 * K + t];
    __syncthreads();
}

// X[c][a][L] = sum_i u_c(i, L) e^{-2 pi i kx_a i / nx}; one block per x line
// NS: cells per thread (lines of up to 256 NS cells)
template <int NS>
__global__ __launch_bounds__(256) void k_fx(Dev D, const double *u, const double *v, const double *w)
{
    __shared__ double sred[4 * 6];
    const long L = blockIdx.x;
    const int t = threadIdx.x;
    double uu[NS][3];
    int ii[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        int i = t + 256 * s;
        ii[s] = i;
        bool ok = i < D.nx;
        long id = L * D.nx + i;
        uu[s][0] = ok ? u[id] : 0.0;
        uu[s][1] = ok ? v[id] : 0.0;
        uu[s][2] = ok ? w[id] : 0.0;
    }
    for (int a = 0; a < D.na; a++) {
        const int kx = D.kv[a];
        double acc[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < NS; s++) {
            if (ii[s] < D.nx) {
                int ph = wrapmod(kx * ii[s], D.nx);
                double cr = D.tw[2 * ph], ci = D.tw[2 * ph + 1];
#pragma unroll
                for (int c = 0; c < 3; c++) { acc[2 * c] += uu[s][c] * cr; acc[2 * c + 1] += uu[s][c] * ci; }
            }
        }
        block_sum<6>(acc, sred);
        if (t < 6) {
            int c = t >> 1, ri = t & 1;
            D.X[(((long)c * D.na + a) * D.NL + L) * 2 + ri] = acc[0];
        }
    }
}

// Y[c][a][b][k] = sum_j X[c][a][j + ny k] e^{-2 pi i ky_b j / ny}; one block per (a, k)
template <int NS>
__global__ __launch_bounds__(256) void k_fy(Dev D)
{
    __shared__ double sred[4 * 6];
    const int a = blockIdx.x % D.na, k = blockIdx.x / D.na;
    const int t = threadIdx.x;
    double xr[NS][3], xi[NS][3];
    int jj[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        int j = t + 256 * s;
        jj[s] = j;
        bool ok = j < D.ny;
        long L = j + (long)D.ny * k;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            long o = (((long)c * D.na + a) * D.NL + L) * 2;
            xr[s][c] = ok ? D.X[o] : 0.0;
            xi[s][c] = ok ? D.X[o + 1] : 0.0;
        }
    }
    const double *twy = D.tw + 2 * D.nx;
    for (int b = 0; b < D.nb; b++) {
        const int ky = D.kv[MAXK + b];
        double acc[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < NS; s++) {
            if (jj[s] < D.ny) {
                int ph = wrapmod(ky * jj[s], D.ny);
                double cr = twy[2 * ph], ci = twy[2 * ph + 1];
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    acc[2 * c] += xr[s][c] * cr - xi[s][c] * ci;
                    acc[2 * c + 1] += xr[s][c] * ci + xi[s][c] * cr;
                }
            }
        }
        block_sum<6>(acc, sred);
        if (t < 6) {
            int c = t >> 1, ri = t & 1;
            D.Y[((((long)c * D.na + a) * D.nb + b) * D.nz + k) * 2 + ri] = acc[0];
        }
    }
}

// z transform at the shell modes, projection, helical mix, band energies,
// theta and the forcing spectrum; one block
__global__ __launch_bounds__(256) void k_fband(Dev D)
{
    __shared__ double ses[MAXMODES], sec[MAXMODES];
    __shared__ double stheta;
    const int t = threadIdx.x;
    const double *twz = D.tw + 2 * (D.nx + D.ny);
    for (int m = t; m < D.nm; m += 256) {
        const int ia = D.modes[4 * m], ib = D.modes[4 * m + 1], ic = D.modes[4 * m + 2];
        const double kx = D.kv[ia], ky = D.kv[MAXK + ib], kz = D.kv[2 * MAXK + ic];
        const int kzi = D.kv[2 * MAXK + ic];
        double ure[3], uim[3];
        for (int c = 0; c < 3; c++) {
            const double *y = D.Y + (((long)c * D.na + ia) * D.nb + ib) * D.nz * 2;
            double zr = 0.0, zi = 0.0;
            for (int k = 0; k < D.nz; k++) {
                int ph = wrapmod(kzi * k, D.nz);
                double cr = twz[2 * ph], ci = twz[2 * ph + 1];
                zr += y[2 * k] * cr - y[2 * k + 1] * ci;
                zi += y[2 * k] * ci + y[2 * k + 1] * cr;
            }
            ure[c] = zr * D.invN;
            uim[c] = zi * D.invN;
        }
        double km = sqrt(kx * kx + ky * ky + kz * kz);
        if (km < 1e-6) km = 1.0;
        const double kk[3] = {kx, ky, kz};
        // longitudinal part (k.u/|k|) k/|k| and the solenoidal remainder
        double kur = (kx * ure[0] + ky * ure[1] + kz * ure[2]) / km;
        double kui = (kx * uim[0] + ky * uim[1] + kz * uim[2]) / km;
        double sr[3], si[3], cr[3], ci[3];
        for (int c = 0; c < 3; c++) {
            sr[c] = ure[c] - kur * kk[c] / km;
            si[c] = uim[c] - kui * kk[c] / km;
            cr[c] = ure[c] - sr[c];
            ci[c] = uim[c] - si[c];
        }
        // helical mix v_s = s + sigma (i k x s)/|k|
        double hr[3], hi[3];
        hr[0] = -(ky * si[2] - kz * si[1]) / km;
        hr[1] = (kx * si[2] - kz * si[0]) / km;
        hr[2] = -(kx * si[1] - ky * si[0]) / km;
        hi[0] = (ky * sr[2] - kz * sr[1]) / km;
        hi[1] = -(kx * sr[2] - kz * sr[0]) / km;
        hi[2] = (kx * sr[1] - ky * sr[0]) / km;
        for (int c = 0; c < 3; c++) {
            sr[c] = sr[c] + D.amp_hel * hr[c];
            si[c] = si[c] + D.amp_hel * hi[c];
        }
        ses[m] = 0.5 * (sr[0] * sr[0] + sr[1] * sr[1] + sr[2] * sr[2] + si[0] * si[0] + si[1] * si[1] + si[2] * si[2]);
        sec[m] = 0.5 * (cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2] + ci[0] * ci[0] + ci[1] * ci[1] + ci[2] * ci[2]);
        for (int c = 0; c < 3; c++) { D.fh[6 * m + 2 * c] = sr[c]; D.fh[6 * m + 2 * c + 1] = si[c]; }
    }
    __syncthreads();
    if (t == 0) {
        double es = 0.0, ec = 0.0;
        for (int m = 0; m < D.nm; m++) { es += ses[m]; ec += sec[m]; }
        double th = sqrt(fabs(D.amp - ec) / es);
        stheta = th;
        D.diag[0] = es; D.diag[1] = ec; D.diag[2] = th;
    }
    __syncthreads();
    const double g = stheta - 1.0;
    for (int m = t; m < D.nm; m += 256) {
        const bool forced = D.modes[4 * m + 3] != 0;
        for (int q = 0; q < 6; q++) D.fh[6 * m + q] = forced ? D.fh[6 * m + q] * g : 0.0;
    }
}

// G[c][a][L] = sum over shell modes with kx = kx_a of fh e^{+2 pi i (ky j/ny + kz k/nz)}
__global__ __launch_bounds__(256) void k_fg(Dev D)
{
    const long L = (long)blockIdx.x * 256 + threadIdx.x;
    if (L >= D.NL) return;
    const int j = (int)(L % D.ny), k = (int)(L / D.ny);
    const double *twy = D.tw + 2 * D.nx, *twz = D.tw + 2 * (D.nx + D.ny);
    for (int a = 0; a < D.na; a++) {
        double acc[6] = {0, 0, 0, 0, 0, 0};
        for (int q = D.astart[a]; q < D.astart[a + 1]; q++) {
            const int m = D.order[q];
            if (!D.modes[4 * m + 3]) continue;
            int py = wrapmod(D.kv[MAXK + D.modes[4 * m + 1]] * j, D.ny);
            int pz = wrapmod(D.kv[2 * MAXK + D.modes[4 * m + 2]] * k, D.nz);
            // conj(tw_y) * conj(tw_z)
            double yr = twy[2 * py], yi = -twy[2 * py + 1], zr = twz[2 * pz], zi = -twz[2 * pz + 1];
            double er = yr * zr - yi * zi, ei = yr * zi + yi * zr;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                double fr = D.fh[6 * m + 2 * c], fi = D.fh[6 * m + 2 * c + 1];
                acc[2 * c] += fr * er - fi * ei;
                acc[2 * c + 1] += fr * ei + fi * er;
            }
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            long o = (((long)c * D.na + a) * D.NL + L) * 2;
            D.X[o] = acc[2 * c];
            D.X[o + 1] = acc[2 * c + 1];
        }
    }
}

// f(i, L) = Re sum_a G[c][a][L] e^{+2 pi i kx_a i/nx}; RHS initialisation and <f.u> partials
struct ApplyOut {
    double *r[5];
};

template <int NS>
__global__ __launch_bounds__(256) void k_fapply(Dev D, const double *u, const double *v, const double *w, ApplyOut O)
{
    __shared__ double sg[3 * MAXK * 2];
    __shared__ double sred[4];
    const long L = blockIdx.x;
    const int t = threadIdx.x;
    for (int q = t; q < 3 * D.na * 2; q += 256) {
        int ri = q & 1, ca = q >> 1;  // ca = c*na + a
        sg[q] = D.X[((long)ca * D.NL + L) * 2 + ri];
    }
    __syncthreads();
    double fu_sum = 0.0;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        int i = t + 256 * s;
        if (i >= D.nx) continue;
        double f[3] = {0.0, 0.0, 0.0};
        for (int a = 0; a < D.na; a++) {
            int ph = wrapmod(D.kv[a] * i, D.nx);
            double tr = D.tw[2 * ph], ti = D.tw[2 * ph + 1];
#pragma unroll
            for (int c = 0; c < 3; c++) f[c] += sg[2 * (c * D.na + a)] * tr + sg[2 * (c * D.na + a) + 1] * ti;
        }
        long id = L * D.nx + i;
        double fu = f[0] * u[id] + f[1] * v[id] + f[2] * w[id];
        O.r[0][id] = 0.0;
        O.r[1][id] = -f[0];
        O.r[2][id] = -f[1];
        O.r[3][id] = -f[2];
        O.r[4][id] = -fu;
        fu_sum += fu;
    }
    double acc[1] = {fu_sum};
    block_sum<1>(acc, sred);
    if (t == 0) D.part[L] = acc[0];
}

__global__ __launch_bounds__(256) void k_fmean(Dev D, double *mean_out)
{
    __shared__ double sred[4];
    double a = 0.0;
    for (long L = threadIdx.x; L < D.NL; L += 256) a += D.part[L];
    double acc[1] = {a};
    block_sum<1>(acc, sred);
    if (threadIdx.x == 0) *mean_out = acc[0] * D.invN;
}

Dev make_dev(const Plan &p)
{
    Dev D;
    D.nx = p.nx; D.ny = p.ny; D.nz = p.nz;
    D.na = p.na; D.nb = p.nb; D.nc = p.nc; D.nm = p.nm;
    D.NL = (long)p.ny * p.nz;
    D.invN = 1.0 / (double)p.N;
    D.amp = p.amp; D.amp_hel = p.amp_hel;
    D.modes = p.modes; D.astart = p.astart; D.order = p.astart + (p.na + 1); D.kv = p.kv;
    D.tw = p.tw; D.X = p.X; D.Y = p.Y; D.fh = p.fh; D.part = p.part; D.diag = p.diag;
    return D;
}

struct HostModes {
    std::vector<int> kx, ky, kz, forced;
};

// shell modes in the order the reference's loops visit them: wave numbers
// folded as index i >= n/2 -> i - n (src/physicsModel.cpp:1483-1484), kmag < 1e-6
// counted as 1 for the energy sums, forced only where the unsubstituted
// |kmag - k_f| < 0.5 (:1534,1560-1580)
HostModes shell_modes(const nuc_config &c)
{
    HostModes h;
    const int n[3] = {c.nx, c.ny, c.nz};
    const int K = (int)std::ceil(c.forcing_k + 0.5) + 1;
    for (int kz = -K; kz <= K; kz++)
        for (int ky = -K; ky <= K; ky++)
            for (int kx = -K; kx <= K; kx++) {
                const int kk[3] = {kx, ky, kz};
                bool same = true;
                for (int d = 0; d < 3; d++) {
                    int i = ((kk[d] % n[d]) + n[d]) % n[d];
                    double k0 = (i >= n[d] / 2.0) ? i - n[d] : i;
                    if (k0 != kk[d]) same = false;
                }
                if (!same) continue;
                double km = std::sqrt((double)kx * kx + (double)ky * ky + (double)kz * kz);
                double kms = km < 1e-6 ? 1.0 : km;
                if (std::fabs(kms - c.forcing_k) < 0.5) {
                    h.kx.push_back(kx); h.ky.push_back(ky); h.kz.push_back(kz);
                    h.forced.push_back(std::fabs(km - c.forcing_k) < 0.5 ? 1 : 0);
                }
            }
    return h;
}

template <typename T>
bool dalloc(T **dst, size_t count, hipStream_t s)
{
    return hipMallocAsync(reinterpret_cast<void **>(dst), std::max<size_t>(1, count) * sizeof(T), s) == hipSuccess;
}

// the host vector must stay alive until the stream has been synchronised
template <typename T>
bool upload(T **dst, const std::vector<T> &src, hipStream_t s)
{
    if (!dalloc(dst, src.size(), s)) return false;
    return src.empty()
           || hipMemcpyAsync(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, s) == hipSuccess;
}

std::vector<int> distinct(const std::vector<int> &v)
{
    std::vector<int> d(v);
    std::sort(d.begin(), d.end());
    d.erase(std::unique(d.begin(), d.end()), d.end());
    return d;
}

}  // namespace

long long plan_bytes(const nuc_config &c)
{
    HostModes h = shell_modes(c);
    long long na = (long long)distinct(h.kx).size(), nb = (long long)distinct(h.ky).size();
    long long nl = (long long)c.ny * c.nz;
    return (3 * na * nl * 2 + 3 * na * nb * c.nz * 2 + (long long)h.kx.size() * 6 + nl + 8
            + 2LL * (c.nx + c.ny + c.nz)) * 8 + 16384;
}

int plan_create(Plan &p, const nuc_config &c, hipStream_t s, std::string &err)
{
    p = Plan();
    p.nx = c.nx; p.ny = c.ny; p.nz = c.nz;
    p.N = (long)c.nx * c.ny * c.nz;
    p.amp = c.forcing_amp; p.amp_hel = c.forcing_amp_hel;
    if (c.nx > 2048 || c.ny > 2048) { err = "forcing: x or y line length above 2048"; return NUC_EINVAL; }
    HostModes h = shell_modes(c);
    p.nm = (int)h.kx.size();
    if (p.nm < 1) { err = "forcing: the shell |k - k_f| < 0.5 holds no resolved mode"; return NUC_EINVAL; }
    if (p.nm > MAXMODES) { err = "forcing: more than 1024 shell modes"; return NUC_EINVAL; }
    std::vector<int> ax = distinct(h.kx), ay = distinct(h.ky), az = distinct(h.kz);
    p.na = (int)ax.size(); p.nb = (int)ay.size(); p.nc = (int)az.size();
    if (p.na > MAXK || p.nb > MAXK || p.nc > MAXK) { err = "forcing: more than 32 distinct wave numbers"; return NUC_EINVAL; }
    std::vector<int> kv(3 * MAXK, 0);
    for (int i = 0; i < p.na; i++) kv[i] = ax[i];
    for (int i = 0; i < p.nb; i++) kv[MAXK + i] = ay[i];
    for (int i = 0; i < p.nc; i++) kv[2 * MAXK + i] = az[i];
    std::vector<int> modes(4 * p.nm);
    for (int m = 0; m < p.nm; m++) {
        modes[4 * m] = (int)(std::lower_bound(ax.begin(), ax.end(), h.kx[m]) - ax.begin());
        modes[4 * m + 1] = (int)(std::lower_bound(ay.begin(), ay.end(), h.ky[m]) - ay.begin());
        modes[4 * m + 2] = (int)(std::lower_bound(az.begin(), az.end(), h.kz[m]) - az.begin());
        modes[4 * m + 3] = h.forced[m];
    }
    // astart (na+1) followed by the mode indices grouped by ia
    std::vector<int> ao(p.na + 1 + p.nm, 0);
    {
        std::vector<int> cnt(p.na, 0);
        for (int m = 0; m < p.nm; m++) cnt[modes[4 * m]]++;
        for (int a = 0; a < p.na; a++) ao[a + 1] = ao[a] + cnt[a];
        std::vector<int> fill(ao.begin(), ao.begin() + p.na);
        for (int m = 0; m < p.nm; m++) ao[p.na + 1 + fill[modes[4 * m]]++] = m;
    }
    // e^{-2 pi i p/n} with the reference's angle reduction p = (k i) mod n (src/fft.cpp)
    const double twopi = 8.0 * std::atan(1.0);
    std::vector<double> tw;
    for (int n : {c.nx, c.ny, c.nz})
        for (int ph = 0; ph < n; ph++) {
            double ang = -twopi * (double)ph / n;
            tw.push_back(std::cos(ang));
            tw.push_back(std::sin(ang));
        }
    const long NL = (long)c.ny * c.nz;
    bool ok = upload(&p.modes, modes, s) && upload(&p.astart, ao, s) && upload(&p.kv, kv, s) && upload(&p.tw, tw, s)
              && dalloc(&p.X, 3L * p.na * NL * 2, s) && dalloc(&p.Y, 3L * p.na * p.nb * c.nz * 2, s)
              && dalloc(&p.fh, (size_t)p.nm * 6, s) && dalloc(&p.part, NL, s) && dalloc(&p.diag, 8, s);
    ok = (hipStreamSynchronize(s) == hipSuccess) && ok;
    if (!ok) {
        plan_destroy(p, s);
        err = "forcing: device allocation failed";
        return NUC_ENOMEM;
    }
    return NUC_OK;
}

void plan_destroy(Plan &p, hipStream_t s)
{
    void *bufs[] = {p.modes, p.astart, p.kv, p.tw, p.X, p.Y, p.fh, p.part, p.diag};
    for (void *b : bufs)
        if (b) (void)hipFreeAsync(b, s);
    (void)hipStreamSynchronize(s);
    p = Plan();
}

void enqueue(const Plan &p, const double *u, const double *v, const double *w, double *const rhs[5], double *mean_out,
             hipStream_t s)
{
    Dev D = make_dev(p);
    const long NL = D.NL;
    // x and y lines of up to 512 cells take 2 cells per thread, longer ones 8
    const bool lx = p.nx > 512, ly = p.ny > 512;
    if (lx) hipLaunchKernelGGL(k_fx<8>, dim3((unsigned)NL), dim3(256), 0, s, D, u, v, w);
    else hipLaunchKernelGGL(k_fx<2>, dim3((unsigned)NL), dim3(256), 0, s, D, u, v, w);
    if (ly) hipLaunchKernelGGL(k_fy<8>, dim3((unsigned)(p.na * p.nz)), dim3(256), 0, s, D);
    else hipLaunchKernelGGL(k_fy<2>, dim3((unsigned)(p.na * p.nz)), dim3(256), 0, s, D);
    hipLaunchKernelGGL(k_fband, dim3(1), dim3(256), 0, s, D);
    hipLaunchKernelGGL(k_fg, dim3((unsigned)((NL + 255) / 256)), dim3(256), 0, s, D);
    ApplyOut O;
    for (int q = 0; q < 5; q++) O.r[q] = rhs[q];
    if (lx) hipLaunchKernelGGL(k_fapply<8>, dim3((unsigned)NL), dim3(256), 0, s, D, u, v, w, O);
    else hipLaunchKernelGGL(k_fapply<2>, dim3((unsigned)NL), dim3(256), 0, s, D, u, v, w, O);
    hipLaunchKernelGGL(k_fmean, dim3(1), dim3(256), 0, s, D, mean_out);
}

}  // namespace nucforce
