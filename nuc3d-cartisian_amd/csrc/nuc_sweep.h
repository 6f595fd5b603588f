// nuc_sweep.h -- the direction sweep (k_sweep) and its launchers, shared by
// the library's main translation unit (which launches them through
// launch_sweep_s / launch_sweep_pack_s) and the per-scheme translation units
// (sweep_inst.hip) that instantiate them, so that the schemes' kernels build
// in parallel.
#pragma once
#include "nuc_device.h"
#include "nuc_lines.h"

#include <hip/hip_runtime.h>
#include <algorithm>
#include <type_traits>

// ---------------------------------------------------------------------------
// kernel arguments
//
// Buffer pointers of the hot-path kernels are declared in the global address
// space for the device (GMEM): the same bytes reach a kernel either as its
// argument or read from a device table (packed launches), and a pointer read
// from memory is otherwise a generic address to the compiler, accessed with
// flat instructions that also hold the LDS wait counter.  The host pass sees
// plain pointers of the same size.
// ---------------------------------------------------------------------------
#if defined(__HIP_DEVICE_COMPILE__)
#define GMEM __attribute__((address_space(1)))
#else
#define GMEM
#endif

namespace nucsweep {
using namespace nucdev;

// a buffer pointer as stored in the argument structs (an explicit cast for
// the device pass, the identity for the host pass)
template <typename T>
__host__ __device__ inline T GMEM *gp(T *p)
{
    return (T GMEM *)p;
}

// whether a sweep applies the RK3 update: never, always, or as A.epilogue says
enum { EPI_NO = 0, EPI_YES = 1, EPI_ARG = 2 };

struct SweepArgs {
    int dir, n, nloc, ts, tsv, vis, riemann;
    int nx, ny, nz;
    long nlines;
    // lines longer than 512 cells: nseg segments of n cells each (n is then the
    // segment length and nfull the line's); otherwise nseg = 1, nfull = n
    int nseg, nfull;
    double inv_h, gamma;
    // state fields: element type S of the kernel (double; float with NUC_FP32)
    const void GMEM *q[5];
    const void GMEM *fv[4];   // viscous flux components (float when fv_f32, else double)
    int fv_f32;
    void GMEM *rhs[5];        // S, as the state
    int stage, epilogue;
    int first;                // the stage's first sweep: writes rhs without reading it
    int accum;                // the first sweep adds to rhs (forcing initialised it)
    int mixed;                // FP32 stencils (host-side dispatch only)
    int fp32;                 // FP32 state and RHS (NUC_FP32; host-side dispatch only)
    const double GMEM *fmean; // last sweep: <f.u> added back to the energy RHS
    const void GMEM *qn[5];
    void GMEM *qout[5];
    const double GMEM *dtp;
    double GMEM *res_part;
};

// Sweep order within a stage: z, y, then x.  The last sweep applies the RK3
// update (reads u^n and the accumulated rhs, writes the new state), and x is
// the direction whose loads are contiguous.
// ---------------------------------------------------------------------------
// the direction sweep
//
// Load mapping (global <-> LDS): for the x sweep thread t owns cells
// lane + 64k of line t/64 (contiguous, coalesced); for y and z the LPB lines
// of a workgroup are consecutive in x and thread t owns cells t/LPB + 64k of
// line t%LPB, so each load instruction reads LPB consecutive doubles per cell.
// Compute mapping: wave w owns line w (both signs, see nuc_lines.h).
// ---------------------------------------------------------------------------
template <int RIE, typename T>
__device__ __forceinline__ void split_params(const T q[5], int d, T gamma, T sp[5], bool ausmp = false)
{
    const T gm1 = gamma - T(1);
    T rho = q[0], ir = T(1) / rho;
    T u = q[1] * ir, vv = q[2] * ir, w = q[3] * ir;
    T p = fabs((q[4] - T(0.5) * rho * (u * u + vv * vv + w * w)) * gm1);
    T cs = sqrt(gamma * p * ir);
    T ud = (d == 0) ? u : (d == 1 ? vv : w);
    if constexpr (RIE == RIE_LF) {
        sp[0] = ud; sp[1] = fabs(ud) + cs; sp[2] = p; sp[3] = T(0); sp[4] = T(0);
    } else {
        // AUSM Mach / pressure splittings, src/physicsModel.cpp:1181-1228
        T mach = ud / cs;
        T mp, mn, pp, pn;
        if (fabs(mach) < T(1)) {
            mp = T(0.25) * sq(mach + T(1)) + T(0.125) * sq(mach * mach - T(1));
            mn = T(-0.25) * sq(mach - T(1)) - T(0.125) * sq(mach * mach - T(1));
            if (ausmp) {
                // AUSM+ pressure splitting, src/physicsModel.cpp:1257-1277
                const T a = T(0.1875) * mach * sq(mach * mach - T(1));
                pp = p * (T(0.25) * sq(mach + T(1)) * (T(2) - mach) + a);
                pn = p * (T(0.25) * sq(mach - T(1)) * (T(2) + mach) - a);
            } else {
                pp = p * (mach + T(1)) / T(2);
                pn = -p * (mach - T(1)) / T(2);
            }
        } else {
            mp = T(0.5) * (mach + fabs(mach));
            mn = T(0.5) * (mach - fabs(mach));
            pp = T(0.5) * p * (mach + fabs(mach)) / mach;
            pn = T(0.5) * p * (mach - fabs(mach)) / mach;
        }
        sp[0] = mp * cs; sp[1] = mn * cs; sp[2] = pp; sp[3] = pn; sp[4] = p;
    }
}

// split fluxes of variable v at one cell (LF src/physicsModel.cpp:540-565,
// AUSM :790-820, per direction d)
template <int RIE, typename T>
__device__ __forceinline__ void split_var(const T *sp, T qv, int v, int d, T &fp, T &fm)
{
    if constexpr (RIE == RIE_LF) {
        T ud = sp[0], lam = sp[1], p = sp[2];
        T F = ud * qv;
        if (v == 1 + d) F = ud * qv + p;
        if (v == 4) F = ud * (qv + p);
        fp = T(0.5) * (F + lam * qv);
        fm = T(0.5) * (F - lam * qv);
    } else {
        fp = sp[0] * qv;
        fm = sp[1] * qv;
        if (v == 1 + d) { fp += sp[2]; fm += sp[3]; }
        if (v == 4) { fp = sp[0] * qv + sp[0] * sp[4]; fm = sp[1] * qv + sp[1] * sp[4]; }
    }
}

// Element of a field at a 32-bit byte offset, through a buffer resource
// (base address and size in scalar registers): the offset is the only
// per-lane operand, and an offset past the field's `bytes` reads 0 and drops
// the store.  Word 3 (0x00020000) selects 32-bit data, the raw-buffer form.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t field_rsrc(const void GMEM *p, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(reinterpret_cast<const void *>(p)), 0, (int)bytes,
                                             0x00020000);
}
__device__ __forceinline__ double ld_at(const double GMEM *p, unsigned ob, unsigned bytes)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(field_rsrc(p, bytes), ob, 0, 0));
}
__device__ __forceinline__ float ld_at(const float GMEM *p, unsigned ob, unsigned bytes)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(field_rsrc(p, bytes), ob, 0, 0));
}
__device__ __forceinline__ void st_at(double GMEM *p, unsigned ob, unsigned bytes, double v)
{
    using u2 = __attribute__((ext_vector_type(2))) unsigned;
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), field_rsrc(p, bytes), ob, 0, 0);
}
__device__ __forceinline__ void st_at(float GMEM *p, unsigned ob, unsigned bytes, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), field_rsrc(p, bytes), ob, 0, 0);
}

// Workgroup index with consecutive indices placed on the same XCD: the
// dispatcher deals workgroups round-robin over the 8 XCDs (each with its own
// L2), so physical block p runs on XCD p%8; giving that XCD the contiguous
// logical range makes neighbouring line groups, which share cache lines in
// the y/z sweeps, meet in one L2.
__device__ __forceinline__ long xcd_block(unsigned p, unsigned nb)
{
    const unsigned q = nb / 8, rem = nb % 8, x = p % 8, i = p / 8;
    return (long)x * q + (x < rem ? x : rem) + i;
}

// Line-buffer synchronisation: in the x sweep every wave loads and computes
// its own line, so a wave-level barrier suffices (LDS operations of one wave
// complete in order); the y/z load mapping spreads a line over all waves.
__device__ __forceinline__ void sync_wave()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void sync_lines(int d)
{
    if (d == 0) sync_wave();
    else __syncthreads();
}

// T: type of the split fluxes, line buffers and stencils (double, or float
// for the mixed path); weights, tridiagonal solves, interface values, the
// flux difference and the RHS/RK3 accumulation are FP64 in both.
// FULL: n = 32R and every workgroup's lines exist, so no bounds checks.
// EPI: EPI_NO / EPI_YES fix at compile time whether this sweep applies the
// RK3 update (the last sweep of a stage); EPI_ARG reads it from A.epilogue.
// The last sweep keeps the current state's values in registers from the load
// phase to the update; the other forms read them again in the final phase,
// so that the first two sweeps do not hold them across the reconstruction.
// blk: the workgroup's logical index within its job (line group blk*LPB ..)
// SEG: a wave owns one segment of a line longer than 512 cells instead of a
// whole line; the 2*HALO cells around the segment are read from the
// neighbouring segments (periodic over the full line) and the segment's n+1
// interfaces are all computed, so that no periodic image is needed.
template <int SCH, int RIE, int R, int LPB, typename T, bool FULL, int EPI, typename S = double, bool SEG = false>
__device__ __forceinline__ void sweep_body(const SweepArgs &A, const long blk)
{
    const bool epi = EPI == EPI_ARG ? A.epilogue != 0 : EPI == EPI_YES;
    constexpr int NP = Skew<R>::pitch(32 * R);
    constexpr int NV = plain_pitch(32 * R);
    constexpr int CPL = (R + 1) / 2;       // cells per thread in the load mapping (odd R: the last slot of
                                           // the upper half-wave is past the line)
    constexpr int NSP = (RIE == RIE_LF) ? 3 : 5;
    constexpr int RAD = Radius<SCH>::value;
    // sP/sM hold the split fluxes (as T) and then the interface values (as
    // double) of the same lines; the phases are separated by barriers
    __shared__ double sP[LPB * NP];
    __shared__ double sM[LPB * NP];
    __shared__ T sV[LPB * NV];
    __shared__ double sRed[LPB];
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int n = A.n, d = A.dir;   // SEG: n is the segment length
    int ll, g;
    if (d == 0) { ll = wave; g = lane; }
    else { ll = t % LPB; g = t / LPB; }
    const long L = blk * LPB + ll;   // SEG: the segment L % nseg of line L / nseg
    const long nll = SEG ? A.nlines * A.nseg : A.nlines;
    const bool lv = FULL || L < nll;
    long line = L;
    int seg = 0;
    if constexpr (SEG) { line = L / A.nseg; seg = (int)(L - line * A.nseg); }
    long base = 0, stride = 1;
    if (d == 0) { base = line * A.nx; stride = 1; }
    else if (d == 1) { long i = line % A.nx, k = line / A.nx; base = i + (long)A.nx * A.ny * k; stride = A.nx; }
    else { long i = line % A.nx, j = line / A.nx; base = i + (long)A.nx * j; stride = (long)A.nx * A.ny; }
    const long lbase = base;          // the full line's first cell
    const int off = SEG ? seg * n : 0;   // the segment's first cell on the line
    base += (long)off * stride;

    // byte offsets of the thread's cells within a field: a field holds at
    // most 2^29 doubles (4 GiB), so 32 bits suffice, and every access is the
    // field's base address (uniform, in scalar registers) plus this offset
    unsigned ob[CPL];
#pragma unroll
    for (int k = 0; k < CPL; k++) ob[k] = (unsigned)(base + (long)(g + 64 * k) * stride) * 8u;
    // bytes of one FP64 field (the buffer's bound: 2^32 - 1 at the largest box)
    const unsigned fb = (unsigned)std::min<long>((long)A.nlines * (SEG ? A.nfull : n) * 8, 0xFFFFFFFFl);
    // state and RHS fields hold S (ob, fb are for 8-byte elements)
    constexpr unsigned SSH = sizeof(S) == 8 ? 0 : 1;
    auto lds = [&](const void GMEM *p, int k) -> double {
        return double(ld_at(static_cast<const S GMEM *>(p), ob[k] >> SSH, fb >> SSH));
    };
    auto sts = [&](void GMEM *p, int k, double val) { st_at(static_cast<S GMEM *>(p), ob[k] >> SSH, fb >> SSH, S(val)); };
    auto ldo = [&](const void GMEM *p, unsigned o) -> double {
        return double(ld_at(static_cast<const S GMEM *>(p), o >> SSH, fb >> SSH));
    };

    // split parameters of the own cells (con2prim, src/physicsModel.cpp:331-344).
    // The compact schemes keep them in LDS (one column per thread, read back
    // by the same thread): their chunk solve needs the registers.  The
    // explicit schemes keep them in CPL*NSP registers: with the LDS copy the
    // workgroup's 53 KB (LF) would allow only three workgroups per CU.
    constexpr bool SP_LDS = SCH == CRWENO5 || SCH == HCCS || SCH == HCCS77;
    __shared__ T sSP[SP_LDS ? CPL * NSP : 1][SP_LDS ? 64 * LPB : 1];
    T spr[SP_LDS ? 1 : CPL][SP_LDS ? 1 : NSP];
    unsigned innermask = 0;   // cd8 inner-form cells (rank-local position)
#pragma unroll
    for (int k = 0; k < CPL; k++) {
        int c = g + 64 * k;
        bool ok = FULL || (lv && c < n);
        T q[5];
#pragma unroll
        for (int v = 0; v < 5; v++) q[v] = ok ? T(lds(A.q[v], k)) : (v == 0 ? T(1) : T(0));
        T tmp[5];
        split_params<RIE, T>(q, d, T(A.gamma), tmp, A.riemann == RIE_AUSMP);
#pragma unroll
        for (int j = 0; j < NSP; j++) {
            if constexpr (SP_LDS) sSP[k * NSP + j][t] = tmp[j];
            else spr[k][j] = tmp[j];
        }
        int cl = (c + off) % A.nloc;
        if (cl >= A.tsv && cl < A.nloc - A.tsv) innermask |= 1u << k;
    }
    // SEG: the thread g < 2*HALO of a line owns its segment's halo cell hc
    // (-HALO..-1 left, n..n+HALO-1 right), a cell of the neighbouring segment
    // (periodic over the full line); its split parameters stay in registers
    const bool hown = SEG && lv && g < 2 * HALO;
    int hc = 0;
    unsigned obh = 0;
    T sph[5];
    if constexpr (SEG) {
        hc = g < HALO ? g - HALO : n + g - HALO;
        long gpos = (long)off + hc;
        if (gpos < 0) gpos += A.nfull;
        else if (gpos >= A.nfull) gpos -= A.nfull;
        obh = (unsigned)(lbase + gpos * stride) * 8u;
        T q[5];
#pragma unroll
        for (int v = 0; v < 5; v++) q[v] = hown ? T(ldo(A.q[v], obh)) : (v == 0 ? T(1) : T(0));
        split_params<RIE, T>(q, d, T(A.gamma), sph, A.riemann == RIE_AUSMP);
    }

    // f^+ / f^- line buffers (skewed, f^- mirrored: Skew<R>::at / mir) and the
    // plain viscous-flux line
    // a line's T view lives inside the bytes of its own double line, so that
    // in the x sweep (wave-level sync only) no wave's writes reach another's line
    T *Pl = reinterpret_cast<T *>(sP + ll * NP), *Ml = reinterpret_cast<T *>(sM + ll * NP);
    T *Vl = sV + ll * NV + HALO;
    double *Pw = sP + wave * NP, *Mw = sM + wave * NP;                 // interface values
    double *Pd = sP + ll * NP, *Md = sM + ll * NP;
    const bool wv = FULL || blk * LPB + wave < nll;
    // the zones of the line this wave reconstructs (line `wave` of the
    // workgroup; in the y/z sweeps not the line of the thread's own cells)
    const int offw = SEG ? (int)((blk * LPB + wave) % A.nseg) * n : 0;
    const Zone z{SEG ? n + 1 : n, A.nloc, A.ts, offw};
    // line-buffer slots of the thread's cells c = g + 64k and of c + 1 (f^+
    // skewed, f^- mirrored).  When R divides 64 they are affine in k (the
    // pads between cells 64 apart are 64/R), so every access is one base
    // register plus an immediate offset instead of one address register per
    // cell held across the variable loop.
    constexpr bool AFF = (64 % R) == 0;
    constexpr int KS = AFF ? 64 + 64 / R : 0;
    const int sa0 = Skew<R>::at(g), sa1 = Skew<R>::at(g + 1), sm0 = Skew<R>::mir(g), sm1 = Skew<R>::mir(g + 1);
    auto slotP = [&](int k) { return AFF ? sa0 + KS * k : Skew<R>::at(g + 64 * k); };
    auto slotP1 = [&](int k) { return AFF ? sa1 + KS * k : Skew<R>::at(g + 64 * k + 1); };
    auto slotM = [&](int k) { return AFF ? sm0 - KS * k : Skew<R>::mir(g + 64 * k); };
    auto slotM1 = [&](int k) { return AFF ? sm1 - KS * k : Skew<R>::mir(g + 64 * k + 1); };
    // compute-mapping geometry of this lane (nuc_lines.h)
    const int h = lane & 31;
    const bool neg = lane >= 32;
    const int chunk = neg ? 31 - h : h;
    const int m0 = chunk * R;
    const T *Lw = reinterpret_cast<const T *>((neg ? sM : sP) + wave * NP) + h * (R + 1);
    const int m0r = neg ? m0 + R - 1 : m0;
    const int mstep = neg ? -1 : 1;
    double res = 0.0;
    const double dt = epi ? *A.dtp : 0.0;

    // global loads of variable v (state and viscous flux) are issued one
    // variable ahead, so that they are in flight during the previous
    // variable's reconstruction
    double qv[CPL];
    T xv[CPL];
    auto fetch = [&](int v, double qo[CPL], T xo[CPL]) {
#pragma unroll
        for (int k = 0; k < CPL; k++) {
            int c = g + 64 * k;
            qo[k] = 0.0;
            xo[k] = T(0);
            if ((FULL || (lv && c < n)) && v < 5) {
                qo[k] = lds(A.q[v], k);
                if (A.vis != VIS_NONE && v > 0)
                    xo[k] = A.fv_f32 ? T(ld_at(static_cast<const float GMEM *>(A.fv[v - 1]), ob[k] >> 1, fb >> 1))
                                     : T(ld_at(static_cast<const double GMEM *>(A.fv[v - 1]), ob[k], fb));
            }
        }
    };
    fetch(0, qv, xv);
    double qh = 0.0;   // SEG: the halo cell's values of the current variable
    T xh = T(0);
    auto fetch_h = [&](int v, double &qo, T &xo) {
        qo = 0.0;
        xo = T(0);
        if (hown && v < 5) {
            qo = ldo(A.q[v], obh);
            if (A.vis != VIS_NONE && v > 0)
                xo = A.fv_f32 ? T(ld_at(static_cast<const float GMEM *>(A.fv[v - 1]), obh >> 1, fb >> 1))
                              : T(ld_at(static_cast<const double GMEM *>(A.fv[v - 1]), obh, fb));
        }
    };
    if constexpr (SEG) fetch_h(0, qh, xh);
#pragma unroll 1
    for (int v = 0; v < 5; v++) {
#pragma unroll
        for (int k = 0; k < CPL; k++) {
            int c = g + 64 * k;
            if (FULL || (lv && c < n)) {
                T fp, fm;
                T spk[NSP];
#pragma unroll
                for (int j = 0; j < NSP; j++) {
                    if constexpr (SP_LDS) spk[j] = sSP[k * NSP + j][t];
                    else spk[j] = spr[k][j];
                }
                split_var<RIE, T>(spk, T(qv[k]), v, d, fp, fm);
                Pl[slotP(k)] = fp; Ml[slotM(k)] = fm;
                if (A.vis != VIS_NONE && v > 0) Vl[c] = xv[k];
                // the cell's periodic image: cells within HALO of an end are
                // copied past the other end, all others to a dummy slot; with
                // n = 32R only the first and last cell of a thread can be near
                // an end, so the other cells skip the copy
                if (!SEG && (!FULL || k == 0 || k == CPL - 1)) {
                    const int cw = (k == 0 && c < HALO) ? c + n : (c >= n - HALO ? c - n : 32 * R + HALO);
                    Pl[Skew<R>::at(cw)] = fp; Ml[Skew<R>::mir(cw)] = fm;
                    if (A.vis != VIS_NONE && v > 0) Vl[cw] = xv[k];
                }
            }
        }
        if constexpr (SEG) {
            if (hown) {
                T fp, fm;
                split_var<RIE, T>(sph, T(qh), v, d, fp, fm);
                Pl[Skew<R>::at(hc)] = fp; Ml[Skew<R>::mir(hc)] = fm;
                if (A.vis != VIS_NONE && v > 0) Vl[hc] = xh;
            }
        }
        double qnx[CPL];
        T xnx[CPL];
        fetch(v + 1, qnx, xnx);   // v + 1 = 5 gives zeros
        double qhn = 0.0;
        T xhn = T(0);
        if constexpr (SEG) fetch_h(v + 1, qhn, xhn);
        sync_lines(d);
        double out[R];
        // weights and rows in T where FP32 keeps them in range; the Z-type
        // weights (WENO5-Z, WENO7-Z and the hccs family) square tau/(is+eps)
        // with eps down to 1e-40 and stay FP64
        constexpr bool zw = SCH == WENO5Z || SCH == WENO7JS || SCH == HCCS || SCH == HCCS77 || SCH == HCCS78;
        using TW = typename std::conditional<zw, double, T>::type;
        if (wv) line_half<SCH, R, T, TW, double>(Lw, m0r, mstep, z, h, neg, out, FULL);
        // the interface values overwrite the wave's own line only: the
        // reconstruction of the other waves is not waited for
        sync_wave();
        if (wv) {
            // interface m = m0r + mstep*r goes to slot at(m) of the f^+ buffer
            // or mir(m) of the f^- buffer; relative to the lane base
            // h*(R+1) these are at(r) and at(r-1), and at(r-1) = at(r)-1
            // for 0 < r < R
            double *Ob = (neg ? Mw : Pw) + h * (R + 1);
            double *Ow = Ob - (neg ? 1 : 0);
            // a segment's interfaces are 0..n, a periodic line's 0..n-1
            const int nif = SEG ? n + 1 : n;
            if (FULL || m0r < nif) Ob[neg ? Skew<R>::at(-1) : Skew<R>::at(0)] = out[0];
#pragma unroll
            for (int r = 1; r < R; r++) {
                int m = m0r + mstep * r;
                if (FULL || m < nif) Ow[Skew<R>::at(r)] = out[r];
            }
            // periodic copy of interface 0 to interface n
            if constexpr (!SEG) {
                if (!neg && h == 0) Pw[Skew<R>::at(n)] = out[0];
                if (neg && h == 31) Mw[Skew<R>::mir(n)] = out[R - 1];
            }
        }
        // the final phase's own global reads (rhs of the previous sweeps; u^n and
        // the current state of the RK update), issued after the reconstruction:
        // held across it they would cost 3*CPL registers, and the other waves
        // of the SIMD cover their latency
        double rin[CPL], unv[CPL], qcur[CPL];
#pragma unroll
        for (int k = 0; k < CPL; k++) {
            int c = g + 64 * k;
            rin[k] = 0.0;
            unv[k] = 0.0;
            qcur[k] = 0.0;
            if (FULL || (lv && c < n)) {
                if (!A.first || A.accum) rin[k] = lds(A.rhs[v], k);
                if (epi && A.stage > 0) unv[k] = lds(A.qn[v], k);
                if constexpr (EPI == EPI_YES) qcur[k] = qv[k];
                else if (epi) qcur[k] = lds(A.q[v], k);
            }
        }
        sync_lines(d);
#pragma unroll
        for (int k = 0; k < CPL; k++) {
            int c = g + 64 * k;
            if (FULL || (lv && c < n)) {
                double r = rin[k];
                // F = f^+ + f^- at the interfaces, then (F[i+1]-F[i])/dx (src/euler3d.cpp:23-56,132-166)
                r += ((Pd[slotP1(k)] + Md[slotM1(k)]) - (Pd[slotP(k)] + Md[slotM(k)])) * A.inv_h;
                if (A.vis != VIS_NONE && v > 0)
                    r -= double(central_diff<T>(Vl, c, (innermask >> k) & 1u, A.vis, T(A.inv_h)));
                if (A.fmean && v == 4) r += *A.fmean;
                if (!epi) {
                    sts(A.rhs[v], k, r);
                } else {
                    // TVD-RK3 stage update, src/TVD-RK3.cpp:59-64 (0), :97-101 (1), :140-144 (2)
                    double rr = -r, nu;
                    if (A.stage == 0) {
                        nu = qcur[k] + rr * dt;
                    } else {
                        double un = unv[k];
                        if (A.stage == 1) nu = 0.75 * un + 0.25 * (rr * dt + qcur[k]);
                        else nu = (1.0 / 3.0) * un + (2.0 / 3.0) * (rr * dt + qcur[k]);
                        nu = double(S(nu));   // the stored value (an FP32 state rounds here)
                        if (A.stage == 2) res += sq(nu - un);
                    }
                    sts(A.qout[v], k, nu);
                }
            }
        }
        sync_lines(d);
        // unconditional, so that the current values are dead after the load
        // phase instead of being carried to the end of the iteration
#pragma unroll
        for (int k = 0; k < CPL; k++) { qv[k] = qnx[k]; xv[k] = xnx[k]; }
        if constexpr (SEG) { qh = qhn; xh = xhn; }
    }
    if (epi && A.stage == 2) {
        // residual partial of this workgroup, fixed reduction order (src/PDEData3d.cpp:93-128)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) res += __shfl_down(res, o);
        if (lane == 0) sRed[wave] = res;
        __syncthreads();
        if (t == 0) {
            double acc = 0.0;
            for (int w = 0; w < LPB; w++) acc += sRed[w];
            A.res_part[blk] = acc;
        }
    }
}

template <int SCH, int RIE, int R, int LPB, typename T, bool FULL, int EPI, typename S = double, bool SEG = false>
__global__ __launch_bounds__(64 * LPB) void k_sweep(SweepArgs A)
{
    sweep_body<SCH, RIE, R, LPB, T, FULL, EPI, S, SEG>(A, xcd_block(blockIdx.x, gridDim.x));
}

constexpr int SWEEP_LPB = 4;
// Schemes whose stencils run in FP32 in the mixed path.  The 7-point
// indicators are evaluated there as sums of squares (w7_sos), which FP32 holds
// to its relative precision.  hccs and hccs77 keep FP64 stencils: their blend
// switches at theta = 0.5 between the upwind compact row and the blended row,
// which do not meet there (src/hccs.cpp:146-183, src/hccs77.cpp:86-163), so an
// FP32 rounding of the indicators that moves theta across 0.5 changes a row by
// the full jump (measured: 7e-4 in the state after 5 steps of 24^3 turbulence
// with hccs77).
template <int SCH>
constexpr bool mixed_fp32()
{
    return SCH != HCCS && SCH != HCCS77;
}

constexpr bool compact_scheme(int sch) { return sch == CRWENO5 || sch == HCCS || sch == HCCS77; }

// segments of lines longer than 512 cells: R = 8 (up to 255 cells), and for
// the compact schemes R = 16 (one rank block of up to 511 cells)
template <int SCH, int RIE, int R>
void launch_sweep_seg(const SweepArgs &A, hipStream_t s)
{
    const long nb = (A.nlines * A.nseg + SWEEP_LPB - 1) / SWEEP_LPB;
    const dim3 grid((unsigned)nb), block(64 * SWEEP_LPB);
    if constexpr (mixed_fp32<SCH>()) {
        if (A.fp32) {
            hipLaunchKernelGGL((k_sweep<SCH, RIE, R, SWEEP_LPB, float, false, EPI_ARG, float, true>), grid, block, 0, s, A);
            return;
        }
        if (A.mixed) {
            hipLaunchKernelGGL((k_sweep<SCH, RIE, R, SWEEP_LPB, float, false, EPI_ARG, double, true>), grid, block, 0, s, A);
            return;
        }
    }
    hipLaunchKernelGGL((k_sweep<SCH, RIE, R, SWEEP_LPB, double, false, EPI_ARG, double, true>), grid, block, 0, s, A);
}

// the kernel forms for one state type S and stencil type T: FULL kernels
// (n = 32R, even R) with and without the RK3 update, else the one that takes
// it as an argument
template <int SCH, int RIE, int R, typename T, typename S>
void launch_sweep_rt(const SweepArgs &A, hipStream_t s, bool full, dim3 grid, dim3 block)
{
    if constexpr (R % 2 == 0) {
        if (full && A.epilogue) {
            hipLaunchKernelGGL((k_sweep<SCH, RIE, R, SWEEP_LPB, T, true, EPI_YES, S>), grid, block, 0, s, A);
            return;
        }
        if (full) {
            hipLaunchKernelGGL((k_sweep<SCH, RIE, R, SWEEP_LPB, T, true, EPI_NO, S>), grid, block, 0, s, A);
            return;
        }
    }
    hipLaunchKernelGGL((k_sweep<SCH, RIE, R, SWEEP_LPB, T, false, EPI_ARG, S>), grid, block, 0, s, A);
}

template <int SCH, int RIE, int R>
void launch_sweep_r(const SweepArgs &A, hipStream_t s)
{
    if (A.nseg > 1) {
        if constexpr (R == 8 || (R == 16 && compact_scheme(SCH))) launch_sweep_seg<SCH, RIE, R>(A, s);
        return;
    }
    long nb = (A.nlines + SWEEP_LPB - 1) / SWEEP_LPB;
    // odd R: the load mapping's 64 (R+1)/2 cells overrun a 32R-cell line, so no FULL form
    const bool full = (R % 2 == 0) && (A.n == 32 * R) && (A.nlines % SWEEP_LPB == 0);
    const dim3 grid((unsigned)nb), block(64 * SWEEP_LPB);
    if constexpr (mixed_fp32<SCH>()) {
        if (A.fp32) {   // NUC_FP32: FP32 state and RHS around the mixed path's arithmetic
            launch_sweep_rt<SCH, RIE, R, float, float>(A, s, full, grid, block);
            return;
        }
        if (A.mixed) {
            launch_sweep_rt<SCH, RIE, R, float, double>(A, s, full, grid, block);
            return;
        }
    }
    launch_sweep_rt<SCH, RIE, R, double, double>(A, s, full, grid, block);
}

template <int SCH, int RIE>
void launch_sweep_sr(int R, const SweepArgs &A, hipStream_t s)
{
    switch (R) {
    case 2: launch_sweep_r<SCH, RIE, 2>(A, s); break;
    case 3: launch_sweep_r<SCH, RIE, 3>(A, s); break;
    case 4: launch_sweep_r<SCH, RIE, 4>(A, s); break;
    case 5: launch_sweep_r<SCH, RIE, 5>(A, s); break;
    case 6: launch_sweep_r<SCH, RIE, 6>(A, s); break;
    case 8: launch_sweep_r<SCH, RIE, 8>(A, s); break;
    case 12: launch_sweep_r<SCH, RIE, 12>(A, s); break;
    default: launch_sweep_r<SCH, RIE, 16>(A, s); break;
    }
}

template <int SCH>
void launch_sweep_s(int R, const SweepArgs &A, hipStream_t s)
{
    if (A.riemann == RIE_LF) launch_sweep_sr<SCH, RIE_LF>(R, A, s);
    else launch_sweep_sr<SCH, RIE_AUSM>(R, A, s);
}

// ---------------------------------------------------------------------------
// packed launches (see nuc3d.hip): the job of workgroup b of a packed grid
// ---------------------------------------------------------------------------
__device__ __forceinline__ int pack_job(const int *off, int nj, long b)
{
    int j = 0;
    while (j + 1 < nj && b >= off[j + 1]) j++;
    return j;
}

template <int SCH, int RIE, int R, int LPB, typename T, bool FULL>
__global__ __launch_bounds__(64 * LPB) void k_sweep_pack(const SweepArgs *__restrict__ tab, const int *__restrict__ off, int nj)
{
    const long b = xcd_block(blockIdx.x, gridDim.x);
    const int j = pack_job(off, nj, b);
    sweep_body<SCH, RIE, R, LPB, T, FULL, EPI_ARG>(tab[j], b - off[j]);
}

struct PackLaunch {
    const SweepArgs *tab;
    const int *off;
    int nj;
    long nblocks;
    bool mixed, full;
};

template <int SCH, int RIE, int R, typename T>
void launch_sweep_pack_rt(const PackLaunch &P, hipStream_t s)
{
    const dim3 grid((unsigned)P.nblocks), block(64 * SWEEP_LPB);
    if constexpr (R % 2 == 0) {
        if (P.full) {
            hipLaunchKernelGGL((k_sweep_pack<SCH, RIE, R, SWEEP_LPB, T, true>), grid, block, 0, s, P.tab, P.off, P.nj);
            return;
        }
    }
    hipLaunchKernelGGL((k_sweep_pack<SCH, RIE, R, SWEEP_LPB, T, false>), grid, block, 0, s, P.tab, P.off, P.nj);
}

template <int SCH, int RIE, int R>
void launch_sweep_pack_r(const PackLaunch &P, hipStream_t s)
{
    if constexpr (mixed_fp32<SCH>()) {
        if (P.mixed) {
            launch_sweep_pack_rt<SCH, RIE, R, float>(P, s);
            return;
        }
    }
    launch_sweep_pack_rt<SCH, RIE, R, double>(P, s);
}

template <int SCH, int RIE>
void launch_sweep_pack_sr(int R, const PackLaunch &P, hipStream_t s)
{
    if (R == 2) launch_sweep_pack_r<SCH, RIE, 2>(P, s);
    else if (R == 3) launch_sweep_pack_r<SCH, RIE, 3>(P, s);
    else launch_sweep_pack_r<SCH, RIE, 4>(P, s);
}

template <int SCH>
void launch_sweep_pack_s(int riemann, int R, const PackLaunch &P, hipStream_t s)
{
    if (riemann == RIE_LF) launch_sweep_pack_sr<SCH, RIE_LF>(R, P, s);
    else launch_sweep_pack_sr<SCH, RIE_AUSM>(R, P, s);
}

}  // namespace nucsweep
