// Per-precision instantiation of the pass kernels; included by
// gla_c64.hip (R = float) and gla_c128.hip (R = double).
#pragma once
#include "gla_launch.hpp"

// kernel-shape defaults (overridable at compile time for variant timing)
#ifndef GLA_EROW
#define GLA_EROW 8
#endif
#ifndef GLA_EROW_F
#define GLA_EROW_F GLA_EROW
#endif
#ifndef GLA_EROW_D
#define GLA_EROW_D GLA_EROW
#endif
// pass-3 schedule: 0 keeps the input column in registers through the even
// parity, 2 splits both parities first (fewer registers for FP64)
#ifndef GLA_ZRELOAD_F
#define GLA_ZRELOAD_F 0
#endif
#ifndef GLA_ZRELOAD_D
#define GLA_ZRELOAD_D 2
#endif
#ifndef GLA_ZBATCH_KB
#define GLA_ZBATCH_KB 64
#endif
// axis-2 passes: consecutive axis-1 columns per workgroup (128-byte rows)
#ifndef GLA_PXY_F
#define GLA_PXY_F 16
#endif
#ifndef GLA_PXY_D
#define GLA_PXY_D 8
#endif
#ifndef GLA_YLDS_KB
#define GLA_YLDS_KB 160
#endif
#ifndef GLA_ZWAVE
#define GLA_ZWAVE 0
#endif
#ifndef GLA_ZMINW
#define GLA_ZMINW 1
#endif
// resident waves per SIMD asked of the compiler for pass 3, per element type
#ifndef GLA_ZMINW_F
#define GLA_ZMINW_F GLA_ZMINW
#endif
#ifndef GLA_ZMINW_D
#define GLA_ZMINW_D GLA_ZMINW
#endif
// pass 3 options (k_zpass GG): bit 0 stages the root table and the split
// phases in LDS, bit 1 pipelines the G^ loads one point ahead, bit 2 issues
// the first point's G^ loads before the forward transform
// (measured at 128^3: FP32 215 -> 189-196 us with 7, FP64 403 -> 355-358 us
// with 3; bit 2 costs FP64 registers)
#ifndef GLA_ZGG_F
#define GLA_ZGG_F 7
#endif
#ifndef GLA_ZGG_D
#define GLA_ZGG_D 3
#endif
#ifndef GLA_ZPADSH
#define GLA_ZPADSH 3
#endif
// pass 3 as a persistent, software-pipelined kernel when S3 <= B3 (off: at
// 128^3 it measured 230-240 us against 211 us for k_zpass, the persistent
// loop costing more registers than the overlap gains)
#ifndef GLA_ZPIPE
#define GLA_ZPIPE 0
#endif
// components of the next tile loaded early in the pipelined pass 3
#ifndef GLA_ZPF_F
#define GLA_ZPF_F 2
#endif
#ifndef GLA_ZPF_D
#define GLA_ZPF_D 2
#endif

namespace gla {

// elements per thread: the divisor of N closest to `target` from below,
// raised until the workgroup fits in `maxtp` threads per pencil
// LDS exchanges of the radix plan for N with E elements per thread
__host__ __device__ constexpr int plan_exchanges(int N, int E)
{
    const Plan p = make_plan(N, E);
    if (p.nst <= 1) return (p.nst == 1 && E % p.rad[0] == 0) ? 0 : 1;
    const bool fr = E % p.rad[0] == 0, lr = E % p.rad[p.nst - 1] == 0;
    return (p.nst - 1) + (fr ? 0 : 1) + (lr ? 0 : 1);
}

// elements per thread: close to `target`, with few LDS exchanges, threads
// per pencil within `maxtp` and, when `wave` is set, dividing 64 (a pencil
// inside one wave)
__host__ __device__ constexpr int pick_e(int N, int target, int maxtp, bool wave = false)
{
    int best = 0, bestscore = 1 << 30;
    for (int e = 1; e <= N; ++e) {
        if (N % e) continue;
        const int tp = N / e;
        if (tp > maxtp) continue;
        const int d = e > target ? 2 * (e - target) : target - e;
        const int score = d + 4 * plan_exchanges(N, e) + ((wave && 64 % tp != 0) ? 1000 : 0);
        if (score < bestscore) { bestscore = score; best = e; }
    }
    return best ? best : N;
}

template <typename R>
struct Cfg {
    static constexpr bool DBL = sizeof(R) == 8;
    static constexpr int PX = DBL ? GLA_PXY_D : GLA_PXY_F;   // axis-2 passes: columns per bundle
    static constexpr int E_ROW = DBL ? GLA_EROW_D : GLA_EROW_F;   // passes 1, 2, 4, 5
    static constexpr int ZSCHED = DBL ? GLA_ZRELOAD_D : GLA_ZRELOAD_F;
    static constexpr int ZMINW = DBL ? GLA_ZMINW_D : GLA_ZMINW_F;
    static constexpr int ZGG = DBL ? GLA_ZGG_D : GLA_ZGG_F;
};


template <typename R, int N>
struct LenK {
    using C = Cfg<R>;
    static constexpr int TP_ROW = N / pick_e(N, C::E_ROW, 64, true);
    static constexpr int PEN = TP_ROW >= 256 ? 1 : 256 / TP_ROW;
    // axis-2 passes: the bundle width, capped so the two parities' LDS
    // buffers and the staged tables stay within GLA_YLDS_KB
    static constexpr int PXY = [] {
        int px = C::PX;
        while (px > 1 && (YK<R>::value * px + (GLA_XYTAB ? 2 : 0)) * N * (int)sizeof(cpx<R>) > GLA_YLDS_KB * 1024) px /= 2;
        return px;
    }();
    static constexpr int TP_Y = N / pick_e(N, C::E_ROW, 1024 / PXY);
    // pass 3: shape from the tuning table (gla_tune.hpp)
    static constexpr int PXZ = ztune(sizeof(cpx<R>), N).px;
    static constexpr int TP_Z = N / pick_e(N, ztune(sizeof(cpx<R>), N).e, 1024 / PXZ, GLA_ZWAVE != 0);
    // pass 3 keeps each pencil inside one wave when its threads divide 64
    static constexpr int WL_Z = (GLA_ZWAVE != 0 && 64 % TP_Z == 0 && (TP_Z * PXZ) % 64 == 0) ? 1 : 0;
    static constexpr int PS_Z = WL_Z ? N + (N >> GLA_ZPADSH) + 1 : N;
    // pass 3 runs its three component transforms together when their LDS
    // buffers fit the budget
    static constexpr int KB_Z = 3 * PXZ * PS_Z * (int)sizeof(cpx<R>) <= GLA_ZBATCH_KB * 1024 ? 3 : 1;
};

template <typename R, int N>
static hipError_t launch_fwd_x(const Geom& g, const DevTab& tb, const void* X, void* W1, int nrhs, hipStream_t st)
{
    using K = LenK<R, N>;
    const int rows = g.S[1] * g.S[2] * 3;
    dim3 grid((rows + K::PEN - 1) / K::PEN, nrhs);
    hipLaunchKernelGGL((k_fwd_x<R, N, K::TP_ROW, K::PEN, 1>), grid, dim3(K::TP_ROW * K::PEN), 0, st,
                       (const cpx<R>*)X, (cpx<R>*)W1, g.S[0], g.S[1], g.S[2], g.Y1, g.Z, g.xs, g.w1s,
                       (const cpx<R>*)tb.phz[0], (const cpx<R>*)tb.tw[0]);
    return hipGetLastError();
}

template <typename R, int N>
static hipError_t launch_inv_x(const Geom& g, const DevTab& tb, const void* W1, void* Y, int nrhs,
                               cpx<R> al, cpx<R> be, int axpby, const Epi& epi, hipStream_t st)
{
    using K = LenK<R, N>;
    const int rows = g.T[1] * g.T[2] * 3;
    dim3 grid((rows + K::PEN - 1) / K::PEN, nrhs);
    hipLaunchKernelGGL((k_inv_x<R, N, K::TP_ROW, K::PEN, 1>), grid, dim3(K::TP_ROW * K::PEN), 0, st,
                       (const cpx<R>*)W1, (cpx<R>*)Y, g.T[0], g.T[1], g.T[2], g.Y1, g.Z, g.w1s, g.ys,
                       (const cpx<R>*)tb.phz[0], (const cpx<R>*)tb.tw[0], al, be, axpby,
                       (const cpx<R>*)epi.x, epi.xs, (const cpx<R>*)epi.chi, epi.chis);
    return hipGetLastError();
}

template <typename R, int N>
static hipError_t launch_fwd_y(const Geom& g, const DevTab& tb, const void* W1, void* W2, int nrhs, hipStream_t st)
{
    using K = LenK<R, N>;
    constexpr int PX = K::PXY;
    dim3 grid((g.L[0] + PX - 1) / PX, g.S[2] * 3, nrhs);
    hipLaunchKernelGGL((k_fwd_y<R, N, K::TP_Y, PX, 1>), grid, dim3(K::TP_Y * PX), 0, st, (const cpx<R>*)W1,
                       (cpx<R>*)W2, g.L[0], g.S[1], g.S[2], g.Y1, g.Z, g.w1s, g.w2s,
                       (const cpx<R>*)tb.phz[1], (const cpx<R>*)tb.tw[1]);
    return hipGetLastError();
}

template <typename R, int N>
static hipError_t launch_inv_y(const Geom& g, const DevTab& tb, const void* W2, void* W1, int nrhs, hipStream_t st)
{
    using K = LenK<R, N>;
    constexpr int PX = K::PXY;
    dim3 grid((g.L[0] + PX - 1) / PX, g.T[2] * 3, nrhs);
    hipLaunchKernelGGL((k_inv_y<R, N, K::TP_Y, PX, 1>), grid, dim3(K::TP_Y * PX), 0, st, (const cpx<R>*)W2,
                       (cpx<R>*)W1, g.L[0], g.T[1], g.T[2], g.Y1, g.Z, g.w2s, g.w1s,
                       (const cpx<R>*)tb.phz[1], (const cpx<R>*)tb.tw[1]);
    return hipGetLastError();
}

template <typename R, int N>
static hipError_t launch_zpass_chunk(const Geom& g, const DevTab& tb, void* W2, int nrhs, hipStream_t st)
{
    using K = LenK<R, N>;
    constexpr int PX = K::PXZ;
    constexpr int SC = Cfg<R>::ZSCHED;
    constexpr int PF = sizeof(R) == 8 ? GLA_ZPF_D : GLA_ZPF_F;
    dim3 grid(8 * tb.zper_xcd * nrhs, 1, 1);   // right-hand side inside the XCD stretch
    dim3 blk(K::TP_Z * PX);
    const int mode = g.S[2] > g.B[2] ? 1 : (g.T[2] > g.B[2] ? 2 : 0);
#define GLA_ZARGS                                                                                   \
    (cpx<R>*)W2, g.L[0], g.L[1], g.S[2], g.T[2], g.Z, g.w2s, (const cpx<R>*)tb.gf, tb.am[0], tb.am[1], \
        tb.am[2], (const cpx<R>*)tb.phz[2], (const cpx<R>*)tb.tw[2], tb.zorder, tb.zper_xcd, nrhs
    if (GLA_ZPIPE && mode != 1 && !K::WL_Z) {
        // one resident wave of workgroups: blocks per CU from the occupancy
        // of this kernel, 8 XCDs of CUs/8 each
        static int bpx = 0;
        if (bpx == 0) {
            int dev = 0, ncu = 0, nb = 0;
            if (hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
                return hipErrorInvalidValue;
            const hipError_t oe = mode == 0
                ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_zpipe<R, N, K::TP_Z, PX, 0, K::KB_Z, PF>, K::TP_Z * PX, 0)
                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_zpipe<R, N, K::TP_Z, PX, 2, K::KB_Z, PF>, K::TP_Z * PX, 0);
            if (oe != hipSuccess) return oe;
            bpx = (nb < 1 ? 1 : nb) * ((ncu + 7) / 8);
        }
        const int nb_xcd = bpx < tb.zper_xcd ? bpx : tb.zper_xcd;
        dim3 pgrid(8 * nb_xcd, 1, nrhs);
        if (mode == 0) hipLaunchKernelGGL((k_zpipe<R, N, K::TP_Z, PX, 0, K::KB_Z, PF>), pgrid, blk, 0, st, GLA_ZARGS);
        else hipLaunchKernelGGL((k_zpipe<R, N, K::TP_Z, PX, 2, K::KB_Z, PF>), pgrid, blk, 0, st, GLA_ZARGS);
    } else if (mode == 0) hipLaunchKernelGGL((k_zpass<R, N, K::TP_Z, PX, 0, Cfg<R>::ZMINW, K::KB_Z, SC, K::WL_Z, Cfg<R>::ZGG, GLA_ZPADSH>), grid, blk, 0, st, GLA_ZARGS);
    else if (mode == 1) hipLaunchKernelGGL((k_zpass<R, N, K::TP_Z, PX, 1, Cfg<R>::ZMINW, K::KB_Z, SC, K::WL_Z, Cfg<R>::ZGG, GLA_ZPADSH>), grid, blk, 0, st, GLA_ZARGS);
    else hipLaunchKernelGGL((k_zpass<R, N, K::TP_Z, PX, 2, Cfg<R>::ZMINW, K::KB_Z, SC, K::WL_Z, Cfg<R>::ZGG, GLA_ZPADSH>), grid, blk, 0, st, GLA_ZARGS);
#undef GLA_ZARGS
    return hipGetLastError();
}

// pass 3 over nrhs right-hand sides: one launch has 8 * per_xcd workgroups per
// right-hand side, so large batches are split to keep the grid within 2^31 - 1
template <typename R, int N>
static hipError_t launch_zpass(const Geom& g, const DevTab& tb, void* W2, int nrhs, hipStream_t st)
{
    const int64_t per = 8LL * tb.zper_xcd;
    const int64_t lim = 2147483647LL / (per > 0 ? per : 1);
    const int maxr = (int)(lim < 1 ? 1 : (lim < nrhs ? lim : nrhs));
    for (int r0 = 0; r0 < nrhs; r0 += maxr) {
        const int cnt = nrhs - r0 < maxr ? nrhs - r0 : maxr;
        const hipError_t e = launch_zpass_chunk<R, N>(g, tb, (cpx<R>*)W2 + (int64_t)r0 * g.w2s, cnt, st);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <typename R, int N>
static hipError_t launch_fwd_z(const Geom& g, const DevTab& tb, void* W2, int nrhs, hipStream_t st)
{
    using K = LenK<R, N>;
    constexpr int PX = K::PXY;
    dim3 grid((g.L[0] + PX - 1) / PX, g.L[1] * 3, nrhs);
    hipLaunchKernelGGL((k_fwd_z<R, N, K::TP_Y, PX, 1>), grid, dim3(K::TP_Y * PX), 0, st, (cpx<R>*)W2, g.L[0], g.L[1],
                       g.S[2], g.Z, g.w2s, (const cpx<R>*)tb.phz[2], (const cpx<R>*)tb.tw[2]);
    return hipGetLastError();
}

// Passes 1 and 2 fused (k_fwd_xy): ComplexF32, B1 = B2 = N in {32, 64, 96,
// 128}, and the plane's rows fit the LDS tile.  Off: the 160 KB tile leaves
// one workgroup per CU, whose load, transform and store phases then run back
// to back -- measured 114 us against 29 + 71 us for the two passes at 128^3
// (84 against 40 us at 96^3), with 1024 threads per workgroup 121 us.
#ifndef GLA_XYFUSE
#define GLA_XYFUSE 0
#endif
#ifndef GLA_XYNT
#define GLA_XYNT 512
#endif
template <int N>
struct XyK {
    static constexpr int TP = N / pick_e(N, GLA_EROW, 64, true);
    // columns per phase-2 group: GLA_XYNT threads when they divide the columns
    static constexpr int CPB = (GLA_XYNT / TP <= N && N % (GLA_XYNT / TP) == 0) ? GLA_XYNT / TP
                                                                                 : ((512 / TP <= N) ? 512 / TP : N);
    static constexpr int NT = TP * CPB;
    static constexpr int PEN1 = NT / TP;
};

template <int N>
static bool xy_fits(const Geom& g)
{
    const int pen = XyK<N>::PEN1;
    const int rows = ((g.S[1] + pen - 1) / pen) * pen;
    return (rows > N ? rows : N) <= XyTile<N>::ROWS;
}

template <typename R>
bool xy_fused(const Geom& g)
{
    if (!GLA_XYFUSE || sizeof(R) != 4 || g.B[0] != g.B[1]) return false;
    switch (g.B[0]) {
        case 32: return xy_fits<32>(g);
        case 64: return xy_fits<64>(g);
        case 96: return xy_fits<96>(g);
        case 128: return xy_fits<128>(g);
        default: return false;
    }
}

template <int N>
static hipError_t launch_fwd_xy(const Geom& g, const DevTab& tb, const void* X, void* W2, int nrhs, hipStream_t st)
{
    using K = XyK<N>;
    dim3 grid(g.S[2] * 3, 2, nrhs);
    hipLaunchKernelGGL((k_fwd_xy<float, N, K::TP, K::TP, K::NT>), grid, dim3(K::NT), 0, st, (const cpx<float>*)X,
                       (cpx<float>*)W2, g.S[0], g.S[1], g.S[2], g.L[0], g.Z, g.xs, g.w2s,
                       (const cpx<float>*)tb.phz[0], (const cpx<float>*)tb.tw[0], (const cpx<float>*)tb.phz[1],
                       (const cpx<float>*)tb.tw[1]);
    return hipGetLastError();
}

template <typename R>
static hipError_t launch_fused_xy(const Geom& g, const DevTab& tb, const void* X, void* W2, int nrhs, hipStream_t st)
{
    if constexpr (sizeof(R) == 4) {
        switch (g.B[0]) {
            case 32: return launch_fwd_xy<32>(g, tb, X, W2, nrhs, st);
            case 64: return launch_fwd_xy<64>(g, tb, X, W2, nrhs, st);
            case 96: return launch_fwd_xy<96>(g, tb, X, W2, nrhs, st);
            case 128: return launch_fwd_xy<128>(g, tb, X, W2, nrhs, st);
            default: break;
        }
    }
    return hipErrorInvalidValue;
}

template <typename R>
hipError_t run_fwd3(const Geom& g, const DevTab& tb, const void* X, void* W1, void* W2, int nrhs, hipStream_t st)
{
    hipError_t e = hipErrorInvalidValue;
#define GLA_CASE_1(n) case n: e = launch_fwd_x<R, n>(g, tb, X, W1, nrhs, st); break;
#define GLA_CASE_2(n) case n: e = launch_fwd_y<R, n>(g, tb, W1, W2, nrhs, st); break;
#define GLA_CASE_3(n) case n: e = launch_fwd_z<R, n>(g, tb, W2, nrhs, st); break;
    switch (g.B[0]) { GLA_FOR_EACH_LEN(GLA_CASE_1) default: return hipErrorInvalidValue; }
    if (e != hipSuccess) return e;
    switch (g.B[1]) { GLA_FOR_EACH_LEN(GLA_CASE_2) default: return hipErrorInvalidValue; }
    if (e != hipSuccess) return e;
    switch (g.B[2]) { GLA_FOR_EACH_LEN(GLA_CASE_3) default: return hipErrorInvalidValue; }
#undef GLA_CASE_1
#undef GLA_CASE_2
#undef GLA_CASE_3
    return e;
}

template <typename R>
hipError_t run_pass(int pass, const Geom& g, const DevTab& tb, const void* X, void* Y, void* W1, void* W2,
                    int nrhs, double ar, double ai, double br, double bi, int axpby, const Epi& epi,
                    hipStream_t st)
{
    hipError_t e = hipErrorInvalidValue;
    const cpx<R> al{R(ar), R(ai)}, be{R(br), R(bi)};
#define GLA_CASE_1(n) case n: e = launch_fwd_x<R, n>(g, tb, X, W1, nrhs, st); break;
#define GLA_CASE_2(n) case n: e = launch_fwd_y<R, n>(g, tb, W1, W2, nrhs, st); break;
#define GLA_CASE_3(n) case n: e = launch_zpass<R, n>(g, tb, W2, nrhs, st); break;
#define GLA_CASE_4(n) case n: e = launch_inv_y<R, n>(g, tb, W2, W1, nrhs, st); break;
#define GLA_CASE_5(n) case n: e = launch_inv_x<R, n>(g, tb, W1, Y, nrhs, al, be, axpby, epi, st); break;
    const bool fused = xy_fused<R>(g);
    switch (pass) {
        case 0:
            if (fused) return launch_fused_xy<R>(g, tb, X, W2, nrhs, st);
            switch (g.B[0]) { GLA_FOR_EACH_LEN(GLA_CASE_1) default: break; }
            break;
        case 1:
            if (fused) return hipSuccess;   // done by the fused pass 1
            switch (g.B[1]) { GLA_FOR_EACH_LEN(GLA_CASE_2) default: break; }
            break;
        case 2: switch (g.B[2]) { GLA_FOR_EACH_LEN(GLA_CASE_3) default: break; } break;
        case 3: switch (g.B[1]) { GLA_FOR_EACH_LEN(GLA_CASE_4) default: break; } break;
        case 4: switch (g.B[0]) { GLA_FOR_EACH_LEN(GLA_CASE_5) default: break; } break;
        default: break;
    }
#undef GLA_CASE_1
#undef GLA_CASE_2
#undef GLA_CASE_3
#undef GLA_CASE_4
#undef GLA_CASE_5
    return e;
}

template <typename R>
hipError_t run_matvec(const Geom& g, const DevTab& tb, const void* X, void* Y, void* W1, void* W2, int nrhs,
                      double ar, double ai, double br, double bi, int axpby, const Epi& epi, hipStream_t st)
{
    for (int p = 0; p < NPASS; ++p) {
        hipError_t e = run_pass<R>(p, g, tb, X, Y, W1, W2, nrhs, ar, ai, br, bi, axpby, epi, st);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace gla
