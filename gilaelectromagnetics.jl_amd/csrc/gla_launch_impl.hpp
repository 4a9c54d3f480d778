// Per-precision instantiation of the pass kernels; included by
// gla_c64.hip (R = float) and gla_c128.hip (R = double).
#pragma once
#include "gla_launch.hpp"

// kernel-shape defaults (overridable at compile time for variant timing); the
// per-length shapes come from the tables of gla_tune.hpp
// threads per workgroup of passes 1 and 5 (rows of one pencil each)
#ifndef GLA_XWG
#define GLA_XWG 256
#endif
#ifndef GLA_ZBATCH_KB
#define GLA_ZBATCH_KB 64
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
// with 3; bit 2 costs FP64 registers).  Per element type and length from
// gtune; GLA_ZGG_F / GLA_ZGG_D force one value in tuning builds.
#ifndef GLA_ZPADSH
#define GLA_ZPADSH 3
#endif

namespace gla {

// elements per thread: the divisor of N closest to `target` from below,
// raised until the workgroup fits in `maxtp` threads per pencil
// LDS exchanges of the radix plan for N with E elements per thread
__host__ __device__ constexpr int plan_exchanges(int N, int E, int pol)
{
    const Plan p = make_plan(N, E, pol);
    if (p.nst <= 1) return (p.nst == 1 && E % p.rad[0] == 0) ? 0 : 1;
    const bool fr = E % p.rad[0] == 0, lr = E % p.rad[p.nst - 1] == 0;
    return (p.nst - 1) + (fr ? 0 : 1) + (lr ? 0 : 1);
}

// elements per thread: close to `target`, with few LDS exchanges, threads
// per pencil within `maxtp` and, when `wave` is set, dividing 64 (a pencil
// inside one wave) -- unless that would leave more than 64 elements per
// thread (lengths with a large odd factor, e.g. 81, 145), in which case the
// pencil spans waves (workgroup barriers) with at most 32 elements per thread
__host__ __device__ constexpr int pick_e_in(int N, int target, int maxtp, bool wave, int emax, int pol)
{
    int best = 0, bestscore = 1 << 30;
    for (int e = 1; e <= N; ++e) {
        if (N % e) continue;
        const int tp = N / e;
        if (tp > maxtp || e > emax) continue;
        const int d = e > target ? 2 * (e - target) : target - e;
        const int score = d + 4 * plan_exchanges(N, e, pol) + ((wave && 64 % tp != 0) ? 1000 : 0);
        if (score < bestscore) { bestscore = score; best = e; }
    }
    return best;
}

__host__ __device__ constexpr int pick_e(int N, int target, int maxtp, bool wave, int pol)
{
    const int e = pick_e_in(N, target, maxtp, wave, N, pol);
    if (e > 0 && e <= 64) return e;
    const int f = pick_e_in(N, target, maxtp, false, 32, pol);
    return f > 0 ? f : (e > 0 ? e : N);
}

template <typename R>
struct Cfg {
    static constexpr bool DBL = sizeof(R) == 8;
    static constexpr int ZMINW = DBL ? GLA_ZMINW_D : GLA_ZMINW_F;
};

// pass-3 shape of length N for the large-grid (SMALL = 0) or small-grid
// class of gla_tune.hpp
template <typename R, int N, int SMALL>
struct ZK {
    static constexpr int CB = (int)sizeof(cpx<R>);
    static constexpr int POL = ptune(CB, N);
    static constexpr int PXZ = ztune(CB, N, SMALL != 0).px;
    static constexpr int TP_Z = N / pick_e(N, ztune(CB, N, SMALL != 0).e, 1024 / PXZ, GLA_ZWAVE != 0, POL);
    // pass 3 keeps each pencil inside one wave when its threads divide 64
    static constexpr int WL_Z = (GLA_ZWAVE != 0 && 64 % TP_Z == 0 && (TP_Z * PXZ) % 64 == 0) ? 1 : 0;
    static constexpr int PS_Z = WL_Z ? N + (N >> GLA_ZPADSH) + 1 : N;
    // pass 3 runs its three component transforms together when their LDS
    // buffers fit the budget and the plan has no direct odd radix above 16
    // (a radix-29 butterfly for three components at once would spill)
    static constexpr int RMAX_Z = [] {
        const Plan p = make_plan(N, N / TP_Z, POL);
        int m = 1;
        for (int i = 0; i < p.nst; ++i) m = p.rad[i] > m ? p.rad[i] : m;
        return m;
    }();
    static constexpr int KB_Z =
        (RMAX_Z <= 16 && 3 * PXZ * PS_Z * (int)sizeof(cpx<R>) <= GLA_ZBATCH_KB * 1024) ? 3 : 1;
    static constexpr int GG = gtune(CB, N);
    static constexpr int RL = ltune(CB, N);
};

template <typename R, int N>
struct LenK {
    static constexpr int CB = (int)sizeof(cpx<R>);
    static constexpr int POL = ptune(CB, N);
    static constexpr RShape RS = rtune(CB, N);
    static constexpr int TP_ROW = N / pick_e(N, RS.ex, 64, true, POL);
    static constexpr int PEN = TP_ROW >= GLA_XWG ? 1 : GLA_XWG / TP_ROW;
    // axis-2 passes: the bundle width, capped so the two parities' LDS
    // buffers and the staged tables stay within GLA_YLDS_KB
    static constexpr int PXY = [] {
        int px = RS.pxy;
        while (px > 1 && (yk_of<R, N>() * px + (GLA_XYTAB ? 2 : 0)) * N * CB > GLA_YLDS_KB * 1024) px /= 2;
        return px;
    }();
    static constexpr int TP_Y = N / pick_e(N, RS.ey, 1024 / PXY, false, POL);
};

template <typename R, int N>
static hipError_t launch_fwd_x(const Geom& g, const DevTab& tb, const void* X, void* W1, int nrhs, hipStream_t st)
{
    using K = LenK<R, N>;
    const int rows = g.S[1] * g.S[2] * 3;
    dim3 grid((rows + K::PEN - 1) / K::PEN, nrhs);
    hipLaunchKernelGGL((k_fwd_x<R, N, K::TP_ROW, K::PEN, 1>), grid, dim3(K::TP_ROW * K::PEN), 0, st,
                       (const cpx<R>*)X, (cpx<R>*)W1, g.S[0], g.S[1], g.S[2], g.Y1, g.Z, g.xs, g.w1s,
                       (const cpx<R>*)tb.phz[0], (const cpx<R>*)tb.tw[0], g.xp);
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
                       (const cpx<R>*)epi.x, epi.xs, (const cpx<R>*)epi.chi, epi.chis, g.yp);
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

template <typename R, int N, int SMALL>
static hipError_t launch_zpass_shape(const Geom& g, const DevTab& tb, void* W2, int nrhs, hipStream_t st)
{
    using K = ZK<R, N, SMALL>;
    constexpr int PX = K::PXZ;
    constexpr int SC = K::RL;
    dim3 grid(8 * tb.zper_xcd * nrhs, 1, 1);   // right-hand side inside the XCD stretch
    dim3 blk(K::TP_Z * PX);
    const int mode = g.S[2] > g.B[2] ? 1 : (g.T[2] > g.B[2] ? 2 : 0);
    // k_zpass addresses a thread's column with a 32-bit byte offset
    // ((column + t * L1 * L2) elements)
    if ((int64_t)K::TP_Z * g.L[0] * g.L[1] * (int64_t)sizeof(cpx<R>) >= ((int64_t)1 << 32)) return hipErrorInvalidValue;
#define GLA_ZARGS                                                                                 \
    (cpx<R>*)W2, g.L[0], g.L[1], g.S[2], g.T[2], g.Z, g.w2s, (const cpx<R>*)tb.gf, tb.am[0], tb.am[1], \
        tb.am[2], (const cpx<R>*)tb.phz[2], (const cpx<R>*)tb.tw[2], tb.zorder, tb.zper_xcd, nrhs
    if (mode == 0) hipLaunchKernelGGL((k_zpass<R, N, K::TP_Z, PX, 0, Cfg<R>::ZMINW, K::KB_Z, SC, K::WL_Z, K::GG, GLA_ZPADSH>), grid, blk, 0, st, GLA_ZARGS);
    else if (mode == 1) hipLaunchKernelGGL((k_zpass<R, N, K::TP_Z, PX, 1, Cfg<R>::ZMINW, K::KB_Z, SC, K::WL_Z, K::GG, GLA_ZPADSH>), grid, blk, 0, st, GLA_ZARGS);
    else hipLaunchKernelGGL((k_zpass<R, N, K::TP_Z, PX, 2, Cfg<R>::ZMINW, K::KB_Z, SC, K::WL_Z, K::GG, GLA_ZPADSH>), grid, blk, 0, st, GLA_ZARGS);
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
        const bool small = (int64_t)g.L[0] * g.L[1] < ZSMALL_COLS;
        cpx<R>* const w = (cpx<R>*)W2 + (int64_t)r0 * g.w2s;
        const hipError_t e = small ? launch_zpass_shape<R, N, 1>(g, tb, w, cnt, st)
                                   : launch_zpass_shape<R, N, 0>(g, tb, w, cnt, st);
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
    switch (pass) {
        case 0: switch (g.B[0]) { GLA_FOR_EACH_LEN(GLA_CASE_1) default: break; } break;
        case 1: switch (g.B[1]) { GLA_FOR_EACH_LEN(GLA_CASE_2) default: break; } break;
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
